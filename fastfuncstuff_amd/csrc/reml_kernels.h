// Shared definitions between the HIP kernels and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace remlk {

// Pairs handled by one wave of the full-grid energy kernel, in two halves;
// the pairs of a half share the AR coefficient a (the pair list is a-major),
// so the AR-filtered sample z_t = y_{t+1} - c_t y_t is computed once per time
// step for each half.  Also the band sets' energy chunk (one pair per group).
constexpr int kEnergyNB = 8;
// Slots per half of an energy chunk: the pairs of a half share a (the two
// halves of a chunk may differ), so the filtered sample is formed once per half.
constexpr int kEnergyHalf = 4;
// Voxels per wave in the energy kernel (one per lane).
constexpr int kEnergyVox = 64;
// Waves per energy-kernel block: they share the block's 64-voxel slab via L1.
constexpr int kEnergyWaves = 8;

// Projection GEMM tiling: BM voxels x BN columns, BK time samples per stage.
constexpr int kProjBM = 128;
constexpr int kProjBK = 32;
// Voxel padding unit of every voxel-major / voxel-minor buffer.
constexpr int kVoxPad = 256;
// Threads per block of the GLS kernel.
constexpr int kGlsBlock = 64;
// Block -> (voxel tile, column tile) map of the projection kernels.  Blocks are
// dispatched round-robin over the 8 XCDs (block id % 8); each XCD takes a
// contiguous group of ceil(n_tiles / 8) column tiles, whose operand stays in
// that XCD's L2, and streams the voxel tiles through it (column tile fastest,
// so the blocks in flight share their residual tile).  The grid holds
// 8 * nvt * group blocks; returns false for blocks past the end.
__device__ inline bool xcd_tile_map(int bid, int nb, int n_tiles, int* vt, int* tile) {
    const int grp = (n_tiles + 7) >> 3;
    const int nvt = nb / (8 * grp);
    const int xcd = bid & 7, local = bid >> 3;
    const int t0 = xcd * grp;
    const int ng = min(grp, n_tiles - t0);
    if (ng <= 0) return false;
    const int v = local / ng;
    if (v >= nvt) return false;
    *vt = v;
    *tile = t0 + local % ng;
    return true;
}

// Largest design width handled by the f64 MFMA GLS kernel (two 16-row tiles).
constexpr int kGlsMfmaMaxP = 32;
// Pair-set search: voxel groups padded to 64-column tiles; blocks of 4 waves.
constexpr int kSetTile = 64;
constexpr int kSetWaves = 4;
// Waves per block of the set energy kernel (one chunk on 8 consecutive tiles).
constexpr int kSetEnergyWaves = 8;

// A family of pair sets (lambda bands, lattice windows).  Set s holds cnt[s]
// pairs, in pair order, in slots [s*sb, s*sb + cnt[s]) (-1 past them); sb is a
// multiple of kEnergyNB and of 4 projection tiles.  Per slot: edge = 1 when
// the pair has a lattice neighbour that the set misses (the set's optimum is
// then not trusted), next = the window to continue with from that pair
// (windows only; -1 = none).
// Energy chunks: chunk m of set s has enb positions in eng groups of enb / eng
// consecutive positions; the pairs of a group share the AR coefficient
// sequence c_t (same a, same lambda == 0 flag), so its filtered sample is
// formed once.  eslot maps each position to a slot of the set (-1: empty);
// coef holds per step the c_t of each group, then (alpha_t, 1/v_{t+1}) of each
// position; s0 = 1/v_0 per position.
struct SetTable {
    int n;
    int sb;
    const int32_t* cnt;    // [n]
    const int32_t* pairs;  // [n][sb]
    const int32_t* edge;   // [n][sb]
    const int32_t* next;   // [n][sb] or null
    int enb, eng, mc;      // positions and groups per energy chunk; chunks per set (stride)
    const int32_t* ecnt;   // [n] energy chunks in use
    const int32_t* eslot;  // [n][mc][enb]
    const float* coef;     // [n][mc][T][eng + 2 enb]
    const float* s0;       // [n][mc][enb]
};

// Scalar arguments of k_tables shared by every design of one launch (one
// timeline, one pair list).
struct KTablesArgs {
    int T, Tpad, p, K;
    const double* ab;          // [K][2]
    const int32_t* white_flag; // [K]
    const int32_t* gap;        // [T]
    int n_task;
    const int32_t* task_idx;   // [n_task]
    int full_f, BN, npt, PG;
    double ridge;
    int nseg;  // > 0: p > 96, pair k's column j in tile k * nseg + j / BN (column j % BN)
};

// Output tables of k_tables for one design (device pointers; null where the
// plan keeps no such table).
struct TablesOut {
    const double* X;
    double *soa, *coef64, *Xw64, *G64, *Lx, *Ainv, *Fm, *work, *ld, *ldR, *ldX;
    float* G32;
    uint16_t *Ghf, *Glf, *Ghb, *Glb;
    float* g16s;
    double *G64p, *coefR;
    // rank-deficient design (null when X has full column rank): columns that
    // depend on the columns before them, and Pi = projection onto the row
    // space of X (p x p)
    const int32_t* dep;
    const double* Pi;
};

// Device-side description of the pair tables of one plan.
struct PlanDev {
    int T;        // retained time points
    int Tpad;     // T rounded up to kProjBK
    int p;        // design columns
    int K;        // grid pairs
    int n_chunks; // energy work chunks (chunk 0 is the lambda==0 chunk when n_white > 0)
    int n_white;  // lambda == 0 pairs
    int BN;       // projection tile width (columns)
    int npt;      // pairs per projection tile
    int n_tiles;  // projection tiles
    int n_a, n_b; // lattice of the hierarchical selector
    int ltop;     // top level of the hierarchical descent
    int zero_a, zero_b;
    int n_task;   // columns in the F statistic (p when no task subset)

    const int32_t* gap;        // [T]
    const double* ab;          // [K][2]
    const double* ld;          // [K] logdet R + logdet X'R^-1X
    const double* coef64;      // [K][T][3]  (sqrt(1/v_t), c_t, alpha_t)
    const double* X;           // [T][p] design
    const double* pinvT;       // [T][p] pinv(X)' (row-major, every p)
    const double* Xw64;        // [K][T][p]  L^-1 X
    const double* G64;         // [K][T][p]  R^-1 X Lx^-T
    const double* Lx;          // [K][p][p]  chol(X'R^-1X)
    const double* Ainv;        // [K][p][p]  (X'R^-1X)^-1
    const double* Fm;          // [K][p][p] F-statistic quadratic form (zero outside the task columns)
    const int32_t* task_idx;   // [n_task]
    const float* G32;          // [n_tiles][Tpad][BN] projection operand
    // split 16-bit projection operand: [n_tiles][Tpad/8][BN][8] hi and lo halves
    const uint16_t* Gh_f16;
    const uint16_t* Gl_f16;
    const uint16_t* Gh_bf16;
    const uint16_t* Gl_bf16;
    const float* g16_unscale;  // [K] 2^-2e_k: the fp16 operand of pair k holds G * 2^e_k
    // GLS on f64 MFMA (p <= kGlsMfmaMaxP): tables padded to PG = 16 * ceil(p / 16) columns.
    int PG;                    // 0 when p > kGlsMfmaMaxP
    int PGA;                   // width of G64p / Xp / pinvTp / XtX: 16 * ceil(p / 16) for every p <= 96
    const double* G64p;        // [K][Tpad][PG]  G, zero outside t < T, j < p
    const double* Xp;          // [Tpad][PG]     X, zero padded
    const double* pinvTp;      // [Tpad][PG]     pinv(X)', zero padded
    const double* coefR;       // [K][Tpad][4]   (sqrt(1/v_t), c_{t-1}, alpha_{t-1}, 0); c = alpha = 0 at run starts
    const double* XtX;         // [p][p]
    // GLS epilogue tables padded to PGA (zero past p; every p <= 96): per
    // pair, for p <= 32 Lx transposed, for 32 < p <= 96 M = Lx^-T (beta = M h,
    // the back substitution with its zero-pivot rule applied to the unit
    // vectors); then the F form and diag (X'R^-1X)^-1; and X'X
    const double* gls_epi;     // [K][2 PGA^2 + PGA]
    const double* XtXp;        // [PGA][PGA]
    // rank-deficient design: Pi = projection onto the row space of X (p x p);
    // null when X has full column rank.  The tables then hold the rank-r
    // forms (zero pivots and Q columns for the dependent columns, Ainv =
    // (X'R^-1X)^+) and the reported beta is Pi times the basic solution.
    const double* Pi;
    int rank;
    int proj_mode;             // 0 fp32 MFMA, 1 bf16 x3, 2 fp16 x3
    // p > 96: segments of BN = 64 columns per pair (pair k's column j in G32
    // tile k * nseg + j / 64, column j % 64; n_tiles = K * nseg); 0 otherwise
    int nseg;
    const float* ecoef;        // [n_chunks][T][2*NB] (alpha_t, 1/v_{t+1}) per chunk slot
    const float* ec;           // [n_chunks][T][2] c_t of the AR coefficient of each half of the chunk
    const float* es0;          // [n_chunks][NB] 1/v_0
    const int32_t* echunk;     // [n_chunks][NB] pair index (-1 = empty slot)
    const int32_t* white;      // [n_white] lambda==0 pairs
    const int32_t* lut;        // [n_a][n_b] pair index or -1
    // Local (5 x 5 window) search tables.
    const float* a_axis;       // [n_a] lattice a values
    const float* b_axis;       // [n_b] lattice b values
    const int32_t* white_flag; // [K]
    const float2* lp;          // [K][T] (alpha_t, 1/v_{t+1})
    const float* lc;           // [K][T] c_t
    const float* ls0;          // [K] 1/v_0
    // Pair sets scored for voxel groups (the full grid's arithmetic per pair):
    // lambda bands (band.n = 0: off) and 5 x 5 lattice windows.
    SetTable band;
    float band_lo, band_h;     // centre of band i = band_lo + i * band_h
    SetTable win;              // set of window centre c = ai * n_b + bi
};

// One pass of the hierarchical descent evaluated on pair sets (k_hier_select):
// pass 0 scores the lattice points that the top level and the level below it
// can visit (one set for every voxel); each later pass scores, per voxel, the
// 3 x 3 neighbourhood of its current point at the pass's level (sets keyed by
// the centre, built only for the centres the descent can reach).
struct HierPass {
    SetTable ST;
    int level;                  // the pass's level (pass 0: ltop, and ltop - 1 when ltop >= 1)
    int first;                  // 1 for pass 0
    const int32_t* point_slot;  // pass 0: [n_a * n_b] slot of each lattice point in the set (-1: absent)
    const int32_t* center_set;  // passes >= 1: [n_a * n_b] set of each centre (-1: unreachable)
    const int32_t* vmap;        // passes >= 1: [ST.n][9] slot of each visit (-1: invalid pair)
};

// Lattice window of the learned search: kWin x kWin pairs around the proposal.
constexpr int kWin = 5;
constexpr int kWinHalf = 2;

}  // namespace remlk
