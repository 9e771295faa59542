// Voxel grouping (bucket sort) and the FP64 GLS refit on f64 MFMA.
// Included by reml_api.hip after reml_kernels.hip (one translation unit).
#pragma once

#include "reml_kernels.h"

namespace remlk {

typedef double f64x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// Bucket sort of voxels by an integer key in [0, nb); other keys are dropped.
// Group b occupies slots [off[b], off[b] + cnt[b]) of perm; off[b] is a
// multiple of `pad` and the unused slots up to off[b+1] hold -1.  Counting is
// block-aggregated in LDS so that a heavily populated key costs one global
// atomic per block instead of one per voxel.  Slot order inside a group is not
// specified (the consumers compute each voxel independently).
// ---------------------------------------------------------------------------
constexpr int kSortBlock = 256;
constexpr int kSortItems = 16;    // keys per thread
constexpr int kSortLdsMax = 8192; // buckets counted in LDS (two int32 arrays)

__device__ __forceinline__ bool key_ok(int b, int nb) { return b >= 0 && b < nb; }

__global__ void __launch_bounds__(kSortBlock) k_bucket_count(const int32_t* __restrict__ keys, int V, int nb,
                                                             int32_t* __restrict__ cnt) {
    extern __shared__ int32_t sh_cnt[];
    for (int i = threadIdx.x; i < nb; i += kSortBlock) sh_cnt[i] = 0;
    __syncthreads();
    const int base = blockIdx.x * kSortBlock * kSortItems + threadIdx.x;
#pragma unroll 4
    for (int it = 0; it < kSortItems; ++it) {
        const int v = base + it * kSortBlock;
        if (v < V) {
            const int b = keys[v];
            if (key_ok(b, nb)) atomicAdd(&sh_cnt[b], 1);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += kSortBlock)
        if (sh_cnt[i]) atomicAdd(&cnt[i], sh_cnt[i]);
}

__global__ void __launch_bounds__(kSortBlock) k_bucket_place(const int32_t* __restrict__ keys, int V, int nb,
                                                             int32_t* __restrict__ cursor, int32_t* __restrict__ perm) {
    extern __shared__ int32_t sh_cnt[];
    int32_t* sh_base = sh_cnt + nb;
    for (int i = threadIdx.x; i < nb; i += kSortBlock) sh_cnt[i] = 0;
    __syncthreads();
    const int base = blockIdx.x * kSortBlock * kSortItems + threadIdx.x;
    int rank[kSortItems], key[kSortItems];
#pragma unroll
    for (int it = 0; it < kSortItems; ++it) {
        const int v = base + it * kSortBlock;
        const int b = v < V ? keys[v] : -1;
        key[it] = key_ok(b, nb) ? b : -1;
        rank[it] = key[it] >= 0 ? atomicAdd(&sh_cnt[key[it]], 1) : -1;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += kSortBlock)
        if (sh_cnt[i]) sh_base[i] = atomicAdd(&cursor[i], sh_cnt[i]);
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kSortItems; ++it)
        if (key[it] >= 0) perm[sh_base[key[it]] + rank[it]] = base + it * kSortBlock;
}

// Many-bucket variants (nb > kSortLdsMax): one global atomic per voxel.
__global__ void k_bucket_count_global(const int32_t* __restrict__ keys, int V, int nb, int32_t* __restrict__ cnt) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V && key_ok(keys[v], nb)) atomicAdd(&cnt[keys[v]], 1);
}
__global__ void k_bucket_place_global(const int32_t* __restrict__ keys, int V, int nb, int32_t* __restrict__ cursor,
                                      int32_t* __restrict__ perm) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V && key_ok(keys[v], nb)) perm[atomicAdd(&cursor[keys[v]], 1)] = v;
}

// Exclusive scan of the padded counts (one block): off[0..nb], cursor = off.
__global__ void __launch_bounds__(1024) k_bucket_scan(const int32_t* __restrict__ cnt, int nb, int pad,
                                                      int32_t* __restrict__ off, int32_t* __restrict__ cursor) {
    __shared__ int32_t part[1024];
    const int tid = threadIdx.x;
    const int per = (nb + 1023) / 1024;
    const int b0 = tid * per, b1 = min(nb, b0 + per);
    int s = 0;
    for (int i = b0; i < b1; ++i) s += (cnt[i] + pad - 1) / pad * pad;
    part[tid] = s;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        const int x = tid >= d ? part[tid - d] : 0;
        __syncthreads();
        part[tid] += x;
        __syncthreads();
    }
    int run = tid ? part[tid - 1] : 0;
    for (int i = b0; i < b1; ++i) {
        off[i] = run;
        cursor[i] = run;
        run += (cnt[i] + pad - 1) / pad * pad;
    }
    if (tid == 1023) off[nb] = part[1023];
}

// Stream compaction of flagged voxels: one global atomic per wave.
__global__ void __launch_bounds__(256) k_compact_flags(const int32_t* __restrict__ flags, int V,
                                                       int32_t* __restrict__ idx, int32_t* __restrict__ count) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    const bool f = v < V && flags[v] != 0;
    const uint64_t m = __ballot(f);
    if (m == 0) return;
    const int lane = threadIdx.x & 63;
    const int n = __popcll(m);
    int base = 0;
    if (lane == __ffsll((unsigned long long)m) - 1) base = atomicAdd(count, n);
    base = __shfl(base, __ffsll((unsigned long long)m) - 1);
    if (f) idx[base + __popcll(m & ((1ull << lane) - 1))] = v;
}

// ---------------------------------------------------------------------------
// Tiles of a padded grouping (groups padded to multiples of `width`): tile t
// covers columns [t*width, (t+1)*width) of the group whose run contains them.
// ---------------------------------------------------------------------------
__global__ void k_group_tiles(const int32_t* __restrict__ off, int ng, int max_tiles, int width,
                              int32_t* __restrict__ tile_group, int32_t* __restrict__ tile_col,
                              int32_t* __restrict__ n_tiles) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int total = off[ng];
    if (t == 0) *n_tiles = total / width;
    if (t >= max_tiles) return;
    const int col = t * width;
    if (col >= total) {
        tile_group[t] = -1;
        tile_col[t] = col;
        return;
    }
    int lo = 0, hi = ng;  // off[lo] <= col < off[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= col)
            lo = mid;
        else
            hi = mid;
    }
    tile_group[t] = lo;
    tile_col[t] = col;
}

// ---------------------------------------------------------------------------
// (iv) FP64 GLS refit on f64 MFMA, one wave per 64 voxels of one pair.
// Per voxel (y normalised by its scale s):
//   h = G_k' y            (G_k = R^-1 X Lx^-T; MFMA, 16-row tiles)
//   q = X' y              (MFMA)
//   E = |L^-1 y|^2        (FP64 innovations recurrence, per lane)
//   beta = Lx^-T h,  ssw = E - |h|^2 (= |L^-1 (y - X beta)|^2),
//   ssres = y'y - 2 beta'q + beta' X'X beta,  sstot = y'y - T ybar^2,
// which are the reference's whitened-residual sums and CoD R^2
// (glm/arma.py:4784-5601) rewritten as sums over the series.
// The voxels arrive grouped by pair with every group padded to 64 slots, so
// the pair and its tables are wave-uniform.  MFMA k-slot q of step m covers
// time t0 + 8q + m, so each lane reads runs of consecutive samples of a voxel row.
// 16x16x4 f64 layouts: A lane l = A[l%16][l/16], B lane l = B[l/16][l%16],
// D lane l element r = D[l/16 + 4r][l%16].
// ---------------------------------------------------------------------------
// Samples [t0, t0 + N) of a row; zero past T.  `full` (wave-uniform) says the
// span lies inside the row and is 16-byte aligned; otherwise each sample is
// read at a clamped index and zeroed past T (no divergent branch).
template <int N>
__device__ __forceinline__ void load_span(const float* __restrict__ row, int T, int t0, bool full, float (&out)[N]) {
    if (full) {
#pragma unroll
        for (int i = 0; i < N / 4; ++i) {
            const f32x4 q = *reinterpret_cast<const f32x4*>(row + t0 + 4 * i);
            out[4 * i] = q.x;
            out[4 * i + 1] = q.y;
            out[4 * i + 2] = q.z;
            out[4 * i + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const float x = row[min(t0 + i, T - 1)];
            out[i] = (t0 + i < T) ? x : 0.f;
        }
    }
}

// Samples of one 16-sample half of a 32-sample block of the GLS accumulation
// (see k_gls_mfma): MFMA samples t0 + 8 kq + 4 hf + m of the four 16-voxel
// groups.  gls_half fetches the rest when it runs: the half's G / X rows
// (L2-resident tables) and the recurrence samples t0 + 16 hf .. + 16 of the
// lane's voxel (lines the MFMA loads of the block have just brought in).
template <int NT>
struct GlsHalf {
    float yb[4][4];
};

template <int NT>
__device__ __forceinline__ void gls_load_half(int t0, int hf, int T, bool full, int j, int kq,
                                              const float* const (&rowB)[4], const float* rowO,
                                              const double* __restrict__ Gk, const double* __restrict__ Xp,
                                              GlsHalf<NT>& h) {
    constexpr int PG = 16 * NT;
#pragma unroll
    for (int g = 0; g < 4; ++g) load_span<4>(rowB[g], T, t0 + 8 * kq + 4 * hf, full, h.yb[g]);
}

template <int NT>
__device__ __forceinline__ void gls_half(int t0, int hf, int j, int kq, int T, bool full, const float* rowO,
                                         const double* __restrict__ Gk, const double* __restrict__ Xp,
                                         const GlsHalf<NT>& h, const double* cbuf, f64x4 (&accG)[4][NT],
                                         f64x4 (&accX)[4][NT], double& E, double& sy, double& syy, double& e,
                                         double& yprev) {
    constexpr int PG = 16 * NT;
    float yo[16];
    load_span<16>(rowO, T, t0 + 16 * hf, full, yo);
    // the half's G / X rows (L2-resident tables), in flight during the recurrence
    double ag[4][NT], ax[4][NT];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int rt = 0; rt < NT; ++rt) {
            const size_t o = (size_t)(t0 + 8 * kq + 4 * hf + m) * PG + 16 * rt + j;
            ag[m][rt] = Gk[o];
            ax[m][rt] = Xp[o];
        }
    // Rows t >= T have zero coefficients and zero samples: they add nothing.
#pragma unroll 4
    for (int i = 0; i < 16; ++i) {
        const int r = 16 * hf + i;
        const double yt = (double)yo[i];
        const double gs = cbuf[4 * r], c = cbuf[4 * r + 1], al = cbuf[4 * r + 2];
        e = fma(al, e, fma(-c, yprev, yt));
        yprev = yt;
        const double w = e * gs;
        E = fma(w, w, E);
        sy += yt;
        syy = fma(yt, yt, syy);
    }
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const double b = (double)h.yb[g][m];
#pragma unroll
            for (int rt = 0; rt < NT; ++rt) {
                accG[g][rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(ag[m][rt], b, accG[g][rt], 0, 0, 0);
                accX[g][rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(ax[m][rt], b, accX[g][rt], 0, 0, 0);
            }
        }
}

// Epilogue tables of the f64 MFMA GLS kernels for pair k (one wave), padded
// to PG x PG (PG = PGA) with zeros, so the epilogues run fixed-size loops whose
// padded terms add exact zeros.  First block: for k_gls_mfma (p <= 32) Lx
// transposed (LT[j][m] = Lx[m][j] for m >= j, zero above); for k_gls_wide
// (32 < p <= 96) M = Lx^-T, column i the back substitution Lx' m = e_i with
// the zero-pivot rule (m_j = 0 where Lx[j][j] = 0), so beta = M h is the
// kernels' basic solution.  Then the F quadratic form and diag (X'R^-1X)^-1.
__global__ void __launch_bounds__(64) k_gls_epi_tables(PlanDev P, double* __restrict__ out) {
    const int k = blockIdx.x;
    if (k >= P.K) return;
    const int p = P.p, PG = P.PGA;
    const double* __restrict__ L = P.Lx + (size_t)k * p * p;
    const double* __restrict__ F = P.Fm + (size_t)k * p * p;
    const double* __restrict__ A = P.Ainv + (size_t)k * p * p;
    double* o = out + (size_t)k * (2 * PG * PG + PG);
    for (int i = threadIdx.x; i < PG * PG; i += 64) {
        const int r = i / PG, c = i % PG;
        const bool in = r < p && c < p;
        if (P.PG) o[i] = in && c >= r ? L[(size_t)c * p + r] : 0.0;  // Lx lower triangle only
        else if (!in) o[i] = 0.0;
        o[PG * PG + i] = in ? F[(size_t)r * p + c] : 0.0;
    }
    for (int j = threadIdx.x; j < PG; j += 64) o[2 * PG * PG + j] = j < p ? A[(size_t)j * p + j] : 0.0;
    if (!P.PG) {
        // column i of M, one lane per column, rows from the last up (each lane
        // reads only the entries of its own column it has written)
        for (int i = threadIdx.x; i < p; i += 64)
            for (int jj = p - 1; jj >= 0; --jj) {
                double acc = jj == i ? 1.0 : 0.0;
                for (int m = jj + 1; m < p; ++m) acc -= L[(size_t)m * p + jj] * o[(size_t)m * PG + i];
                const double ljj = L[(size_t)jj * p + jj];
                o[(size_t)jj * PG + i] = ljj != 0.0 ? acc / ljj : 0.0;
            }
    }
}

template <int NT>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) k_gls_mfma(PlanDev P, const float* __restrict__ Y, int64_t ldY,
                                                 const float* __restrict__ scale, const int32_t* __restrict__ perm,
                                                 const int32_t* __restrict__ best, float* __restrict__ betas,
                                                 float* __restrict__ tstats, float* __restrict__ fstats,
                                                 float* __restrict__ sigma2, float* __restrict__ r2,
                                                 const PlanDev* __restrict__ plans,
                                                 const int32_t* __restrict__ vgroup) {
    constexpr int PG = 16 * NT;
    constexpr int TP = PG + 1;  // transpose pitch (doubles)
    __shared__ double cbuf[32 * 4];
    __shared__ double tbuf[64 * TP];
    __shared__ double bes[64 * TP];
    const int lane = threadIdx.x;
    const int64_t slot0 = (int64_t)blockIdx.x * 64;
    const int v0 = __builtin_amdgcn_readfirstlane(perm[slot0]);
    if (v0 < 0) return;  // padding or beyond the last group
    if (plans) P = plans[vgroup[v0]];  // grouped: one (design, pair) group per 64 slots
    const int k = __builtin_amdgcn_readfirstlane(best[v0]);
    const int v = perm[slot0 + lane];
    const int T = P.T, Tpad = P.Tpad, p = P.p;
    const int j = lane & 15, kq = lane >> 4;
    const double* __restrict__ Gk = P.G64p + (size_t)k * Tpad * PG;
    const double* __restrict__ Xp = P.Xp;
    const double* __restrict__ cR = P.coefR + (size_t)k * Tpad * 4;
    const bool vec4 = (ldY & 3) == 0 && (reinterpret_cast<uintptr_t>(Y) & 15) == 0;
    const float* rowB[4];
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        const int vb = perm[slot0 + 16 * g + j];
        rowB[g] = Y + (size_t)(vb >= 0 ? vb : v0) * ldY;  // padding columns are computed and discarded
    }
    const float* rowO = Y + (size_t)(v >= 0 ? v : v0) * ldY;

    f64x4 accG[4][NT], accX[4][NT];
#pragma unroll
    for (int g = 0; g < 4; ++g)
#pragma unroll
        for (int rt = 0; rt < NT; ++rt) {
            accG[g][rt] = f64x4{0.0, 0.0, 0.0, 0.0};
            accX[g][rt] = f64x4{0.0, 0.0, 0.0, 0.0};
        }
    double E = 0.0, sy = 0.0, syy = 0.0, e = 0.0, yprev = 0.0;
    // 32-sample blocks, each two halves; the operands of the next half (and
    // the next block's recurrence coefficients) are fetched before the
    // current half is computed (one wave per block: the loads must overlap).
    GlsHalf<NT> hn;
    gls_load_half<NT>(0, 0, T, vec4 && 32 <= T, j, kq, rowB, rowO, Gk, Xp, hn);
    double c0 = cR[2 * lane], c1 = cR[2 * lane + 1];
    for (int t0 = 0; t0 < Tpad; t0 += 32) {
        const bool full = vec4 && t0 + 32 <= T;
        __syncthreads();  // the previous block's coefficient reads are done
        cbuf[2 * lane] = c0;
        cbuf[2 * lane + 1] = c1;
        __syncthreads();
        if (t0 + 32 < Tpad) {
            c0 = cR[(size_t)(t0 + 32) * 4 + 2 * lane];
            c1 = cR[(size_t)(t0 + 32) * 4 + 2 * lane + 1];
        }
#pragma unroll
        for (int hf = 0; hf < 2; ++hf) {
            const GlsHalf<NT> h = hn;
            if (hf == 0)
                gls_load_half<NT>(t0, 1, T, full, j, kq, rowB, rowO, Gk, Xp, hn);
            else if (t0 + 32 < Tpad)
                gls_load_half<NT>(t0 + 32, 0, T, vec4 && t0 + 64 <= T, j, kq, rowB, rowO, Gk, Xp, hn);
            gls_half<NT>(t0, hf, j, kq, T, full, rowO, Gk, Xp, h, cbuf, accG, accX, E, sy, syy, e, yprev);
        }
    }
    // Accumulators -> one voxel per lane through LDS: h first (beta, |h|^2),
    // then q (read back element by element).
#pragma unroll
    for (int g = 0; g < 4; ++g)
#pragma unroll
        for (int rt = 0; rt < NT; ++rt)
#pragma unroll
            for (int r = 0; r < 4; ++r) tbuf[(16 * g + j) * TP + 16 * rt + kq + 4 * r] = accG[g][rt][r];
    __syncthreads();
    const double s = (v >= 0 && scale) ? (double)scale[v] : 1.0;
    const double inv = 1.0 / s, inv2 = inv * inv;
    E *= inv2;
    sy *= inv;
    syy *= inv2;
    // beta = Lx^-T h (back substitution on the upper factor Lx'), h = G'y / s.
    // beta lives in the lane's LDS row, zero-filled to PG; the inner loops run
    // over the padded width PG with the epilogue tables (Lx' zero below its
    // diagonal and past p, X'X and F zero past p), so every padded term adds
    // an exact zero and the sums equal those over p columns, while the loads of
    // a row are issued together.  The outer loops stay runtime loops (their
    // index addresses the LDS row).
    const double* __restrict__ ep = P.gls_epi + (size_t)k * (2 * PG * PG + PG);
    const double* __restrict__ LT = ep;
    const double* __restrict__ Fq = ep + PG * PG;
    const double* __restrict__ Ad = ep + 2 * PG * PG;
    double* be = bes + lane * TP;
#pragma unroll
    for (int m = 0; m < PG; ++m) be[m] = 0.0;
    double hh = 0.0;
#pragma unroll 1
    for (int jj = p - 1; jj >= 0; --jj) {
        const double hj = tbuf[lane * TP + jj] * inv;
        hh = fma(hj, hj, hh);
        double acc = hj;
#pragma unroll
        for (int m = 0; m < PG; ++m) acc -= LT[jj * PG + m] * be[m];  // m <= jj: zero terms
        const double ljj = LT[jj * PG + jj];
        be[jj] = ljj != 0.0 ? acc / ljj : 0.0;  // zero pivot: basic solution
    }
    __syncthreads();
#pragma unroll
    for (int g = 0; g < 4; ++g)
#pragma unroll
        for (int rt = 0; rt < NT; ++rt)
#pragma unroll
            for (int r = 0; r < 4; ++r) tbuf[(16 * g + j) * TP + 16 * rt + kq + 4 * r] = accX[g][rt][r];
    __syncthreads();
    if (v < 0) return;
    double bq = 0.0, bxb = 0.0;
    const double* __restrict__ XtX = P.XtXp;
#pragma unroll 1
    for (int a = 0; a < p; ++a) {
        bq = fma(be[a], tbuf[lane * TP + a] * inv, bq);
        double ra = 0.0;
#pragma unroll
        for (int b2 = 0; b2 < PG; ++b2) ra = fma(XtX[a * PG + b2], be[b2], ra);
        bxb = fma(be[a], ra, bxb);
    }
    const double ssw = E - hh;
    const double ssres = syy - 2.0 * bq + bxb;
    const double sstot = syy - sy * (sy / T);
    const int dof = max(1, T - p);
    const double sig = ssw / dof;
#pragma unroll 1
    for (int jj = 0; jj < p; ++jj) {
        const double se = sqrt(sig * Ad[jj]);
        tstats[(size_t)v * p + jj] = (float)(be[jj] / (se + 1e-10));
        betas[(size_t)v * p + jj] = (float)(be[jj] * s);
    }
    double quad = 0.0;
#pragma unroll 1
    for (int a = 0; a < p; ++a) {
        double ra = 0.0;
#pragma unroll
        for (int b2 = 0; b2 < PG; ++b2) ra += Fq[a * PG + b2] * be[b2];
        quad += be[a] * ra;
    }
    fstats[v] = (float)(quad / (sig * P.n_task + 1e-10));
    sigma2[v] = (float)(sig * s * s);
    r2[v] = (float)(sstot <= 1e-10 ? 0.0 : fmin(1.0 - ssres / sstot, 1.0));
}

// Rank-deficient designs (plan Pi set): the GLS kernels leave the basic
// solution (beta 0 on the dependent columns); the reported beta is the
// minimum-norm solution Pi beta_basic, the solution of the reference's pinv
// fallback (glm/arma.py:3410-3416), with t from diag (X'R^-1X)^+ and F from
// the plan's quadratic form on that beta.  Fitted values, residuals, sigma^2
// and R^2 are the same for every least-squares solution and stay as written.
// One wave per voxel; grouped fits pass plans / vgroup.
__global__ void __launch_bounds__(64) k_gls_rank(PlanDev P, int V, const float* __restrict__ scale,
                                                 const int32_t* __restrict__ best, float* __restrict__ betas,
                                                 float* __restrict__ tstats, float* __restrict__ fstats,
                                                 const float* __restrict__ sigma2, const PlanDev* __restrict__ plans,
                                                 const int32_t* __restrict__ vgroup) {
    const int v = blockIdx.x;
    if (v >= V) return;
    if (plans) P = plans[vgroup[v]];
    if (!P.Pi) return;
    const int k = best[v];
    if (k < 0 || k >= P.K) return;
    const int p = P.p, tid = threadIdx.x;
    extern __shared__ double rank_lds[];  // 2 p doubles
    double* bb = rank_lds;
    double* bm = rank_lds + p;
    const double s = scale ? (double)scale[v] : 1.0;
    for (int j = tid; j < p; j += 64) bb[j] = (double)betas[(size_t)v * p + j] / s;
    __syncthreads();
    for (int i = tid; i < p; i += 64) {
        double acc = 0.0;
        for (int j = 0; j < p; ++j) acc = fma(P.Pi[i * p + j], bb[j], acc);
        bm[i] = acc;
    }
    __syncthreads();
    const double sig = (double)sigma2[v] / (s * s);
    const double* __restrict__ Ai = P.Ainv + (size_t)k * p * p;
    const double* __restrict__ F = P.Fm + (size_t)k * p * p;
    double part = 0.0;
    for (int i = tid; i < p; i += 64) {
        const double se = sqrt(sig * Ai[i * p + i]);
        tstats[(size_t)v * p + i] = (float)(bm[i] / (se + 1e-10));
        betas[(size_t)v * p + i] = (float)(bm[i] * s);
        double r = 0.0;
        for (int j = 0; j < p; ++j) r = fma(F[i * p + j], bm[j], r);
        part = fma(bm[i], r, part);
    }
    for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o);
    if (tid == 0) fstats[v] = (float)(part / (sig * P.n_task + 1e-10));
}

// NaN statistics for voxels whose pair index is not a valid pair.
__global__ void k_gls_invalid(const int32_t* __restrict__ best, int V, int K, int p, float* __restrict__ betas,
                              float* __restrict__ tstats, float* __restrict__ fstats, float* __restrict__ sigma2,
                              float* __restrict__ r2) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const int k = best[v];
    if (k >= 0 && k < K) return;
    for (int j = 0; j < p; ++j) {
        betas[(size_t)v * p + j] = NAN;
        tstats[(size_t)v * p + j] = NAN;
    }
    fstats[v] = sigma2[v] = r2[v] = NAN;
}

}  // namespace remlk
