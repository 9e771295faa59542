// Pair-set search: voxel groups scored on a set of pairs with the arithmetic of
// the full grid.  Included by reml_api.hip after reml_proj16.hip (one
// translation unit).
//
// Two families of sets use it:
//   windows  the 5 x 5 lattice window around a voxel's proposal (learned path);
//   bands    the pairs whose lag-1 correlation
//            lambda = (1 + ab)(a + b) / (1 + 2ab + b^2) lies near the
//            proposal's.  The REML surface of a weakly correlated series is a
//            long valley along a lambda contour: the proposal network places
//            such a voxel's lambda well but not its position along the valley.
// Voxels are grouped by set (groups padded to 64 columns of a time-major
// residual buffer) and every pair of their set is scored so that each score
// equals the full-grid surface entry bitwise:
//   energy      k_set_energy: the k_energy recurrence and accumulation
//               schedule, 8 slots per wave; the AR-filtered sample is formed per
//               pair, which for pairs sharing a is the value k_energy shares;
//   projection  k_set_project: one wave per 8-pair tile of the set, the same
//               MFMA sequence per output column as k_project (fp32) or
//               k_project16b (split 16-bit), B gathered column by column from
//               the full-grid operand;
//   argmin      k_set_select: first minimum in pair order, lambda == 0 pairs
//               mapped to the first of them (the global first-minimum rule).
#pragma once

#include "reml_kernels.h"

namespace remlk {

// Energy chunk coefficients of a set table, one block per (set, chunk): per
// step the c_t of each group (taken from its first pair: the pairs of a group
// share it), then (alpha_t, 1/v_{t+1}) of each position, zero for empty
// positions; s0 = 1/v_0.  The same fp32 values as the full-grid chunk tables.
__global__ void k_set_tables(PlanDev P, SetTable ST, float* __restrict__ coef, float* __restrict__ s0) {
    const int cb = blockIdx.x;
    if (cb >= ST.n * ST.mc) return;
    const int set = cb / ST.mc, m = cb % ST.mc;
    if (m >= ST.ecnt[set]) return;
    const int T = P.T, NB = ST.enb, NG = ST.eng, GS = NB / NG, W = NG + 2 * NB;
    const int32_t* pos = ST.eslot + (size_t)cb * NB;
    const int32_t* slots = ST.pairs + (size_t)set * ST.sb;
    for (int t = threadIdx.x; t < T; t += blockDim.x) {
        float* o = coef + ((size_t)cb * T + t) * W;
        for (int g = 0; g < NG; ++g) {
            float c = 0.f;
            for (int i = GS - 1; i >= 0; --i) {
                const int q = pos[g * GS + i];
                if (q >= 0) c = P.lc[(size_t)slots[q] * T + t];
            }
            o[g] = c;
        }
        for (int j = 0; j < NB; ++j) {
            const int q = pos[j];
            float al = 0.f, sn = 0.f;
            if (q >= 0) {
                const float2 v = P.lp[(size_t)slots[q] * T + t];
                al = v.x;
                sn = v.y;
            }
            o[NG + 2 * j] = al;
            o[NG + 2 * j + 1] = sn;
        }
    }
    for (int j = threadIdx.x; j < NB; j += blockDim.x) {
        const int q = pos[j];
        s0[(size_t)cb * NB + j] = q >= 0 ? P.ls0[slots[q]] : 0.f;
    }
}

__device__ inline float axis_clamp(const float* axis, int n, float x) {
    float lo = axis[0], hi = axis[0];
    for (int i = 1; i < n; ++i) {
        lo = fminf(lo, axis[i]);
        hi = fmaxf(hi, axis[i]);
    }
    return fminf(fmaxf(x, lo), hi);
}

__device__ inline int nearest_index(const float* axis, int n, float x) {
    int best = 0;
    float bd = fabsf(axis[0] - x);
    for (int i = 1; i < n; ++i) {
        const float d = fabsf(axis[i] - x);
        if (d < bd) {
            bd = d;
            best = i;
        }
    }
    return best;
}

// Window centre kept inside the lattice.
__device__ inline int clamp_center(const PlanDev& P, int ai, int bi) {
    if (P.n_a >= kWin) ai = min(max(ai, kWinHalf), P.n_a - 1 - kWinHalf);
    if (P.n_b >= kWin) bi = min(max(bi, kWinHalf), P.n_b - 1 - kWinHalf);
    return ai * P.n_b + bi;
}

// Window of each voxel from its proposal (a_hat, b_hat, c): the window around
// the nearest lattice point, kept inside the lattice.  Voxels whose confidence
// c < conf_min or whose proposal is not finite get need = 2 and no window
// (key -1); the others need = 0.
__global__ void k_win_key(PlanDev P, const float* __restrict__ props, int V, float conf_min,
                          int32_t* __restrict__ key, int32_t* __restrict__ need) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const float ah = props[3 * v], bh = props[3 * v + 1], c = props[3 * v + 2];
    const bool finite = isfinite(ah) && isfinite(bh) && isfinite(c);
    if (!finite || !(c >= conf_min)) {
        key[v] = -1;
        need[v] = 2;
        return;
    }
    key[v] = clamp_center(P, nearest_index(P.a_axis, P.n_a, ah), nearest_index(P.b_axis, P.n_b, bh));
    need[v] = 0;
}

// Band of each voxel still flagged (need[v] != 0) whose proposal is finite:
// lambda of the proposal clamped to the lattice box, nearest band centre; -1
// for every other voxel and for empty bands.
__global__ void k_band_key(PlanDev P, const float* __restrict__ props, int V, const int32_t* __restrict__ need,
                           int32_t* __restrict__ key) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    int out = -1;
    if (need[v] != 0) {
        const float ah = props[3 * v], bh = props[3 * v + 1];
        if (isfinite(ah) && isfinite(bh)) {
            const double a = axis_clamp(P.a_axis, P.n_a, ah), b = axis_clamp(P.b_axis, P.n_b, bh);
            const double den = 1.0 + 2.0 * a * b + b * b;
            double lam = den > 0.0 ? (1.0 + a * b) * (a + b) / den : 0.0;
            lam = fmin(fmax(lam, -1.0), 1.0);
            int bin = (int)rint((lam - (double)P.band_lo) / (double)P.band_h);
            bin = min(max(bin, 0), P.band.n - 1);
            if (P.band.cnt[bin] > 0) out = bin;
        }
    }
    key[v] = out;
}

// Work items of a set pass are split into 8 contiguous tile ranges, one per
// XCD (block id % 8), whose blocks stride over them: the items of one voxel
// tile share that XCD's L2.
struct SetItems {
    int lo, n, lb, nlb;
};
__device__ inline SetItems set_items(int n_tiles, int per_tile) {
    SetItems r;
    const int xcd = blockIdx.x & 7;
    const int per = (n_tiles + 7) >> 3;
    r.lo = xcd * per;
    r.n = max(0, min(n_tiles, r.lo + per) - r.lo) * per_tile;
    r.lb = blockIdx.x >> 3;
    r.nlb = gridDim.x >> 3;
    return r;
}

// Energies E[col][slot] of every live 64-column tile and every energy chunk of
// its set: the k_energy recurrence and fp64 flush schedule, so each energy
// equals the full-grid one bitwise.  NB positions per chunk in NG groups; per
// step the filtered sample z = y_{t+1} - c_t y_t of each group, then per
// position e <- alpha e + z, E += e^2 / v.  Items are (octet of consecutive
// tiles, chunk), chunk fastest, each XCD (block id % 8) taking a contiguous
// eighth: the kSetEnergyWaves waves of a block run one chunk on the octet's
// tiles, which mostly share a set and so its coefficient stream (scalar-cache
// hits), and the chunks of an octet run together on one XCD, so its residual
// columns come from that XCD's L2 after the first.  The grid is sized for the
// largest tile count; blocks past the live items return at once.
template <int NB, int NG>
__global__ void __launch_bounds__(64 * kSetEnergyWaves) k_set_energy(PlanDev P, SetTable ST,
                                                                     const float* __restrict__ Yt, int64_t ldY,
                                                                     const int32_t* __restrict__ tset,
                                                                     const int32_t* __restrict__ n_tiles,
                                                                     float* __restrict__ Eb) {
    constexpr int GS = NB / NG, W = NG + 2 * NB;
    static_assert(GS * NG == NB, "groups must split the chunk evenly");
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nt = *n_tiles, MC = ST.mc;
    const int n_xo = (nt + kSetEnergyWaves - 1) / kSetEnergyWaves;
    const int n_items = n_xo * MC;
    const int per = (n_items + 7) >> 3;
    const int local = blockIdx.x >> 3;
    const int item = (blockIdx.x & 7) * per + local;
    if (local >= per || item >= n_items) return;
    const int xo = item / MC, chunk = item % MC;
    const int tt = xo * kSetEnergyWaves + w;
    if (tt >= nt) return;
    const int set = tset[tt];
    if (chunk >= ST.ecnt[set]) return;
    const int T = P.T;
    const size_t cb = (size_t)set * MC + chunk;
    const int col = tt * kSetTile + lane;
    const float* __restrict__ y = Yt + col;
    const float* __restrict__ co = ST.coef + cb * T * W;
    const float* __restrict__ s0 = ST.s0 + cb * NB;
    float e[NB], part[NB];
    double acc[NB];
    const float y0 = y[0];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        e[j] = y0;
        part[j] = y0 * y0 * s0[j];
        acc[j] = 0.0;
    }
    float yc = y0;
    auto step = [&](float yn, const float* c) __attribute__((always_inline)) {
        float z[NG];
#pragma unroll
        for (int g = 0; g < NG; ++g) z[g] = fmaf(-c[g], yc, yn);
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            e[j] = fmaf(c[NG + 2 * j], e[j], z[j / GS]);
            part[j] = fmaf(e[j] * c[NG + 2 * j + 1], e[j], part[j]);
        }
        yc = yn;
    };
    constexpr int U = 8;  // steps per fp64 flush
    constexpr int G = 2;  // steps whose coefficients are fetched together
    const int nsteps = T - 1;
    const int nfull = nsteps / U * U;
    // samples of the next block, fetched one block ahead (indices clamped: the
    // values past T are never used)
    float ybuf[U];
#pragma unroll
    for (int i = 0; i < U; ++i) ybuf[i] = y[(size_t)min(1 + i, T - 1) * ldY];
    for (int t0 = 0; t0 < nfull; t0 += U) {
        float ycur[U];
#pragma unroll
        for (int i = 0; i < U; ++i) ycur[i] = ybuf[i];
#pragma unroll
        for (int i = 0; i < U; ++i) ybuf[i] = y[(size_t)min(t0 + U + 1 + i, T - 1) * ldY];
#pragma unroll
        for (int g0 = 0; g0 < U; g0 += G) {
            float cg[G][W];
#pragma unroll
            for (int i = 0; i < G; ++i)
#pragma unroll
                for (int q = 0; q < W; ++q) cg[i][q] = co[(size_t)(t0 + g0 + i) * W + q];
#pragma unroll
            for (int i = 0; i < G; ++i) step(ycur[g0 + i], cg[i]);
        }
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            acc[j] += part[j];
            part[j] = 0.f;
        }
    }
    for (int t = nfull; t < nsteps; ++t) step(y[(size_t)(t + 1) * ldY], co + (size_t)t * W);
    float* out = Eb + (size_t)col * ST.sb;
    const int32_t* pos = ST.eslot + cb * NB;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        const int q = pos[j];
        if (q >= 0) out[q] = (float)(acc[j] + part[j]);
    }
}

// Projection + score of every live 64-column tile over its set: items
// (tile, quad of 4 set tiles); wave w takes set tile 4 quad + w (npt pairs,
// BN columns) for all 64 columns.  MODE 0: fp32 MFMA 32x32x2 over the fp32
// residual rows and the fp32 operand (k_project's k-order); MODE 1 / 2: split
// bf16 / fp16 halves on 32x32x16 (k_project16b's order: per 16-sample k-step
// al*bh, ah*bl, ah*bh).  Scores go over the energies in Sb[col][slot].
// SEG: p > 96, each pair's columns in P.nseg segments of BN (one pair per set
// tile); a separate instantiation, so the narrow ones keep their schedule.
template <int BN, int MODE, bool SEG>
__global__ void __launch_bounds__(64 * kSetWaves) k_set_project(PlanDev P, SetTable ST, const float* __restrict__ Yt,
                                                                int64_t ldY, const int32_t* __restrict__ tset,
                                                                const int32_t* __restrict__ n_tiles,
                                                                float* __restrict__ Sb) {
    constexpr int NBLK = BN / 32;
    constexpr int C_LD = BN + 1;
    __shared__ float patch[kSetWaves][32 * C_LD];
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int SB = ST.sb, npt = P.npt, p = P.p;
    const int QT = SB / (kSetWaves * npt);  // quads per set
    const SetItems I = set_items(*n_tiles, QT);
    const double tp = (double)(P.T - p);
    float* Cw = patch[w];
    for (int it = I.lb; it < I.n; it += I.nlb) {
        const int tt = I.lo + it / QT, quad = it % QT;
        const int set = tset[tt];
        if (quad * kSetWaves * npt >= ST.cnt[set]) continue;  // block-uniform
        const int stile = quad * kSetWaves + w;
        const int32_t* __restrict__ kp = ST.pairs + (size_t)set * SB + stile * npt;
        const int col0 = tt * kSetTile;
        // p > 96 (P.nseg segments of BN columns, one pair per set tile): the
        // products of each segment in turn, each row's squares summed across
        // them (lanes 0..31, one sum per row block), the score after the last
        const int nsg = SEG ? P.nseg : 1;
        double spp[2] = {0.0, 0.0};
        for (int sg = 0; sg < nsg; ++sg) {
        // this lane's B column in n-block n: source tile and column of its pair
        int bst[NBLK], bcc[NBLK];
        bool bok[NBLK];
#pragma unroll
        for (int n = 0; n < NBLK; ++n) {
            const int c = n * 32 + (lane & 31);
            if (SEG) {
                const int k = kp[0];
                bok[n] = k >= 0 && sg * BN + c < p;
                bst[n] = k >= 0 ? k * P.nseg + sg : 0;
                bcc[n] = c;
                continue;
            }
            const int q = c / p, j = c - q * p;
            const int k = q < npt ? kp[q] : -1;
            bok[n] = k >= 0;
            bst[n] = k >= 0 ? k / npt : 0;
            bcc[n] = k >= 0 ? (k - bst[n] * npt) * p + j : 0;
        }
        f32x16 acc[2][NBLK];
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int n = 0; n < NBLK; ++n)
#pragma unroll
                for (int q = 0; q < 16; ++q) acc[r][n][q] = 0.f;
        if (MODE == 0) {
            const float* __restrict__ G = P.G32;
            size_t boff[NBLK];
#pragma unroll
            for (int n = 0; n < NBLK; ++n) boff[n] = (size_t)bst[n] * P.Tpad * BN + bcc[n];
            const int ar = lane >> 5;
            const float* __restrict__ ya = Yt + col0 + (lane & 31);
            for (int t0 = 0; t0 < P.Tpad; t0 += 8) {
                float av[4][2], bv[4][NBLK];
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const int t = t0 + 2 * s + ar;
#pragma unroll
                    for (int r = 0; r < 2; ++r) av[s][r] = ya[(size_t)t * ldY + 32 * r];
#pragma unroll
                    for (int n = 0; n < NBLK; ++n) {
                        const float x = G[boff[n] + (size_t)t * BN];
                        bv[s][n] = bok[n] ? x : 0.f;
                    }
                }
#pragma unroll
                for (int s = 0; s < 4; ++s)
#pragma unroll
                    for (int n = 0; n < NBLK; ++n)
#pragma unroll
                        for (int r = 0; r < 2; ++r)
                            acc[r][n] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[s][r], bv[s][n], acc[r][n], 0, 0, 0);
            }
        } else {
            // A (the block's 64 residual columns) is shared by the 4 waves:
            // staged through LDS per 32-sample stage (double-buffered); each
            // wave's B columns come straight from the operand into registers,
            // one stage ahead.
            constexpr bool BF16 = MODE == 1;
            const uint16_t* __restrict__ gh = BF16 ? P.Gh_bf16 : P.Gh_f16;
            const uint16_t* __restrict__ gl = BF16 ? P.Gl_bf16 : P.Gl_f16;
            const uint16_t* __restrict__ Ash = reinterpret_cast<const uint16_t*>(Yt + (size_t)P.Tpad * ldY);
            const uint16_t* __restrict__ Asl = Ash + (size_t)P.Tpad * ldY;
            uint16_t* As = reinterpret_cast<uint16_t*>(&patch[0][0]);  // [buf][plane][tb][col][8]
            constexpr int A_BUF = 2 * 4 * kSetTile * 8;                // halves per buffer
            size_t boff[NBLK];
#pragma unroll
            for (int n = 0; n < NBLK; ++n) boff[n] = (size_t)bst[n] * P.Tpad * BN + (size_t)bcc[n] * 8;
            const int nst = P.Tpad / 32;
            const int tid = threadIdx.x;
            auto load_a = [&](int kt, u32x4 (&ra)[2]) __attribute__((always_inline)) {
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const int q = tid + u * 256;  // plane = q / 256, t-block = (q / 64) % 4, column = q % 64
                    const size_t o = ((size_t)(kt * 4 + ((q >> 6) & 3)) * ldY + col0 + (q & 63)) * 8;
                    ra[u] = *reinterpret_cast<const u32x4*>(((q >> 8) ? Asl : Ash) + o);
                }
            };
            auto store_a = [&](int buf, const u32x4 (&ra)[2]) __attribute__((always_inline)) {
#pragma unroll
                for (int u = 0; u < 2; ++u)
                    *reinterpret_cast<u32x4*>(As + buf * A_BUF + (tid + u * 256) * 8) = ra[u];
            };
            // B of one 16-sample k-step (this lane's columns)
            auto load_b = [&](int ks, u32x4 (&bh)[NBLK], u32x4 (&bl)[NBLK]) __attribute__((always_inline)) {
                const int tb = 2 * ks + (lane >> 5);
#pragma unroll
                for (int n = 0; n < NBLK; ++n) {
                    const size_t o = boff[n] + (size_t)tb * BN * 8;
                    const u32x4 xh = *reinterpret_cast<const u32x4*>(gh + o);
                    const u32x4 xl = *reinterpret_cast<const u32x4*>(gl + o);
                    bh[n] = bok[n] ? xh : u32x4{0u, 0u, 0u, 0u};
                    bl[n] = bok[n] ? xl : u32x4{0u, 0u, 0u, 0u};
                }
            };
            __syncthreads();  // the patch (aliased by the A stages) is free
            // A stage st + 1 is fetched into registers during stage st - 1 and
            // stored to LDS at the end of stage st (two stages of latency cover)
            u32x4 ra[2], ra2[2];
            load_a(0, ra);
            store_a(0, ra);
            if (nst > 1) load_a(1, ra);
            u32x4 bh[NBLK], bl[NBLK];
            load_b(0, bh, bl);
            __syncthreads();
            const int nks = 2 * nst;
            for (int ks = 0; ks < nks; ++ks) {
                const int st = ks >> 1, buf = st & 1, half = ks & 1;
                if (half == 0 && st + 2 < nst) load_a(st + 2, ra2);
                u32x4 bhc[NBLK], blc[NBLK];
#pragma unroll
                for (int n = 0; n < NBLK; ++n) {
                    bhc[n] = bh[n];
                    blc[n] = bl[n];
                }
                if (ks + 1 < nks) load_b(ks + 1, bh, bl);
                const int tb = 2 * half + (lane >> 5);
                u32x4 ahc[2], alc[2];
#pragma unroll
                for (int r = 0; r < 2; ++r) {
                    const int c = 32 * r + (lane & 31);
                    ahc[r] = *reinterpret_cast<const u32x4*>(As + buf * A_BUF + ((0 * 4 + tb) * kSetTile + c) * 8);
                    alc[r] = *reinterpret_cast<const u32x4*>(As + buf * A_BUF + ((1 * 4 + tb) * kSetTile + c) * 8);
                }
                // per accumulator: al*bh, then ah*bl, then ah*bh
#pragma unroll
                for (int n = 0; n < NBLK; ++n)
#pragma unroll
                    for (int r = 0; r < 2; ++r) acc[r][n] = mfma16u<BF16>(alc[r], bhc[n], acc[r][n]);
#pragma unroll
                for (int n = 0; n < NBLK; ++n)
#pragma unroll
                    for (int r = 0; r < 2; ++r) acc[r][n] = mfma16u<BF16>(ahc[r], blc[n], acc[r][n]);
#pragma unroll
                for (int n = 0; n < NBLK; ++n)
#pragma unroll
                    for (int r = 0; r < 2; ++r) acc[r][n] = mfma16u<BF16>(ahc[r], bhc[n], acc[r][n]);
                if (half == 1) {
                    if (st + 1 < nst) store_a(buf ^ 1, ra);
#pragma unroll
                    for (int u = 0; u < 2; ++u) ra[u] = ra2[u];
                    __syncthreads();
                }
            }
        }
        // scores, one 32-row block at a time through the wave's patch
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            __syncthreads();  // the patch is free
#pragma unroll
            for (int n = 0; n < NBLK; ++n)
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    const int row = 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
                    Cw[row * C_LD + n * 32 + (lane & 31)] = acc[r][n][q];
                }
            __syncthreads();
            if (SEG) {
                if (lane < 32 && kp[0] >= 0) {
                    const float* cr = Cw + lane * C_LD;
                    const int nj = min(BN, p - sg * BN);
                    for (int j = 0; j < nj; ++j) {
                        const double cv = (double)cr[j];
                        spp[r] = fma(cv, cv, spp[r]);
                    }
                }
                continue;
            }
            for (int i = lane; i < 32 * npt; i += 64) {
                const int row = i / npt, qq = i % npt;
                const int k = kp[qq];
                if (k < 0) continue;
                const float* cr = Cw + row * C_LD + qq * p;
                double pp = 0.0;
                for (int j = 0; j < p; ++j) {
                    const double cv = (double)cr[j];
                    pp = fma(cv, cv, pp);
                }
                if (MODE == 2) pp *= (double)P.g16_unscale[k];
                float* sp = Sb + (size_t)(col0 + 32 * r + row) * SB + stile * npt + qq;
                const double rss = (double)(*sp) - pp;
                *sp = (float)fma(tp, log(rss + 1e-10), P.ld[k]);
            }
        }
        __syncthreads();
        }  // segments
        if (SEG && lane < 32 && kp[0] >= 0) {
            const int k = kp[0];
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                const double pp = MODE == 2 ? spp[r] * (double)P.g16_unscale[k] : spp[r];
                float* sp = Sb + (size_t)(col0 + 32 * r + lane) * SB + stile;
                const double rss = (double)(*sp) - pp;
                *sp = (float)fma(tp, log(rss + 1e-10), P.ld[k]);
            }
        }
    }
}

// First minimum over the set's slots (pair order), lambda == 0 pairs mapped
// to the first of them; need = 1 when that pair's edge flag is set or no slot
// has a finite score, else 0; next_key (nullable) = the window to continue
// with (edge) or -1.
__global__ void k_set_select(PlanDev P, SetTable ST, const int32_t* __restrict__ rows,
                             const int32_t* __restrict__ tset, const int32_t* __restrict__ n_tiles,
                             const float* __restrict__ Sb, int32_t* __restrict__ best, float* __restrict__ best_score,
                             int32_t* __restrict__ need, int32_t* __restrict__ next_key) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= *n_tiles * kSetTile) return;
    const int v = rows[c];
    if (v < 0) return;
    const int set = tset[c / kSetTile];
    const int SB = ST.sb, nb = ST.cnt[set];
    const float* sr = Sb + (size_t)c * SB;
    float bs = INFINITY;
    int bslot = -1;
    for (int s = 0; s < nb; ++s) {
        const float sc = sr[s];
        if (sc < bs) {
            bs = sc;
            bslot = s;
        }
    }
    if (bslot < 0) {
        best[v] = -1;
        best_score[v] = INFINITY;
        need[v] = 1;
        if (next_key) next_key[v] = -1;
        return;
    }
    const size_t o = (size_t)set * SB + bslot;
    int k = ST.pairs[o];
    const int edge = ST.edge[o];
    if (P.white_flag[k]) k = P.white[0];
    best[v] = k;
    best_score[v] = bs;
    need[v] = edge ? 1 : 0;
    if (next_key) next_key[v] = edge && ST.next ? ST.next[o] : -1;
}

// ---------------------------------------------------------------------------
// Hierarchical descent (glm/arma.py:2224-2322) on demand: the levels of the
// descent that k_select mode 1 runs on the full surface, run on pair sets
// whose scores equal the surface bitwise, so the selected pair and score are
// the same.  The state of each voxel between passes is its current lattice
// point (ba, bb) and score cur.
// ---------------------------------------------------------------------------
// Set of each voxel for a pass >= 1: the one keyed by its current point.
__global__ void k_hier_key(PlanDev P, const int32_t* __restrict__ center_set, const int32_t* __restrict__ hba,
                           const int32_t* __restrict__ hbb, int V, int32_t* __restrict__ key) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    key[v] = center_set[hba[v] * P.n_b + hbb[v]];
}

// One pass: per voxel column, the level(s) of the pass with k_select mode 1's
// visit order (top level: b-outer / a-inner over the multiples of the step;
// below: the clamped 3 x 3 neighbourhood, b-offset outer), first strict
// minimum, a level with a NaN candidate leaving the point unchanged, and the
// move only to a score below the current one.  last: write the voxel's pair
// and score.
__global__ void k_hier_select(PlanDev P, HierPass H, const int32_t* __restrict__ rows,
                              const int32_t* __restrict__ tset, const int32_t* __restrict__ n_tiles,
                              const float* __restrict__ Sb, int32_t* __restrict__ hba, int32_t* __restrict__ hbb,
                              float* __restrict__ hcur, int last, int32_t* __restrict__ best,
                              float* __restrict__ best_score) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= *n_tiles * kSetTile) return;
    const int v = rows[c];
    if (v < 0) return;
    const int set = tset[c / kSetTile];
    const float* sr = Sb + (size_t)c * H.ST.sb;
    int ba, bb;
    float cur;
    auto run_level = [&](int lev, auto score_of) {
        const int step = 1 << lev;
        float mn = 0.f;
        int sa = 0, sb = 0, i = 0;
        bool first = true, nan = false;
        auto visit = [&](int x, int y) {
            const float s = score_of(i++, x, y);
            if (s != s) nan = true;
            if (first || s < mn) {
                mn = s;
                sa = x;
                sb = y;
                first = false;
            }
        };
        if (lev == P.ltop) {
            for (int y = 0; y < P.n_b; y += step)
                for (int x = 0; x < P.n_a; x += step) visit(x, y);
        } else {
            for (int oy = -1; oy <= 1; ++oy) {
                const int y = min(max(bb + oy * step, 0), P.n_b - 1);
                for (int ox = -1; ox <= 1; ++ox) visit(min(max(ba + ox * step, 0), P.n_a - 1), y);
            }
        }
        if (!nan && mn < cur) {
            cur = mn;
            ba = sa;
            bb = sb;
        }
    };
    if (H.first) {
        auto by_point = [&](int, int x, int y) {
            const int sl = H.point_slot[x * P.n_b + y];
            return sl >= 0 ? sr[sl] : INFINITY;
        };
        ba = P.zero_a;
        bb = P.zero_b;
        cur = by_point(0, ba, bb);
        run_level(P.ltop, by_point);
        if (P.ltop >= 1) run_level(P.ltop - 1, by_point);
    } else {
        ba = hba[v];
        bb = hbb[v];
        cur = hcur[v];
        const int32_t* vm = H.vmap + (size_t)set * 9;
        run_level(H.level, [&](int i, int, int) {
            const int sl = vm[i];
            return sl >= 0 ? sr[sl] : INFINITY;
        });
    }
    hba[v] = ba;
    hbb[v] = bb;
    hcur[v] = cur;
    if (last) {
        best[v] = P.lut[ba * P.n_b + bb];
        best_score[v] = cur;
    }
}

}  // namespace remlk
