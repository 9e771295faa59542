// Normalise + residualise on f64 MFMA (p <= 32).  Included by reml_api.hip
// after reml_sort_gls.hip (one translation unit).
#pragma once

#include "reml_kernels.h"

namespace remlk {

// Waves per 64-column block (each takes a contiguous range of time steps).
constexpr int kPrepWaves = 4;
// Steps of operands in flight per wave: as deep as the register file allows
// two waves per SIMD.
constexpr int prep_depth(int NT, int NF) { return NT == 1 && NF == 0 ? 2 : 1; }

__device__ __forceinline__ int prep_voxel(int c, int ncols, const int32_t* __restrict__ rows, int V) {
    const int v = c < ncols ? (rows ? rows[c] : c) : -1;
    return v < V ? v : -1;
}

// ---------------------------------------------------------------------------
// One block of kPrepWaves waves per 64 output columns (voxel of column
// c = rows ? rows[c] : c), following glm/arma.py:1554-1571 (std
// normalisation, projection on the design's pseudo-inverse, residual).  The
// 16-sample time steps are split into kPrepWaves contiguous ranges, one per
// wave, so that each wave's chain of dependent loads is short:
//   pass 1: partial c = pinv(X) y on MFMA (k-slot q of step m = time
//           t0 + 4q + m) and partial sums of y, y^2 over the wave's range,
//           added in wave order through LDS (every wave forms the same sums);
//           s = max(float(sd), 1e-8) (normalize bit 0; bit 1 keeps c = 0,
//           i.e. no residualisation);
//   pass 2: X c on MFMA for the wave's 16-row time tiles, r = y / s - X c / s
//           written to Yt[t][column] (rows T..Tpad-1 come out zero); optional
//           lag-1..NF autocorrelations of r through the wave's LDS tile (a
//           wave first replays the step before its range to seed the lags),
//           added in wave order.
// The A operand of pass 2 permutes the tile rows so that a lane's four output
// rows are four consecutive time points.  Every voxel's arithmetic is the same
// whatever column it lands in.  CI: pass 1 is replaced by coef_in (c of each
// voxel, as coef_out stores it) and scale (read, not written).
// ---------------------------------------------------------------------------
template <int NT, int NF, bool CI>
__global__ void __launch_bounds__(64 * kPrepWaves) k_prep_mfma(PlanDev P, const float* __restrict__ Y, int64_t ldY,
                                                               int V, const int32_t* __restrict__ rows, int ncols,
                                                               int ncols_pad, int normalize, float* __restrict__ Yt,
                                                               int64_t ldYt, float* __restrict__ scale,
                                                               float* __restrict__ feats,
                                                               const int32_t* __restrict__ ncols_live,
                                                               uint16_t* __restrict__ Ysh, uint16_t* __restrict__ Ysl,
                                                               int split_bf16, const double* __restrict__ coef_in,
                                                               double* __restrict__ coef_out,
                                                               const PlanDev* __restrict__ plans,
                                                               const int32_t* __restrict__ grp) {
    if (plans) P = plans[max(grp[blockIdx.x], 0)];  // grouped layout: the design of this 64-column block
    constexpr int PG = 16 * NT;
    constexpr int RB = 65;  // feature tile pitch (floats)
    constexpr int W = kPrepWaves;
    constexpr int RN = 16 * NT + 8;  // partial sums per lane: c (4 x NT x 4), then s1, s2 (4 each)
    static_assert(NF <= 16 && NF + 1 <= RN, "feature lags must fit one step and the reduction buffer");
    __shared__ double red[W][RN][64];
    __shared__ float rbuf[NF > 0 ? W * 16 * RB : 1];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int col0 = blockIdx.x * 64;
    if (col0 >= ncols_pad) return;
    if (ncols_live && col0 >= *ncols_live) return;  // beyond the columns in use (device-side count)
    const int T = P.T, Tpad = P.Tpad;
    const int nst = Tpad / 16;
    const int tb = 16 * (w * nst / W), te = 16 * ((w + 1) * nst / W);  // this wave's time range
    const int j = lane & 15, kq = lane >> 4;
    const bool vec4 = (ldY & 3) == 0 && (reinterpret_cast<uintptr_t>(Y) & 15) == 0;
    int vB[4];
    const float* rowB[4];
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        vB[g] = prep_voxel(col0 + 16 * g + j, ncols, rows, V);
        rowB[g] = Y + (size_t)(vB[g] >= 0 ? vB[g] : 0) * ldY;
    }
    const double* __restrict__ Pp = P.pinvTp;
    const double* __restrict__ Xp = P.Xp;

    // ---- pass 1 --------------------------------------------------------
    f64x4 acc[4][NT];
#pragma unroll
    for (int g = 0; g < 4; ++g)
#pragma unroll
        for (int rt = 0; rt < NT; ++rt) acc[g][rt] = f64x4{0.0, 0.0, 0.0, 0.0};
    double s1[4] = {0.0, 0.0, 0.0, 0.0}, s2[4] = {0.0, 0.0, 0.0, 0.0};
    // operands of one 16-sample step, fetched D steps ahead
    auto load1 = [&](int t0, float (&yb)[4][4], double (&a)[4][NT]) __attribute__((always_inline)) {
        const bool full = vec4 && t0 + 16 <= T;
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            load_span<4>(rowB[g], T, t0 + 4 * kq, full, yb[g]);
#pragma unroll
            for (int m = 0; m < 4; ++m) yb[g][m] = vB[g] >= 0 ? yb[g][m] : 0.f;
        }
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int rt = 0; rt < NT; ++rt) a[m][rt] = Pp[(size_t)(t0 + 4 * kq + m) * PG + 16 * rt + j];
    };
    constexpr int D = prep_depth(NT, NF);
    if (!CI) {
        float yq[D][4][4];
        double aq[D][4][NT];
#pragma unroll
        for (int u = 0; u < D; ++u)
            if (tb + 16 * u < te) load1(tb + 16 * u, yq[u], aq[u]);
        for (int t0b = tb; t0b < te; t0b += 16 * D) {
#pragma unroll
            for (int u = 0; u < D; ++u) {
                const int t0 = t0b + 16 * u;
                if (t0 >= te) break;
                float yb[4][4];
                double a[4][NT];
#pragma unroll
                for (int g = 0; g < 4; ++g)
#pragma unroll
                    for (int m = 0; m < 4; ++m) yb[g][m] = yq[u][g][m];
#pragma unroll
                for (int m = 0; m < 4; ++m)
#pragma unroll
                    for (int rt = 0; rt < NT; ++rt) a[m][rt] = aq[u][m][rt];
                if (t0 + 16 * D < te) load1(t0 + 16 * D, yq[u], aq[u]);
#pragma unroll
                for (int m = 0; m < 4; ++m)
#pragma unroll
                    for (int g = 0; g < 4; ++g) {
                        const double b = (double)yb[g][m];
                        s1[g] += b;
                        s2[g] = fma(b, b, s2[g]);
#pragma unroll
                        for (int rt = 0; rt < NT; ++rt)
                            acc[g][rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[m][rt], b, acc[g][rt], 0, 0, 0);
                    }
            }
        }
    }
    double inv[4];
    // partial sums of the block's waves, added in wave order by every wave
    // (CI: the coefficients and scales of an earlier call are given instead)
    if (!CI) {
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            s1[g] += __shfl_xor(s1[g], 16);
            s1[g] += __shfl_xor(s1[g], 32);
            s2[g] += __shfl_xor(s2[g], 16);
            s2[g] += __shfl_xor(s2[g], 32);
        }
        {
            int q = 0;
#pragma unroll
            for (int g = 0; g < 4; ++g)
#pragma unroll
                for (int rt = 0; rt < NT; ++rt)
#pragma unroll
                    for (int i = 0; i < 4; ++i) red[w][q++][lane] = acc[g][rt][i];
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                red[w][q++][lane] = s1[g];
                red[w][q++][lane] = s2[g];
            }
        }
        __syncthreads();
        {
            auto sum = [&](int q) __attribute__((always_inline)) {
                double x = red[0][q][lane];
#pragma unroll
                for (int ww = 1; ww < W; ++ww) x += red[ww][q][lane];
                return x;
            };
            int q = 0;
#pragma unroll
            for (int g = 0; g < 4; ++g)
#pragma unroll
                for (int rt = 0; rt < NT; ++rt)
#pragma unroll
                    for (int i = 0; i < 4; ++i) acc[g][rt][i] = sum(q++);
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                s1[g] = sum(q++);
                s2[g] = sum(q++);
            }
        }
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            double s = 1.0;
            if (normalize & 1) {
                const double mean = s1[g] / T;
                const double var = T > 1 ? (s2[g] - T * mean * mean) / (T - 1) : 0.0;
                const double sd = var > 0.0 ? sqrt(var) : 0.0;
                s = (double)fmaxf((float)sd, 1e-8f);
            }
            inv[g] = 1.0 / s;
            if (w == 0 && kq == 0 && vB[g] >= 0 && scale) scale[vB[g]] = (float)s;
        }
        if (coef_out) {
            // coefficient 16 rt + 4 kq + i of the voxel in column 16 g + j; wave g % W writes group g
#pragma unroll
            for (int g = 0; g < 4; ++g)
                if (g % W == w && vB[g] >= 0)
#pragma unroll
                    for (int rt = 0; rt < NT; ++rt)
#pragma unroll
                        for (int i = 0; i < 4; ++i) coef_out[(size_t)vB[g] * PG + 16 * rt + 4 * kq + i] = acc[g][rt][i];
        }
    } else {
        // c and s of an earlier call with the same data (bitwise the values it
        // computed: s is a float widened to double)
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const bool ok = vB[g] >= 0;
#pragma unroll
            for (int rt = 0; rt < NT; ++rt)
#pragma unroll
                for (int i = 0; i < 4; ++i) acc[g][rt][i] = ok ? coef_in[(size_t)vB[g] * PG + 16 * rt + 4 * kq + i] : 0.0;
            inv[g] = ok ? 1.0 / (double)scale[vB[g]] : 1.0;
        }
    }
    if (!Yt && NF == 0) return;
    if (normalize & 2) {  // keep the design component (no residualisation)
#pragma unroll
        for (int g = 0; g < 4; ++g)
#pragma unroll
            for (int rt = 0; rt < NT; ++rt) acc[g][rt] = f64x4{0.0, 0.0, 0.0, 0.0};
    }

    // ---- pass 2 --------------------------------------------------------
    // A row i of a tile holds time tt + 4 (i % 4) + i / 4, so output element r
    // of lane (j, kq) is time tt + 4 kq + r.
    const int arow = 4 * (j & 3) + (j >> 2);
    float hist[NF > 0 ? NF : 1];
    double acf[NF > 0 ? NF : 1];
    double acf0 = 0.0;
#pragma unroll
    for (int k = 0; k < (NF > 0 ? NF : 1); ++k) {
        hist[k] = 0.f;
        acf[k] = 0.0;
    }
    float* rb = rbuf + (NF > 0 ? w * 16 * RB : 0);
    auto load2 = [&](int tt, double (&xa)[NT][4], float (&y4)[4][4]) __attribute__((always_inline)) {
#pragma unroll
        for (int rt = 0; rt < NT; ++rt)
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) xa[rt][kk] = Xp[(size_t)(tt + arow) * PG + 16 * rt + kq + 4 * kk];
        const bool full = vec4 && tt + 16 <= T;
#pragma unroll
        for (int g = 0; g < 4; ++g) load_span<4>(rowB[g], T, tt + 4 * kq, full, y4[g]);
    };
    // residuals of one step; emit = false only updates the lag history
    auto step2 = [&](int tt, const double (&xa)[NT][4], const float (&y4c)[4][4], bool emit)
                     __attribute__((always_inline)) {
        f64x4 d[4];
#pragma unroll
        for (int g = 0; g < 4; ++g) d[g] = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int rt = 0; rt < NT; ++rt)
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const double a = xa[rt][kk];
#pragma unroll
                for (int g = 0; g < 4; ++g) d[g] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, acc[g][rt][kk], d[g], 0, 0, 0);
            }
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            float r4[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const double y = vB[g] >= 0 ? (double)y4c[g][r] : 0.0;
                const float rr = (float)((y - d[g][r]) * inv[g]);
                r4[r] = rr;
                const int t = tt + 4 * kq + r;
                if (emit && Yt) Yt[(size_t)t * ldYt + col0 + 16 * g + j] = rr;
                if (NF > 0) rb[(4 * kq + r) * RB + 16 * g + j] = rr;
            }
            if (emit && Ysh) {
                // split halves, [t/8][column][8] (this lane: one 8-byte run of 4 samples)
                uint16_t h[4], l[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    if (split_bf16)
                        split1<true>(r4[r], h[r], l[r]);
                    else
                        split1<false>(r4[r], h[r], l[r]);
                }
                const int t4 = tt + 4 * kq;
                const size_t o = ((size_t)(t4 >> 3) * ldYt + col0 + 16 * g + j) * 8 + (t4 & 7);
                *reinterpret_cast<uint2*>(Ysh + o) = make_uint2(h[0] | (uint32_t)h[1] << 16, h[2] | (uint32_t)h[3] << 16);
                *reinterpret_cast<uint2*>(Ysl + o) = make_uint2(l[0] | (uint32_t)l[1] << 16, l[2] | (uint32_t)l[3] << 16);
            }
        }
        if (NF > 0) {
            // the tile is the wave's own: order its LDS writes and reads within the wave
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // fp32 sums over the step's 16 samples, added to the fp64 totals once per step
            float p0 = 0.f, pk[NF > 0 ? NF : 1];
#pragma unroll
            for (int k = 0; k < NF; ++k) pk[k] = 0.f;
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const float x = rb[i * RB + lane];  // zero past T
                p0 = fmaf(x, x, p0);
#pragma unroll
                for (int k = 0; k < NF; ++k) pk[k] = fmaf(x, hist[k], pk[k]);
#pragma unroll
                for (int k = NF - 1; k > 0; --k) hist[k] = hist[k - 1];
                hist[0] = x;
            }
            if (emit) {
                acf0 += (double)p0;
#pragma unroll
                for (int k = 0; k < NF; ++k) acf[k] += (double)pk[k];
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    };
    if (NF > 0 && tb > 0 && tb < te) {  // seed the lags with the step before the range
        double xa[NT][4];
        float y4[4][4];
        load2(tb - 16, xa, y4);
        step2(tb - 16, xa, y4, false);
    }
    {
        double xq[D][NT][4];
        float y4q[D][4][4];
#pragma unroll
        for (int u = 0; u < D; ++u)
            if (tb + 16 * u < te) load2(tb + 16 * u, xq[u], y4q[u]);
        for (int ttb = tb; ttb < te; ttb += 16 * D) {
#pragma unroll
            for (int u = 0; u < D; ++u) {
                const int tt = ttb + 16 * u;
                if (tt >= te) break;
                double xa[NT][4];
                float y4c[4][4];
#pragma unroll
                for (int rt = 0; rt < NT; ++rt)
#pragma unroll
                    for (int kk = 0; kk < 4; ++kk) xa[rt][kk] = xq[u][rt][kk];
#pragma unroll
                for (int g = 0; g < 4; ++g)
#pragma unroll
                    for (int r = 0; r < 4; ++r) y4c[g][r] = y4q[u][g][r];
                if (tt + 16 * D < te) load2(tt + 16 * D, xq[u], y4q[u]);
                step2(tt, xa, y4c, true);
            }
        }
    }
    if (NF > 0) {
        __syncthreads();  // every wave has read the pass-1 partials
        red[w][0][lane] = acf0;
#pragma unroll
        for (int k = 0; k < NF; ++k) red[w][1 + k][lane] = acf[k];
        __syncthreads();
        if (w == 0 && feats) {
            const int v = prep_voxel(col0 + lane, ncols, rows, V);
            if (v >= 0) {
                double a0 = red[0][0][lane];
#pragma unroll
                for (int ww = 1; ww < W; ++ww) a0 += red[ww][0][lane];
#pragma unroll
                for (int k = 0; k < NF; ++k) {
                    double ak = red[0][1 + k][lane];
#pragma unroll
                    for (int ww = 1; ww < W; ++ww) ak += red[ww][1 + k][lane];
                    feats[(size_t)v * NF + k] = a0 > 0.0 ? (float)(ak / a0) : 0.f;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Normalise + residualise on f64 MFMA for 32 < p <= 96 (PGA = 16 NT columns,
// NT = 3..6): k_prep_mfma's two passes organised for many design columns.
// Pass 1: wave w owns the column tiles w, w + 4 of c = pinv(X) y over the
// whole time range (no cross-wave sum), with the block's 64 columns staged
// through LDS 16 samples at a time (double-buffered, the next step's samples
// in flight during the MFMAs); wave 0 also forms the sums of y and y^2.  c and
// 1/s go to LDS.  Pass 2: each wave computes r = y / s - X c / s for a quarter
// of the time steps, c read from LDS as the B operand, and writes Yt[t][column]
// (rows T..Tpad-1 come out zero) and, when asked, the split halves as
// k_prep_mfma does.  No features, no stored coefficients.
// ---------------------------------------------------------------------------
template <int NT>
__global__ void __launch_bounds__(64 * kPrepWaves) k_prep_mfma_w(PlanDev P, const float* __restrict__ Y, int64_t ldY,
                                                                 int V, const int32_t* __restrict__ rows, int ncols,
                                                                 int ncols_pad, int normalize, float* __restrict__ Yt,
                                                                 int64_t ldYt, float* __restrict__ scale,
                                                                 const int32_t* __restrict__ ncols_live,
                                                                 uint16_t* __restrict__ Ysh, uint16_t* __restrict__ Ysl,
                                                                 int split_bf16, const PlanDev* __restrict__ plans,
                                                                 const int32_t* __restrict__ grp) {
    if (plans) P = plans[max(grp[blockIdx.x], 0)];  // grouped layout: the design of this 64-column block
    constexpr int PG = 16 * NT;
    constexpr int W = kPrepWaves;
    constexpr int RW = (NT + W - 1) / W;  // column tiles per wave
    constexpr int CP = PG + 1;            // LDS row pitch of c (doubles)
    __shared__ float ys[2][16][64];       // one step's samples, [t][column]
    __shared__ double cL[64 * CP];        // c of each column
    __shared__ double invL[64];
    const int tid = threadIdx.x, lane = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int col0 = blockIdx.x * 64;
    if (col0 >= ncols_pad) return;
    if (ncols_live && col0 >= *ncols_live) return;  // beyond the columns in use (device-side count)
    const int T = P.T, Tpad = P.Tpad;
    const int nst = Tpad / 16;
    const int j = lane & 15, kq = lane >> 4;
    const bool vec4 = (ldY & 3) == 0 && (reinterpret_cast<uintptr_t>(Y) & 15) == 0;
    const double* __restrict__ Pp = P.pinvTp;
    const double* __restrict__ Xp = P.Xp;

    // staging: thread tid fetches samples t0 + 4 (tid & 3) .. + 4 of column tid / 4
    const int sv = prep_voxel(col0 + (tid >> 2), ncols, rows, V);
    const float* srow = Y + (size_t)(sv >= 0 ? sv : 0) * ldY;
    auto fetch = [&](int t0, float (&q)[4]) __attribute__((always_inline)) {
        load_span<4>(srow, T, t0 + 4 * (tid & 3), vec4 && t0 + 16 <= T, q);
#pragma unroll
        for (int m = 0; m < 4; ++m) q[m] = sv >= 0 ? q[m] : 0.f;
    };
    auto store = [&](int buf, const float (&q)[4]) __attribute__((always_inline)) {
#pragma unroll
        for (int m = 0; m < 4; ++m) ys[buf][4 * (tid & 3) + m][tid >> 2] = q[m];
    };

    // ---- pass 1 --------------------------------------------------------
    f64x4 acc[4][RW];
#pragma unroll
    for (int g = 0; g < 4; ++g)
#pragma unroll
        for (int u = 0; u < RW; ++u) acc[g][u] = f64x4{0.0, 0.0, 0.0, 0.0};
    double s1[4] = {0.0, 0.0, 0.0, 0.0}, s2[4] = {0.0, 0.0, 0.0, 0.0};
    const bool resid = !(normalize & 2);
    float q[4];
    fetch(0, q);
    store(0, q);
    __syncthreads();
#pragma unroll 1
    for (int st = 0; st < nst; ++st) {
        const int buf = st & 1, t0 = 16 * st;
        if (st + 1 < nst) fetch(t0 + 16, q);
        // MFMA k-slot kq of step m covers time t0 + 4 kq + m; a wave past the
        // last tile repeats it (its sums are not stored), so no MFMA is branched
        double a[4][RW];
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int u = 0; u < RW; ++u) {
                const int rt = min(w + W * u, NT - 1);
                const double x = Pp[(size_t)(t0 + 4 * kq + m) * PG + 16 * rt + j];
                a[m][u] = resid ? x : 0.0;
            }
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const double b = (double)ys[buf][4 * kq + m][16 * g + j];
                s1[g] += b;
                s2[g] = fma(b, b, s2[g]);
#pragma unroll
                for (int u = 0; u < RW; ++u)
                    acc[g][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[m][u], b, acc[g][u], 0, 0, 0);
            }
        if (st + 1 < nst) store(buf ^ 1, q);
        __syncthreads();
    }
    int vB[4];
    const float* rowB[4];
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        vB[g] = prep_voxel(col0 + 16 * g + j, ncols, rows, V);
        rowB[g] = Y + (size_t)(vB[g] >= 0 ? vB[g] : 0) * ldY;
    }
    if (w == 0) {
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            s1[g] += __shfl_xor(s1[g], 16);
            s1[g] += __shfl_xor(s1[g], 32);
            s2[g] += __shfl_xor(s2[g], 16);
            s2[g] += __shfl_xor(s2[g], 32);
            double s = 1.0;
            if (normalize & 1) {
                const double mean = s1[g] / T;
                const double var = T > 1 ? (s2[g] - T * mean * mean) / (T - 1) : 0.0;
                const double sd = var > 0.0 ? sqrt(var) : 0.0;
                s = (double)fmaxf((float)sd, 1e-8f);
            }
            if (kq == 0) {
                invL[16 * g + j] = 1.0 / s;
                if (vB[g] >= 0 && scale) scale[vB[g]] = (float)s;
            }
        }
    }
    // c of column 16 g + j, design column 16 rt + kq + 4 r (the f64 MFMA D layout)
#pragma unroll
    for (int g = 0; g < 4; ++g)
#pragma unroll
        for (int u = 0; u < RW; ++u) {
            const int rt = w + W * u;
            if (rt < NT)
#pragma unroll
                for (int r = 0; r < 4; ++r) cL[(16 * g + j) * CP + 16 * rt + kq + 4 * r] = acc[g][u][r];
        }
    __syncthreads();
    if (!Yt) return;

    // ---- pass 2 --------------------------------------------------------
    // A row i of a tile holds time tt + 4 (i % 4) + i / 4, so output element r
    // of lane (j, kq) is time tt + 4 kq + r.
    const int arow = 4 * (j & 3) + (j >> 2);
    double inv[4];
#pragma unroll
    for (int g = 0; g < 4; ++g) inv[g] = invL[16 * g + j];
    const int tb = 16 * (w * nst / W), te = 16 * ((w + 1) * nst / W);  // this wave's time range
#pragma unroll 1
    for (int tt = tb; tt < te; tt += 16) {
        float y4[4][4];
        const bool full = vec4 && tt + 16 <= T;
#pragma unroll
        for (int g = 0; g < 4; ++g) load_span<4>(rowB[g], T, tt + 4 * kq, full, y4[g]);
        f64x4 d[4];
#pragma unroll
        for (int g = 0; g < 4; ++g) d[g] = f64x4{0.0, 0.0, 0.0, 0.0};
        if (resid) {
#pragma unroll 1
            for (int rt = 0; rt < NT; ++rt)
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    const double a = Xp[(size_t)(tt + arow) * PG + 16 * rt + kq + 4 * kk];
#pragma unroll
                    for (int g = 0; g < 4; ++g)
                        d[g] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, cL[(16 * g + j) * CP + 16 * rt + kq + 4 * kk],
                                                                    d[g], 0, 0, 0);
                }
        }
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            float r4[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const double y = vB[g] >= 0 ? (double)y4[g][r] : 0.0;
                r4[r] = (float)((y - d[g][r]) * inv[g]);
                Yt[(size_t)(tt + 4 * kq + r) * ldYt + col0 + 16 * g + j] = r4[r];
            }
            if (Ysh) {
                // split halves, [t/8][column][8] (this lane: one 8-byte run of 4 samples)
                uint16_t h[4], l[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    if (split_bf16)
                        split1<true>(r4[r], h[r], l[r]);
                    else
                        split1<false>(r4[r], h[r], l[r]);
                }
                const int t4 = tt + 4 * kq;
                const size_t o = ((size_t)(t4 >> 3) * ldYt + col0 + 16 * g + j) * 8 + (t4 & 7);
                *reinterpret_cast<uint2*>(Ysh + o) = make_uint2(h[0] | (uint32_t)h[1] << 16, h[2] | (uint32_t)h[3] << 16);
                *reinterpret_cast<uint2*>(Ysl + o) = make_uint2(l[0] | (uint32_t)l[1] << 16, l[2] | (uint32_t)l[3] << 16);
            }
        }
    }
}

}  // namespace remlk
