// Designs wider than 96 columns (96 < p <= kSegMaxP): a pair's p columns span
// nseg = ceil(p / 64) projection tiles of 64 columns ("segments"), tile
// k * nseg + s holding columns 64 s .. 64 s + 63 of pair k (zero past p).
// Included by reml_api.hip after reml_kernels.hip (one translation unit).
#pragma once

#include "reml_kernels.h"

namespace remlk {

constexpr int kSegBN = 64;          // columns per segment
constexpr int kSegMaxP = 768;       // widest design of the segmented path
constexpr int kSegPrepVox = 8;      // voxels per block of k_prep_seg
constexpr int kSegPrepChunk = 64;   // samples staged per step of k_prep_seg
constexpr int kSegGlsWaves = 4;     // voxels (one per wave) per block of k_gls_seg

// LDS bytes of the two segmented kernels that size their shared memory at launch.
inline size_t seg_prep_lds(int p) { return sizeof(double) * ((size_t)kSegPrepVox * (p + 1) + kSegPrepChunk * kSegPrepVox); }
inline size_t seg_gls_lds(int p) { return sizeof(double) * 2 * (size_t)p * kSegGlsWaves; }

// ---------------------------------------------------------------------------
// Standardisation and residualisation (glm/arma.py:3947-3954, 1554-1571),
// p > 96: per voxel c = pinv(X) y in FP64 (t ascending), s = the sample
// standard deviation of y floored at 1e-8, r_t = y_t / s - sum_j X[t][j] c_j / s
// (j ascending), written to the time-major buffer Yt[t][col].
// One block of 256 threads per kSegPrepVox consecutive output columns.
// Pass 1: each thread owns columns j = tid + 256 m of c for all the block's
// voxels; the samples are staged through LDS kSegPrepChunk at a time and the
// rows of pinv(X)' are read coalesced.  c / s goes to LDS.  Pass 2: each
// thread takes (t, column) outputs, consecutive threads consecutive columns.
// bit 0 of normalize: standardise; bit 1: no residualisation.
// ---------------------------------------------------------------------------
template <int MJ>
__global__ void __launch_bounds__(256) k_prep_seg(int T, int p, const float* __restrict__ Y, int64_t ldY, int V,
                                                  const int32_t* __restrict__ rows, int ncols, int ncols_pad,
                                                  const double* __restrict__ pinvT, const double* __restrict__ X,
                                                  int normalize, float* __restrict__ Yt, int64_t ldYt,
                                                  float* __restrict__ scale) {
    constexpr int NV = kSegPrepVox, CH = kSegPrepChunk;
    extern __shared__ double seg_lds[];
    double* cs = seg_lds;                       // [NV][p + 1]  c / s
    double* ys = cs + (size_t)NV * (p + 1);     // [CH][NV]     staged samples
    __shared__ double st_inv[NV];
    __shared__ int st_v[NV];
    const int tid = threadIdx.x;
    const int col0 = blockIdx.x * NV;
    if (col0 >= ncols_pad) return;  // block-uniform
    if (tid < NV) {
        const int col = col0 + tid;
        int v = col < ncols ? (rows ? rows[col] : col) : -1;
        st_v[tid] = (v >= V) ? -1 : v;
    }
    __syncthreads();
    double acc[MJ][NV];
#pragma unroll
    for (int m = 0; m < MJ; ++m)
#pragma unroll
        for (int u = 0; u < NV; ++u) acc[m][u] = 0.0;
    double s1 = 0.0, s2 = 0.0;  // threads tid < NV: sums of voxel tid
    for (int t0 = 0; t0 < T; t0 += CH) {
        const int n = min(CH, T - t0);
        for (int e = tid; e < CH * NV; e += 256) {
            const int u = e / CH, i = e % CH;
            const int v = st_v[u];
            ys[i * NV + u] = (v >= 0 && i < n) ? (double)Y[(size_t)v * ldY + t0 + i] : 0.0;
        }
        __syncthreads();
        if (tid < NV)
            for (int i = 0; i < n; ++i) {
                const double yt = ys[i * NV + tid];
                s1 += yt;
                s2 += yt * yt;
            }
        for (int i = 0; i < n; ++i) {
            const double* pr = pinvT + (size_t)(t0 + i) * p;
            const double* yr = ys + i * NV;
#pragma unroll
            for (int m = 0; m < MJ; ++m) {
                const int j = tid + 256 * m;
                const double pj = j < p ? pr[j] : 0.0;
#pragma unroll
                for (int u = 0; u < NV; ++u) acc[m][u] += pj * yr[u];
            }
        }
        __syncthreads();
    }
    if (tid < NV) {
        double s = 1.0;
        if (normalize & 1) {
            const double mean = s1 / T;
            const double var = T > 1 ? (s2 - T * mean * mean) / (T - 1) : 0.0;
            const double sd = var > 0.0 ? sqrt(var) : 0.0;
            s = (double)fmaxf((float)sd, 1e-8f);
        }
        st_inv[tid] = 1.0 / s;
        const int v = st_v[tid];
        if (v >= 0 && scale) scale[v] = (float)s;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < MJ; ++m) {
        const int j = tid + 256 * m;
        if (j < p)
#pragma unroll
            for (int u = 0; u < NV; ++u) cs[u * (p + 1) + j] = (normalize & 2) ? 0.0 : acc[m][u] * st_inv[u];
    }
    __syncthreads();
    if (!Yt) return;
    for (int e = tid; e < T * NV; e += 256) {
        const int t = e / NV, u = e % NV;
        const int v = st_v[u];
        float rf = 0.f;
        if (v >= 0) {
            double r = (double)Y[(size_t)v * ldY + t] * st_inv[u];
            const double* xr = X + (size_t)t * p;
            const double* cr = cs + u * (p + 1);
            for (int j = 0; j < p; ++j) r -= xr[j] * cr[j];
            rf = (float)r;
        }
        Yt[(size_t)t * ldYt + col0 + u] = rf;
    }
}

// ---------------------------------------------------------------------------
// (ii) Projection for p > 96 on fp32 MFMA (32x32x2), with the REML score
// epilogue (glm/arma.py:1988-2004): item (128-voxel tile, pair k) runs the
// GEMM of each of the pair's segments in turn (tiling and staging of
// k_project<64>), and each row's thread adds the segment's squared columns
// into one FP64 sum, j ascending over the pair's p columns:
//   P = sum_j (G_k' y)_j^2, score = ld_k + (T - p) log(E - P + 1e-10),
// written over E in the surface.  Items follow xcd_tile_map with the pairs
// as the column tiles.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_project_seg(PlanDev P, const float* __restrict__ Yt, int64_t ldY, int V,
                                                     float* __restrict__ S, int64_t ldS,
                                                     const int32_t* __restrict__ n_live) {
    constexpr int BM = kProjBM, BN = kSegBN, BK = kProjBK, NBLK = BN / 32;
    constexpr int A_ELEMS = BK * BM, B_ELEMS = BK * BN;
    constexpr int STAGE = A_ELEMS + B_ELEMS;
    constexpr int C_LD = BN + 1;
    constexpr int LDS_FLOATS = (2 * STAGE > BM * C_LD) ? 2 * STAGE : BM * C_LD;
    __shared__ float lds[LDS_FLOATS];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int w = tid >> 6;
    int vt, k;
    if (!xcd_tile_map(blockIdx.x, gridDim.x, P.K, &vt, &k)) return;
    if (n_live && vt * BM >= *n_live) return;  // past the live voxels
    const int v0 = vt * BM;
    const float* __restrict__ Ag = Yt + v0;
    const int p = P.p, nseg = P.nseg;
    constexpr int A_V4 = A_ELEMS / 4 / 256;
    constexpr int B_V4 = (B_ELEMS / 4 + 255) / 256;
    const int nk = P.Tpad / BK;
    double pp = 0.0;  // threads tid < BM: voxel row tid
    for (int sg = 0; sg < nseg; ++sg) {
        const float* __restrict__ Bg = P.G32 + ((size_t)k * nseg + sg) * P.Tpad * BN;
        f32x16 acc[NBLK];
#pragma unroll
        for (int i = 0; i < NBLK; ++i)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][r] = 0.f;
        f32x4 ra[A_V4], rb[B_V4];
        proj_gload<BM, BN, BK, A_V4, B_V4>(ra, rb, Ag, Bg, ldY, 0, tid);
        proj_sstore<BM, BN, BK, A_V4, B_V4>(ra, rb, lds, tid);
        __syncthreads();
        for (int kt = 0; kt < nk; ++kt) {
            const int cur = kt & 1;
            if (kt + 1 < nk) proj_gload<BM, BN, BK, A_V4, B_V4>(ra, rb, Ag, Bg, ldY, kt + 1, tid);
            const float* As = lds + cur * STAGE;
            const float* Bs = As + A_ELEMS;
            const int ar = lane >> 5, ac = lane & 31;
#pragma unroll
            for (int kk = 0; kk < BK / 2; ++kk) {
                const float av = As[(2 * kk + ar) * BM + w * 32 + ac];
#pragma unroll
                for (int n = 0; n < NBLK; ++n) {
                    const float bv = Bs[(2 * kk + ar) * BN + n * 32 + ac];
                    acc[n] = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv, acc[n], 0, 0, 0);
                }
            }
            if (kt + 1 < nk) proj_sstore<BM, BN, BK, A_V4, B_V4>(ra, rb, lds + (cur ^ 1) * STAGE, tid);
            __syncthreads();
        }
        // accumulators to LDS (32x32 layout: col = lane % 32,
        // row = 8 (r / 4) + 4 (lane / 32) + r % 4), then each row's squares
        float* Cs = lds;
#pragma unroll
        for (int n = 0; n < NBLK; ++n)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = w * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                Cs[row * C_LD + n * 32 + (lane & 31)] = acc[n][r];
            }
        __syncthreads();
        if (tid < BM) {
            const int nj = min(BN, p - sg * BN);
            const float* cr = Cs + tid * C_LD;
            for (int j = 0; j < nj; ++j) {
                const double cv = (double)cr[j];
                pp = fma(cv, cv, pp);
            }
        }
        __syncthreads();
    }
    if (tid < BM) {
        const int v = v0 + tid;
        if (v < V) {
            float* sp = S + (size_t)v * ldS + k;
            const double rss = (double)(*sp) - pp;
            *sp = (float)fma((double)(P.T - p), log(rss + 1e-10), P.ld[k]);
        }
    }
}

// Sum over the 64 lanes of a wave (every lane gets the total).
__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

// Whitened samples yw_t = sqrt(1/v_t) e_t of one 64-sample chunk, lane i
// holding t = t0 + i: e_t = (y_t - c_{t-1} y_{t-1}) + alpha_{t-1} e_{t-1}
// (c = alpha = 0 at run starts: the shifted coefficient rows of coefR) as an
// inclusive scan of the affine maps e -> alpha e + z over the chunk, started
// from the previous chunk's last e and y (ecarry, ycarry; updated).  Lanes
// past T give 0.
__device__ __forceinline__ double seg_whiten_chunk(const double* __restrict__ cR, const float* __restrict__ row,
                                                   double inv, int T, int t0, double& ecarry, double& ycarry,
                                                   double& yt_out) {
    const int lane = threadIdx.x & 63;
    const int t = t0 + lane;
    const bool in = t < T;
    const double yt = in ? (double)row[t] * inv : 0.0;
    double yprev = __shfl_up(yt, 1);
    if (lane == 0) yprev = ycarry;
    const double* c4 = cR + 4 * (size_t)(in ? t : 0);
    double a = in ? c4[2] : 0.0;
    double z = in ? yt - c4[1] * yprev : 0.0;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const double ap = __shfl_up(a, off), zp = __shfl_up(z, off);
        if (lane >= off) {
            z = fma(a, zp, z);
            a = a * ap;
        }
    }
    const double e = fma(a, ecarry, z);
    ecarry = __shfl(e, 63);
    ycarry = __shfl(yt, 63);
    yt_out = yt;
    return in ? e * c4[0] : 0.0;
}

// ---------------------------------------------------------------------------
// (iv) GLS refit for p > 96 (glm/arma.py:4784-5601, 5827-5862): one wave per
// voxel, in the given order (perm, -1 = empty slot) or voxel order.
//   yw = L^-1 y (the innovations recurrence, by chunk scans), q = X_w' yw,
//   beta = (X'R^-1X)^-1 q  (the plan's Ainv; for a rank-deficient design
//   its pseudo-inverse, which gives the minimum-norm solution),
//   r = yw - X_w beta, sigma2 = r'r / max(1, T - p), t_j = beta_j / sqrt(sigma2 Ainv_jj),
//   F = beta' Fm beta / (sigma2 n_task + 1e-10), R^2 = CoD(y, X beta),
// with the outputs of k_gls (betas and sigma2 unscaled; optional series:
// whitened residuals (FP64), residuals and fitted values).  Voxels whose pair
// is not valid are left to k_gls_invalid.  q and beta of each wave live in
// LDS; all waves reach the block barriers.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64 * kSegGlsWaves) k_gls_seg(PlanDev P, const float* __restrict__ Y, int64_t ldY,
                                                               const float* __restrict__ scale,
                                                               const int32_t* __restrict__ perm,
                                                               const int32_t* __restrict__ best, int n_slots, int V,
                                                               float* __restrict__ betas, float* __restrict__ tstats,
                                                               float* __restrict__ fstats, float* __restrict__ sigma2,
                                                               float* __restrict__ r2, double* __restrict__ resid_w,
                                                               int64_t ldR, float* __restrict__ resid,
                                                               float* __restrict__ pred_out) {
    constexpr int MJ = (kSegMaxP + 63) / 64;
    extern __shared__ double seg_lds[];
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int T = P.T, p = P.p;
    double* q = seg_lds + (size_t)w * 2 * p;
    double* be = q + p;
    const int i = blockIdx.x * kSegGlsWaves + w;
    int v = i < n_slots ? (perm ? perm[i] : i) : -1;
    if (v >= V) v = -1;
    int k = v >= 0 ? best[v] : -1;
    const bool active = v >= 0 && k >= 0 && k < P.K;  // wave-uniform
    if (!active) k = 0;
    const double s = (active && scale) ? (double)scale[v] : 1.0;
    const double inv = 1.0 / s;
    const double* __restrict__ Xw = P.Xw64 + (size_t)k * T * p;
    const double* __restrict__ cR = P.coefR + (size_t)k * P.Tpad * 4;
    const float* __restrict__ row = Y + (size_t)(active ? v : 0) * ldY;
    double sy = 0.0;
    if (active) {
        double qa[MJ];
#pragma unroll
        for (int m = 0; m < MJ; ++m) qa[m] = 0.0;
        double ec = 0.0, yc = 0.0;
        for (int t0 = 0; t0 < T; t0 += 64) {
            double yt;
            const double yw = seg_whiten_chunk(cR, row, inv, T, t0, ec, yc, yt);
            sy += yt;
            const int n = min(64, T - t0);
            for (int ii = 0; ii < n; ++ii) {
                const double ywi = __shfl(yw, ii);
                const double* xr = Xw + (size_t)(t0 + ii) * p;
#pragma unroll
                for (int m = 0; m < MJ; ++m) {
                    const int j = lane + 64 * m;
                    if (j < p) qa[m] = fma(xr[j], ywi, qa[m]);
                }
            }
        }
#pragma unroll
        for (int m = 0; m < MJ; ++m) {
            const int j = lane + 64 * m;
            if (j < p) q[j] = qa[m];
        }
    }
    __syncthreads();
    const double* __restrict__ Ai = P.Ainv + (size_t)k * p * p;
    if (active) {
        // beta_i = sum_j Ainv[i][j] q_j, read as row j of the symmetric Ainv (coalesced)
        double ba[MJ];
#pragma unroll
        for (int m = 0; m < MJ; ++m) ba[m] = 0.0;
        for (int j = 0; j < p; ++j) {
            const double qj = q[j];
            const double* ar = Ai + (size_t)j * p;
#pragma unroll
            for (int m = 0; m < MJ; ++m) {
                const int c = lane + 64 * m;
                if (c < p) ba[m] = fma(ar[c], qj, ba[m]);
            }
        }
#pragma unroll
        for (int m = 0; m < MJ; ++m) {
            const int c = lane + 64 * m;
            if (c < p) be[c] = ba[m];
        }
    }
    __syncthreads();
    if (!active) return;  // no barrier below
    sy = wave_sum(sy);
    const double mean = sy / T;
    double bl[MJ];
#pragma unroll
    for (int m = 0; m < MJ; ++m) {
        const int c = lane + 64 * m;
        bl[m] = c < p ? be[c] : 0.0;
    }
    // residual pass: per sample, the two dot products over j (lanes over j,
    // then a wave sum); lane ii of the chunk keeps sample t0 + ii
    double ssw = 0.0, ssres = 0.0, sstot = 0.0;
    double ec = 0.0, yc = 0.0;
    for (int t0 = 0; t0 < T; t0 += 64) {
        double yt;
        const double yw = seg_whiten_chunk(cR, row, inv, T, t0, ec, yc, yt);
        const int n = min(64, T - t0);
        double dw = 0.0, dx = 0.0;
        for (int ii = 0; ii < n; ++ii) {
            const double* xwr = Xw + (size_t)(t0 + ii) * p;
            const double* xr = P.X + (size_t)(t0 + ii) * p;
            double aw = 0.0, ax = 0.0;
#pragma unroll
            for (int m = 0; m < MJ; ++m) {
                const int j = lane + 64 * m;
                if (j < p) {
                    aw = fma(xwr[j], bl[m], aw);
                    ax = fma(xr[j], bl[m], ax);
                }
            }
            aw = wave_sum(aw);
            ax = wave_sum(ax);
            if (lane == ii) {
                dw = aw;
                dx = ax;
            }
        }
        if (lane < n) {
            const int t = t0 + lane;
            const double r = yw - dw;
            const double dr = yt - dx, dm = yt - mean;
            ssw += r * r;
            ssres += dr * dr;
            sstot += dm * dm;
            if (resid_w) resid_w[(size_t)v * ldR + t] = r * s;
            if (resid) resid[(size_t)v * ldR + t] = (float)(dr * s);
            if (pred_out) pred_out[(size_t)v * ldR + t] = (float)(dx * s);
        }
    }
    ssw = wave_sum(ssw);
    ssres = wave_sum(ssres);
    sstot = wave_sum(sstot);
    const int dof = max(1, T - p);
    const double sig = ssw / dof;
    // t statistics and betas; u_c = sum_a beta_a Fm[a][c] for the F form
    const double* __restrict__ F = P.Fm + (size_t)k * p * p;
    double ua[MJ];
#pragma unroll
    for (int m = 0; m < MJ; ++m) ua[m] = 0.0;
    for (int a = 0; a < p; ++a) {
        const double ba = be[a];
        const double* fr = F + (size_t)a * p;
#pragma unroll
        for (int m = 0; m < MJ; ++m) {
            const int c = lane + 64 * m;
            if (c < p) ua[m] = fma(fr[c], ba, ua[m]);
        }
    }
    double quad = 0.0;
#pragma unroll
    for (int m = 0; m < MJ; ++m) {
        const int c = lane + 64 * m;
        if (c < p) {
            quad = fma(ua[m], bl[m], quad);
            const double se = sqrt(sig * Ai[(size_t)c * p + c]);
            tstats[(size_t)v * p + c] = (float)(bl[m] / (se + 1e-10));
            betas[(size_t)v * p + c] = (float)(bl[m] * s);
        }
    }
    quad = wave_sum(quad);
    if (lane == 0) {
        fstats[v] = (float)(quad / (sig * P.n_task + 1e-10));
        sigma2[v] = (float)(sig * s * s);
        r2[v] = (float)(sstot <= 1e-10 ? 0.0 : fmin(1.0 - ssres / sstot, 1.0));
    }
}

}  // namespace remlk
