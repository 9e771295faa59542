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
constexpr int kSegGlsWaves = 4;     // waves per block of k_gls_seg
constexpr int kSegGlsVox = 4;       // voxels of one pair per wave of k_gls_seg (pair-grouped order)

// LDS bytes of k_prep_seg (sized at launch).
inline size_t seg_prep_lds(int p) { return sizeof(double) * ((size_t)kSegPrepVox * (p + 1) + kSegPrepChunk * kSegPrepVox); }

// ---------------------------------------------------------------------------
// Standardisation and residualisation (glm/arma.py:3947-3954, 1554-1571),
// p > 96: per voxel c = pinv(X) y in FP64 (t ascending), s = the sample
// standard deviation of y floored at 1e-8, r_t = y_t / s - sum_j X[t][j] c_j / s
// (j ascending), written to the time-major buffer Yt[t][col] (rows T..Tpad-1
// zero); in the split projection modes also its hi / lo 16-bit halves (Ysh,
// Ysl: [t/8][col][8], the layout the split projection reads).
// One block of 256 threads per kSegPrepVox consecutive output columns.
// Pass 1: each thread owns columns j = tid + 256 m of c for all the block's
// voxels; the samples are staged through LDS kSegPrepChunk at a time and the
// rows of pinv(X)' are read coalesced.  c / s goes to LDS.  Pass 2: each
// thread takes (t, column) outputs, consecutive threads consecutive columns.
// bit 0 of normalize: standardise; bit 1: no residualisation.
// ---------------------------------------------------------------------------
template <int MJ>
__global__ void __launch_bounds__(256) k_prep_seg(int T, int Tpad, int p, const float* __restrict__ Y, int64_t ldY,
                                                  int V, const int32_t* __restrict__ rows, int ncols, int ncols_pad,
                                                  const double* __restrict__ pinvT, const double* __restrict__ X,
                                                  int normalize, float* __restrict__ Yt, int64_t ldYt,
                                                  float* __restrict__ scale, uint16_t* __restrict__ Ysh,
                                                  uint16_t* __restrict__ Ysl, int split_bf16) {
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
    for (int e = tid; e < Tpad * NV; e += 256) {
        const int t = e / NV, u = e % NV;
        const int v = st_v[u];
        float rf = 0.f;
        if (v >= 0 && t < T) {
            double r = (double)Y[(size_t)v * ldY + t] * st_inv[u];
            const double* xr = X + (size_t)t * p;
            const double* cr = cs + u * (p + 1);
            for (int j = 0; j < p; ++j) r -= xr[j] * cr[j];
            rf = (float)r;
        }
        const int col = col0 + u;
        Yt[(size_t)t * ldYt + col] = rf;
        if (Ysh) {
            uint16_t h, l;
            if (split_bf16)
                split1<true>(rf, h, l);
            else
                split1<false>(rf, h, l);
            const size_t o = ((size_t)(t >> 3) * ldYt + col) * 8 + (t & 7);
            Ysh[o] = h;
            Ysl[o] = l;
        }
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

// ---------------------------------------------------------------------------
// (ii') Split 16-bit projection for p > 96: C = Ah Bh + Ah Bl + Al Bh on
// 32x32x16 MFMA (f16 or bf16 halves, fp32 accumulation, the product order and
// operand layouts of k_project16b), item (256-voxel tile, pair k) over the
// pair's segments in turn.  Each wave owns 64 voxel rows (two 32-row blocks);
// after a segment, lanes 0..31 add the squares of their row's columns into
// FP64 sums (one per row block, j ascending over the pair's p columns), then
// P = sum * 2^-2e for fp16 and the score epilogue of k_project_seg.
// ---------------------------------------------------------------------------
template <bool BF16>
__global__ void __launch_bounds__(256) k_project16_seg(PlanDev P, const float* __restrict__ Yt, int64_t ldY, int V,
                                                       float* __restrict__ S, int64_t ldS,
                                                       const int32_t* __restrict__ n_live) {
    constexpr int BN = kSegBN;
    using L = Proj16Lds<BN>;
    constexpr int BM = kProj16BM, NBLK = BN / 32, NTB = L::NTB, RP = L::RP, C_LD = L::C_LD;
    constexpr int B_N = 2 * NTB * BN;  // 16-byte B units per stage
    constexpr int B_U = B_N / 256;
    static_assert(B_N % 256 == 0, "B stage must split evenly over the block");
    __shared__ __attribute__((aligned(16))) unsigned char lds_raw[L::BYTES];
    uint16_t* Bh = reinterpret_cast<uint16_t*>(lds_raw);
    uint16_t* Bl = Bh + L::B_H;

    int vt, k;
    if (!xcd_tile_map(blockIdx.x, gridDim.x, P.K, &vt, &k)) return;
    if (n_live && vt * BM >= *n_live) return;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int v0 = vt * BM;
    const int p = P.p, nseg = P.nseg;
    const uint16_t* __restrict__ gh = BF16 ? P.Gh_bf16 : P.Gh_f16;
    const uint16_t* __restrict__ gl = BF16 ? P.Gl_bf16 : P.Gl_f16;
    const uint16_t* __restrict__ Ash = reinterpret_cast<const uint16_t*>(Yt + (size_t)P.Tpad * ldY);
    const uint16_t* __restrict__ Asl = Ash + (size_t)P.Tpad * ldY;
    const int a_col = v0 + 64 * w + (lane & 31);
    const int a_tb = lane >> 5;
    const int nk = P.Tpad / kProjBK;
    double pp[2] = {0.0, 0.0};  // lanes 0..31: row (block r, lane)
    for (int sg = 0; sg < nseg; ++sg) {
        const size_t tile = (size_t)k * nseg + sg;
        size_t ub[B_U];
#pragma unroll
        for (int u = 0; u < B_U; ++u) {
            const int rr = (tid + u * 256) % (NTB * BN);
            ub[u] = tile * P.Tpad * BN + ((size_t)(rr / BN) * BN + rr % BN) * 8;
        }
        u32x4 ah[2][2], al[2][2];
        u32x4 rb[B_U];
        auto load_a = [&](int kt) __attribute__((always_inline)) {
#pragma unroll
            for (int r = 0; r < 2; ++r)
#pragma unroll
                for (int ks = 0; ks < 2; ++ks) {
                    const size_t o = ((size_t)(kt * NTB + 2 * ks + a_tb) * ldY + a_col + 32 * r) * 8;
                    ah[r][ks] = *reinterpret_cast<const u32x4*>(Ash + o);
                    al[r][ks] = *reinterpret_cast<const u32x4*>(Asl + o);
                }
        };
        auto load_b = [&](int kt) __attribute__((always_inline)) {
#pragma unroll
            for (int u = 0; u < B_U; ++u) {
                const int plane = (tid + u * 256) / (NTB * BN);
                rb[u] = *reinterpret_cast<const u32x4*>((plane ? gl : gh) + ub[u] + (size_t)kt * NTB * BN * 8);
            }
        };
        f32x16 acc[2][NBLK];
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int i = 0; i < NBLK; ++i)
#pragma unroll
                for (int q = 0; q < 16; ++q) acc[r][i][q] = 0.f;
        load_a(0);
        load_b(0);
        for (int kt = 0; kt < nk; ++kt) {
#pragma unroll
            for (int u = 0; u < B_U; ++u) {
                const int q = tid + u * 256;
                const int plane = q / (NTB * BN), rr = q % (NTB * BN);
                const int tb = rr / BN, col = rr % BN;
                *reinterpret_cast<u32x4*>((plane ? Bl : Bh) + col * RP + tb * 8) = rb[u];
            }
            u32x4 ahi[2][2], alo[2][2];
#pragma unroll
            for (int r = 0; r < 2; ++r)
#pragma unroll
                for (int ks = 0; ks < 2; ++ks) {
                    ahi[r][ks] = ah[r][ks];
                    alo[r][ks] = al[r][ks];
                }
            __syncthreads();
            if (kt + 1 < nk) {
                load_a(kt + 1);
                load_b(kt + 1);
            }
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) {
                const int tb = 2 * ks + (lane >> 5);
                u32x4 bh[NBLK], bl[NBLK];
#pragma unroll
                for (int n = 0; n < NBLK; ++n) {
                    const int col = n * 32 + (lane & 31);
                    bh[n] = *reinterpret_cast<const u32x4*>(Bh + col * RP + tb * 8);
                    bl[n] = *reinterpret_cast<const u32x4*>(Bl + col * RP + tb * 8);
                }
#pragma unroll
                for (int n = 0; n < NBLK; ++n)
#pragma unroll
                    for (int r = 0; r < 2; ++r) acc[r][n] = mfma16u<BF16>(alo[r][ks], bh[n], acc[r][n]);
#pragma unroll
                for (int n = 0; n < NBLK; ++n)
#pragma unroll
                    for (int r = 0; r < 2; ++r) acc[r][n] = mfma16u<BF16>(ahi[r][ks], bl[n], acc[r][n]);
#pragma unroll
                for (int n = 0; n < NBLK; ++n)
#pragma unroll
                    for (int r = 0; r < 2; ++r) acc[r][n] = mfma16u<BF16>(ahi[r][ks], bh[n], acc[r][n]);
            }
            __syncthreads();
        }
        // squares of this segment's columns through a wave-private patch
        float* Cw = reinterpret_cast<float*>(lds_raw) + w * 32 * C_LD;
        const int nj = min(BN, p - sg * BN);
#pragma unroll
        for (int r = 0; r < 2; ++r) {
#pragma unroll
            for (int n = 0; n < NBLK; ++n)
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    const int row = 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
                    Cw[row * C_LD + n * 32 + (lane & 31)] = acc[r][n][q];
                }
            __syncthreads();
            if (lane < 32) {
                const float* cr = Cw + lane * C_LD;
                for (int j = 0; j < nj; ++j) {
                    const double cv = (double)cr[j];
                    pp[r] = fma(cv, cv, pp[r]);
                }
            }
            __syncthreads();
        }
    }
    if (lane < 32) {
        const double tp = (double)(P.T - p);
        const double un = BF16 ? 1.0 : (double)P.g16_unscale[k];
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int v = v0 + 64 * w + 32 * r + lane;
            if (v >= V) continue;
            float* sp = S + (size_t)v * ldS + k;
            const double rss = (double)(*sp) - pp[r] * un;
            *sp = (float)fma(tp, log(rss + 1e-10), P.ld[k]);
        }
    }
}

// Columns of a prepared time-major residual buffer gathered into another
// column order (dst column c = src column rows[c], zero where rows[c] < 0):
// the Tpad fp32 rows, and with halves the split hi / lo planes that follow
// them ([t/8][col][8] 16-bit units, 8 samples of a column per 16 bytes).  The
// residual of a voxel does not depend on its column, so the gathered buffer
// equals a prep of the voxels in the new order.  Threads: consecutive dst
// columns, then 8-sample blocks.
__global__ void k_gather_cols(const float* __restrict__ src, int64_t lds, float* __restrict__ dst, int64_t ldd,
                              const int32_t* __restrict__ rows, int ncols, int Tpad, int halves) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = (int)(i % ncols);
    const int tb = (int)(i / ncols);
    if (tb >= Tpad / 8) return;
    const int v = rows[c];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int t = 8 * tb + e;
        dst[(size_t)t * ldd + c] = v >= 0 ? src[(size_t)t * lds + v] : 0.f;
    }
    if (!halves) return;
    const uint16_t* sh = reinterpret_cast<const uint16_t*>(src + (size_t)Tpad * lds);
    uint16_t* dh = reinterpret_cast<uint16_t*>(dst + (size_t)Tpad * ldd);
#pragma unroll
    for (int plane = 0; plane < 2; ++plane) {
        const uint16_t* sp = sh + (size_t)plane * Tpad * lds;
        uint16_t* dp = dh + (size_t)plane * Tpad * ldd;
        const u32x4 x = v >= 0 ? *reinterpret_cast<const u32x4*>(sp + ((size_t)tb * lds + v) * 8) : u32x4{0u, 0u, 0u, 0u};
        *reinterpret_cast<u32x4*>(dp + ((size_t)tb * ldd + c) * 8) = x;
    }
}

// Sum over the 64 lanes of a wave (every lane gets the total).
__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

// Whitened samples yw_t = sqrt(1/v_t) e_t of one 64-sample chunk of a series
// y (lane i holds y_t, t = t0 + i; 0 past T): e_t = (y_t - c_{t-1} y_{t-1}) + alpha_{t-1} e_{t-1}
// (c = alpha = 0 at run starts: the shifted coefficient rows of coefR) as an
// inclusive scan of the affine maps e -> alpha e + z over the chunk, started
// from the previous chunk's last e and y (ecarry, ycarry; updated).  Lanes
// past T give 0.
__device__ __forceinline__ double seg_whiten_vals(const double* __restrict__ cR, double yt, int T, int t0,
                                                  double& ecarry, double& ycarry) {
    const int lane = threadIdx.x & 63;
    const int t = t0 + lane;
    const bool in = t < T;
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
    return in ? e * c4[0] : 0.0;
}

// The same for the samples of a series row (float, scaled by inv); yt_out =
// the lane's scaled sample (0 past T).
__device__ __forceinline__ double seg_whiten_chunk(const double* __restrict__ cR, const float* __restrict__ row,
                                                   double inv, int T, int t0, double& ecarry, double& ycarry,
                                                   double& yt_out) {
    const int t = t0 + (threadIdx.x & 63);
    const double yt = t < T ? (double)row[t] * inv : 0.0;
    yt_out = yt;
    return seg_whiten_vals(cR, yt, T, t0, ecarry, ycarry);
}

// ---------------------------------------------------------------------------
// (iv) GLS refit for p > 96 (glm/arma.py:4784-5601, 5827-5862): one wave per
// NV voxels of one pair (the pair-grouped order padded to NV slots per group,
// -1 = empty slot; NV = 1 for the voxel order), so each row of the pair's
// tables is read once for the NV voxels.  Per voxel:
//   yw = L^-1 y (the innovations recurrence, by chunk scans), q = X_w' yw,
//   beta = (X'R^-1X)^-1 q  (the plan's Ainv; for a rank-deficient design
//   its pseudo-inverse, which gives the minimum-norm solution),
//   r = L^-1 (y - X beta) (= yw - X_w beta), sigma2 = r'r / max(1, T - p), t_j = beta_j / sqrt(sigma2 Ainv_jj),
//   F = beta' Fm beta / (sigma2 n_task + 1e-10), R^2 = CoD(y, X beta),
// with the outputs of k_gls (betas and sigma2 unscaled; optional series:
// whitened residuals (FP64), residuals and fitted values).  Lane l holds
// columns l + 64 m (m < MJ) of q, beta and the F form; a column's value
// reaches every lane by a shuffle, so the kernel uses no LDS and no block
// barrier.  Voxels whose pair is not valid are left to k_gls_invalid.
// ---------------------------------------------------------------------------
template <int MJ, int NV>
__global__ void __launch_bounds__(64 * kSegGlsWaves) k_gls_seg(PlanDev P, const float* __restrict__ Y, int64_t ldY,
                                                               const float* __restrict__ scale,
                                                               const int32_t* __restrict__ perm,
                                                               const int32_t* __restrict__ best, int n_slots, int V,
                                                               float* __restrict__ betas, float* __restrict__ tstats,
                                                               float* __restrict__ fstats, float* __restrict__ sigma2,
                                                               float* __restrict__ r2, double* __restrict__ resid_w,
                                                               int64_t ldR, float* __restrict__ resid,
                                                               float* __restrict__ pred_out) {
    const int lane = threadIdx.x & 63;
    const int wv = blockIdx.x * kSegGlsWaves + (threadIdx.x >> 6);
    const int T = P.T, p = P.p;
    int vox[NV];
    bool act[NV];
    int k = -1;
#pragma unroll
    for (int u = 0; u < NV; ++u) {
        const int i = wv * NV + u;
        int v = i < n_slots ? (perm ? perm[i] : i) : -1;
        if (v >= V) v = -1;
        const int kv = v >= 0 ? best[v] : -1;
        act[u] = v >= 0 && kv >= 0 && kv < P.K;
        vox[u] = act[u] ? v : 0;
        if (act[u] && k < 0) k = kv;  // the slots of a wave share one pair
    }
    if (k < 0) return;  // wave-uniform: no active slot
    double sc[NV], inv[NV];
    const float* rows[NV];
#pragma unroll
    for (int u = 0; u < NV; ++u) {
        sc[u] = (act[u] && scale) ? (double)scale[vox[u]] : 1.0;
        inv[u] = 1.0 / sc[u];
        rows[u] = Y + (size_t)vox[u] * ldY;
    }
    const double* __restrict__ Xw = P.Xw64 + (size_t)k * T * p;
    const double* __restrict__ cR = P.coefR + (size_t)k * P.Tpad * 4;
    const double* __restrict__ Ai = P.Ainv + (size_t)k * p * p;
    const double* __restrict__ F = P.Fm + (size_t)k * p * p;

    // pass 1: q = X_w' yw (t ascending) and the sums of y for the means
    double acc[MJ][NV], sy[NV];
#pragma unroll
    for (int m = 0; m < MJ; ++m)
#pragma unroll
        for (int u = 0; u < NV; ++u) acc[m][u] = 0.0;
    {
        double ec[NV], yc[NV];
#pragma unroll
        for (int u = 0; u < NV; ++u) sy[u] = ec[u] = yc[u] = 0.0;
        for (int t0 = 0; t0 < T; t0 += 64) {
            double yw[NV];
#pragma unroll
            for (int u = 0; u < NV; ++u) {
                double yt;
                yw[u] = seg_whiten_chunk(cR, rows[u], inv[u], T, t0, ec[u], yc[u], yt);
                sy[u] += yt;
            }
            const int n = min(64, T - t0);
            for (int ii = 0; ii < n; ++ii) {
                const double* xr = Xw + (size_t)(t0 + ii) * p;
                double xv[MJ];
#pragma unroll
                for (int m = 0; m < MJ; ++m) {
                    const int j = lane + 64 * m;
                    xv[m] = j < p ? xr[j] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < NV; ++u) {
                    const double ywi = __shfl(yw[u], ii);
#pragma unroll
                    for (int m = 0; m < MJ; ++m) acc[m][u] = fma(xv[m], ywi, acc[m][u]);
                }
            }
        }
    }
    // beta_c = sum_j Ainv[j][c] q_j (row j of the symmetric Ainv, coalesced)
    double bl[MJ][NV];
#pragma unroll
    for (int m = 0; m < MJ; ++m)
#pragma unroll
        for (int u = 0; u < NV; ++u) bl[m][u] = 0.0;
#pragma unroll
    for (int m2 = 0; m2 < MJ; ++m2) {
        const int nj = min(64, p - 64 * m2);
        for (int ii = 0; ii < nj; ++ii) {
            const double* ar = Ai + (size_t)(64 * m2 + ii) * p;
            double av[MJ];
#pragma unroll
            for (int m = 0; m < MJ; ++m) {
                const int c = lane + 64 * m;
                av[m] = c < p ? ar[c] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < NV; ++u) {
                const double qj = __shfl(acc[m2][u], ii);
#pragma unroll
                for (int m = 0; m < MJ; ++m) bl[m][u] = fma(av[m], qj, bl[m][u]);
            }
        }
    }
    double mean[NV];
#pragma unroll
    for (int u = 0; u < NV; ++u) mean[u] = wave_sum(sy[u]) / T;
    // residual pass: per sample, the fitted value x_t' beta of each voxel
    // (lanes over j, then a wave sum; lane ii of the chunk keeps sample
    // t0 + ii), then r = L^-1 (y - X beta) = yw - X_w beta by the chunk scan
    // of the residual (the rows of the design X are shared by every voxel)
    double ssw[NV], ssres[NV], sstot[NV];
    {
        double ec[NV], dc[NV];
#pragma unroll
        for (int u = 0; u < NV; ++u) ssw[u] = ssres[u] = sstot[u] = ec[u] = dc[u] = 0.0;
        for (int t0 = 0; t0 < T; t0 += 64) {
            const int t = t0 + lane;
            double yt[NV], dx[NV];
#pragma unroll
            for (int u = 0; u < NV; ++u) {
                yt[u] = t < T ? (double)rows[u][t] * inv[u] : 0.0;
                dx[u] = 0.0;
            }
            const int n = min(64, T - t0);
            for (int ii = 0; ii < n; ++ii) {
                const double* xr = P.X + (size_t)(t0 + ii) * p;
                double xv[MJ];
#pragma unroll
                for (int m = 0; m < MJ; ++m) {
                    const int j = lane + 64 * m;
                    xv[m] = j < p ? xr[j] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < NV; ++u) {
                    double ax = 0.0;
#pragma unroll
                    for (int m = 0; m < MJ; ++m) ax = fma(xv[m], bl[m][u], ax);
                    ax = wave_sum(ax);
                    if (lane == ii) dx[u] = ax;
                }
            }
#pragma unroll
            for (int u = 0; u < NV; ++u) {
                const double dr = t < T ? yt[u] - dx[u] : 0.0;
                const double r = seg_whiten_vals(cR, dr, T, t0, ec[u], dc[u]);
                if (t < T && act[u]) {
                    const double dm = yt[u] - mean[u];
                    ssw[u] += r * r;
                    ssres[u] += dr * dr;
                    sstot[u] += dm * dm;
                    const size_t o = (size_t)vox[u] * ldR + t;
                    if (resid_w) resid_w[o] = r * sc[u];
                    if (resid) resid[o] = (float)(dr * sc[u]);
                    if (pred_out) pred_out[o] = (float)(dx[u] * sc[u]);
                }
            }
        }
    }
    // F form: u_c = sum_a beta_a Fm[a][c] (row a of Fm, coalesced)
#pragma unroll
    for (int m = 0; m < MJ; ++m)
#pragma unroll
        for (int u = 0; u < NV; ++u) acc[m][u] = 0.0;
#pragma unroll
    for (int m2 = 0; m2 < MJ; ++m2) {
        const int na = min(64, p - 64 * m2);
        for (int ii = 0; ii < na; ++ii) {
            const double* fr = F + (size_t)(64 * m2 + ii) * p;
            double fv[MJ];
#pragma unroll
            for (int m = 0; m < MJ; ++m) {
                const int c = lane + 64 * m;
                fv[m] = c < p ? fr[c] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < NV; ++u) {
                const double ba = __shfl(bl[m2][u], ii);
#pragma unroll
                for (int m = 0; m < MJ; ++m) acc[m][u] = fma(fv[m], ba, acc[m][u]);
            }
        }
    }
    const int dof = max(1, T - p);
#pragma unroll
    for (int u = 0; u < NV; ++u) {
        const double w2 = wave_sum(ssw[u]), rs = wave_sum(ssres[u]), tt = wave_sum(sstot[u]);
        const double sig = w2 / dof;
        double quad = 0.0;
#pragma unroll
        for (int m = 0; m < MJ; ++m) {
            const int c = lane + 64 * m;
            if (c < p) {
                quad = fma(acc[m][u], bl[m][u], quad);
                if (act[u]) {
                    const double se = sqrt(sig * Ai[(size_t)c * p + c]);
                    tstats[(size_t)vox[u] * p + c] = (float)(bl[m][u] / (se + 1e-10));
                    betas[(size_t)vox[u] * p + c] = (float)(bl[m][u] * sc[u]);
                }
            }
        }
        quad = wave_sum(quad);
        if (lane == 0 && act[u]) {
            fstats[vox[u]] = (float)(quad / (sig * P.n_task + 1e-10));
            sigma2[vox[u]] = (float)(sig * sc[u] * sc[u]);
            r2[vox[u]] = (float)(tt <= 1e-10 ? 0.0 : fmin(1.0 - rs / tt, 1.0));
        }
    }
}

}  // namespace remlk
