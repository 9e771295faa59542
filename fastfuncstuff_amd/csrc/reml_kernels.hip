// gfx950 kernels of the ARMA(1,1) REML search and GLS refit.
//
//   k_tables      FP64 per-pair tables: innovation coefficients, L^-1 X,
//                 chol(X'R^-1X), its inverse, G = R^-1 X Lx^-T      (one block per pair)
//   k_prep        per-voxel std normalisation + residualisation (+ ACF features)
//   k_energy      (i)  ||L^-1 y||^2 for every (voxel, pair) by the innovations recurrence
//   k_project     (ii) MFMA  Y (V x T) . G (T x K p), fused REML score epilogue
//   k_select      (iii) exhaustive / hierarchical argmin over the surface
//   k_gls         (iv) FP64 GLS refit + t / F / R^2 / sigma^2
//   k_ljung_box   whitened-residual Ljung-Box statistic
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define FFS_HD __host__ __device__
#include "reml_kernels.h"
#include "reml_math.h"

namespace remlk {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// FP64 per-pair tables.  One 64-thread block per pair.
//   scratch: [K][3][T]  s2 / c / alpha (SoA, kept for the energy tables)
//   work:    [K][2*p*p] Gauss-Jordan / Cholesky-inverse workspace
// ---------------------------------------------------------------------------
// Block-cooperative forms of cholesky_inplace / cholesky_inverse /
// gauss_jordan_inverse (reml_math.h) for the k_tables block (64 or 256 lanes).  Every
// element is computed by one lane with the serial routine's expression and
// summation order; only independent rows / columns are spread over lanes, with
// a barrier between dependent steps.  `flag` is block-shared and starts at 0;
// every lane reads it after the same barrier, so all lanes leave together.

// A (p x p, row-major, symmetric) -> lower factor L (upper part zeroed).
// Returns -1 (flag = 1) when a pivot is not positive.  dep (optional): the
// columns marked there get a zero pivot and a zero column below it (the
// rank-r factor of a design whose column j depends on columns 0..j-1).
__device__ int block_cholesky(int p, double* A, int* flag, const int32_t* dep = nullptr) {
    const int tid = threadIdx.x;
    for (int j = 0; j < p; ++j) {
        const bool skip = dep && dep[j];
        if (tid == 0) {
            if (skip) {
                A[j * p + j] = 0.0;
            } else {
                double d = A[j * p + j];
                for (int k = 0; k < j; ++k) d -= A[j * p + k] * A[j * p + k];
                if (!(d > 0.0))
                    *flag = 1;
                else
                    A[j * p + j] = sqrt(d);
            }
        }
        __syncthreads();
        if (*flag) return -1;
        const double d = A[j * p + j];
        for (int i = j + 1 + tid; i < p; i += (int)blockDim.x) {
            if (skip) {
                A[i * p + j] = 0.0;
                continue;
            }
            double s = A[i * p + j];
            for (int k = 0; k < j; ++k) s -= A[i * p + k] * A[j * p + k];
            A[i * p + j] = s / d;
        }
        for (int k = j + 1 + tid; k < p; k += (int)blockDim.x) A[j * p + k] = 0.0;
        __syncthreads();
    }
    return 0;
}

// Ainv = (L L')^-1 = L^-T L^-1; tmp (p x p) receives L^-1, one column per lane.
// Zero pivots (dependent columns of a rank-r factor) get zero rows and
// columns in L^-1, so Ainv is the rank-r inverse embedded at the independent
// columns.
__device__ void block_cholesky_inverse(int p, const double* L, double* Ainv, double* tmp) {
    const int tid = threadIdx.x;
    for (int j = tid; j < p; j += (int)blockDim.x) {
        for (int i = 0; i < p; ++i) tmp[i * p + j] = 0.0;
        if (L[j * p + j] == 0.0) continue;
        tmp[j * p + j] = 1.0 / L[j * p + j];
        for (int i = j + 1; i < p; ++i) {
            if (L[i * p + i] == 0.0) continue;
            double s = 0.0;
            for (int k = j; k < i; ++k) s += L[i * p + k] * tmp[k * p + j];
            tmp[i * p + j] = -s / L[i * p + i];
        }
    }
    __syncthreads();
    for (int idx = tid; idx < p * p; idx += (int)blockDim.x) {
        const int i = idx / p, j = idx % p;
        if (j > i) continue;
        double s = 0.0;
        for (int k = i; k < p; ++k) s += tmp[k * p + i] * tmp[k * p + j];
        Ainv[i * p + j] = s;
        Ainv[j * p + i] = s;
    }
    __syncthreads();
}

// A <- Pi A Pi' (p x p, A symmetric); W (p x p) is workspace.
__device__ void block_sandwich(int p, const double* Pi, double* A, double* W) {
    const int tid = threadIdx.x;
    for (int idx = tid; idx < p * p; idx += (int)blockDim.x) {
        const int i = idx / p, j = idx % p;
        double s = 0.0;
        for (int q = 0; q < p; ++q) s = fma(Pi[i * p + q], A[q * p + j], s);
        W[idx] = s;
    }
    __syncthreads();
    for (int idx = tid; idx < p * p; idx += (int)blockDim.x) {
        const int i = idx / p, j = idx % p;
        if (j > i) continue;
        double s = 0.0;
        for (int q = 0; q < p; ++q) s = fma(W[i * p + q], Pi[j * p + q], s);
        A[i * p + j] = s;
        A[j * p + i] = s;
    }
    __syncthreads();
}

// General inverse with partial pivoting; -1 (flag = fail_code) if singular.
__device__ int block_gauss_jordan_inverse(int p, const double* A, double* inv, double* work, int* piv_sh, int* flag,
                                          int fail_code) {
    const int tid = threadIdx.x;
    for (int i = tid; i < p * p; i += (int)blockDim.x) {
        work[i] = A[i];
        inv[i] = (i / p == i % p) ? 1.0 : 0.0;
    }
    __syncthreads();
    for (int col = 0; col < p; ++col) {
        if (tid == 0) {
            int piv = col;
            double best = fabs(work[col * p + col]);
            for (int r = col + 1; r < p; ++r)
                if (fabs(work[r * p + col]) > best) {
                    best = fabs(work[r * p + col]);
                    piv = r;
                }
            if (!(best > 0.0)) *flag = fail_code;
            *piv_sh = piv;
        }
        __syncthreads();
        if (*flag) return -1;
        const int piv = *piv_sh;
        if (piv != col)
            for (int j = tid; j < p; j += (int)blockDim.x) {
                double t = work[col * p + j];
                work[col * p + j] = work[piv * p + j];
                work[piv * p + j] = t;
                t = inv[col * p + j];
                inv[col * p + j] = inv[piv * p + j];
                inv[piv * p + j] = t;
            }
        __syncthreads();
        const double d = work[col * p + col];
        __syncthreads();
        for (int j = tid; j < p; j += (int)blockDim.x) {
            work[col * p + j] /= d;
            inv[col * p + j] /= d;
        }
        __syncthreads();
        for (int r = tid; r < p; r += (int)blockDim.x) {
            if (r == col) continue;
            const double f = work[r * p + col];
            if (f == 0.0) continue;
            for (int j = 0; j < p; ++j) {
                work[r * p + j] -= f * work[col * p + j];
                inv[r * p + j] -= f * inv[col * p + j];
            }
        }
        __syncthreads();
    }
    return 0;
}

__global__ void __launch_bounds__(256) k_tables(KTablesArgs A, const TablesOut* __restrict__ outs,
                                               int* __restrict__ status_all) {
    // one block per (pair, design): blockIdx.x = pair, blockIdx.y = design
    const TablesOut o = outs[blockIdx.y];
    int* status = status_all + blockIdx.y;
    const int T = A.T, Tpad = A.Tpad, p = A.p, K = A.K, n_task = A.n_task, full_f = A.full_f, BN = A.BN,
              npt = A.npt, PG = A.PG;
    const double ridge = A.ridge;
    const double* __restrict__ ab = A.ab;
    const int32_t* __restrict__ white_flag = A.white_flag;
    const int32_t* __restrict__ gap = A.gap;
    const int32_t* __restrict__ task_idx = A.task_idx;
    const double* __restrict__ X = o.X;
    double* __restrict__ soa = o.soa;
    double* __restrict__ coef64 = o.coef64;
    double* __restrict__ Xw64 = o.Xw64;
    double* __restrict__ G64 = o.G64;
    double* __restrict__ Lx = o.Lx;
    double* __restrict__ Ainv = o.Ainv;
    double* __restrict__ Fm = o.Fm;
    double* __restrict__ work = o.work;
    double* __restrict__ ld = o.ld;
    double* __restrict__ ldR_out = o.ldR;
    double* __restrict__ ldX_out = o.ldX;
    float* __restrict__ G32 = o.G32;
    uint16_t* __restrict__ Ghf = o.Ghf;
    uint16_t* __restrict__ Glf = o.Glf;
    uint16_t* __restrict__ Ghb = o.Ghb;
    uint16_t* __restrict__ Glb = o.Glb;
    float* __restrict__ g16_unscale = o.g16s;
    double* __restrict__ G64p = o.G64p;
    double* __restrict__ coefR = o.coefR;
    const int k = blockIdx.x;
    if (k >= K) return;
    const int tid = threadIdx.x, nth = blockDim.x;  // 64, or 256 for wide designs
    const double a = ab[2 * k], b = ab[2 * k + 1];
    double* s2 = soa + (size_t)k * 3 * T;
    double* cc = s2 + T;
    double* al = cc + T;
    double* xw = Xw64 + (size_t)k * T * p;
    double* g = G64 + (size_t)k * T * p;
    double* L = Lx + (size_t)k * p * p;
    double* Ai = Ainv + (size_t)k * p * p;
    double* wk = work + (size_t)k * 2 * p * p;
    __shared__ double sh_ldR, sh_ldX;
    __shared__ int sh_fail, sh_piv;

    if (tid == 0) {
        sh_ldR = pair_coefficients(a, b, T, gap, white_flag[k], s2, cc, al);
        sh_fail = 0;
    }
    __syncthreads();
    double* cf = coef64 + (size_t)k * T * 3;
    for (int t = tid; t < T; t += nth) {
        cf[3 * t + 0] = sqrt(s2[t]);
        cf[3 * t + 1] = cc[t];
        cf[3 * t + 2] = al[t];
    }
    // X_w = L^-1 X, one column per thread.
    for (int j = tid; j < p; j += nth) whiten_forward(T, s2, cc, al, gap, X + j, p, xw + j, p);
    __syncthreads();
    // A = X_w' X_w  (into L, full symmetric)
    for (int idx = tid; idx < p * p; idx += nth) {
        const int i = idx / p, j = idx % p;
        if (j > i) continue;
        double s = 0.0;
        for (int t = 0; t < T; ++t) s += xw[(size_t)t * p + i] * xw[(size_t)t * p + j];
        L[i * p + j] = s;
        L[j * p + i] = s;
    }
    __syncthreads();
    // F-statistic quadratic form, embedded in p x p (zero outside the task columns).
    // The factorisations below run on the whole block (block_* helpers); each
    // element gets the arithmetic of the one-lane reml_math.h routines in the
    // same order, so the tables are identical to a serial build.
    double* F = Fm + (size_t)k * p * p;
    for (int i = tid; i < p * p; i += nth) F[i] = full_f ? L[i] : 0.0;
    __syncthreads();
    // ridge plans (glm/arma.py:1160-1176, 2724-2755): A + ridge I in the
    // log-determinant and in the solve for beta
    if (ridge > 0.0)
        for (int j = tid; j < p; j += nth) L[j * p + j] += ridge;
    __syncthreads();
    // rank-deficient designs (never ridge plans): the dependent columns get
    // zero pivots (glm/arma.py:1493-1500 falls back to QR, whose R has a
    // rounding-level diagonal there)
    if (block_cholesky(p, L, &sh_fail, ridge > 0.0 ? nullptr : o.dep) == 0) {
        if (tid == 0) {
            double s = 0.0;
            if (ridge > 0.0)
                for (int j = 0; j < p; ++j) s += 2.0 * log(L[j * p + j]);
            else
                for (int j = 0; j < p; ++j) s += 2.0 * log(fabs(L[j * p + j]) + 1e-10);
            sh_ldX = s;
        }
        block_cholesky_inverse(p, L, Ai, wk);
        // (X'R^-1X)^+ = Pi (rank-r inverse) Pi'
        if (o.Pi && ridge <= 0.0) block_sandwich(p, o.Pi, Ai, wk);
        if (ridge > 0.0) {
            // |y_w - X_w beta|^2 = E - c' W c, c = X_w' y_w, M = (A + rI)^-1,
            // W = M (A + 2rI) M = M + r M M; L <- chol(W)
            double* Wm = wk;
            for (int idx = tid; idx < p * p; idx += nth) {
                const int i = idx / p, j = idx % p;
                double mm = 0.0;
                for (int q = 0; q < p; ++q) mm += Ai[i * p + q] * Ai[q * p + j];
                Wm[idx] = Ai[idx] + ridge * mm;
            }
            __syncthreads();
            for (int i = tid; i < p * p; i += nth) L[i] = Wm[i];
            __syncthreads();
            block_cholesky(p, L, &sh_fail);
        }
        if (!full_f && !sh_fail) {
            // inv(A^-1[task, task] + 1e-8 I)   (glm/arma.py:4888-4896)
            // At (wk) -> inverse in F's storage (work: W2) -> copied to W2 ->
            // F zeroed and the block embedded at the task positions.
            double* At = wk;
            double* W2 = wk + p * p;
            for (int idx = tid; idx < n_task * n_task; idx += nth) {
                const int i = idx / n_task, j = idx % n_task;
                At[idx] = Ai[task_idx[i] * p + task_idx[j]] + (i == j ? 1e-8 : 0.0);
            }
            __syncthreads();
            if (block_gauss_jordan_inverse(n_task, At, F, W2, &sh_piv, &sh_fail, 2) == 0) {
                for (int q = tid; q < n_task * n_task; q += nth) W2[q] = F[q];
                __syncthreads();
                for (int q = tid; q < p * p; q += nth) F[q] = 0.0;
                __syncthreads();
                for (int idx = tid; idx < n_task * n_task; idx += nth) {
                    const int a2 = idx / n_task, b2 = idx % n_task;
                    F[task_idx[a2] * p + task_idx[b2]] = W2[idx];
                }
            }
        }
    }
    __syncthreads();
    if (sh_fail) {
        if (tid == 0) atomicMax(status, sh_fail);
        return;
    }
    // Q rows: q_t = Lx^-1 xw_t (forward substitution), stored in g; ridge
    // plans: q_t = chol(W)' xw_t.
    for (int t = tid; t < T; t += nth) {
        const double* x = xw + (size_t)t * p;
        double* q = g + (size_t)t * p;
        if (ridge > 0.0) {
            for (int j = 0; j < p; ++j) {
                double s = 0.0;
                for (int i = j; i < p; ++i) s += L[i * p + j] * x[i];
                q[j] = s;
            }
            continue;
        }
        for (int i = 0; i < p; ++i) {
            double s = x[i];
            for (int m = 0; m < i; ++m) s -= L[i * p + m] * q[m];
            q[i] = L[i * p + i] != 0.0 ? s / L[i * p + i] : 0.0;  // zero pivot: dependent column
        }
    }
    __syncthreads();
    // G = L^-T Q, in place, one column per thread.
    for (int j = tid; j < p; j += nth) whiten_adjoint(T, s2, cc, al, gap, g + j, p, g + j, p);
    __syncthreads();
    // fp32 projection operand, tile-major: G32[tile][t][slot*p + j]
    const int tile = k / npt, slot = k % npt;
    if (G32 && A.nseg > 0) {  // p > 96: the pair's segments of BN columns
        for (int idx = tid; idx < T * p; idx += nth) {
            const int t = idx / p, j = idx % p;
            G32[((size_t)k * A.nseg + j / BN) * Tpad * BN + (size_t)t * BN + j % BN] = (float)g[idx];
        }
    } else if (G32) {  // null for a pair-list (GLS-only) plan
        float* Gt = G32 + (size_t)tile * Tpad * BN + slot * p;
        for (int idx = tid; idx < T * p; idx += nth) {
            const int t = idx / p, j = idx % p;
            Gt[(size_t)t * BN + j] = (float)g[idx];
        }
    }
    // split 16-bit operands, layout [tile][t/8][BN][8].  The fp16 halves hold
    // G * 2^e with 2^e * max|G| in [2^13, 2^14): the high half keeps headroom
    // below the fp16 maximum and the low half stays out of the subnormal range.
    // The epilogue multiplies the sum of squares by 2^-2e (exact).
    if (Ghf) {
        __shared__ double sh_max[256];
        double m = 0.0;
        for (int idx = tid; idx < T * p; idx += nth) m = fmax(m, fabs(g[idx]));
        sh_max[tid] = m;
        __syncthreads();
        for (int off = nth >> 1; off > 0; off >>= 1) {
            if (tid < off) sh_max[tid] = fmax(sh_max[tid], sh_max[tid + off]);
            __syncthreads();
        }
        int e2 = 0;
        if (sh_max[0] > 0.0) frexp(sh_max[0], &e2);  // max in [2^(e2-1), 2^e2)
        const int e = min(max(14 - e2, -60), 60);
        if (tid == 0) g16_unscale[k] = (float)ldexp(1.0, -2 * e);
        const size_t tbase = (size_t)tile * Tpad * BN;
        for (int idx = tid; idx < T * p; idx += nth) {
            const int t = idx / p, j = idx % p;
            // p > 96: column j % BN of the pair's segment tile k * nseg + j / BN
            const size_t o = A.nseg > 0
                                 ? ((size_t)k * A.nseg + j / BN) * Tpad * BN + ((size_t)(t >> 3) * BN + j % BN) * 8 + (t & 7)
                                 : tbase + ((size_t)(t >> 3) * BN + slot * p + j) * 8 + (t & 7);
            const double gv = g[idx];
            const double gs = ldexp(gv, e);
            const _Float16 hf = (_Float16)gs;
            const _Float16 lf = (_Float16)(gs - (double)hf);
            const __bf16 hb = (__bf16)(float)gv;
            const __bf16 lb = (__bf16)(float)(gv - (double)(float)hb);
            Ghf[o] = __builtin_bit_cast(uint16_t, hf);
            Glf[o] = __builtin_bit_cast(uint16_t, lf);
            Ghb[o] = __builtin_bit_cast(uint16_t, hb);
            Glb[o] = __builtin_bit_cast(uint16_t, lb);
        }
    }
    // f64 MFMA GLS operand: G padded to PG columns (p <= 32 only).
    if (G64p) {
        double* gp = G64p + (size_t)k * Tpad * PG;
        for (int idx = tid; idx < T * p; idx += nth) {
            const int t = idx / p, j = idx % p;
            gp[(size_t)t * PG + j] = g[idx];
        }
    }
    // Recurrence coefficients for every p (f64 GLS, -dsort and the whiteners),
    // shifted so row t holds what step t needs (zeros past T and at run starts).
    {
        double* cr = coefR + (size_t)k * Tpad * 4;
        for (int t = tid; t < T; t += nth) {
            const bool start = (t == 0) || gap[t - 1] <= 0;
            cr[4 * t] = sqrt(s2[t]);
            cr[4 * t + 1] = start ? 0.0 : cc[t - 1];
            cr[4 * t + 2] = start ? 0.0 : al[t - 1];
            cr[4 * t + 3] = 0.0;
        }
    }
    if (tid == 0) {
        ld[k] = sh_ldR + sh_ldX;
        ldR_out[k] = sh_ldR;
        ldX_out[k] = sh_ldX;
    }
}

// Energy-kernel tables from the FP64 SoA coefficients.
__global__ void k_energy_tables(int T, int n_chunks, const int32_t* __restrict__ echunk,
                                const double* __restrict__ soa, float* __restrict__ ecoef,
                                float* __restrict__ ec, float* __restrict__ es0) {
    const int chunk = blockIdx.x;
    if (chunk >= n_chunks) return;
    const int32_t* slots = echunk + chunk * kEnergyNB;
    int first[2] = {-1, -1};  // first pair of each half (every pair of a half has the same a)
    for (int h = 0; h < 2; ++h)
        for (int j = 0; j < kEnergyHalf; ++j)
            if (first[h] < 0 && slots[h * kEnergyHalf + j] >= 0) first[h] = slots[h * kEnergyHalf + j];
    for (int t = threadIdx.x; t < T; t += blockDim.x) {
        for (int h = 0; h < 2; ++h)
            ec[((size_t)chunk * T + t) * 2 + h] = first[h] >= 0 ? (float)soa[(size_t)first[h] * 3 * T + T + t] : 0.f;
        for (int j = 0; j < kEnergyNB; ++j) {
            const int k = slots[j];
            float alpha = 0.f, s2n = 0.f;
            if (k >= 0) {
                const double* s2 = soa + (size_t)k * 3 * T;
                alpha = (float)s2[2 * T + t];
                s2n = (t + 1 < T) ? (float)s2[t + 1] : 0.f;
            }
            ecoef[((size_t)chunk * T + t) * 2 * kEnergyNB + 2 * j] = alpha;
            ecoef[((size_t)chunk * T + t) * 2 * kEnergyNB + 2 * j + 1] = s2n;
        }
    }
    if (threadIdx.x < kEnergyNB) {
        const int k = slots[threadIdx.x];
        es0[chunk * kEnergyNB + threadIdx.x] = k >= 0 ? (float)soa[(size_t)k * 3 * T] : 0.f;
    }
}

// Per-pair fp32 coefficients of the local search (same roundings as the chunk tables).
__global__ void k_local_tables(int T, int K, const double* __restrict__ soa, float2* __restrict__ lp,
                               float* __restrict__ lc, float* __restrict__ ls0) {
    const int k = blockIdx.x;
    if (k >= K) return;
    const double* s2 = soa + (size_t)k * 3 * T;
    for (int t = threadIdx.x; t < T; t += blockDim.x) {
        lp[(size_t)k * T + t] = make_float2((float)s2[2 * T + t], (t + 1 < T) ? (float)s2[t + 1] : 0.f);
        lc[(size_t)k * T + t] = (float)s2[T + t];
    }
    if (threadIdx.x == 0) ls0[k] = (float)s2[0];
}

// ---------------------------------------------------------------------------
// Prep: normalise by the unbiased std (glm/arma.py:3947-3954), residualise
// against the design with its pseudo-inverse (glm/arma.py:1554-1571) and
// write the residual time-major (Yt[t][v]) for the search kernels.
// Optional: lag-1..nfeat autocorrelation of the residual (proposal features).
// ---------------------------------------------------------------------------
constexpr int kStageChunk = 32;

// Samples [t0, t0+kStageChunk) of one row into registers.  Full chunks of a
// 16-byte aligned row use unconditional float4 loads, all issued before use;
// the tail chunk and unaligned rows use guarded scalar loads.
__device__ __forceinline__ void load_chunk(const float* __restrict__ row, int T, int t0, bool valid, bool vec4,
                                           float (&yv)[kStageChunk]) {
    if (valid && vec4 && t0 + kStageChunk <= T) {
        const float4* r4 = reinterpret_cast<const float4*>(row + t0);
        float4 q[kStageChunk / 4];
#pragma unroll
        for (int i = 0; i < kStageChunk / 4; ++i) q[i] = r4[i];
#pragma unroll
        for (int i = 0; i < kStageChunk / 4; ++i) {
            yv[4 * i] = q[i].x;
            yv[4 * i + 1] = q[i].y;
            yv[4 * i + 2] = q[i].z;
            yv[4 * i + 3] = q[i].w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < kStageChunk; ++i) yv[i] = (valid && t0 + i < T) ? row[t0 + i] : 0.f;
    }
}

// Stage rows [t0, t0+kStageChunk) of a row-major (T x w) float64 table into
// dst[i*w + j] (all lanes of the wave cooperate; coalesced; w <= WMAX); loads
// are issued before the LDS writes.
template <int WMAX>
__device__ inline void stage_table(const double* __restrict__ src, int T, int w, int t0, double* dst) {
    const int lane = threadIdx.x & 63;
    const int n = min(kStageChunk, T - t0) * w;
    const double* s0 = src + (size_t)t0 * w;
    constexpr int NI = (kStageChunk * WMAX + 63) / 64;
    double vals[NI];
#pragma unroll
    for (int i = 0; i < NI; ++i) {
        const int q = lane + 64 * i;
        vals[i] = q < n ? s0[q] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < NI; ++i) {
        const int q = lane + 64 * i;
        if (q < n) dst[q] = vals[i];
    }
}

// One wave per 64 output columns; the rows are staged through LDS.
template <int PM, int NF>
__global__ void __launch_bounds__(64) k_prep(int T, int p, const float* __restrict__ Y, int64_t ldY, int V,
                                             const int32_t* __restrict__ rows, int ncols, int ncols_pad,
                                             const double* __restrict__ pinvT, const double* __restrict__ X,
                                             int normalize, float* __restrict__ Yt, int64_t ldYt,
                                             float* __restrict__ scale, float* __restrict__ feats,
                                             const PlanDev* __restrict__ plans, const int32_t* __restrict__ grp) {
    __shared__ double tab[kStageChunk * PM];
    const int lane = threadIdx.x & 63;
    const int col = blockIdx.x * 64 + lane;
    if (blockIdx.x * 64 >= ncols_pad) return;
    if (plans) {  // grouped layout: the design of this 64-column block
        const int g = max(grp[blockIdx.x], 0);
        pinvT = plans[g].pinvT;
        X = plans[g].X;
    }
    int v = (col < ncols) ? (rows ? rows[col] : col) : -1;
    if (v >= V) v = -1;
    const float* __restrict__ row = Y + (size_t)(v >= 0 ? v : 0) * ldY;
    const bool vec4 = (ldY & 3) == 0;
    float yrow[kStageChunk];
    double s1 = 0.0, s2 = 0.0, c[PM];
#pragma unroll
    for (int j = 0; j < PM; ++j) c[j] = 0.0;
    for (int t0 = 0; t0 < T; t0 += kStageChunk) {
        load_chunk(row, T, t0, v >= 0, vec4, yrow);
        __syncthreads();
        stage_table<PM>(pinvT, T, p, t0, tab);
        __syncthreads();
        const int n = min(kStageChunk, T - t0);
#pragma unroll
        for (int i = 0; i < kStageChunk; ++i) {
            if (i >= n) break;
            const double yt = (double)yrow[i];
            s1 += yt;
            s2 += yt * yt;
            const double* pr = tab + i * p;
#pragma unroll
            for (int j = 0; j < PM; ++j)
                if (j < p) c[j] += pr[j] * yt;
        }
    }
    double s = 1.0;
    if (normalize & 1) {
        const double mean = s1 / T;
        double var = T > 1 ? (s2 - T * mean * mean) / (T - 1) : 0.0;
        double sd = var > 0.0 ? sqrt(var) : 0.0;
        s = (double)fmaxf((float)sd, 1e-8f);
    }
    const double inv = 1.0 / s;
#pragma unroll
    for (int j = 0; j < PM; ++j) c[j] = (normalize & 2) ? 0.0 : c[j] * inv;  // bit 1: no residualisation
    float hist[NF > 0 ? NF : 1];
    double acf[NF > 0 ? NF : 1];
    double acf0 = 0.0;
#pragma unroll
    for (int k = 0; k < (NF > 0 ? NF : 1); ++k) {
        hist[k] = 0.f;
        acf[k] = 0.0;
    }
    if (Yt || NF > 0) {
        for (int t0 = 0; t0 < T; t0 += kStageChunk) {
            load_chunk(row, T, t0, v >= 0, vec4, yrow);
            __syncthreads();
            stage_table<PM>(X, T, p, t0, tab);
            __syncthreads();
            const int n = min(kStageChunk, T - t0);
#pragma unroll
            for (int i = 0; i < kStageChunk; ++i) {
                if (i >= n) break;
                const int t = t0 + i;
                double r = (double)yrow[i] * inv;
                const double* xr = tab + i * p;
#pragma unroll
                for (int j = 0; j < PM; ++j)
                    if (j < p) r -= xr[j] * c[j];
                const float rf = v >= 0 ? (float)r : 0.f;
                if (Yt && col < ncols_pad) Yt[(size_t)t * ldYt + col] = rf;
                if (NF > 0) {
                    acf0 += (double)rf * rf;
#pragma unroll
                    for (int k = 0; k < NF; ++k) acf[k] += (double)rf * hist[k];
#pragma unroll
                    for (int k = NF - 1; k > 0; --k) hist[k] = hist[k - 1];
                    hist[0] = rf;
                }
            }
        }
    }
    if (v < 0) return;
    if (scale) scale[v] = (float)s;
    if (NF > 0 && feats) {
#pragma unroll
        for (int k = 0; k < NF; ++k) feats[(size_t)v * NF + k] = acf0 > 0.0 ? (float)(acf[k] / acf0) : 0.f;
    }
}

// Energy of the lambda == 0 pairs: sum y_t^2 with the accumulation schedule of
// k_energy (first sample, then groups of 8 steps flushed into fp64).
__device__ inline float white_energy(const float* __restrict__ y, int64_t ldY, int T) {
    const float y0 = y[0];
    float part = y0 * y0;
    double acc = 0.0;
    for (int t0 = 0; t0 < T - 1; t0 += 8) {
        for (int i = 0; i < 8; ++i) {
            const int t = t0 + i;
            if (t < T - 1) {
                const float e = y[(size_t)(t + 1) * ldY];
                part = fmaf(e * 1.0f, e, part);
            }
        }
        acc += part;
        part = 0.f;
    }
    return (float)(acc + part);
}

// ---------------------------------------------------------------------------
// (i) Energy ||L^-1 y||^2 for every (voxel, pair).
// Block: kEnergyWaves waves on one chunk (kEnergyNB pairs, each half of
// kEnergyHalf sharing the AR coefficient); wave w takes the block's w-th
// 64-voxel tile.
// Per step and half: z = y_{t+1} - c_t y_t; per pair: e <- alpha e + z ;
// E += e^2 / v   (3 VALU ops), fp32 partial sums flushed into fp64 every 8 steps.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kEnergyVox* kEnergyWaves) k_energy(PlanDev P, const float* __restrict__ Yt,
                                                                     int64_t ldY, int V, float* __restrict__ E,
                                                                     int64_t ldE, const int32_t* __restrict__ n_live) {
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // Item (voxel octet, chunk) of this block: its kEnergyWaves waves take the
    // chunk on kEnergyWaves consecutive 64-voxel tiles, so they read one
    // coefficient stream (scalar-cache hits for all but one wave).  Items are
    // numbered octet-major, chunk fastest, and each XCD (block id % 8) takes a
    // contiguous eighth of them: the chunks of an octet run together on one
    // XCD (at most two at a range boundary), so its residual columns are
    // fetched once into that XCD's L2, and any voxel count spreads evenly.
    // With a live count the items cover the live voxels only (the grid is
    // sized for V).
    if (n_live) V = min(V, *n_live);
    const int n_xt = (V + kEnergyVox - 1) / kEnergyVox;
    const int n_xo = (n_xt + kEnergyWaves - 1) / kEnergyWaves;
    const int n_items = n_xo * P.n_chunks;
    const int per = (n_items + 7) >> 3;  // items per XCD
    const int bid = blockIdx.x, local = bid >> 3;
    const int item = (bid & 7) * per + local;
    if (local >= per || item >= n_items) return;
    const int xo = item / P.n_chunks, chunk = item % P.n_chunks;
    const int xt = xo * kEnergyWaves + w;
    if (xt >= n_xt) return;
    const int v = xt * kEnergyVox + lane;
    const int T = P.T;
    const float* __restrict__ y = Yt + v;

    if (P.n_white > 0 && chunk == 0) {
        // lambda == 0: R = I, so the energy is sum y^2, shared by every such pair.
        // Same accumulation schedule as energy_accumulate (bitwise equal to it
        // with identity coefficients).
        const float acc = white_energy(y, ldY, T);
        if (v < V)
            for (int i = 0; i < P.n_white; ++i) E[(size_t)v * ldE + P.white[i]] = acc;
        return;
    }
    const float* __restrict__ cz = P.ec + (size_t)chunk * T * 2;
    const float* __restrict__ co = P.ecoef + (size_t)chunk * T * 2 * kEnergyNB;
    const float* __restrict__ s0 = P.es0 + chunk * kEnergyNB;
    float e[kEnergyNB], part[kEnergyNB];
    double acc[kEnergyNB];
    const float y0 = y[0];
#pragma unroll
    for (int j = 0; j < kEnergyNB; ++j) {
        e[j] = y0;
        part[j] = y0 * y0 * s0[j];
        acc[j] = 0.0;
    }
    float yc = y0;
    // One step t -> t+1 for every pair of the chunk (coefficients wave-uniform):
    // the AR-filtered sample of each half, shared by its pairs.
    auto step = [&](float yn, float cz0, float cz1, const float* c) __attribute__((always_inline)) {
        const float z0 = fmaf(-cz0, yc, yn);
        const float z1 = fmaf(-cz1, yc, yn);
#pragma unroll
        for (int j = 0; j < kEnergyNB; ++j) {
            e[j] = fmaf(c[2 * j], e[j], j < kEnergyHalf ? z0 : z1);
            part[j] = fmaf(e[j] * c[2 * j + 1], e[j], part[j]);
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
    for (int i = 0; i < U; ++i) ybuf[i] = Yt[(size_t)min(1 + i, T - 1) * ldY + v];
    for (int t0 = 0; t0 < nfull; t0 += U) {
        float ycur[U];
#pragma unroll
        for (int i = 0; i < U; ++i) ycur[i] = ybuf[i];
#pragma unroll
        for (int i = 0; i < U; ++i) ybuf[i] = Yt[(size_t)min(t0 + U + 1 + i, T - 1) * ldY + v];
#pragma unroll
        for (int g0 = 0; g0 < U; g0 += G) {
            float cg[G][2 * kEnergyNB], zg[G][2];
#pragma unroll
            for (int i = 0; i < G; ++i) {
                zg[i][0] = cz[2 * (t0 + g0 + i)];
                zg[i][1] = cz[2 * (t0 + g0 + i) + 1];
#pragma unroll
                for (int q = 0; q < 2 * kEnergyNB; ++q) cg[i][q] = co[(size_t)(t0 + g0 + i) * 2 * kEnergyNB + q];
            }
#pragma unroll
            for (int i = 0; i < G; ++i) step(ycur[g0 + i], zg[i][0], zg[i][1], cg[i]);
        }
#pragma unroll
        for (int j = 0; j < kEnergyNB; ++j) {
            acc[j] += part[j];
            part[j] = 0.f;
        }
    }
    // tail (fewer than U steps); the flush matches the blocked schedule
    for (int t = nfull; t < nsteps; ++t)
        step(Yt[(size_t)(t + 1) * ldY + v], cz[2 * t], cz[2 * t + 1], co + (size_t)t * 2 * kEnergyNB);
    if (v < V) {
        const int32_t* slots = P.echunk + chunk * kEnergyNB;
#pragma unroll
        for (int j = 0; j < kEnergyNB; ++j) {
            const int k = slots[j];
            if (k >= 0) E[(size_t)v * ldE + k] = (float)(acc[j] + part[j]);
        }
    }
}

// Global -> register stage of the projection tiles (A: BK x BM from the
// time-major residual, B: BK x BN from the tile-major operand).
template <int BM, int BN, int BK, int A_V4, int B_V4>
__device__ __forceinline__ void proj_gload(f32x4 (&ra)[A_V4], f32x4 (&rb)[B_V4], const float* __restrict__ Ag,
                                           const float* __restrict__ Bg, int64_t ldY, int kt, int tid) {
    const int t0 = kt * BK;
#pragma unroll
    for (int i = 0; i < A_V4; ++i) {
        const int idx = tid + i * 256;
        const int r = idx / (BM / 4), c4 = idx % (BM / 4);
        ra[i] = *reinterpret_cast<const f32x4*>(Ag + (size_t)(t0 + r) * ldY + c4 * 4);
    }
#pragma unroll
    for (int i = 0; i < B_V4; ++i) {
        const int idx0 = tid + i * 256;
        const int idx = idx0 < BK * BN / 4 ? idx0 : BK * BN / 4 - 1;  // clamped: unconditional load
        const int r = idx / (BN / 4), c4 = idx % (BN / 4);
        rb[i] = *reinterpret_cast<const f32x4*>(Bg + (size_t)(t0 + r) * BN + c4 * 4);
    }
}

template <int BM, int BN, int BK, int A_V4, int B_V4>
__device__ __forceinline__ void proj_sstore(const f32x4 (&ra)[A_V4], const f32x4 (&rb)[B_V4], float* stage,
                                            int tid) {
    float* As = stage;
    float* Bs = stage + BK * BM;
#pragma unroll
    for (int i = 0; i < A_V4; ++i) *reinterpret_cast<f32x4*>(As + (tid + i * 256) * 4) = ra[i];
    constexpr bool FULL = (BK * BN / 4) % 256 == 0;
#pragma unroll
    for (int i = 0; i < B_V4; ++i) {
        const int idx = tid + i * 256;
        if (FULL || idx < BK * BN / 4) *reinterpret_cast<f32x4*>(Bs + idx * 4) = rb[i];
    }
}

// ---------------------------------------------------------------------------
// (ii) Projection GEMM on MFMA (fp32 32x32x2) with the REML score epilogue.
// Block tile: kProjBM voxels x BN columns (npt pairs of p columns), 4 waves,
// each wave 32 voxel rows x BN columns.  LDS double-buffered BK-deep stages.
// Epilogue: P = sum_j (G_k' y)_j^2, rss = E - P, score = ld_k + (T-p) log(rss+1e-10)
// (glm/arma.py:1988-2004), written over E in the surface.
// ---------------------------------------------------------------------------
template <int BN>
__global__ void __launch_bounds__(256) k_project(PlanDev P, const float* __restrict__ Yt, int64_t ldY, int V,
                                                 float* __restrict__ S, int64_t ldS, const int32_t* __restrict__ n_live,
                                                 const PlanDev* __restrict__ plans, const int32_t* __restrict__ grp) {
    constexpr int BM = kProjBM, BK = kProjBK, NBLK = BN / 32;
    constexpr int A_ELEMS = BK * BM, B_ELEMS = BK * BN;
    constexpr int STAGE = A_ELEMS + B_ELEMS;
    constexpr int C_LD = BN + 1;
    constexpr int LDS_FLOATS = (2 * STAGE > BM * C_LD) ? 2 * STAGE : BM * C_LD;
    __shared__ float lds[LDS_FLOATS];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int w = tid >> 6;
    int vt, tile;
    if (!xcd_tile_map(blockIdx.x, gridDim.x, P.n_tiles, &vt, &tile)) return;
    if (n_live && vt * BM >= *n_live) return;  // past the live voxels
    if (plans) {  // grouped layout: the design of this voxel tile
        const int g = grp[vt];
        if (g < 0) return;
        P = plans[g];
    }
    const int v0 = vt * BM;
    const float* __restrict__ Ag = Yt + v0;
    const float* __restrict__ Bg = P.G32 + (size_t)tile * P.Tpad * BN;

    f32x16 acc[NBLK];
#pragma unroll
    for (int i = 0; i < NBLK; ++i)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[i][r] = 0.f;

    constexpr int A_V4 = A_ELEMS / 4 / 256;  // float4 per thread
    constexpr int B_V4 = (B_ELEMS / 4 + 255) / 256;
    f32x4 ra[A_V4], rb[B_V4];
    const int nk = P.Tpad / BK;
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
    // Scatter the accumulators (32x32 layout: col = lane%32,
    // row = 8*(r/4) + 4*(lane/32) + r%4) to LDS.
    float* Cs = lds;
#pragma unroll
    for (int n = 0; n < NBLK; ++n)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = w * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            const int col = n * 32 + (lane & 31);
            Cs[row * C_LD + col] = acc[n][r];
        }
    __syncthreads();
    const int p = P.p, npt = P.npt, K = P.K;
    const double tp = (double)(P.T - p);
    for (int cidx = tid; cidx < BM * npt; cidx += 256) {
        const int row = cidx / npt, q = cidx % npt;
        const int k = tile * npt + q;
        const int v = v0 + row;
        if (k >= K || v >= V) continue;
        const float* cr = Cs + row * C_LD + q * p;
        double pp = 0.0;
        for (int j = 0; j < p; ++j) {
            const double cv = (double)cr[j];
            pp = fma(cv, cv, pp);
        }
        float* sp = S + (size_t)v * ldS + k;
        const double rss = (double)(*sp) - pp;
        *sp = (float)fma(tp, log(rss + 1e-10), P.ld[k]);
    }
}

// ---------------------------------------------------------------------------
// (ii') Split 16-bit projection: C = Ah Bh + Ah Bl + Al Bh on 32x32x16 MFMA
// (f16 or bf16 halves, fp32 accumulation).  Same tiling, XCD order and score
// epilogue as k_project.  A (the residual) is split on the way into LDS and
// stored [voxel][t] so each lane's 8 time samples are one 16-byte read; B
// halves come precomputed from the FP64 tables in the [t/8][col][8] layout.
// ---------------------------------------------------------------------------
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// x = hi + lo in two 16-bit halves (round-to-nearest hi, remainder lo).
template <bool BF16>
__device__ __forceinline__ void split1(float x, uint16_t& h, uint16_t& l) {
    if (BF16) {
        const __bf16 hb = (__bf16)x;
        const __bf16 lb = (__bf16)(x - (float)hb);
        h = __builtin_bit_cast(uint16_t, hb);
        l = __builtin_bit_cast(uint16_t, lb);
    } else {
        const _Float16 hf = (_Float16)x;
        const _Float16 lf = (_Float16)(x - (float)hf);
        h = __builtin_bit_cast(uint16_t, hf);
        l = __builtin_bit_cast(uint16_t, lf);
    }
}

template <bool BF16>
__device__ __forceinline__ void split8(const float (&x)[8], uint4& hi, uint4& lo) {
    uint16_t h[8], l[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) split1<BF16>(x[e], h[e], l[e]);
    hi = make_uint4(h[0] | (uint32_t)h[1] << 16, h[2] | (uint32_t)h[3] << 16, h[4] | (uint32_t)h[5] << 16,
                    h[6] | (uint32_t)h[7] << 16);
    lo = make_uint4(l[0] | (uint32_t)l[1] << 16, l[2] | (uint32_t)l[3] << 16, l[4] | (uint32_t)l[5] << 16,
                    l[6] | (uint32_t)l[7] << 16);
}

template <bool BF16>
__device__ __forceinline__ f32x16 mfma16(uint4 a, uint4 b, f32x16 c) {
    if (BF16)
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                       0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c,
                                                      0, 0, 0);
}

template <int BM, int BN, int NTB, int A_U, int B_U>
__device__ __forceinline__ void proj16_gload(float (&ra)[A_U][8], u32x4 (&rb)[B_U], const float* __restrict__ Yt,
                                             int64_t ldY, int v0, const uint16_t* __restrict__ gh,
                                             const uint16_t* __restrict__ gl, int kt, int tid) {
    const int t0 = kt * NTB * 8;
#pragma unroll
    for (int u = 0; u < A_U; ++u) {
        const int q = tid + u * 256;
        const int vv = q % BM, tb = q / BM;
        const float* src = Yt + (size_t)(t0 + tb * 8) * ldY + v0 + vv;
#pragma unroll
        for (int e = 0; e < 8; ++e) ra[u][e] = src[(size_t)e * ldY];
    }
    constexpr int B_N = 2 * NTB * BN;
#pragma unroll
    for (int u = 0; u < B_U; ++u) {
        const int q0 = tid + u * 256;
        const int q = q0 < B_N ? q0 : B_N - 1;
        const int plane = q / (NTB * BN), rr = q % (NTB * BN);
        const int tb = rr / BN, col = rr % BN;
        const size_t o = ((size_t)(kt * NTB + tb) * BN + col) * 8;
        rb[u] = *reinterpret_cast<const u32x4*>((plane ? gl : gh) + o);
    }
}

template <int BM, int BN, int NTB, int A_U, int B_U, bool BF16>
__device__ __forceinline__ void proj16_sstore(const float (&ra)[A_U][8], const u32x4 (&rb)[B_U], uint16_t* Ah,
                                              uint16_t* Al, uint16_t* Bh, uint16_t* Bl, int RP, int tid) {
#pragma unroll
    for (int u = 0; u < A_U; ++u) {
        const int q = tid + u * 256;
        const int vv = q % BM, tb = q / BM;
        uint4 hi, lo;
        split8<BF16>(ra[u], hi, lo);
        *reinterpret_cast<uint4*>(Ah + vv * RP + tb * 8) = hi;
        *reinterpret_cast<uint4*>(Al + vv * RP + tb * 8) = lo;
    }
    constexpr int B_N = 2 * NTB * BN;
    constexpr bool FULL = B_N % 256 == 0;
#pragma unroll
    for (int u = 0; u < B_U; ++u) {
        const int q = tid + u * 256;
        if (FULL || q < B_N) {
            const int plane = q / (NTB * BN), rr = q % (NTB * BN);
            const int tb = rr / BN, col = rr % BN;
            *reinterpret_cast<u32x4*>((plane ? Bl : Bh) + col * RP + tb * 8) = rb[u];
        }
    }
}

template <int BN, bool BF16>
__global__ void __launch_bounds__(256) k_project16(PlanDev P, const float* __restrict__ Yt, int64_t ldY, int V,
                                                   float* __restrict__ S, int64_t ldS,
                                                   const int32_t* __restrict__ n_live,
                                                   const PlanDev* __restrict__ plans,
                                                   const int32_t* __restrict__ grp) {
    constexpr int BM = kProjBM, BK = kProjBK, NBLK = BN / 32, NTB = BK / 8;
    constexpr int RP = NTB * 8 + 8;             // row pitch in halves (padded 16 B)
    constexpr int A_H = BM * RP, B_H = BN * RP;  // halves per hi (or lo) plane
    constexpr int C_LD = BN + 1;
    constexpr int TILE_BYTES = 2 * (A_H + B_H) * 2;
    constexpr int LDS_BYTES = TILE_BYTES > BM * C_LD * 4 ? TILE_BYTES : BM * C_LD * 4;
    __shared__ __attribute__((aligned(16))) unsigned char lds_raw[LDS_BYTES];
    uint16_t* Ah = reinterpret_cast<uint16_t*>(lds_raw);
    uint16_t* Al = Ah + A_H;
    uint16_t* Bh = Al + A_H;
    uint16_t* Bl = Bh + B_H;

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    int vt, tile;
    if (!xcd_tile_map(blockIdx.x, gridDim.x, P.n_tiles, &vt, &tile)) return;
    if (n_live && vt * BM >= *n_live) return;  // past the live voxels
    if (plans) {  // grouped layout: the design of this voxel tile
        const int g = grp[vt];
        if (g < 0) return;
        P = plans[g];
    }
    const int v0 = vt * BM;
    const uint16_t* __restrict__ gh = (BF16 ? P.Gh_bf16 : P.Gh_f16) + (size_t)tile * P.Tpad * BN;
    const uint16_t* __restrict__ gl = (BF16 ? P.Gl_bf16 : P.Gl_f16) + (size_t)tile * P.Tpad * BN;

    // per stage: A units (voxel, t-block) = BM * NTB = 512 -> 2 per thread;
    // B units (plane, t-block, col) = 2 * NTB * BN -> B_U per thread (clamped, unconditional).
    constexpr int A_U = BM * NTB / 256;
    constexpr int B_N = 2 * NTB * BN;
    constexpr int B_U = (B_N + 255) / 256;
    float ra[A_U][8];
    u32x4 rb[B_U];
    f32x16 acc[NBLK];
#pragma unroll
    for (int i = 0; i < NBLK; ++i)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[i][r] = 0.f;
    const int nk = P.Tpad / BK;
    proj16_gload<BM, BN, NTB, A_U, B_U>(ra, rb, Yt, ldY, v0, gh, gl, 0, tid);
    for (int kt = 0; kt < nk; ++kt) {
        proj16_sstore<BM, BN, NTB, A_U, B_U, BF16>(ra, rb, Ah, Al, Bh, Bl, RP, tid);
        __syncthreads();
        if (kt + 1 < nk) proj16_gload<BM, BN, NTB, A_U, B_U>(ra, rb, Yt, ldY, v0, gh, gl, kt + 1, tid);
        const int row = w * 32 + (lane & 31);
#pragma unroll
        for (int ks = 0; ks < BK / 16; ++ks) {
            const int tb = 2 * ks + (lane >> 5);
            const uint4 ah = *reinterpret_cast<const uint4*>(Ah + row * RP + tb * 8);
            const uint4 al = *reinterpret_cast<const uint4*>(Al + row * RP + tb * 8);
#pragma unroll
            for (int n = 0; n < NBLK; ++n) {
                const int col = n * 32 + (lane & 31);
                const uint4 bh = *reinterpret_cast<const uint4*>(Bh + col * RP + tb * 8);
                const uint4 bl = *reinterpret_cast<const uint4*>(Bl + col * RP + tb * 8);
                acc[n] = mfma16<BF16>(al, bh, acc[n]);
                acc[n] = mfma16<BF16>(ah, bl, acc[n]);
                acc[n] = mfma16<BF16>(ah, bh, acc[n]);
            }
        }
        __syncthreads();
    }
    float* Cs = reinterpret_cast<float*>(lds_raw);
#pragma unroll
    for (int n = 0; n < NBLK; ++n)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = w * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            const int col = n * 32 + (lane & 31);
            Cs[row * C_LD + col] = acc[n][r];
        }
    __syncthreads();
    const int p = P.p, npt = P.npt, K = P.K;
    const double tp = (double)(P.T - p);
    for (int cidx = tid; cidx < BM * npt; cidx += 256) {
        const int row = cidx / npt, q = cidx % npt;
        const int k = tile * npt + q;
        const int v = v0 + row;
        if (k >= K || v >= V) continue;
        const float* cr = Cs + row * C_LD + q * p;
        double pp = 0.0;
        for (int j = 0; j < p; ++j) {
            const double cv = (double)cr[j];
            pp = fma(cv, cv, pp);
        }
        if (!BF16) pp *= (double)P.g16_unscale[k];
        float* sp = S + (size_t)v * ldS + k;
        const double rss = (double)(*sp) - pp;
        *sp = (float)fma(tp, log(rss + 1e-10), P.ld[k]);
    }
}

// ---------------------------------------------------------------------------
// (iii) Selection over the (V, K) surface.
//   mode 0: global argmin, strict '<' in pair order (glm/arma.py:2006-2010)
//   mode 2: subset descent (glm/arma.py:2021-2221), see below
//   mode 1: hierarchical descent (glm/arma.py:2224-2322): start (0,0); top
//           level = coarse lattice; lower levels = clamped 3x3 at +/-step;
//           candidates b-outer / a-inner, first minimum; a NaN candidate
//           makes the level's minimum NaN (torch.min), which never improves.
// ---------------------------------------------------------------------------
template <typename ST>
__device__ inline ST lat_val(const PlanDev& P, const ST* row, int ai, int bi) {
    const int k = P.lut[ai * P.n_b + bi];
    return k >= 0 ? row[k] : (ST)INFINITY;
}

// (iii) mode 0, one wave per voxel: lane l takes pairs l, l + 64, ... with a
// strict '<' (its first minimum), then the wave keeps the smallest score and,
// among equal scores, the smallest pair index: the first minimum over all K
// pairs, as the sequential scan (glm/arma.py:2008-2018).  Pairs whose score
// is never below +inf (NaN, inf) are never chosen; a voxel with none takes
// the (0, 0) pair.  n_live (nullable, device) bounds the voxels processed.
template <typename ST>
__global__ void __launch_bounds__(256) k_select_first(PlanDev P, const ST* __restrict__ S, int64_t ldS, int V,
                                                      const int32_t* __restrict__ n_live, int32_t* __restrict__ best,
                                                      ST* __restrict__ best_score) {
    const int lane = threadIdx.x & 63;
    const int v = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int Vl = n_live ? min(*n_live, V) : V;
    if (v >= Vl) return;
    const ST* row = S + (size_t)v * ldS;
    ST bs = (ST)INFINITY;
    int bk = -1;
    for (int k = lane; k < P.K; k += 64) {
        const ST s = row[k];
        if (s < bs) {
            bs = s;
            bk = k;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const ST os = __shfl_xor(bs, o);
        const int ok = __shfl_xor(bk, o);
        if (ok >= 0 && (os < bs || (os == bs && (bk < 0 || ok < bk)))) {
            bs = os;
            bk = ok;
        }
    }
    if (lane == 0) {
        best[v] = bk >= 0 ? bk : P.lut[P.zero_a * P.n_b + P.zero_b];
        best_score[v] = bs;
    }
}

template <typename ST>
__global__ void __launch_bounds__(256) k_select(PlanDev P, const ST* __restrict__ S, int64_t ldS, int V, int mode,
                                                int32_t* __restrict__ best, ST* __restrict__ best_score) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const ST* row = S + (size_t)v * ldS;
    const int k00 = P.lut[P.zero_a * P.n_b + P.zero_b];
    if (mode == 0) {
        ST bs = (ST)INFINITY;
        int bk = -1;
        for (int k = 0; k < P.K; ++k) {
            const ST s = row[k];
            if (s < bs) {
                bs = s;
                bk = k;
            }
        }
        best[v] = bk >= 0 ? bk : k00;
        best_score[v] = bs;
        return;
    }
    if (mode == 2) {
        // subset descent (glm/arma.py:2021-2221): per level the lattice points
        // that are multiples of step inside the +/-step window around the best
        // point at the level's start (whole lattice at the top level), b-outer
        // / a-inner, each compared with a sequential strict '<' (the best point
        // moves during the sweep); invalid pairs are skipped.
        int ba = P.zero_a, bb = P.zero_b;
        ST bst = (ST)INFINITY;
        {
            const ST s0 = lat_val(P, row, ba, bb);
            if (s0 < bst) bst = s0;
        }
        for (int lev = P.ltop; lev >= 0; --lev) {
            const int step = 1 << lev;
            int alo = 0, ahi = P.n_a - 1, blo = 0, bhi = P.n_b - 1;
            if (lev != P.ltop) {
                alo = max(ba - step, 0);
                ahi = min(ba + step, P.n_a - 1);
                blo = max(bb - step, 0);
                bhi = min(bb + step, P.n_b - 1);
            }
            for (int y = (blo + step - 1) / step * step; y <= bhi; y += step)
                for (int x = (alo + step - 1) / step * step; x <= ahi; x += step) {
                    const int k = P.lut[x * P.n_b + y];
                    if (k < 0) continue;
                    const ST sv = row[k];
                    if (sv < bst) {
                        bst = sv;
                        ba = x;
                        bb = y;
                    }
                }
        }
        best[v] = P.lut[ba * P.n_b + bb];
        best_score[v] = bst;
        return;
    }
    int ba = P.zero_a, bb = P.zero_b;
    ST cur = lat_val(P, row, ba, bb);
    for (int lev = P.ltop; lev >= 0; --lev) {
        const int step = 1 << lev;
        ST mn = 0;
        int sa = 0, sb = 0;
        bool first = true, nan = false;
        if (lev == P.ltop) {
            for (int y = 0; y < P.n_b; y += step)
                for (int x = 0; x < P.n_a; x += step) {
                    const ST s = lat_val(P, row, x, y);
                    if (s != s) nan = true;
                    if (first || s < mn) {
                        mn = s;
                        sa = x;
                        sb = y;
                        first = false;
                    }
                }
        } else {
            for (int oy = -1; oy <= 1; ++oy) {
                const int y = min(max(bb + oy * step, 0), P.n_b - 1);
                for (int ox = -1; ox <= 1; ++ox) {
                    const int x = min(max(ba + ox * step, 0), P.n_a - 1);
                    const ST s = lat_val(P, row, x, y);
                    if (s != s) nan = true;
                    if (first || s < mn) {
                        mn = s;
                        sa = x;
                        sb = y;
                        first = false;
                    }
                }
            }
        }
        if (!nan && mn < cur) {
            cur = mn;
            ba = sa;
            bb = sb;
        }
    }
    best[v] = P.lut[ba * P.n_b + bb];
    best_score[v] = cur;
}

// ---------------------------------------------------------------------------
// (iv) FP64 GLS refit (glm/arma.py:4784-5601, X'X branch):
//   beta = (X'R^-1X)^-1 X'R^-1 y = Lx^-T (G' y);  whitened residual
//   r = L^-1 y - X_w beta;  sigma2 = sum r^2 / dof;  t = beta / (sqrt(sigma2 diag A^-1) + 1e-10);
//   F = beta_t' Fm beta_t / (sigma2 n_task + 1e-10);  R^2 = CoD(y, X beta).
// y is the std-normalised series (scale = 1 for the FP64 variant); betas and
// sigma2 are unscaled on output (glm/arma.py:5827-5862).
// One wave per 64 consecutive voxels of the pair-sorted order; their rows are
// staged through LDS in 64-sample chunks with coalesced loads.  When the 64
// voxels share one pair (the usual case after the sort) the pair's tables are
// read through wave-uniform (scalar) loads.
// ---------------------------------------------------------------------------
constexpr int kGlsChunk = kStageChunk;

template <int PM, bool UNIFORM>
__device__ inline void gls_body(const PlanDev& P, const float* __restrict__ Y, int64_t ldY, int v, int k, double s,
                                double* tab, float* __restrict__ betas, float* __restrict__ tstats,
                                float* __restrict__ fstats, float* __restrict__ sigma2, float* __restrict__ r2,
                                double* __restrict__ resid_w, int64_t ldR, float* __restrict__ resid,
                                float* __restrict__ pred_out) {
    const int lane = threadIdx.x & 63;
    const int T = P.T, p = P.p;
    const bool active = v >= 0;
    const int kk = UNIFORM ? __builtin_amdgcn_readfirstlane(k) : (active ? k : 0);
    const double inv = 1.0 / s;
    const double* __restrict__ G = P.G64 + (size_t)kk * T * p;
    const double* __restrict__ Xw = P.Xw64 + (size_t)kk * T * p;
    const double* __restrict__ cf = P.coef64 + (size_t)kk * T * 3;
    const double* __restrict__ L = P.Lx + (size_t)kk * p * p;
    const float* __restrict__ row = Y + (size_t)(active ? v : 0) * ldY;
    const bool vec4 = (ldY & 3) == 0;
    float yv[kGlsChunk];
    double h[PM], be[PM];
#pragma unroll
    for (int j = 0; j < PM; ++j) h[j] = 0.0;
    double sy = 0.0;
    for (int t0 = 0; t0 < T; t0 += kGlsChunk) {
        load_chunk(row, T, t0, active, vec4, yv);
        __syncthreads();
        if (UNIFORM) stage_table<PM>(G, T, p, t0, tab);
        __syncthreads();
        const int n = min(kGlsChunk, T - t0);
#pragma unroll
        for (int i = 0; i < kGlsChunk; ++i) {
            if (i >= n) break;
            const double yt = (double)yv[i] * inv;
            sy += yt;
            const double* gr = UNIFORM ? tab + i * p : G + (size_t)(t0 + i) * p;
#pragma unroll
            for (int j = 0; j < PM; ++j)
                if (j < p) h[j] = fma(gr[j], yt, h[j]);
        }
    }
    // beta = Lx^-T h  (back substitution on the upper factor Lx')
#pragma unroll
    for (int jj = PM - 1; jj >= 0; --jj) {
        if (jj >= p) {
            be[jj] = 0.0;
            continue;
        }
        double acc = h[jj];
#pragma unroll
        for (int m = jj + 1; m < PM; ++m)
            if (m < p) acc -= L[m * p + jj] * be[m];
        const double ljj = L[jj * p + jj];
        be[jj] = ljj != 0.0 ? acc / ljj : 0.0;  // zero pivot: basic solution
    }
    const double mean = sy / T;
    double ssw = 0.0, ssres = 0.0, sstot = 0.0;
    double yp = 0.0, e = 0.0;
    double* tx = tab + kGlsChunk * PM;       // X rows
    double* tc = tab + 2 * kGlsChunk * PM;   // coefficient rows (3 per t)
    for (int t0 = 0; t0 < T; t0 += kGlsChunk) {
        load_chunk(row, T, t0, active, vec4, yv);
        __syncthreads();
        stage_table<PM>(P.X, T, p, t0, tx);
        if (UNIFORM) {
            stage_table<PM>(Xw, T, p, t0, tab);
            stage_table<3>(cf, T, 3, t0, tc);
        }
        __syncthreads();
        const int n = min(kGlsChunk, T - t0);
#pragma unroll
        for (int i = 0; i < kGlsChunk; ++i) {
            if (i >= n) break;
            const int t = t0 + i;
            const double yt = (double)yv[i] * inv;
            // coefficients of step t-1 (previous chunk's last row when i == 0)
            const double* cprev = (UNIFORM && i > 0) ? tc + 3 * (i - 1) : cf + 3 * (t - 1);
            const double* ccur = UNIFORM ? tc + 3 * i : cf + 3 * t;
            if (t == 0 || P.gap[t - 1] <= 0)
                e = yt;
            else
                e = (yt - cprev[1] * yp) + cprev[2] * e;
            yp = yt;
            double r = e * ccur[0];
            const double* xw = UNIFORM ? tab + i * p : Xw + (size_t)t * p;
            const double* xr = tx + i * p;
            double pred = 0.0;
#pragma unroll
            for (int j = 0; j < PM; ++j)
                if (j < p) {
                    r -= xw[j] * be[j];
                    pred += xr[j] * be[j];
                }
            ssw += r * r;
            const double dr = yt - pred, dm = yt - mean;
            ssres += dr * dr;
            sstot += dm * dm;
            if (active) {
                if (resid_w) resid_w[(size_t)v * ldR + t] = r * s;
                if (resid) resid[(size_t)v * ldR + t] = (float)(dr * s);
                if (pred_out) pred_out[(size_t)v * ldR + t] = (float)(pred * s);
            }
        }
    }
    if (!active) return;
    const int dof = max(1, T - p);
    const double sig = ssw / dof;
    const double* Ai = P.Ainv + (size_t)kk * p * p;
#pragma unroll
    for (int j = 0; j < PM; ++j)
        if (j < p) {
            const double se = sqrt(sig * Ai[j * p + j]);
            tstats[(size_t)v * p + j] = (float)(be[j] / (se + 1e-10));
            betas[(size_t)v * p + j] = (float)(be[j] * s);
        }
    const int nt = P.n_task;
    const double* F = P.Fm + (size_t)kk * p * p;
    double quad = 0.0;
#pragma unroll
    for (int a = 0; a < PM; ++a) {
        if (a >= p) continue;
        double ra = 0.0;
#pragma unroll
        for (int b2 = 0; b2 < PM; ++b2)
            if (b2 < p) ra += F[a * p + b2] * be[b2];
        quad += be[a] * ra;
    }
    fstats[v] = (float)(quad / (sig * nt + 1e-10));
    sigma2[v] = (float)(sig * s * s);
    r2[v] = (float)(sstot <= 1e-10 ? 0.0 : fmin(1.0 - ssres / sstot, 1.0));
}

template <int PM>
__global__ void __launch_bounds__(64) k_gls(PlanDev P, const float* __restrict__ Y, int64_t ldY,
                                            const float* __restrict__ scale, const int32_t* __restrict__ perm,
                                            const int32_t* __restrict__ best, int V, float* __restrict__ betas,
                                            float* __restrict__ tstats, float* __restrict__ fstats,
                                            float* __restrict__ sigma2, float* __restrict__ r2,
                                            double* __restrict__ resid_w, int64_t ldR, float* __restrict__ resid,
                                            float* __restrict__ pred_out, const PlanDev* __restrict__ plans,
                                            const int32_t* __restrict__ vgroup) {
    __shared__ double tab[kGlsChunk * (2 * PM + 3)];
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 64 + lane;
    int v = i < V ? (perm ? perm[i] : i) : -1;
    if (plans) {  // grouped: the 64 slots of a block hold one (design, pair) group
        const int v0 = __builtin_amdgcn_readfirstlane(perm[blockIdx.x * 64]);
        if (v0 < 0) return;
        P = plans[vgroup[v0]];
    }
    int k = v >= 0 ? best[v] : -1;
    if (v >= 0 && (k < 0 || k >= P.K)) {  // no valid pair: NaN statistics
        for (int j = 0; j < P.p; ++j) {
            betas[(size_t)v * P.p + j] = NAN;
            tstats[(size_t)v * P.p + j] = NAN;
        }
        fstats[v] = sigma2[v] = r2[v] = NAN;
        v = -1;
    }
    const double s = (v >= 0 && scale) ? (double)scale[v] : 1.0;
    // all active lanes on one pair?
    const int k0 = __builtin_amdgcn_readfirstlane(v >= 0 ? k : -1);
    const bool same = (v < 0) || (k == k0);
    const bool uniform = __all(same) && k0 >= 0;
    if (uniform)
        gls_body<PM, true>(P, Y, ldY, v, k0, s, tab, betas, tstats, fstats, sigma2, r2, resid_w, ldR, resid,
                           pred_out);
    else
        gls_body<PM, false>(P, Y, ldY, v, k, s, tab, betas, tstats, fstats, sigma2, r2, resid_w, ldR, resid,
                            pred_out);
}

// ---------------------------------------------------------------------------
// Ljung-Box over whitened residuals (glm/arma.py:6059-6141), one wave per
// voxel.  Lag pairs are pre-binned by time lag on the host: pairs[bin_off[d] ..
// bin_off[d+1]) hold index pairs (i, j) whose tau lag is d; wbin[d] = 1/n_d for
// bins with n_d > 1, else 0.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_ljung_box(const double* __restrict__ rw, int64_t ldR, int V, int n,
                                                   int nbins, const int32_t* __restrict__ bin_off,
                                                   const int32_t* __restrict__ pairs,
                                                   const double* __restrict__ wbin, float* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int v = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (v >= V) return;
    const double* r = rw + (size_t)v * ldR;
    double s0 = 0.0;
    for (int t = lane; t < n; t += 64) s0 += r[t] * r[t];
    for (int o = 32; o > 0; o >>= 1) s0 += __shfl_xor(s0, o);
    const bool ok = s0 >= 1e-10;
    const double sum0 = ok ? s0 : 1.0;
    double g = 0.0;
    for (int d = 1; d <= nbins; ++d) {
        const double wd = wbin[d];
        if (wd == 0.0) continue;
        double sk = 0.0;
        for (int q = bin_off[d] + lane; q < bin_off[d + 1]; q += 64) sk += r[pairs[2 * q]] * r[pairs[2 * q + 1]];
        for (int o = 32; o > 0; o >>= 1) sk += __shfl_xor(sk, o);
        const double ck = sk / sum0;
        g += ck * ck * wd;
    }
    if (lane == 0) {
        g *= (double)n * (n + 2.0);
        out[v] = ok ? (float)fmin(g, 1.0e10) : 0.f;
    }
}

__global__ void k_fill_i32(int32_t* __restrict__ p, int64_t n, int32_t val) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = val;
}

// Zero rows [r0, r1) of a row-major float buffer with row stride ld.
__global__ void k_zero_rows(float* __restrict__ p, int64_t ld, int r0, int r1, int64_t ncols) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = (int64_t)(r1 - r0) * ncols;
    if (i < n) p[(int64_t)(r0 + i / ncols) * ld + i % ncols] = 0.f;
}

// Compaction of flagged voxels (order within the list is not significant).
// n_live (nullable, device) bounds n.
__global__ void k_scatter_best(const int32_t* __restrict__ idx, int n, const int32_t* __restrict__ n_live,
                               const int32_t* __restrict__ best_sub, const float* __restrict__ score_sub,
                               int32_t* __restrict__ best, float* __restrict__ score) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (n_live ? min(*n_live, n) : n)) return;
    best[idx[i]] = best_sub[i];
    score[idx[i]] = score_sub[i];
}

}  // namespace remlk
