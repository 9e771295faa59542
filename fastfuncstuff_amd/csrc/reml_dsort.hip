// Whitening of per-voxel columns by each voxel's pair (the -dsort pass).
// Included by reml_api.hip (one translation unit).
#pragma once

#include "reml_kernels.h"

namespace remlk {

// out[v][j][t] = (L_k^-1 x)_t for column x = cols[v][j][0..T) and k = pair[v]:
// the innovations recurrence of the GLS tables, e_t = alpha_{t-1} e_{t-1} +
// x_t - c_{t-1} x_{t-1} (c = alpha = 0 at run starts), w_t = e_t / sqrt(v_t),
// in fp64.  One lane per (voxel, column); columns of a voxel are adjacent
// lanes, so the coefficient reads of a step are shared.  A voxel whose pair
// is outside [0, K) gets zero columns.
__global__ void k_whiten_cols(PlanDev P, const int32_t* __restrict__ pair, int V, int m,
                              const double* __restrict__ cols, int64_t ld, double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)V * m) return;
    const int v = (int)(i / m);
    const int k = pair[v];
    const double* __restrict__ x = cols + i * ld;
    double* __restrict__ w = out + i * ld;
    const int T = P.T;
    if (k < 0 || k >= P.K) {
        for (int t = 0; t < T; ++t) w[t] = 0.0;
        return;
    }
    const double* __restrict__ cR = P.coefR + (size_t)k * P.Tpad * 4;
    double e = 0.0, xp = 0.0;
    for (int t = 0; t < T; ++t) {
        const double xt = x[t];
        const double gs = cR[4 * t], c = cR[4 * t + 1], al = cR[4 * t + 2];
        e = fma(al, e, fma(-c, xp, xt));
        xp = xt;
        w[t] = e * gs;
    }
}

}  // namespace remlk

namespace remlk {

// -dsort per-voxel GLS (glm/arma.py:3248-3539): the final fit redone on the
// extended design E = [X | D_v] (p + q columns, D_v the voxel's dsort series)
// at the voxel's pair.  One 64-lane block per voxel:
//   pass 1  the whitened columns of [E | y] in chunks of kDsortChunk samples
//           (X_w rows from the plan, D_v and y whitened by lanes 0..q with the
//           plan's recurrence) into the lower triangle of their Gram matrix
//           (LDS, one entry per lane, fixed order over t);
//   solve   A = E_w'E_w + 1e-8 I: Cholesky (block helpers, FP64, workspace
//           in global memory), Ai = A^-1, beta = Ai E_w'y_w, whitened RSS =
//           y_w'y_w - 2 beta'c + beta'E_w'E_w beta, sigma^2 = RSS / dof,
//           t = beta / (sqrt(sigma^2 diag Ai) + 1e-10), F over the columns
//           fci: quad = b' inv(Ai[fci, fci] + 1e-8 I) b / (sigma^2 n_f + 1e-30);
//   pass 2  pred = E beta in original space, R^2 = CoD(y, pred) (0 when
//           sum (y - mean)^2 <= 1e-10), optional pred / resid / whitened
//           residual (the recurrence on y - pred, since whitening is linear).
// info[v] = 1 when the Cholesky fails (the reference then takes pinv(A)).
constexpr int kDsortChunk = 16;
constexpr int kDsortMaxM = 128;  // p + q + 1

struct DsortArgs {
    const int32_t* pair;
    int V, q;
    const double* D;  // [V][q][T]
    const double* Y;  // [V][T], normalised
    int n_f;
    int32_t fci[kDsortMaxM];  // F columns (by value: no device copy)
    double *beta, *tstat, *sigma2, *fstat, *r2;  // [V][pe], [V][pe], [V] x 3
    double* Ai_out;                              // [V][pe][pe] or null
    double *rw, *resid, *pred;                   // [V][T] or null
    int32_t* info;                               // [V]
    double* work;                                // [V][dsort_work_doubles]
    int64_t ws;                                  // doubles of work per voxel
};

__host__ __device__ inline int64_t dsort_work_doubles(int pe, int n_f, int T) {
    return 3 * (int64_t)pe * pe + 3 * (int64_t)n_f * n_f + 2 * (int64_t)pe + T;
}

__global__ void __launch_bounds__(64) k_dsort_gls(PlanDev P, DsortArgs A) {
    extern __shared__ __attribute__((aligned(16))) double dsm[];
    const int v = blockIdx.x;
    if (v >= A.V) return;
    const int tid = threadIdx.x;
    const int T = P.T, p = P.p, q = A.q, pe = p + q, m = pe + 1;
    const int NE = m * (m + 1) / 2;
    double* acc = dsm;            // [NE] lower triangle, row-major packed
    double* W = dsm + NE;         // [kDsortChunk][m]
    __shared__ int sh_flag, sh_piv;
    const int k = A.pair[v];
    if (k < 0 || k >= P.K) {
        if (tid == 0) A.info[v] = 2;  // no valid pair: left to the caller (NaN statistics)
        return;
    }
    if (tid == 0) sh_flag = 0;
    for (int e = tid; e < NE; e += 64) acc[e] = 0.0;
    const double* __restrict__ Xw = P.Xw64 + (size_t)k * T * p;
    const double* __restrict__ cR = P.coefR + (size_t)k * P.Tpad * 4;
    const double* __restrict__ yv = A.Y + (size_t)v * T;
    const double* __restrict__ Dv = A.D + (size_t)v * q * T;
    double e_rec = 0.0, xprev = 0.0, sy = 0.0;
    __syncthreads();
    for (int t0 = 0; t0 < T; t0 += kDsortChunk) {
        const int n = min(kDsortChunk, T - t0);
        if (tid <= q) {
            const double* x = tid < q ? Dv + (size_t)tid * T : yv;
            for (int i = 0; i < n; ++i) {
                const int t = t0 + i;
                const double xt = x[t];
                const double gs = cR[4 * t], c = cR[4 * t + 1], al = cR[4 * t + 2];
                e_rec = fma(al, e_rec, fma(-c, xprev, xt));
                xprev = xt;
                W[i * m + p + tid] = e_rec * gs;
                if (tid == q) sy += xt;
            }
        }
        for (int idx = tid; idx < n * p; idx += 64) {
            const int i = idx / p, j = idx % p;
            W[i * m + j] = Xw[(size_t)(t0 + i) * p + j];
        }
        __syncthreads();
        for (int e = tid; e < NE; e += 64) {
            int i = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
            while (i * (i + 1) / 2 > e) --i;
            while ((i + 1) * (i + 2) / 2 <= e) ++i;
            const int j = e - i * (i + 1) / 2;
            double s = 0.0;
            for (int r = 0; r < n; ++r) s = fma(W[r * m + i], W[r * m + j], s);
            acc[e] += s;
        }
        __syncthreads();
    }
    // solve
    double* Am = A.work + (size_t)v * A.ws;  // [pe][pe]: A, then its factor
    double* Ai = Am + (size_t)pe * pe;       // [pe][pe]
    double* tmp = Ai + (size_t)pe * pe;      // [pe][pe]
    double* bt = tmp + (size_t)pe * pe;      // [pe] beta
    double* cv = bt + pe;                    // [pe] E_w'y_w
    double* At = cv + pe;                    // [n_f][n_f] x 3
    double* Ati = At + (size_t)A.n_f * A.n_f;
    double* Atw = Ati + (size_t)A.n_f * A.n_f;
    double* rr = Atw + (size_t)A.n_f * A.n_f;  // [T] raw residual
    auto G = [&](int i, int j) { return i >= j ? acc[i * (i + 1) / 2 + j] : acc[j * (j + 1) / 2 + i]; };
    for (int idx = tid; idx < pe * pe; idx += 64) {
        const int i = idx / pe, j = idx % pe;
        Am[idx] = G(i, j) + (i == j ? 1e-8 : 0.0);
    }
    for (int i = tid; i < pe; i += 64) cv[i] = G(pe, i);
    __syncthreads();
    if (block_cholesky(pe, Am, &sh_flag) != 0) {
        if (tid == 0) A.info[v] = 1;
        return;
    }
    block_cholesky_inverse(pe, Am, Ai, tmp);
    for (int i = tid; i < pe; i += 64) {
        double s = 0.0;
        for (int j = 0; j < pe; ++j) s = fma(Ai[i * pe + j], cv[j], s);
        bt[i] = s;
    }
    __syncthreads();
    // whitened RSS from the Gram matrix (no 1e-8 term)
    double part = 0.0;
    for (int i = tid; i < pe; i += 64) {
        double gb = 0.0;
        for (int j = 0; j < pe; ++j) gb = fma(G(i, j), bt[j], gb);
        part += bt[i] * (gb - 2.0 * cv[i]);
    }
    for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o);
    const double rss = G(pe, pe) + part;
    const int dof = max(1, T - pe);
    const double sig = rss / dof;
    for (int i = tid; i < pe; i += 64) {
        A.beta[(size_t)v * pe + i] = bt[i];
        A.tstat[(size_t)v * pe + i] = bt[i] / (sqrt(sig * Ai[i * pe + i]) + 1e-10);
    }
    if (A.Ai_out)
        for (int idx = tid; idx < pe * pe; idx += 64) A.Ai_out[(size_t)v * pe * pe + idx] = Ai[idx];
    // F over the columns fci
    const int nf = A.n_f;
    for (int idx = tid; idx < nf * nf; idx += 64) {
        const int i = idx / nf, j = idx % nf;
        At[idx] = Ai[A.fci[i] * pe + A.fci[j]] + (i == j ? 1e-8 : 0.0);
    }
    __syncthreads();
    double quad = NAN;
    if (block_gauss_jordan_inverse(nf, At, Ati, Atw, &sh_piv, &sh_flag, 1) == 0) {
        double qp = 0.0;
        for (int i = tid; i < nf; i += 64) {
            double s = 0.0;
            for (int j = 0; j < nf; ++j) s = fma(Ati[i * nf + j], bt[A.fci[j]], s);
            qp = fma(bt[A.fci[i]], s, qp);
        }
        for (int o = 32; o > 0; o >>= 1) qp += __shfl_xor(qp, o);
        quad = qp;
    }
    // pass 2: prediction in original space
    const double mean = __shfl(sy, q) / T;
    const double* __restrict__ X = P.X;
    double ssr = 0.0, sst = 0.0;
    for (int t = tid; t < T; t += 64) {
        double pr = 0.0;
        for (int j = 0; j < p; ++j) pr = fma(X[(size_t)t * p + j], bt[j], pr);
        for (int d = 0; d < q; ++d) pr = fma(Dv[(size_t)d * T + t], bt[p + d], pr);
        const double y = yv[t], r = y - pr;
        ssr = fma(r, r, ssr);
        sst = fma(y - mean, y - mean, sst);
        rr[t] = r;
        if (A.pred) A.pred[(size_t)v * T + t] = pr;
        if (A.resid) A.resid[(size_t)v * T + t] = r;
    }
    for (int o = 32; o > 0; o >>= 1) {
        ssr += __shfl_xor(ssr, o);
        sst += __shfl_xor(sst, o);
    }
    if (A.rw) {
        __syncthreads();
        if (tid == 0) {
            double e = 0.0, xp = 0.0;
            for (int t = 0; t < T; ++t) {
                const double xt = rr[t];
                e = fma(cR[4 * t + 2], e, fma(-cR[4 * t + 1], xp, xt));
                xp = xt;
                A.rw[(size_t)v * T + t] = e * cR[4 * t];
            }
        }
    }
    if (tid == 0) {
        A.sigma2[v] = sig;
        A.fstat[v] = quad / (sig * nf + 1e-30);
        A.r2[v] = sst <= 1e-10 ? 0.0 : fmin(1.0 - ssr / sst, 1.0);
        A.info[v] = 0;
    }
}

}  // namespace remlk
