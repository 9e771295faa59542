// REML score for a caller-supplied dense covariance (one block).  Included by
// reml_api.hip (one translation unit).
#pragma once

#include "reml_kernels.h"

namespace remlk {

// glm/arma.py:1106-1186 compute_reml_likelihood for one dense R (T x T):
//   L = chol(R); term1 = 2 sum log(diag L + 1e-10);
//   X_w = L^-1 X, Y_w = L^-1 Y (T x m);  A = X_w' X_w + 1e-6 I;
//   term2 = log det A (1e10 when A is not positive definite);
//   beta = A^-1 X_w' Y_w;  rss = sum (Y_w - X_w beta)^2 (all m columns);
//   out = term1 + term2 + (T - p) log(rss + 1e-10); 1e10 when chol(R) fails.
// work: T*T + T*(p+m) + p*p + p*m doubles.  Row-major inputs.
constexpr int kDenseThreads = 256;

__global__ void __launch_bounds__(kDenseThreads) k_dense_reml(const double* __restrict__ X,
                                                              const double* __restrict__ Y, int m,
                                                              const double* __restrict__ R, int T, int p,
                                                              double* __restrict__ work, double* __restrict__ out) {
    const int tid = threadIdx.x;
    double* L = work;                       // T x T
    double* Xw = L + (size_t)T * T;         // T x p
    double* Yw = Xw + (size_t)T * p;        // T x m
    double* A = Yw + (size_t)T * m;         // p x p
    double* C = A + (size_t)p * p;          // p x m (X_w' Y_w, then beta)
    __shared__ int fail;
    __shared__ double red[kDenseThreads];
    for (size_t i = tid; i < (size_t)T * T; i += kDenseThreads) L[i] = R[i];
    if (tid == 0) fail = 0;
    __syncthreads();
    // right-looking Cholesky (lower triangle)
    for (int j = 0; j < T; ++j) {
        if (tid == 0) {
            const double d = L[(size_t)j * T + j];
            if (!(d > 0.0))
                fail = 1;
            else
                L[(size_t)j * T + j] = sqrt(d);
        }
        __syncthreads();
        if (fail) break;
        const double djj = L[(size_t)j * T + j];
        for (int i = j + 1 + tid; i < T; i += kDenseThreads) L[(size_t)i * T + j] /= djj;
        __syncthreads();
        const int n = T - j - 1;
        for (int q = tid; q < n * n; q += kDenseThreads) {
            const int i = j + 1 + q / n, k = j + 1 + q % n;
            if (k <= i) L[(size_t)i * T + k] -= L[(size_t)i * T + j] * L[(size_t)k * T + j];
        }
        __syncthreads();
    }
    if (fail) {
        if (tid == 0) *out = 1e10;
        return;
    }
    // term1
    double s = 0.0;
    for (int j = tid; j < T; j += kDenseThreads) s += log(L[(size_t)j * T + j] + 1e-10);
    red[tid] = s;
    __syncthreads();
    for (int o = kDenseThreads / 2; o > 0; o >>= 1) {
        if (tid < o) red[tid] += red[tid + o];
        __syncthreads();
    }
    const double term1 = 2.0 * red[0];
    __syncthreads();
    // forward substitution, one right-hand side column per thread
    for (int c = tid; c < p + m; c += kDenseThreads) {
        const bool isx = c < p;
        const int cc = isx ? c : c - p;
        const int ld = isx ? p : m;
        const double* src = isx ? X : Y;
        double* dst = isx ? Xw : Yw;
        for (int i = 0; i < T; ++i) {
            double v = src[(size_t)i * ld + cc];
            for (int k = 0; k < i; ++k) v -= L[(size_t)i * T + k] * dst[(size_t)k * ld + cc];
            dst[(size_t)i * ld + cc] = v / L[(size_t)i * T + i];
        }
    }
    __syncthreads();
    for (int q = tid; q < p * p + p * m; q += kDenseThreads) {
        double v = 0.0;
        if (q < p * p) {
            const int i = q / p, j = q % p;
            for (int t = 0; t < T; ++t) v += Xw[(size_t)t * p + i] * Xw[(size_t)t * p + j];
            A[q] = v + (i == j ? 1e-6 : 0.0);
        } else {
            const int i = (q - p * p) / m, j = (q - p * p) % m;
            for (int t = 0; t < T; ++t) v += Xw[(size_t)t * p + i] * Yw[(size_t)t * m + j];
            C[(size_t)i * m + j] = v;
        }
    }
    __syncthreads();
    __shared__ double term2;
    if (tid == 0) {
        // A = LA LA' in place; beta = A^-1 C
        if (cholesky_inplace(p, A) != 0) {
            term2 = 1e10;
            fail = 1;
        } else {
            double ld = 0.0;
            for (int j = 0; j < p; ++j) ld += 2.0 * log(A[j * p + j]);
            term2 = ld;
            for (int c = 0; c < m; ++c) {
                for (int i = 0; i < p; ++i) {
                    double v = C[(size_t)i * m + c];
                    for (int k = 0; k < i; ++k) v -= A[i * p + k] * C[(size_t)k * m + c];
                    C[(size_t)i * m + c] = v / A[i * p + i];
                }
                for (int i = p - 1; i >= 0; --i) {
                    double v = C[(size_t)i * m + c];
                    for (int k = i + 1; k < p; ++k) v -= A[k * p + i] * C[(size_t)k * m + c];
                    C[(size_t)i * m + c] = v / A[i * p + i];
                }
            }
        }
    }
    __syncthreads();
    double rs = 0.0;
    if (!fail)
        for (int q = tid; q < T * m; q += kDenseThreads) {
            const int t = q / m, c = q % m;
            double r = Yw[(size_t)t * m + c];
            for (int j = 0; j < p; ++j) r -= Xw[(size_t)t * p + j] * C[(size_t)j * m + c];
            rs += r * r;
        }
    red[tid] = rs;
    __syncthreads();
    for (int o = kDenseThreads / 2; o > 0; o >>= 1) {
        if (tid < o) red[tid] += red[tid + o];
        __syncthreads();
    }
    if (tid == 0) *out = term1 + term2 + (fail ? 0.0 : (double)(T - p) * log(red[0] + 1e-10));
}

}  // namespace remlk
