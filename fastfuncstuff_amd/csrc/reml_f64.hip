// FP64 REML surface: the reference's float64 search (use_double,
// glm/arma.py:3869: no std normalisation, no residualisation, float64 GEMMs
// and sums), scored per (voxel, pair) as
//   S[v, k] = logdet R_k + logdet X'R_k^-1X + (T - p) log(||L_k^-1 y||^2 - ||G_k' y||^2 + 1e-10)
// (glm/arma.py:1986-2004) with ||L_k^-1 y||^2 from the innovations recurrence
// and G_k' y = Q_k' L_k^-1 y from dot products with the pair's FP64 table,
// all in float64.  One wave per (64-voxel tile, pair), pairs fastest so that
// the K blocks of a tile find its rows in L2; the pair's table rows and
// recurrence coefficients are staged through LDS in 32-sample chunks.
// Included by reml_api.hip (uses reml_kernels.hip's staging helpers).
#include "reml_kernels.h"

namespace remlk {

template <int PM, typename TY>
__global__ void __launch_bounds__(64) k_surface_f64(PlanDev P, const TY* __restrict__ Y, int64_t ldY,
                                                    const float* __restrict__ scale, int V, double* __restrict__ S,
                                                    int64_t ldS) {
    __shared__ double tab[kStageChunk * PM];
    __shared__ double tc[kStageChunk * 3];
    const int K = P.K, T = P.T, p = P.p;
    const int vt = blockIdx.x / K, k = blockIdx.x - vt * K;
    const int lane = threadIdx.x;
    const int v = vt * 64 + lane;
    const bool active = v < V;
    const double inv = (active && scale) ? 1.0 / (double)scale[v] : 1.0;
    const double* __restrict__ G = P.G64 + (size_t)k * T * p;
    const double* __restrict__ cf = P.coef64 + (size_t)k * T * 3;  // sqrt(1/v_t), c_t, alpha_t
    const TY* __restrict__ row = Y + (size_t)(active ? v : 0) * ldY;
    double h[PM];
#pragma unroll
    for (int j = 0; j < PM; ++j) h[j] = 0.0;
    double E = 0.0, e = 0.0, yp = 0.0, c_prev = 0.0, a_prev = 0.0;
    for (int t0 = 0; t0 < T; t0 += kStageChunk) {
        __syncthreads();
        stage_table<PM>(G, T, p, t0, tab);
        stage_table<3>(cf, T, 3, t0, tc);
        __syncthreads();
        const int n = min(kStageChunk, T - t0);
        for (int i = 0; i < n; ++i) {
            const int t = t0 + i;
            const double yt = active ? (double)row[t] * inv : 0.0;
            // e_t = y_t at a run start, else (y_t - c_{t-1} y_{t-1}) + alpha_{t-1} e_{t-1}
            e = (t == 0 || P.gap[t - 1] <= 0) ? yt : (yt - c_prev * yp) + a_prev * e;
            const double w = e * tc[3 * i];
            E = fma(w, w, E);
            yp = yt;
            c_prev = tc[3 * i + 1];
            a_prev = tc[3 * i + 2];
            const double* gr = tab + i * p;
#pragma unroll
            for (int j = 0; j < PM; ++j)
                if (j < p) h[j] = fma(gr[j], yt, h[j]);
        }
    }
    if (!active) return;
    double q = 0.0;
#pragma unroll
    for (int j = 0; j < PM; ++j)
        if (j < p) q = fma(h[j], h[j], q);
    S[(size_t)v * ldS + k] = P.ld[k] + (double)(T - p) * log(E - q + 1e-10);
}

// The same score for designs wider than 96 columns: G_k' y in chunks of 32
// columns (each chunk a pass over t with its columns of the pair's rows staged
// through LDS), the squares added into one sum with j ascending, so every
// term has the arithmetic of the one-pass kernel; the energy is formed in the
// first pass.
template <typename TY>
__global__ void __launch_bounds__(64) k_surface_f64_seg(PlanDev P, const TY* __restrict__ Y, int64_t ldY,
                                                        const float* __restrict__ scale, int V,
                                                        double* __restrict__ S, int64_t ldS) {
    constexpr int PM = 32;
    __shared__ double tab[kStageChunk * PM];
    __shared__ double tc[kStageChunk * 3];
    const int K = P.K, T = P.T, p = P.p;
    const int vt = blockIdx.x / K, k = blockIdx.x - vt * K;
    const int lane = threadIdx.x;
    const int v = vt * 64 + lane;
    const bool active = v < V;
    const double inv = (active && scale) ? 1.0 / (double)scale[v] : 1.0;
    const double* __restrict__ G = P.G64 + (size_t)k * T * p;
    const double* __restrict__ cf = P.coef64 + (size_t)k * T * 3;  // sqrt(1/v_t), c_t, alpha_t
    const TY* __restrict__ row = Y + (size_t)(active ? v : 0) * ldY;
    double E = 0.0, q = 0.0;
    for (int j0 = 0; j0 < p; j0 += PM) {
        const int nj = min(PM, p - j0);
        double h[PM];
#pragma unroll
        for (int j = 0; j < PM; ++j) h[j] = 0.0;
        double e = 0.0, yp = 0.0, c_prev = 0.0, a_prev = 0.0;
        for (int t0 = 0; t0 < T; t0 += kStageChunk) {
            const int n = min(kStageChunk, T - t0);
            __syncthreads();
            for (int q2 = lane; q2 < n * PM; q2 += 64) {
                const int i = q2 / PM, jj = q2 % PM;
                tab[q2] = jj < nj ? G[(size_t)(t0 + i) * p + j0 + jj] : 0.0;
            }
            if (j0 == 0) stage_table<3>(cf, T, 3, t0, tc);
            __syncthreads();
            for (int i = 0; i < n; ++i) {
                const int t = t0 + i;
                const double yt = active ? (double)row[t] * inv : 0.0;
                if (j0 == 0) {
                    e = (t == 0 || P.gap[t - 1] <= 0) ? yt : (yt - c_prev * yp) + a_prev * e;
                    const double w = e * tc[3 * i];
                    E = fma(w, w, E);
                    yp = yt;
                    c_prev = tc[3 * i + 1];
                    a_prev = tc[3 * i + 2];
                }
                const double* gr = tab + i * PM;
#pragma unroll
                for (int j = 0; j < PM; ++j) h[j] = fma(gr[j], yt, h[j]);
            }
        }
#pragma unroll
        for (int j = 0; j < PM; ++j)
            if (j < nj) q = fma(h[j], h[j], q);
    }
    if (!active) return;
    S[(size_t)v * ldS + k] = P.ld[k] + (double)(T - p) * log(E - q + 1e-10);
}

}  // namespace remlk
