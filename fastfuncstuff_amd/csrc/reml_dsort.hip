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
