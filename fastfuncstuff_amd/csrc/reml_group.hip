// Helper kernels of the grouped engine (one fit over voxel groups that each
// have their own design, glm/arma.py:2836-3183): the voxels of every group
// sit in a padded column layout (each group padded to 256 columns), so every
// 64-column prep block and every projection voxel tile belongs to one group.
// Included by reml_api.hip.
#include "reml_kernels.h"

namespace remlk {

// Padded column c -> voxel rows[c] (-1: padding): the selected pair and its
// score in voxel order, and the GLS grouping key design * K + pair (-1 for an
// invalid pair, which the GLS reports as NaN statistics).
__global__ void k_group_scatter(const int32_t* __restrict__ rows, int ncols, const int32_t* __restrict__ best_col,
                                const float* __restrict__ score_col, const int32_t* __restrict__ vgroup, int K,
                                int32_t* __restrict__ best, float* __restrict__ score, int32_t* __restrict__ key) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncols) return;
    const int v = rows[c];
    if (v < 0) return;
    const int b = best_col[c];
    best[v] = b;
    score[v] = score_col[c];
    key[v] = (b >= 0 && b < K) ? vgroup[v] * K + b : -1;
}

// Rows of the padded surface to voxel order: out[v * ldO + k] = S[c * ldS + k]
// for v = rows[c] (one block per column).
__global__ void k_group_gather_rows(const int32_t* __restrict__ rows, int ncols, const float* __restrict__ S,
                                    int64_t ldS, int K, float* __restrict__ out, int64_t ldO) {
    const int c = blockIdx.x;
    if (c >= ncols) return;
    const int v = rows[c];
    if (v < 0) return;
    for (int k = threadIdx.x; k < K; k += blockDim.x) out[(size_t)v * ldO + k] = S[(size_t)c * ldS + k];
}

}  // namespace remlk
