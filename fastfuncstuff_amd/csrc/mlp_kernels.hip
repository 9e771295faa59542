// (a, b) proposal MLP: 10 -> 64 GELU -> 64 GELU -> 3, forward, backward and
// Adam on gfx950.  Parameters are one flat float32 buffer:
//   W1[10][64] b1[64] W2[64][64] b2[64] W3[64][3] b3[3]   (row-major [in][out])
// Loss per sample: 0.5 e^c ((a_hat-a)^2 + (b_hat-b)^2) - c, c clipped to
// [-10, 15], averaged over the batch.  Gradients: one wave per 64 samples runs
// forward and backward (a sample per lane), keeps the per-sample activations
// and deltas in LDS and forms the block's weight gradients there (h1' d2 on
// 32x32x2 f32 MFMA, the small blocks on the vector ALU, each in a fixed sample
// order); a second kernel adds the block partials in block order.  The result
// is deterministic.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <string>

#include "../../include/ffs_mlp.h"

namespace mlpk {

constexpr int NIN = 10, NH = 64, NOUT = 3;
constexpr int OW1 = 0, OB1 = OW1 + NIN * NH, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH,
              OB3 = OW3 + NH * NOUT, NPARAM = OB3 + NOUT;
constexpr float C_MIN = -10.f, C_MAX = 15.f;

__device__ inline float gelu(float x) { return 0.5f * x * (1.f + erff(x * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float x) {
    return 0.5f * (1.f + erff(x * 0.70710678118654752f)) + x * expf(-0.5f * x * x) * 0.39894228040143268f;
}

__device__ inline void load_params(const float* __restrict__ params, float* sp) {
    for (int i = threadIdx.x; i < NPARAM; i += blockDim.x) sp[i] = params[i];
    __syncthreads();
}

// Layer-1 pre-activation of unit i (recomputed where needed: 10 FMAs).
__device__ inline float z1_of(const float* sp, const float* x, int i) {
    float s = sp[OB1 + i];
#pragma unroll
    for (int k = 0; k < NIN; ++k) s = fmaf(x[k], sp[OW1 + k * NH + i], s);
    return s;
}

// z2 = b2 + gelu(z1) W2 ; optionally stores h1 = gelu(z1) to h1_out.
__device__ inline void layer2(const float* sp, const float* x, float* z2, float* h1_out) {
#pragma unroll
    for (int j = 0; j < NH; ++j) z2[j] = sp[OB2 + j];
#pragma unroll 1
    for (int i = 0; i < NH; ++i) {
        const float h = gelu(z1_of(sp, x, i));
        if (h1_out) h1_out[i] = h;
#pragma unroll
        for (int j = 0; j < NH; ++j) z2[j] = fmaf(h, sp[OW2 + i * NH + j], z2[j]);
    }
}

// Inference, one sample per lane.  Layer-1 weights are staged transposed in
// LDS (a unit's 10 inputs in one broadcast row); layer-2 and -3 weights are
// read straight from global memory at wave-uniform addresses (scalar loads,
// scalar FMA operands).  Same operation order as layer2 / z1_of.
constexpr int W1T_LD = 12;  // padded row of the transposed layer-1 weights
__global__ void __launch_bounds__(256) k_mlp_forward(const float* __restrict__ params, const float* __restrict__ X,
                                                     const int32_t* __restrict__ idx, int64_t B,
                                                     float* __restrict__ out) {
    __shared__ float w1t[NH * W1T_LD];
    for (int q = threadIdx.x; q < NH * NIN; q += blockDim.x) {
        const int k = q / NH, i = q % NH;
        w1t[i * W1T_LD + k] = params[OW1 + k * NH + i];
    }
    __syncthreads();
    const float* __restrict__ sp = params;
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const int64_t r = idx ? idx[b] : b;
    float x[NIN];
#pragma unroll
    for (int i = 0; i < NIN; ++i) x[i] = X[r * NIN + i];
    float z2[NH];
#pragma unroll
    for (int j = 0; j < NH; ++j) z2[j] = sp[OB2 + j];
#pragma unroll 2
    for (int i = 0; i < NH; ++i) {
        const float* wr = w1t + i * W1T_LD;
        float zs = sp[OB1 + i];
#pragma unroll
        for (int k = 0; k < NIN; ++k) zs = fmaf(x[k], wr[k], zs);
        const float h = gelu(zs);
#pragma unroll
        for (int j = 0; j < NH; ++j) z2[j] = fmaf(h, sp[OW2 + i * NH + j], z2[j]);
    }
    float o[NOUT];
#pragma unroll
    for (int k = 0; k < NOUT; ++k) o[k] = sp[OB3 + k];
#pragma unroll
    for (int j = 0; j < NH; ++j) {
        const float g = gelu(z2[j]);
#pragma unroll
        for (int k = 0; k < NOUT; ++k) o[k] = fmaf(g, sp[OW3 + j * NOUT + k], o[k]);
    }
#pragma unroll
    for (int k = 0; k < NOUT; ++k) out[b * NOUT + k] = o[k];
}

// ---------------------------------------------------------------------------
// Batch gradient.  k_mlp_block_grad: one block of four waves per kGB = 64
// samples.  The layers run as 64 x 64 GEMMs on 32x32x2 f32 MFMA, wave w on
// output tile (w / 2, w % 2) (32 rows x 32 columns; A lane l = A[l % 32][k],
// B lane l = B[k][l % 32] with k = 2 s + l / 32; accumulator element q of
// lane l = row 8 (q / 4) + 4 (l / 32) + q % 4, column l % 32):
//   Z1 = X W1 + b1, H1 = gelu(Z1);  Z2 = H1 W2 + b2, H2 = gelu(Z2);
//   O = H2 W3 + b3 and the loss delta d3 per sample (vector ALU);
//   D2 = (d3 W3') . gelu'(Z2);  D1 = (D2 W2') . gelu'(Z1);
//   dW2 = H1' D2, dW1 = X' D1 (MFMA), dW3 = H2' d3 and the biases (vector
//   ALU), each summed over the block's samples in a fixed order.
// partial[blk][0..NPARAM) = the block's gradient, lossp[blk] = its loss sum.
// k_mlp_grad_finalize: grad[q] = sum over blocks in block order; loss = mean.
// ---------------------------------------------------------------------------
constexpr int kGB = 64;
constexpr int kLD = NH + 1;
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ int acc_row(int q, int lane) { return 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3); }

__global__ void __launch_bounds__(256) k_mlp_block_grad(const float* __restrict__ params, const float* __restrict__ X,
                                                        const float* __restrict__ target,
                                                        const int32_t* __restrict__ idx, int64_t B,
                                                        float* __restrict__ partial, double* __restrict__ lossp) {
    __shared__ float sx[kGB][NIN + 1];
    __shared__ float sz1[kGB][kLD], sh1[kGB][kLD], sz2[kGB][kLD], sh2[kGB][kLD], sd2[kGB][kLD], sd1[kGB][kLD];
    __shared__ float sd3[kGB][NOUT];
    __shared__ double sl[kGB];
    const float* __restrict__ W1 = params + OW1;
    const float* __restrict__ W2 = params + OW2;
    const float* __restrict__ W3 = params + OW3;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int ti = w >> 1, tj = w & 1;  // this wave's output tile
    const int64_t b0 = (int64_t)blockIdx.x * kGB;
    const float invB = 1.f / (float)B;
    for (int q = tid; q < kGB * NIN; q += 256) {
        const int s = q / NIN, i = q % NIN;
        const int64_t b = b0 + s;
        const int64_t r = b < B ? (idx ? idx[b] : b) : -1;
        sx[s][i] = r >= 0 ? X[r * NIN + i] : 0.f;
    }
    __syncthreads();
    const int rl = lane & 31, kh = lane >> 5;
    f32x16 acc;
    // Z1 = X W1 + b1 (K = NIN, padded to an even count with zeros)
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = 0.f;
#pragma unroll
    for (int s = 0; s < (NIN + 1) / 2; ++s) {
        const int k = 2 * s + kh;
        const float a = k < NIN ? sx[32 * ti + rl][k] : 0.f;
        const float bb = k < NIN ? W1[k * NH + 32 * tj + rl] : 0.f;
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, bb, acc, 0, 0, 0);
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int r = 32 * ti + acc_row(q, lane), c = 32 * tj + rl;
        const float z = acc[q] + params[OB1 + c];
        sz1[r][c] = z;
        sh1[r][c] = gelu(z);
    }
    __syncthreads();
    // Z2 = H1 W2 + b2
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = 0.f;
#pragma unroll 8
    for (int s = 0; s < NH / 2; ++s) {
        const int k = 2 * s + kh;
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(sh1[32 * ti + rl][k], W2[k * NH + 32 * tj + rl], acc, 0, 0, 0);
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int r = 32 * ti + acc_row(q, lane), c = 32 * tj + rl;
        const float z = acc[q] + params[OB2 + c];
        sz2[r][c] = z;
        sh2[r][c] = gelu(z);
    }
    __syncthreads();
    // output layer, loss and d3: one sample per lane of wave 0
    if (w == 0) {
        const int s = lane;
        const int64_t b = b0 + s;
        const bool live = b < B;
        float o[NOUT];
#pragma unroll
        for (int k = 0; k < NOUT; ++k) o[k] = params[OB3 + k];
        for (int j = 0; j < NH; ++j) {
            const float h = sh2[s][j];
#pragma unroll
            for (int k = 0; k < NOUT; ++k) o[k] = fmaf(h, W3[j * NOUT + k], o[k]);
        }
        const int64_t r = live ? (idx ? idx[b] : b) : 0;
        const float c = fminf(fmaxf(o[2], C_MIN), C_MAX);
        const float ec = expf(c);
        const float da = live ? o[0] - target[r * 2] : 0.f, db = live ? o[1] - target[r * 2 + 1] : 0.f;
        const float sse = da * da + db * db;
        sd3[s][0] = live ? ec * da * invB : 0.f;
        sd3[s][1] = live ? ec * db * invB : 0.f;
        sd3[s][2] = (live && o[2] > C_MIN && o[2] < C_MAX) ? (0.5f * ec * sse - 1.f) * invB : 0.f;
        sl[s] = live ? (double)(0.5f * ec * sse - c) : 0.0;
    }
    __syncthreads();
    // D2 = (d3 W3') . gelu'(Z2), 16 elements per thread
    for (int q = tid; q < kGB * NH; q += 256) {
        const int s = q / NH, j = q % NH;
        float t = 0.f;
#pragma unroll
        for (int k = 0; k < NOUT; ++k) t = fmaf(sd3[s][k], W3[j * NOUT + k], t);
        sd2[s][j] = t * gelu_grad(sz2[s][j]);
    }
    __syncthreads();
    // D1 = (D2 W2') . gelu'(Z1)
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = 0.f;
#pragma unroll 8
    for (int s = 0; s < NH / 2; ++s) {
        const int k = 2 * s + kh;
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(sd2[32 * ti + rl][k], W2[(32 * tj + rl) * NH + k], acc, 0, 0, 0);
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int r = 32 * ti + acc_row(q, lane), c = 32 * tj + rl;
        sd1[r][c] = acc[q] * gelu_grad(sz1[r][c]);
    }
    __syncthreads();
    float* __restrict__ out = partial + (size_t)blockIdx.x * NPARAM;
    // dW2 = H1' D2 (tile rows = input units, columns = output units, K = samples)
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = 0.f;
#pragma unroll 8
    for (int s = 0; s < kGB / 2; ++s) {
        const int k = 2 * s + kh;
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(sh1[k][32 * ti + rl], sd2[k][32 * tj + rl], acc, 0, 0, 0);
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) out[OW2 + (32 * ti + acc_row(q, lane)) * NH + 32 * tj + rl] = acc[q];
    // dW1 = X' D1 (rows = the NIN inputs, padded to 32): waves 0 and 1
    if (ti == 0) {
#pragma unroll
        for (int q = 0; q < 16; ++q) acc[q] = 0.f;
#pragma unroll 8
        for (int s = 0; s < kGB / 2; ++s) {
            const int k = 2 * s + kh;
            const float a = rl < NIN ? sx[k][rl] : 0.f;
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, sd1[k][32 * tj + rl], acc, 0, 0, 0);
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int r = acc_row(q, lane);
            if (r < NIN) out[OW1 + r * NH + 32 * tj + rl] = acc[q];
        }
    }
    // dW3, the biases and the loss: one output unit per thread of wave 2 / 3
    if (w == 2) {
        const int j = lane;
        float w3[NOUT] = {0.f, 0.f, 0.f}, b1 = 0.f, b2 = 0.f;
        for (int s = 0; s < kGB; ++s) {
            const float h2 = sh2[s][j];
#pragma unroll
            for (int k = 0; k < NOUT; ++k) w3[k] = fmaf(h2, sd3[s][k], w3[k]);
            b1 += sd1[s][j];
            b2 += sd2[s][j];
        }
#pragma unroll
        for (int k = 0; k < NOUT; ++k) out[OW3 + j * NOUT + k] = w3[k];
        out[OB1 + j] = b1;
        out[OB2 + j] = b2;
    } else if (w == 3) {
        if (lane < NOUT) {
            float t = 0.f;
            for (int s = 0; s < kGB; ++s) t += sd3[s][lane];
            out[OB3 + lane] = t;
        }
        double l = sl[lane];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
        if (lane == 0) lossp[blockIdx.x] = l;
    }
}

__global__ void __launch_bounds__(256) k_mlp_grad_finalize(const float* __restrict__ partial,
                                                           const double* __restrict__ lossp, int nblk, int64_t B,
                                                           float* __restrict__ grad, float* __restrict__ loss) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < NPARAM) {
        float acc = 0.f;
        for (int k = 0; k < nblk; ++k) acc += partial[(size_t)k * NPARAM + q];
        grad[q] = acc;
    } else if (q == NPARAM && loss) {
        double l = 0.0;
        for (int k = 0; k < nblk; ++k) l += lossp[k];
        *loss = (float)(l / (double)B);
    }
}

// k_mlp_grad_finalize followed by the Adam update of each parameter (one
// thread per parameter; the arithmetic of k_mlp_grad_finalize and k_mlp_adam).
__global__ void __launch_bounds__(256) k_mlp_finalize_adam(const float* __restrict__ partial,
                                                           const double* __restrict__ lossp, int nblk, int64_t B,
                                                           float* __restrict__ params, float* __restrict__ m,
                                                           float* __restrict__ v, float lr, float beta1, float beta2,
                                                           float eps, float c1, float c2, float* __restrict__ loss) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < NPARAM) {
        float g = 0.f;
        for (int k = 0; k < nblk; ++k) g += partial[(size_t)k * NPARAM + q];
        const float mi = beta1 * m[q] + (1.f - beta1) * g;
        const float vi = beta2 * v[q] + (1.f - beta2) * g * g;
        m[q] = mi;
        v[q] = vi;
        params[q] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    } else if (q == NPARAM && loss) {
        double l = 0.0;
        for (int k = 0; k < nblk; ++k) l += lossp[k];
        *loss = (float)(l / (double)B);
    }
}

__global__ void k_mlp_adam(float* __restrict__ params, const float* __restrict__ grad, float* __restrict__ m,
                           float* __restrict__ v, int64_t n, float lr, float beta1, float beta2, float eps,
                           float c1, float c2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = grad[i];
    const float mi = beta1 * m[i] + (1.f - beta1) * g;
    const float vi = beta2 * v[i] + (1.f - beta2) * g * g;
    m[i] = mi;
    v[i] = vi;
    params[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace mlpk

using namespace mlpk;

static thread_local std::string g_mlp_err;

static int mlp_fail(int code, const char* msg) {
    g_mlp_err = msg;
    return code;
}

#define MLP_TRY(expr)                                            \
    do {                                                         \
        hipError_t _e = (expr);                                  \
        if (_e != hipSuccess) return mlp_fail(-1, hipGetErrorString(_e)); \
    } while (0)

extern "C" int ffs_mlp_param_count(void) { return NPARAM; }
// Workspace: per 64-sample block a double loss partial (two floats) and
// NPARAM gradient partials.
extern "C" int ffs_mlp_workspace_floats(int64_t B) {
    const int64_t nblk = (std::max<int64_t>(B, 0) + kGB - 1) / kGB;
    return (int)(nblk * (NPARAM + 2));
}
extern "C" const char* ffs_mlp_last_error(void) { return g_mlp_err.c_str(); }

extern "C" int ffs_mlp_forward(const float* params, const float* X, const int32_t* idx, int64_t B, float* out,
                               void* stream) {
    if (B <= 0) return 0;
    hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, (hipStream_t)stream, params,
                       X, idx, B, out);
    MLP_TRY(hipGetLastError());
    return 0;
}

extern "C" int ffs_mlp_grad(const float* params, const float* X, const float* target, const int32_t* idx,
                            int64_t B, float* grad, float* loss, float* workspace, void* stream) {
    if (B <= 0) return mlp_fail(-3, "mlp_grad: empty batch");
    hipStream_t s = (hipStream_t)stream;
    const int64_t nblk = (B + kGB - 1) / kGB;
    if (nblk > 0x7fffffff) return mlp_fail(-3, "mlp_grad: batch too large");
    // [nblk] double loss partials, then [nblk][NPARAM] float gradient partials
    double* lossp = reinterpret_cast<double*>(workspace);
    float* partial = workspace + 2 * nblk;
    hipLaunchKernelGGL(k_mlp_block_grad, dim3((unsigned)nblk), dim3(256), 0, s, params, X, target, idx, B, partial,
                       lossp);
    MLP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_mlp_grad_finalize, dim3((NPARAM + 1 + 255) / 256), dim3(256), 0, s, partial, lossp,
                       (int)nblk, B, grad, loss);
    MLP_TRY(hipGetLastError());
    return 0;
}

// n_steps Adam steps over consecutive batches perm[s * batch .. + batch) of
// (X, target), steps step0 + 1 .. step0 + n_steps of a run of total_steps with
// the learning rate lr (lr_final / lr)^((step - 1) / (total_steps - 1)); the
// loss of step s goes to losses[s] (nullable).  Launches two kernels per step
// from here, so a host loop never sits between steps.
extern "C" int ffs_mlp_train_steps(float* params, float* m, float* v, const float* X, const float* target,
                                   const int32_t* perm, int64_t batch, int32_t n_steps, int32_t step0,
                                   int32_t total_steps, float lr, float lr_final, float beta1, float beta2, float eps,
                                   float* losses, float* workspace, void* stream) {
    if (batch <= 0 || n_steps < 0 || step0 < 0 || total_steps < step0 + n_steps || !(lr > 0.f) ||
        !(lr_final > 0.f))
        return mlp_fail(-3, "mlp_train_steps: invalid batch, step counts or learning rates");
    if (!params || !m || !v || !X || !target || !perm || !workspace)
        return mlp_fail(-3, "mlp_train_steps: null argument");
    hipStream_t s = (hipStream_t)stream;
    const int64_t nblk = (batch + kGB - 1) / kGB;
    if (nblk > 0x7fffffff) return mlp_fail(-3, "mlp_train_steps: batch too large");
    double* lossp = reinterpret_cast<double*>(workspace);
    float* partial = workspace + 2 * nblk;
    for (int i = 0; i < n_steps; ++i) {
        const int step = step0 + i + 1;
        const float frac = (float)(step - 1) / (float)std::max(1, total_steps - 1);
        const float lr_t = lr * powf(lr_final / lr, frac);
        const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
        hipLaunchKernelGGL(k_mlp_block_grad, dim3((unsigned)nblk), dim3(256), 0, s, params, X, target,
                           perm + (size_t)i * batch, batch, partial, lossp);
        hipLaunchKernelGGL(k_mlp_finalize_adam, dim3((NPARAM + 1 + 255) / 256), dim3(256), 0, s, partial, lossp,
                           (int)nblk, batch, params, m, v, lr_t, beta1, beta2, eps, c1, c2,
                           losses ? losses + i : nullptr);
    }
    MLP_TRY(hipGetLastError());
    return 0;
}

extern "C" int ffs_mlp_adam(float* params, const float* grad, float* m, float* v, int64_t n, float lr, float beta1,
                            float beta2, float eps, int32_t step, void* stream) {
    if (step < 1) return mlp_fail(-3, "mlp_adam: step must be >= 1");
    const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
    hipLaunchKernelGGL(k_mlp_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, params, grad,
                       m, v, n, lr, beta1, beta2, eps, c1, c2);
    MLP_TRY(hipGetLastError());
    return 0;
}
