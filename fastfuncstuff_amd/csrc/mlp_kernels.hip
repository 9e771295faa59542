// (a, b) proposal MLP: 10 -> 64 GELU -> 64 GELU -> 3, forward, backward and
// Adam on gfx950.  Parameters are one flat float32 buffer:
//   W1[10][64] b1[64] W2[64][64] b2[64] W3[64][3] b3[3]   (row-major [in][out])
// Loss per sample: 0.5 e^c ((a_hat-a)^2 + (b_hat-b)^2) - c, c clipped to
// [-10, 15], averaged over the batch.  The gradient reduction over the batch is
// done in a fixed order (one thread per parameter), so it is deterministic.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <string>

#include "../../include/ffs_mlp.h"

namespace mlpk {

constexpr int NIN = 10, NH = 64, NOUT = 3;
constexpr int OW1 = 0, OB1 = OW1 + NIN * NH, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH,
              OB3 = OW3 + NH * NOUT, NPARAM = OB3 + NOUT;
// Per-sample workspace: h1, h2, d1, d2 (64 each), d3 (3), loss (1).
constexpr int WS = 4 * NH + 4;
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

// Forward + per-sample backward; stores h1, h2, d1, d2, d3 and the sample loss.
__global__ void __launch_bounds__(256) k_mlp_sample_grad(const float* __restrict__ params, const float* __restrict__ X,
                                                         const float* __restrict__ target,
                                                         const int32_t* __restrict__ idx, int64_t B,
                                                         float* __restrict__ ws) {
    __shared__ float sp[NPARAM];
    load_params(params, sp);
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const int64_t r = idx ? idx[b] : b;
    float* w = ws + b * WS;
    float x[NIN], z2[NH];
#pragma unroll
    for (int i = 0; i < NIN; ++i) x[i] = X[r * NIN + i];
    layer2(sp, x, z2, w);
    float o[NOUT];
#pragma unroll
    for (int k = 0; k < NOUT; ++k) o[k] = sp[OB3 + k];
#pragma unroll
    for (int j = 0; j < NH; ++j) {
        const float h = gelu(z2[j]);
        w[NH + j] = h;
#pragma unroll
        for (int k = 0; k < NOUT; ++k) o[k] = fmaf(h, sp[OW3 + j * NOUT + k], o[k]);
    }
    const float c = fminf(fmaxf(o[2], C_MIN), C_MAX);
    const float ec = expf(c);
    const float da = o[0] - target[r * 2], db = o[1] - target[r * 2 + 1];
    const float sse = da * da + db * db;
    const float invB = 1.f / (float)B;
    float d3[NOUT];
    d3[0] = ec * da * invB;
    d3[1] = ec * db * invB;
    d3[2] = (o[2] > C_MIN && o[2] < C_MAX) ? (0.5f * ec * sse - 1.f) * invB : 0.f;
#pragma unroll
    for (int k = 0; k < NOUT; ++k) w[4 * NH + k] = d3[k];
    w[4 * NH + 3] = 0.5f * ec * sse - c;
    // d2 = (d3 W3') * gelu'(z2)   (z2 is overwritten by d2)
#pragma unroll
    for (int j = 0; j < NH; ++j) {
        float s = 0.f;
#pragma unroll
        for (int k = 0; k < NOUT; ++k) s = fmaf(d3[k], sp[OW3 + j * NOUT + k], s);
        z2[j] = s * gelu_grad(z2[j]);
        w[3 * NH + j] = z2[j];
    }
    // d1 = (d2 W2') * gelu'(z1)
#pragma unroll 1
    for (int i = 0; i < NH; ++i) {
        float s = 0.f;
#pragma unroll
        for (int j = 0; j < NH; ++j) s = fmaf(z2[j], sp[OW2 + i * NH + j], s);
        w[2 * NH + i] = s * gelu_grad(z1_of(sp, x, i));
    }
}

// One thread per parameter (plus one for the loss), summing over the batch in order.
__global__ void __launch_bounds__(256) k_mlp_reduce(const float* __restrict__ X, const int32_t* __restrict__ idx,
                                                    int64_t B, const float* __restrict__ ws,
                                                    float* __restrict__ grad, float* __restrict__ loss) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q > NPARAM) return;
    float acc = 0.f;
    if (q == NPARAM) {
        double l = 0.0;
        for (int64_t b = 0; b < B; ++b) l += ws[b * WS + 4 * NH + 3];
        if (loss) *loss = (float)(l / (double)B);
        return;
    }
    if (q < OB1) {
        const int i = q / NH, j = q % NH;
        for (int64_t b = 0; b < B; ++b) {
            const int64_t r = idx ? idx[b] : b;
            acc = fmaf(X[r * NIN + i], ws[b * WS + 2 * NH + j], acc);
        }
    } else if (q < OW2) {
        const int j = q - OB1;
        for (int64_t b = 0; b < B; ++b) acc += ws[b * WS + 2 * NH + j];
    } else if (q < OB2) {
        const int i = (q - OW2) / NH, j = (q - OW2) % NH;
        for (int64_t b = 0; b < B; ++b) acc = fmaf(ws[b * WS + i], ws[b * WS + 3 * NH + j], acc);
    } else if (q < OW3) {
        const int j = q - OB2;
        for (int64_t b = 0; b < B; ++b) acc += ws[b * WS + 3 * NH + j];
    } else if (q < OB3) {
        const int j = (q - OW3) / NOUT, k = (q - OW3) % NOUT;
        for (int64_t b = 0; b < B; ++b) acc = fmaf(ws[b * WS + NH + j], ws[b * WS + 4 * NH + k], acc);
    } else {
        const int k = q - OB3;
        for (int64_t b = 0; b < B; ++b) acc += ws[b * WS + 4 * NH + k];
    }
    grad[q] = acc;
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
extern "C" int ffs_mlp_workspace_floats(int64_t B) { return (int)(B * WS); }
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
    hipLaunchKernelGGL(k_mlp_sample_grad, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, s, params, X, target, idx,
                       B, workspace);
    MLP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_mlp_reduce, dim3((NPARAM + 1 + 255) / 256), dim3(256), 0, s, X, idx, B, workspace, grad,
                       loss);
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
