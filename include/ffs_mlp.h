/*
 * ffs_mlp -- C ABI of the (a, b) proposal network (10 residual
 * autocorrelation lags -> 64 GELU -> 64 GELU -> (a_hat, b_hat, log-confidence)).
 *
 * The reference has no learned component on this path: these entry points
 * have no counterpart in fastfuncstuff.  They serve the learned-proposal
 * search that narrows the reference's exhaustive / hierarchical grid search
 * (glm/arma.py:1886-2322) to a local window.  Pointers are device pointers;
 * work is enqueued on the given hipStream_t (void*).  Status 0 = success.
 */
#ifndef FFS_MLP_H
#define FFS_MLP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Number of float32 parameters (5059). */
int ffs_mlp_param_count(void);

/* Floats of workspace ffs_mlp_grad needs for a batch of B samples. */
int ffs_mlp_workspace_floats(int64_t B);

const char* ffs_mlp_last_error(void);

/* out[b*3 + k] for samples X[idx[b]*10 ..] (idx nullable = identity). */
int ffs_mlp_forward(const float* params, const float* X, const int32_t* idx, int64_t B, float* out, void* stream);

/* Mean loss (nullable) and its gradient over the batch; target is (n, 2). */
int ffs_mlp_grad(const float* params, const float* X, const float* target, const int32_t* idx, int64_t B,
                 float* grad, float* loss, float* workspace, void* stream);

/* n_steps training steps (gradient as ffs_mlp_grad, then the Adam update of
 * ffs_mlp_adam) over consecutive batches perm[s*batch .. +batch) (device
 * int32), steps step0+1 .. step0+n_steps of a run of total_steps, learning
 * rate lr * (lr_final / lr)^((step-1) / max(1, total_steps-1)); per-step mean
 * loss to losses[s] (device, nullable).  workspace: ffs_mlp_workspace_floats(batch). */
int ffs_mlp_train_steps(float* params, float* m, float* v, const float* X, const float* target, const int32_t* perm,
                        int64_t batch, int32_t n_steps, int32_t step0, int32_t total_steps, float lr, float lr_final,
                        float beta1, float beta2, float eps, float* losses, float* workspace, void* stream);

/* In-place Adam update with bias correction for step >= 1. */
int ffs_mlp_adam(float* params, const float* grad, float* m, float* v, int64_t n, float lr, float beta1,
                 float beta2, float eps, int32_t step, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* FFS_MLP_H */
