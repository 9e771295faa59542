/*
 * ffs_reml -- C ABI of the MI355X ARMA(1,1) REML engine.
 *
 * Every entry point returns 0 on success and a negative status on failure;
 * ffs_reml_last_error() returns the calling thread's last error message.
 * Device pointers are caller-owned (e.g. torch tensors on the current HIP
 * device); the library allocates only the workspace that a plan owns and
 * frees it in ffs_reml_plan_destroy.  All work is enqueued on the given
 * hipStream_t (passed as void*); nothing synchronises except
 * ffs_reml_plan_create (which returns only once the tables are built) and
 * ffs_reml_plan_logdets (a host copy).
 *
 * Reference interfaces replaced (paths relative to fastfuncstuff/):
 *   ffs_reml_plan_create   precompute_reml_grid           glm/arma.py:1371-1531
 *                          (+ precompute_autocorr_grid    glm/arma.py:1306-1368,
 *                             arma11_whiteners            glm/whiten.py:197-271)
 *   ffs_reml_prep          std normalisation              glm/arma.py:3947-3954
 *                          _residualize_reml_search_data  glm/arma.py:1554-1571
 *   ffs_reml_energy        ||Y_w||^2 of search_voxels_precomputed_grid  glm/arma.py:1983-1995
 *   ffs_reml_project       ||Q'Y_w||^2 + score of the same loop         glm/arma.py:1986-2004
 *   ffs_reml_select        strict-< update (glm/arma.py:2006-2010) and
 *                          select_arma_params_hierarchical_from_surface glm/arma.py:2224-2322
 *   ffs_reml_gls           per-(a,b)-group GLS refit      glm/arma.py:4784-5601
 *   ffs_reml_ljung_box     _ljung_box_batched             glm/arma.py:6059-6141
 */
#ifndef FFS_REML_H
#define FFS_REML_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ffs_reml_plan ffs_reml_plan;

/* Library version string. */
const char* ffs_reml_version(void);

/* Last error message of the calling thread ("" when none). */
const char* ffs_reml_last_error(void);

/*
 * Build the per-pair FP64 tables on the current device.
 *   T, p        retained time points and design columns (p <= 768).  Designs of
 *               more than 96 columns keep each pair's columns in segments of 64
 *               projection columns: they support prep, energy, project, surface,
 *               select and the GLS (the split 16-bit projection modes read the
 *               halves that prep wrote in the same mode, so ffs_reml_surface runs
 *               them with n_chunks = 1), the hierarchical descent on demand and
 *               surface_f64; the windows, bands, subset search, prep_features and
 *               the grouped engine return -4 on them
 *   X           host, T x p row-major float64 design
 *   pinvT       host, T x p row-major float64 transpose of the design pseudo-inverse
 *   gap         host, T int32: tau[t+1]-tau[t] inside a run, 0 at the last point of a run
 *   K, ab       host, K x 2 float64 (a, b) pairs in search order
 *   white       host, K int32: 1 where lambda == 0 (R is exactly the identity)
 *   n_a, n_b, lut   host lattice of the hierarchical selector: lut[ai*n_b+bi] = pair or -1;
 *               n_a = n_b = 0 (lut NULL) makes a pair-list plan for GLS / whitening only
 *               (the -load_Rvar re-use path, glm/arma.py:4388-4447, which builds one
 *               whitener per unique pair): no search operands, windows or selector
 *               tables are built, so memory is linear in K, and the search entry
 *               points return -4 on it
 *   zero_a, zero_b  lattice coordinates of (0, 0)
 *   ltop        top level of the hierarchical descent
 *   n_task, task_idx  host; F statistic over these columns (n_task = 0: all columns, F = b'Ab)
 *   stream      hipStream_t
 * Rank-deficient designs (glm/arma.py:1493-1500 falls back to QR when the
 * Cholesky of X_w'X_w fails; the GLS to inv, :4864-4873, or pinv, :3410-3416)
 * build: a column whose residual against the columns before it has squared
 * norm <= 1e-13 |x_j|^2 is dependent (whitening keeps the column space, so
 * this is decided on X).  Its pivot is zero, so logdet X'R^-1X carries
 * 2 log(1e-10) for it (the QR branch's |R_jj| + 1e-10 with a rounding-level
 * R_jj), the projection is onto the rank-r column space, (X'R^-1X)^-1 becomes
 * the pseudo-inverse and ffs_reml_gls reports the minimum-norm beta (the pinv
 * solution).  Returns -2 only when a pivot of an independent column is not
 * positive after whitening (or X is not finite).
 */
int ffs_reml_plan_create(ffs_reml_plan** out, int32_t T, int32_t p, const double* X, const double* pinvT,
                         const int32_t* gap, int32_t K, const double* ab, const int32_t* white, int32_t n_a,
                         int32_t n_b, const int32_t* lut, int32_t zero_a, int32_t zero_b, int32_t ltop,
                         int32_t n_task, const int32_t* task_idx, void* stream);

int ffs_reml_plan_destroy(ffs_reml_plan* plan);

/* Column rank of the plan's design as decided at creation (p = full rank);
 * -1 for a null plan.  No reference counterpart: the reference reports the
 * QR fallback only as a verbose message (glm/arma.py:1507-1508). */
int32_t ffs_reml_plan_rank(const ffs_reml_plan* plan);

/* G plans of one timeline and pair list with G designs (X and pinvT: G x T x p,
 * host, row-major per design), for grouped fits (glm/arma.py:2836-3183 fits
 * one design per voxel group on one shared autocorrelation grid,
 * glm/arma.py:2971-2989).  The per-pair tables of all designs are built in one
 * launch over (pair, design).  Same other arguments as ffs_reml_plan_create;
 * out[0..G) receive the plans (each destroyed with ffs_reml_plan_destroy).
 * Returns -2 naming the first rank-deficient design. */
int ffs_reml_plan_create_many(ffs_reml_plan** out, int32_t G, int32_t T, int32_t p, const double* X,
                              const double* pinvT, const int32_t* gap, int32_t K, const double* ab,
                              const int32_t* white, int32_t n_a, int32_t n_b, const int32_t* lut, int32_t zero_a,
                              int32_t zero_b, int32_t ltop, int32_t n_task, const int32_t* task_idx, void* stream);

/* Search-only plan whose score uses A + ridge I (A = X'R^-1X) in the
 * log-determinant and in the solve for beta, with the whitened residual of
 * that beta: the score of compute_reml_likelihood / reml_grid_search_batched
 * (glm/arma.py:1160-1176, 2724-2755; ridge = 1e-6 there).  Same arguments as
 * ffs_reml_plan_create without the task subset.  ffs_reml_gls refuses it. */
int ffs_reml_plan_create_ridge(ffs_reml_plan** out, int32_t T, int32_t p, const double* X, const double* pinvT,
                               const int32_t* gap, int32_t K, const double* ab, const int32_t* white, int32_t n_a,
                               int32_t n_b, const int32_t* lut, int32_t zero_a, int32_t zero_b, int32_t ltop,
                               double ridge, void* stream);

/* Host copies of logdet R and logdet X'R^-1X per pair (synchronous). */
int ffs_reml_plan_logdets(const ffs_reml_plan* plan, double* logdet_R, double* logdet_XRX);

/* Host copy (synchronous) of (X'R^-1X)^-1 per pair, K x p x p row-major: the
 * XwTXw_inv of glm/arma.py:5030-5043 from which GLT variances c' A c follow. */
int ffs_reml_plan_ainv(const ffs_reml_plan* plan, double* ainv);

/* Padded sizes the search buffers must have: Tpad rows of Vpad floats for the
 * time-major residual buffer (rows >= T must be zero), and the voxel padding unit. */
int ffs_reml_plan_dims(const ffs_reml_plan* plan, int32_t* Tpad, int32_t* vox_pad, int32_t* K);

/* Rows the Yt buffer of ffs_reml_prep / ffs_reml_project needs: Tpad, or
 * 2 Tpad in the split 16-bit projection modes with p <= 32, where prep also
 * writes the residual's hi / lo halves into rows Tpad..2 Tpad-1 and the
 * projection reads them (prep and project must run in the same mode). */
int ffs_reml_plan_yt_rows(const ffs_reml_plan* plan, int32_t* rows);

/*
 * Normalise + residualise Y (V x T float32, row stride ldY) into the
 * time-major buffer Yt (Tpad x ldYt float32).  Column c of Yt holds voxel
 * rows[c] (rows nullable = identity, then ncols = V); a -1 entry, and every
 * column from ncols up to the next multiple of the voxel padding unit, is zero,
 * as are rows T..Tpad-1.  ldYt >= ncols rounded up to the padding unit.
 * normalize bit 0: divide by scale[v] = unbiased std clamped at 1e-8 (else
 * scale = 1); bit 1: skip the residualisation (Yt holds y / scale).  scale is
 * indexed by voxel; feats (nullable, V x 10) receives lag-1..10
 * autocorrelations of the residual.  Yt may be null (statistics / features only).
 */
int ffs_reml_prep(const ffs_reml_plan* plan, const float* Y, int64_t ldY, int64_t V, const int32_t* rows,
                  int64_t ncols, float* Yt, int64_t ldYt, float* scale, int32_t normalize, float* feats,
                  void* stream);

/* Features of every voxel for the learned proposal (no counterpart in the
 * reference): scale[v] and feats[v*10 + k] as ffs_reml_prep with rows = NULL
 * and Yt = NULL, and, when coef is not NULL, each voxel's residualisation
 * coefficients pinv(X) y / scale (normalised units) in
 * coef[v*stride .. v*stride + stride), stride from ffs_reml_plan_coef_stride.
 * The window, band and subset preparations given that coef and scale (same Y,
 * normalize = 1) skip their first pass and produce bitwise the residuals they
 * would compute themselves.  Returns -4 with coef when stride is 0 (p > 32). */
int ffs_reml_prep_features(const ffs_reml_plan* plan, const float* Y, int64_t ldY, int64_t V, float* scale,
                           int32_t normalize, float* feats, double* coef, void* stream);
/* Doubles per voxel of the coefficient array of ffs_reml_prep_features (16 or
 * 32; 0 when the plan's design is wider than 32 columns). */
int ffs_reml_plan_coef_stride(const ffs_reml_plan* plan, int32_t* stride);

/* (i) E[v*ldE + k] = ||L_k^-1 y_v||^2 for every voxel and pair. */
int ffs_reml_energy(const ffs_reml_plan* plan, const float* Yt, int64_t ldYt, int64_t V, float* E, int64_t ldE,
                    void* stream);

/* (ii) S[v*ldS + k]: in = energy from ffs_reml_energy, out = REML score
 * logdet R + logdet X'R^-1X + (T-p) log(E - ||G_k' y||^2 + 1e-10). */
int ffs_reml_project(const ffs_reml_plan* plan, const float* Yt, int64_t ldYt, int64_t V, float* S, int64_t ldS,
                     void* stream);

/* Precision of the full-grid projection (ii): 0 = fp32 MFMA (default),
 * 1 = split bf16 (hi + lo, three products, fp32 accumulation),
 * 2 = split fp16 (same, fp16 halves).  The local window search uses the same
 * mode, so its scores stay equal to the full-grid surface. */
int ffs_reml_plan_set_projection(ffs_reml_plan* plan, int32_t mode);

/* (i)+(ii) over n_chunks voxel chunks (1..16), the energy of one chunk
 * overlapping the projection of the previous one on a second stream owned by
 * the plan; same result as ffs_reml_energy then ffs_reml_project.  stream is
 * ordered after all of it on return. */
int ffs_reml_surface(const ffs_reml_plan* plan, const float* Yt, int64_t ldYt, int64_t V, float* S, int64_t ldS,
                     int32_t n_chunks, void* stream);

/*
 * Hierarchical descent on demand (glm/arma.py:2224-2322, the reference's GPU
 * default selector): the levels that ffs_reml_select mode 1 runs on the full
 * surface, run on pair sets holding only the lattice points the descent can
 * visit (pass 0: the top level and the level below it, the same points for
 * every voxel; then one pass per level, each voxel scored on the 3 x 3
 * neighbourhood of its current point).  The set scores equal the surface
 * bitwise, so best / best_score are those of ffs_reml_prep + ffs_reml_surface
 * + ffs_reml_select(mode 1) for the same Y (normalize = 1), at about 45 pair
 * evaluations per voxel instead of K.  scale (out): each voxel's std, as
 * ffs_reml_prep writes it.  ffs_reml_plan_build_hier builds the pass tables
 * once per plan (about 60 MB at K = 561 on a 33 x 33 lattice, T = 380).
 */
int ffs_reml_plan_build_hier(ffs_reml_plan* plan, void* stream);
int ffs_reml_hier_workspace(const ffs_reml_plan* plan, int64_t V, int64_t* bytes);
int ffs_reml_hier_search(const ffs_reml_plan* plan, const float* Y, int64_t ldY, int64_t V, int32_t* best,
                         float* best_score, float* scale, void* workspace, int64_t ws_bytes, void* stream);

/* (iii) mode 0: global argmin (first minimum), glm/arma.py:1886-2018;
 * mode 1: hierarchical descent on the surface, glm/arma.py:2224-2322 (the
 * reference's GPU default); mode 2: subset descent, glm/arma.py:2021-2221 (the
 * reference's CPU default), evaluated on the surface. */
int ffs_reml_select(const ffs_reml_plan* plan, const float* S, int64_t ldS, int64_t V, int32_t mode, int32_t* best,
                    float* best_score, void* stream);

/*
 * FP64 search: the reference's float64 search (use_double, glm/arma.py:3869;
 * no residualisation, float64 whitening, projection and sums,
 * glm/arma.py:1886-2018).  S[v*ldS + k] = logdet R_k + logdet X'R_k^-1X +
 * (T-p) log(||L_k^-1 y||^2 - ||G_k' y||^2 + 1e-10) in float64 for y = row v of
 * Y (V x T, row stride ldY; float64 when y_is_f64, else float32 samples read
 * exactly into float64) divided by scale[v] (nullable = 1).  The energy runs
 * the innovations recurrence of pair k in float64.  Lattice plans only (-4 on
 * a pair-list plan); on a ridge plan the score is that plan's ridge score.
 */
int ffs_reml_surface_f64(const ffs_reml_plan* plan, const void* Y, int32_t y_is_f64, int64_t ldY, int64_t V,
                         const float* scale, double* S, int64_t ldS, void* stream);
/* ffs_reml_select on a float64 surface (same modes and tie rules), float64 scores out. */
int ffs_reml_select_f64(const ffs_reml_plan* plan, const double* S, int64_t ldS, int64_t V, int32_t mode,
                        int32_t* best, double* best_score, void* stream);

/*
 * (iv) FP64 GLS refit of Y (V x T float32, original units) at pair best[v].
 * scale (nullable) = per-voxel std used to normalise before the fit.
 * Outputs: betas, tstats (V x p), fstats, sigma2, r2 (V).  Optional (nullable,
 * row stride ldR): resid_w (float64 whitened residuals), resid (float32
 * Y - X beta) and pred (float32 X beta), all in original units.
 * work: caller-provided int32 workspace of ffs_reml_gls_workspace(V) elements
 * (voxels grouped by pair; nullable = no grouping, slower).  Without series
 * outputs and with p <= 32 the fit runs on f64 MFMA over pair-uniform waves.
 * On a rank-deficient plan beta is the minimum-norm solution (zero
 * component in the null space of X), t uses diag (X'R^-1X)^+.
 */
/* int32 elements of the GLS workspace for V voxels (glm/arma.py:4784-4811 groups
 * voxels by unique (a, b); the workspace holds that grouping). */
int ffs_reml_gls_workspace(const ffs_reml_plan* plan, int64_t V, int64_t* n_int32);

int ffs_reml_gls(const ffs_reml_plan* plan, const float* Y, int64_t ldY, const float* scale, const int32_t* best,
                 int64_t V, float* betas, float* tstats, float* fstats, float* sigma2, float* r2, double* resid_w,
                 int64_t ldR, float* resid, float* pred, int32_t* work, void* stream);

/*
 * Grouped fits (fit_glm_arma11_grouped, glm/arma.py:2836-3183, which loops
 * over the groups one fit at a time): one launch per stage over the voxels of
 * all groups.  ffs_reml_group_create takes G plans of one timeline, pair
 * list, lattice, width and projection mode (ffs_reml_plan_create_many) and
 * the group of every voxel (host, V int32 in [0, G)); it lays the voxels out
 * group by group, each group padded to 256 columns, so every prep block and
 * projection voxel tile holds one design.  The group borrows the plans: keep
 * them alive while it is used.
 * ffs_reml_group_fit: prep (normalise + residualise on each voxel's design),
 * energy, projection with the score epilogue, selection (mode as
 * ffs_reml_select), then the FP64 GLS of every voxel at its pair with its
 * design (voxels grouped by (design, pair)).  Outputs in voxel order: best
 * pair, score, scale (std used to normalise), betas, tstats (V x p), fstats,
 * sigma2, r2 (V), and, when S_out is given, the surface rows (V x K, stride
 * ldS_out).  The stages are captured once into a HIP graph on the group's own
 * streams and replayed on later calls with the same pointers (the caller's
 * stream waits for the replay); other pointers capture again.
 * ffs_reml_group_tune: the autotuner; times `reps` replays of each candidate
 * launch configuration on the given data and keeps the fastest; `report`
 * (nullable) receives a JSON description of the candidates and the choice.
 */
typedef struct ffs_reml_group ffs_reml_group;
int ffs_reml_group_create(ffs_reml_group** out, int32_t G, ffs_reml_plan* const* plans, int64_t V,
                          const int32_t* group_of_voxel, void* stream);
int ffs_reml_group_destroy(ffs_reml_group* grp);
int ffs_reml_group_fit(ffs_reml_group* grp, const float* Y, int64_t ldY, int32_t mode, int32_t* best, float* score,
                       float* scale, float* betas, float* tstats, float* fstats, float* sigma2, float* r2,
                       float* S_out, int64_t ldS_out, void* stream);
int ffs_reml_group_tune(ffs_reml_group* grp, const float* Y, int64_t ldY, int32_t mode, int32_t* best, float* score,
                        float* scale, float* betas, float* tstats, float* fstats, float* sigma2, float* r2,
                        int32_t reps, char* report, int64_t report_len, void* stream);

/*
 * Allocation-free variant of ffs_reml_gls.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as ffs_reml_gls without `work`: the voxel grouping is
 * kept in a library-owned device array of FFS_REML_GLS_NOALLOC_MAX_INT32 int32
 * (8 MiB, one copy per GPU, held until the process ends), so nothing is
 * allocated or freed.  A call needs 3 K + 1 + 64 ceil((V + 63 min(V, K)) / 64)
 * int32 of it (K = plan pairs); larger V returns -3 with the limit in the
 * message (about 2.0 M voxels at K = 561).  Every call on one GPU shares the
 * array: calls that overlap on one GPU (different streams or threads) race.
 */
#define FFS_REML_GLS_NOALLOC_MAX_INT32 (1 << 21)
int ffs_reml_gls_noalloc(const ffs_reml_plan* plan, const float* Y, int64_t ldY, const float* scale,
                         const int32_t* best, int64_t V, float* betas, float* tstats, float* fstats, float* sigma2,
                         float* r2, double* resid_w, int64_t ldR, float* resid, float* pred, void* stream);

/*
 * Ljung-Box of whitened residuals (V x ldR float64).  Lag pairs are pre-binned
 * by time lag: pairs[2q], pairs[2q+1] for q in [bin_off[d], bin_off[d+1]) have
 * lag d (1 <= d <= nbins); wbin[d] = 1/n_d for usable bins, else 0.  All three
 * arrays are device pointers.
 */
int ffs_reml_ljung_box(const double* resid_w, int64_t ldR, int64_t V, int32_t n, int32_t nbins,
                       const int32_t* bin_off, const int32_t* pairs, const double* wbin, float* out, void* stream);

/*
 * Learned-proposal local search (no counterpart in the reference, which always
 * searches the whole grid, glm/arma.py:1886-2322).  props (V x 3, device) =
 * (a_hat, b_hat, log-confidence) from the proposal network.  Each voxel is
 * scored on the 5 x 5 lattice window around its snapped proposal (kept inside
 * the lattice; the table is ffs_reml_plan_window_info) with the same
 * arithmetic as the full grid, so the scores equal the full-grid surface
 * entries bitwise (in every projection mode).  Outputs per voxel: best pair
 * (-1 if no valid pair), its score, need_full = 1 when the optimum is on a
 * window side that the lattice continues past, 2 when the confidence
 * is below conf_min or the proposal is not finite (no window is searched; best
 * and score are left as they were); scale = per-voxel std, or, when coef
 * (nullable) is given, the scale written by ffs_reml_prep_features with that
 * coef, which is then read instead of recomputed.  The split
 * projection modes need p <= 32 (returns -4 otherwise).  workspace:
 * ffs_reml_local_workspace bytes for the plan's current projection mode.
 */
int ffs_reml_local_workspace(const ffs_reml_plan* plan, int64_t V, int64_t* bytes);
/* The three stages of ffs_reml_local_search (same workspace, same stream):
 * window centres + grouping + grouped residualisation; window energies;
 * window projection + score + local argmin. */
int ffs_reml_local_prepare(const ffs_reml_plan* plan, const float* Y, int64_t ldY, int64_t V, const float* props,
                           float conf_min, int32_t* need_full, float* scale, const double* coef, void* workspace,
                           int64_t ws_bytes, void* stream);
int ffs_reml_local_energy(const ffs_reml_plan* plan, int64_t V, void* workspace, int64_t ws_bytes, void* stream);
int ffs_reml_local_project(const ffs_reml_plan* plan, int64_t V, int32_t* best, float* best_score,
                           int32_t* need_full, void* workspace, int64_t ws_bytes, void* stream);
/* One more window step for the voxels whose last window had its optimum on an
 * edge (need_full = 1): each is re-grouped and residualised around a window
 * centred on that optimum (kept inside the lattice); follow with
 * ffs_reml_local_energy and ffs_reml_local_project, which overwrite best,
 * best_score and need_full of those voxels only.  Same workspace and stream
 * as the preceding ffs_reml_local_project. */
int ffs_reml_local_recenter(const ffs_reml_plan* plan, const float* Y, int64_t ldY, int64_t V, float* scale,
                            const double* coef, void* workspace, int64_t ws_bytes, void* stream);
int ffs_reml_local_search(const ffs_reml_plan* plan, const float* Y, int64_t ldY, int64_t V, const float* props,
                          float conf_min, int32_t* best, float* best_score, int32_t* need_full, float* scale,
                          const double* coef, void* workspace, int64_t ws_bytes, void* stream);

/*
 * Band search (no counterpart in the reference, which searches the whole grid,
 * glm/arma.py:1886-2018).  Band i holds, in pair order, the pairs whose lag-1
 * correlation lambda = (1 + ab)(a + b) / (1 + 2ab + b^2) lies within
 * half_width of -1 + i * bin_width.  A voxel with need_full[v] != 0 and a
 * finite proposal (a_hat, b_hat) is scored on every pair of the band nearest
 * lambda(a_hat, b_hat) (clamped to the lattice box), with the full grid's
 * arithmetic, so its scores equal the full-grid surface entries bitwise; best
 * = first minimum in pair order (lambda == 0 pairs mapped to the first of
 * them), need_full = 1 when that pair has a lattice neighbour outside the band,
 * else 0.  Other voxels are left as they were.  Returns -4 without bands, or
 * in a split projection mode with p > 32; scale = per-voxel std of the band
 * voxels.  half_width <= 0 removes the bands.  The workspace size depends on
 * the projection mode.
 */
int ffs_reml_plan_set_bands(ffs_reml_plan* plan, double half_width, double bin_width, void* stream);
/* Host copies (synchronous; each pointer nullable): band count, slots per band,
 * pairs [n_bands][slots] (-1 = empty), counts [n_bands], edge flags [n_bands][slots]. */
int ffs_reml_plan_band_info(const ffs_reml_plan* plan, int32_t* n_bands, int32_t* slots, int32_t* pairs,
                            int32_t* counts, int32_t* edge);
int ffs_reml_band_workspace(const ffs_reml_plan* plan, int64_t V, int64_t* bytes);
/* Host copies (synchronous; each pointer nullable) of the window table of the
 * learned search, one window per lattice cell c = ai * n_b + bi: window count,
 * slots per window, pairs [n_windows][slots] (-1 = empty), counts
 * [n_windows], edge flags [n_windows][slots]. */
int ffs_reml_plan_window_info(const ffs_reml_plan* plan, int32_t* n_windows, int32_t* slots, int32_t* pairs,
                              int32_t* counts, int32_t* edge);
/* Stages of ffs_reml_band_search (same workspace and stream): band choice,
 * grouping and grouped residualisation; band energies; band projection, score
 * and band argmin. */
int ffs_reml_band_prepare(const ffs_reml_plan* plan, const float* Y, int64_t ldY, int64_t V, const float* props,
                          const int32_t* need_full, float* scale, const double* coef, void* workspace,
                          int64_t ws_bytes, void* stream);
int ffs_reml_band_energy(const ffs_reml_plan* plan, int64_t V, void* workspace, int64_t ws_bytes, void* stream);
int ffs_reml_band_project(const ffs_reml_plan* plan, int64_t V, int32_t* best, float* best_score,
                          int32_t* need_full, void* workspace, int64_t ws_bytes, void* stream);
int ffs_reml_band_search(const ffs_reml_plan* plan, const float* Y, int64_t ldY, int64_t V, const float* props,
                         int32_t* best, float* best_score, int32_t* need_full, float* scale, const double* coef,
                         void* workspace, int64_t ws_bytes, void* stream);

/*
 * Full-grid search of the voxels idx[0 .. min(*n_live, cap)) (idx and n_live
 * device memory, e.g. from ffs_reml_compact) without the host reading the
 * count: normalise + residualise, energies and projection over every pair in
 * the plan's projection mode, first minimum over the grid (the exhaustive
 * rule of glm/arma.py:1886-2018); writes best[idx[i]], best_score[idx[i]] and
 * scale[idx[i]] (read instead, with coef, as in ffs_reml_local_search).
 * Voxels past cap are left untouched (the caller compares
 * *n_live with cap afterwards).  workspace: ffs_reml_subset_workspace(cap) bytes.
 */
int ffs_reml_subset_workspace(const ffs_reml_plan* plan, int64_t cap, int64_t* bytes);
int ffs_reml_subset_search(const ffs_reml_plan* plan, const float* Y, int64_t ldY, int64_t V, const int32_t* idx,
                           const int32_t* n_live, int64_t cap, int32_t* best, float* best_score, float* scale,
                           const double* coef, void* workspace, int64_t ws_bytes, void* stream);

/* Whitening of per-voxel columns for the -dsort pass (glm/arma.py:3248-3539
 * whitens the base design and each voxel's dsort columns with
 * arma11_whitener(a, b).cols): out[(v*m + j)*ld + t] = (L_k^-1 x)_t, t < T, for
 * x = cols[(v*m + j)*ld + 0 .. T) and k = pair[v] (device arrays, fp64; a pair
 * outside [0, K) gives zeros).  L_k is the plan's whitener of pair k (censored
 * samples and run starts included), the one ffs_reml_gls applies; any design
 * width. */
int ffs_reml_whiten_cols(const ffs_reml_plan* plan, const int32_t* pair, int64_t V, int32_t m, const double* cols,
                         int64_t ld, double* out, void* stream);

/*
 * -dsort per-voxel GLS (glm/arma.py:3248-3539, the final fit redone on the
 * extended design [X | D_v] at the pair selected without the dsort columns):
 * one batched small solve per voxel on the device.  pair (V int32), D (V x q
 * x T float64: each voxel's q dsort series, run-expanded), Y (V x T float64,
 * the normalised series) and every output are device arrays; fci (host,
 * n_f int32 in [0, p + q)) are the F columns.  Per voxel, with E = [X | D_v]
 * and pe = p + q: A = E_w'E_w + 1e-8 I by Cholesky, beta = A^-1 E_w'y_w,
 * sigma2 = |y_w - E_w beta|^2 / max(1, T - pe), tstat = beta /
 * (sqrt(sigma2 diag A^-1) + 1e-10), fstat = b' inv(A^-1[fci, fci] + 1e-8 I) b /
 * (sigma2 n_f + 1e-30) over b = beta[fci], r2 = CoD(y, E beta) in original
 * space.  Outputs (float64): beta, tstat (V x pe), sigma2, fstat, r2 (V);
 * nullable: Ainv (V x pe x pe), resid_w, resid, pred (V x T).  info[v] = 0,
 * 1 when the Cholesky fails (the reference then solves with pinv(A); the
 * caller does that for these voxels), 2 for a pair outside [0, K) (no
 * outputs).  work: ffs_reml_dsort_workspace doubles.  q <= 63, p + q <= 127.
 */
int ffs_reml_dsort_workspace(const ffs_reml_plan* plan, int64_t V, int32_t q, int32_t n_f, int64_t* n_double);
int ffs_reml_dsort_gls(const ffs_reml_plan* plan, const int32_t* pair, int64_t V, int32_t q, const double* D,
                       const double* Y, const int32_t* fci, int32_t n_f, double* beta, double* tstat, double* sigma2,
                       double* fstat, double* r2, double* Ainv, double* resid_w, double* resid, double* pred,
                       int32_t* info, double* work, int64_t work_doubles, void* stream);

/* idx[0..*count) = voxels with flags[v] != 0 (order unspecified); count is device memory. */
int ffs_reml_compact(const int32_t* flags, int64_t V, int32_t* idx, int32_t* count, void* stream);

/* best[idx[i]] = best_sub[i], score[idx[i]] = score_sub[i] for i < n. */
int ffs_reml_scatter_best(const int32_t* idx, int64_t n, const int32_t* best_sub, const float* score_sub,
                          int32_t* best, float* score, void* stream);

/* REML score of one dense covariance (compute_reml_likelihood,
 * glm/arma.py:1106-1186): X (T x p), Y (T x m), R (T x T), row-major float64
 * device arrays; work >= T*T + T*(p+m) + p*p + p*m doubles; *out (device) =
 * 2 sum log(diag chol R + 1e-10) + log det(X_w'X_w + 1e-6 I)
 * + (T - p) log(sum of squared whitened residuals + 1e-10), or 1e10 when R is
 * not positive definite. */
int ffs_reml_dense_likelihood(const double* X, const double* Y, int32_t m, const double* R, int32_t T, int32_t p,
                              double* work, int64_t work_doubles, double* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* FFS_REML_H */
