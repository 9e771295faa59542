// C ABI of the REML engine: plan ownership, launch geometry, error strings.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ffs_reml.h"
#include "reml_kernels.h"

// The kernels are compiled in this translation unit so that the templated
// launches below instantiate them.
#include "reml_kernels.hip"
#include "reml_sort_gls.hip"
#include "reml_prep_mfma.hip"
#include "reml_proj16.hip"
#include "reml_sets.hip"
#include "reml_dense.hip"
#include "reml_dsort.hip"
#include "reml_f64.hip"
#include "reml_group.hip"
#include "reml_gls_wide.hip"
#include "reml_seg.hip"

using namespace remlk;

static thread_local std::string g_err;

static int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

// Search entry points need the lattice and search operands a pair-list plan lacks.
#define FFS_NEED_SEARCH(P, what)                                                                   \
    do {                                                                                           \
        if ((P)->gls_only)                                                                         \
            return fail(-4, std::string(what) + ": pair-list plan (no lattice) supports GLS only"); \
    } while (0)

// Entry points built on the pair sets or the grouped engine hold a pair's
// columns in one projection tile: designs of at most 96 columns.
#define FFS_NEED_NARROW(P, what)                                                                          \
    do {                                                                                                  \
        if ((P)->d.nseg)                                                                                  \
            return fail(-4, std::string(what) +                                                           \
                                ": designs wider than 96 columns support the full-grid search, the selectors " \
                                "and the GLS only");                                                      \
    } while (0)

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess) return fail(-1, std::string(#expr ": ") + hipGetErrorString(_e)); \
    } while (0)

struct ffs_reml_plan {
    PlanDev d{};
    int device = 0;
    std::vector<void*> allocs;
    const double* pinvT = nullptr;
    const double* ldR = nullptr;
    const double* ldX = nullptr;
    double ridge = 0.0;  // > 0: search-only plan with the ridge-regularised score
    // second stream + events of ffs_reml_surface (created on first use)
    static constexpr int kMaxSurfaceChunks = 16;
    mutable hipStream_t aux = nullptr;
    mutable hipEvent_t ev[kMaxSurfaceChunks + 1] = {};
    // Yt buffer whose rows Tpad..2 Tpad-1 the last ffs_reml_prep filled with
    // split halves (projection reads them only for this buffer)
    mutable const float* split_yt = nullptr;
    // no lattice: a pair-list plan for the GLS / whitening entry points only
    // (no search operands, windows or selector tables are built)
    bool gls_only = false;
    // host copies of the pair list and lattice (band tables are built from them)
    std::vector<double> h_ab;
    std::vector<int32_t> h_lut;
    std::vector<int32_t> h_white;
    // passes of the hierarchical descent on pair sets (ffs_reml_plan_build_hier)
    std::vector<HierPass> hier;
    template <class Tp>
    int alloc(Tp** out, size_t n) {
        void* p = nullptr;
        if (n == 0) n = 1;
        hipError_t e = hipMalloc(&p, n * sizeof(Tp));
        if (e != hipSuccess) return fail(-1, std::string("hipMalloc: ") + hipGetErrorString(e));
        allocs.push_back(p);
        *out = static_cast<Tp*>(p);
        return 0;
    }
    ~ffs_reml_plan() {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        if (aux) (void)hipStreamDestroy(aux);
        for (void* p : allocs) (void)hipFree(p);
    }
};

extern "C" const char* ffs_reml_version(void) { return "ffs_reml 0.1 gfx950"; }
extern "C" const char* ffs_reml_last_error(void) { return g_err.c_str(); }

template <class Tp>
static int upload(ffs_reml_plan* P, Tp** dst, const Tp* src, size_t n, hipStream_t s) {
    if (int rc = P->alloc(dst, n)) return rc;
    if (n) HIP_TRY(hipMemcpyAsync(*dst, src, n * sizeof(Tp), hipMemcpyHostToDevice, s));
    return 0;
}

static int pick_bn(int p, int* npt) {
    // Widest supported tile that holds at least one pair; prefer the one with
    // the least padding per pair.
    const int cands[2] = {64, 96};
    int best = -1;
    double waste = 1e9;
    for (int bn : cands) {
        if (p > bn) continue;
        const int n = bn / p;
        const double w = (double)(bn - n * p) / (n * p);
        if (w < waste - 1e-12) {
            waste = w;
            best = bn;
        }
    }
    if (best < 0) return -1;
    *npt = best / p;
    return best;
}

static int plan_create_impl(ffs_reml_plan** out, int32_t T, int32_t p, const double* X, const double* pinvT,
                            const int32_t* gap, int32_t K, const double* ab, const int32_t* white, int32_t n_a,
                            int32_t n_b, const int32_t* lut, int32_t zero_a, int32_t zero_b, int32_t ltop,
                            int32_t n_task, const int32_t* task_idx, double ridge, void* stream);
namespace {
int build_windows(ffs_reml_plan* P, hipStream_t s);
size_t align256(size_t x) { return (x + 255) / 256 * 256; }
}  // namespace

extern "C" int ffs_reml_plan_create(ffs_reml_plan** out, int32_t T, int32_t p, const double* X, const double* pinvT,
                                    const int32_t* gap, int32_t K, const double* ab, const int32_t* white,
                                    int32_t n_a, int32_t n_b, const int32_t* lut, int32_t zero_a, int32_t zero_b,
                                    int32_t ltop, int32_t n_task, const int32_t* task_idx, void* stream) {
    return plan_create_impl(out, T, p, X, pinvT, gap, K, ab, white, n_a, n_b, lut, zero_a, zero_b, ltop, n_task,
                            task_idx, 0.0, stream);
}

extern "C" int ffs_reml_plan_create_ridge(ffs_reml_plan** out, int32_t T, int32_t p, const double* X,
                                          const double* pinvT, const int32_t* gap, int32_t K, const double* ab,
                                          const int32_t* white, int32_t n_a, int32_t n_b, const int32_t* lut,
                                          int32_t zero_a, int32_t zero_b, int32_t ltop, double ridge, void* stream) {
    if (!(ridge > 0.0)) return fail(-3, "plan_create_ridge: ridge must be > 0");
    return plan_create_impl(out, T, p, X, pinvT, gap, K, ab, white, n_a, n_b, lut, zero_a, zero_b, ltop, 0, nullptr,
                            ridge, stream);
}

// Column rank structure of a design (glm/arma.py:1493-1500 falls back to QR
// when the Cholesky of X_w'X_w fails).  Whitening by the invertible L^-1
// keeps the column space, so which columns depend on the columns before them
// is a property of X alone: column j is dependent when its squared residual
// against columns 0..j-1 (the pivot of a Cholesky of X'X) is at most
// kRankTol |x_j|^2.  Returns the rank r; for r < p fills dep (1 = dependent)
// and Pi = M'(MM')^-1 M, the projection onto the row space of X, where
// X = X_I M with X_I the independent columns (M = (X_I'X_I)^-1 X_I'X).
// FP64 on the host: p^3 work once per design.
constexpr double kRankTol = 1e-13;
static int design_rank(int p, const std::vector<double>& A, std::vector<int32_t>& dep, std::vector<double>& Pi) {
    std::vector<double> L((size_t)p * p, 0.0);
    std::vector<int> I;
    dep.assign(p, 0);
    for (int j = 0; j < p; ++j) {
        const double ajj = A[(size_t)j * p + j];
        double d = ajj;
        for (int k = 0; k < j; ++k) d -= L[(size_t)j * p + k] * L[(size_t)j * p + k];
        if (std::isfinite(d) && !(d > kRankTol * ajj)) {
            dep[j] = 1;
            continue;
        }
        const double l = sqrt(d);  // NaN designs fail later in the device Cholesky
        L[(size_t)j * p + j] = l;
        for (int i = j + 1; i < p; ++i) {
            double s = A[(size_t)i * p + j];
            for (int k = 0; k < j; ++k) s -= L[(size_t)i * p + k] * L[(size_t)j * p + k];
            L[(size_t)i * p + j] = s / l;
        }
        I.push_back(j);
    }
    const int r = (int)I.size();
    if (r == p) return p;
    Pi.assign((size_t)p * p, 0.0);
    if (r == 0) return 0;
    // M (r x p): L_I z = A_I,c, L_I' m = z per column c (L_I: rows / cols I of L)
    std::vector<double> M((size_t)r * p), z(r);
    for (int c = 0; c < p; ++c) {
        for (int a = 0; a < r; ++a) {
            double s = A[(size_t)I[a] * p + c];
            for (int b = 0; b < a; ++b) s -= L[(size_t)I[a] * p + I[b]] * z[b];
            z[a] = s / L[(size_t)I[a] * p + I[a]];
        }
        for (int a = r - 1; a >= 0; --a) {
            double s = z[a];
            for (int b = a + 1; b < r; ++b) s -= L[(size_t)I[b] * p + I[a]] * M[(size_t)b * p + c];
            M[(size_t)a * p + c] = s / L[(size_t)I[a] * p + I[a]];
        }
    }
    // Pi = Z'Z with Z = C^-1 M, C = chol(MM') (MM' >= I: M holds I_r)
    std::vector<double> C((size_t)r * r, 0.0), Z((size_t)r * p);
    for (int a = 0; a < r; ++a)
        for (int b = 0; b <= a; ++b) {
            double s = 0.0;
            for (int c = 0; c < p; ++c) s += M[(size_t)a * p + c] * M[(size_t)b * p + c];
            for (int k = 0; k < b; ++k) s -= C[(size_t)a * r + k] * C[(size_t)b * r + k];
            C[(size_t)a * r + b] = (a == b) ? sqrt(s) : s / C[(size_t)b * r + b];
        }
    for (int c = 0; c < p; ++c)
        for (int a = 0; a < r; ++a) {
            double s = M[(size_t)a * p + c];
            for (int b = 0; b < a; ++b) s -= C[(size_t)a * r + b] * Z[(size_t)b * p + c];
            Z[(size_t)a * p + c] = s / C[(size_t)a * r + a];
        }
    for (int i = 0; i < p; ++i)
        for (int j = 0; j < p; ++j) {
            double s = 0.0;
            for (int a = 0; a < r; ++a) s += Z[(size_t)a * p + i] * Z[(size_t)a * p + j];
            Pi[(size_t)i * p + j] = s;
        }
    return r;
}

// Work left after the allocation phase of one plan (see plan_alloc).
struct PlanPending {
    ffs_reml_plan* P = nullptr;
    bool gls_only = false;
    double *soa = nullptr, *ldR = nullptr, *ldX = nullptr, *dpinv = nullptr;
    float *ecoef = nullptr, *ec = nullptr, *es0 = nullptr, *lc = nullptr, *ls0 = nullptr;
    float2* lp = nullptr;
    int32_t* dchunk = nullptr;
    double ridge = 0.0;
};

// Validation, host tables, device allocations and uploads of one plan; fills
// `to` with the k_tables outputs of the plan and `pend` with what the finish
// step needs.  Launches nothing that reads the tables.
static int plan_alloc(ffs_reml_plan* P, int32_t T, int32_t p, const double* X, const double* pinvT, const int32_t* gap,
                      int32_t K, const double* ab, const int32_t* white, int32_t n_a, int32_t n_b, const int32_t* lut,
                      int32_t zero_a, int32_t zero_b, int32_t ltop, int32_t n_task, const int32_t* task_idx,
                      double ridge, hipStream_t s, TablesOut* to, KTablesArgs* ka, PlanPending* pend) {
    if (T < 2 || p < 1 || K < 1) return fail(-3, "plan_create: need T >= 2, p >= 1, K >= 1");
    if (p > kSegMaxP)
        return fail(-3, "plan_create: p > " + std::to_string(kSegMaxP) + " is not supported");
    if (gap[T - 1] != 0) return fail(-3, "plan_create: gap[T-1] must be 0 (end of the last run)");
    if ((n_a <= 0) != (n_b <= 0)) return fail(-3, "plan_create: n_a and n_b must both be > 0 or both 0");
    const bool gls_only = n_a <= 0;
    if (gls_only && ridge > 0.0) return fail(-3, "plan_create_ridge: a ridge plan needs a lattice (it is search-only)");
    if (gls_only) n_a = n_b = 0;
    for (int t = 0; t < T; ++t)
        if (gap[t] < 0) return fail(-3, "plan_create: negative time gap");
    HIP_TRY(hipGetDevice(&P->device));
    PlanDev& d = P->d;
    P->gls_only = gls_only;
    P->h_ab.assign(ab, ab + (size_t)2 * K);
    P->h_white.assign(white, white + K);
    if (!gls_only) P->h_lut.assign(lut, lut + (size_t)n_a * n_b);
    d.T = T;
    d.Tpad = (T + kProjBK - 1) / kProjBK * kProjBK;
    d.p = p;
    d.K = K;
    d.n_a = n_a;
    d.n_b = n_b;
    d.ltop = ltop;
    d.zero_a = zero_a;
    d.zero_b = zero_b;
    int npt = 1;
    const bool seg = p > 96;  // pairs span segments of kSegBN columns
    const int BN = seg ? kSegBN : pick_bn(p, &npt);
    if (BN < 0) return fail(-3, "plan_create: no projection tile fits p");
    d.BN = BN;
    d.npt = npt;
    d.nseg = seg ? (p + kSegBN - 1) / kSegBN : 0;
    d.n_tiles = seg ? K * d.nseg : (K + npt - 1) / npt;
    const int full_f = (n_task <= 0);
    std::vector<int32_t> tidx;
    if (full_f) {
        for (int j = 0; j < p; ++j) tidx.push_back(j);
    } else {
        for (int j = 0; j < n_task; ++j) {
            if (task_idx[j] < 0 || task_idx[j] >= p) return fail(-3, "plan_create: task index out of range");
            tidx.push_back(task_idx[j]);
        }
    }
    d.n_task = (int)tidx.size();

    // Energy chunks: lambda == 0 pairs form chunk 0; the rest are grouped by
    // identical a (consecutive in the a-major pair order) into halves of
    // kEnergyHalf slots, two halves (of one or two values of a) per chunk, so a
    // row of the lattice leaves at most kEnergyHalf - 1 empty slots.
    std::vector<int32_t> whites, chunks;
    for (int k = 0; k < K; ++k)
        if (white[k]) whites.push_back(k);
    if (!whites.empty()) chunks.insert(chunks.end(), kEnergyNB, -1);
    int k = 0;
    while (k < K) {
        int e = k;
        while (e < K && ab[2 * e] == ab[2 * k]) ++e;
        std::vector<int32_t> grp;
        for (int q = k; q < e; ++q)
            if (!white[q]) grp.push_back(q);
        for (size_t q = 0; q < grp.size(); q += kEnergyHalf)
            for (int j = 0; j < kEnergyHalf; ++j) chunks.push_back(q + j < grp.size() ? grp[q + j] : -1);
        k = e;
    }
    if (chunks.size() % kEnergyNB) chunks.insert(chunks.end(), kEnergyHalf, -1);
    d.n_chunks = (int)chunks.size() / kEnergyNB;
    d.n_white = (int)whites.size();

    int rc = 0;
    int32_t *dgap, *dwhiteflag, *dtask, *dwhite, *dchunk, *dlut;
    double *dab, *dX, *dpinv;
    if ((rc = upload(P, &dgap, gap, T, s)) || (rc = upload(P, &dwhiteflag, white, K, s)) ||
        (rc = upload(P, &dtask, tidx.data(), tidx.size(), s)) ||
        (rc = upload(P, &dwhite, whites.data(), whites.size(), s)) ||
        (rc = upload(P, &dchunk, chunks.data(), chunks.size(), s)) ||
        (rc = upload(P, &dlut, lut, (size_t)n_a * n_b, s)) || (rc = upload(P, &dab, ab, (size_t)2 * K, s)) ||
        (rc = upload(P, &dX, X, (size_t)T * p, s)) || (rc = upload(P, &dpinv, pinvT, (size_t)T * p, s)))
        return rc;
    double *soa, *coef64, *Xw64, *G64, *Lx, *Ainv, *Fm, *work, *ld, *ldR, *ldX;
    float *G32, *ecoef, *ec, *es0;
    if ((rc = P->alloc(&soa, (size_t)K * 3 * T)) || (rc = P->alloc(&coef64, (size_t)K * 3 * T)) ||
        (rc = P->alloc(&Xw64, (size_t)K * T * p)) || (rc = P->alloc(&G64, (size_t)K * T * p)) ||
        (rc = P->alloc(&Lx, (size_t)K * p * p)) || (rc = P->alloc(&Ainv, (size_t)K * p * p)) ||
        (rc = P->alloc(&Fm, (size_t)K * p * p)) || (rc = P->alloc(&work, (size_t)K * 2 * p * p)) ||
        (rc = P->alloc(&ld, K)) || (rc = P->alloc(&ldR, K)) || (rc = P->alloc(&ldX, K)))
        return rc;
    G32 = ecoef = ec = es0 = nullptr;
    if (!gls_only) {
        if ((rc = P->alloc(&G32, (size_t)d.n_tiles * d.Tpad * BN)) ||
            (rc = P->alloc(&ecoef, (size_t)d.n_chunks * T * 2 * kEnergyNB)) ||
            (rc = P->alloc(&ec, (size_t)d.n_chunks * T * 2)) || (rc = P->alloc(&es0, (size_t)d.n_chunks * kEnergyNB)))
            return rc;
        HIP_TRY(hipMemsetAsync(G32, 0, sizeof(float) * d.n_tiles * d.Tpad * BN, s));
    }
    // Local-search tables: lattice axes, per-pair fp32 coefficients, pair-padded operand.
    std::vector<float> a_axis(n_a, 0.f), b_axis(n_b, 0.f);
    for (int ai = 0; ai < n_a; ++ai)
        for (int bi = 0; bi < n_b; ++bi) {
            const int q = lut[ai * n_b + bi];
            if (q >= 0) {
                a_axis[ai] = (float)ab[2 * q];
                b_axis[bi] = (float)ab[2 * q + 1];
            }
        }
    float *da_axis, *db_axis, *lc = nullptr, *ls0 = nullptr;
    float2* lp = nullptr;
    if ((rc = upload(P, &da_axis, a_axis.data(), a_axis.size(), s)) ||
        (rc = upload(P, &db_axis, b_axis.data(), b_axis.size(), s)))
        return rc;
    if (!gls_only && ((rc = P->alloc(&lp, (size_t)K * T)) || (rc = P->alloc(&lc, (size_t)K * T)) ||
                      (rc = P->alloc(&ls0, K))))
        return rc;
    uint16_t *Ghf = nullptr, *Glf = nullptr, *Ghb = nullptr, *Glb = nullptr;
    float* g16s;
    if ((rc = P->alloc(&g16s, K))) return rc;
    if (!gls_only) {  // split 16-bit operands (per segment tile for p > 96)
        const size_t n16 = (size_t)d.n_tiles * d.Tpad * BN;
        if ((rc = P->alloc(&Ghf, n16)) || (rc = P->alloc(&Glf, n16)) || (rc = P->alloc(&Ghb, n16)) ||
            (rc = P->alloc(&Glb, n16)))
            return rc;
        HIP_TRY(hipMemsetAsync(Ghf, 0, 2 * n16, s));
        HIP_TRY(hipMemsetAsync(Glf, 0, 2 * n16, s));
        HIP_TRY(hipMemsetAsync(Ghb, 0, 2 * n16, s));
        HIP_TRY(hipMemsetAsync(Glb, 0, 2 * n16, s));
    }
    // X'X of the design (f64 MFMA GLS table and the rank structure)
    std::vector<double> XtX((size_t)p * p, 0.0);
    for (int a = 0; a < p; ++a)
        for (int b = 0; b < p; ++b) {
            double acc = 0.0;
            for (int t = 0; t < T; ++t) acc += X[(size_t)t * p + a] * X[(size_t)t * p + b];
            XtX[(size_t)a * p + b] = acc;
        }
    int32_t* ddep = nullptr;
    double* dPi = nullptr;
    d.rank = p;
    if (ridge <= 0.0) {
        std::vector<int32_t> dep;
        std::vector<double> Pi;
        d.rank = design_rank(p, XtX, dep, Pi);
        if (d.rank < p &&
            ((rc = upload(P, &ddep, dep.data(), dep.size(), s)) || (rc = upload(P, &dPi, Pi.data(), Pi.size(), s))))
            return rc;
    }
    d.Pi = dPi;
    // f64 MFMA GLS tables
    double *G64p = nullptr, *coefR = nullptr, *dXp = nullptr, *dXtX = nullptr, *dPp = nullptr;
    // f64 MFMA tables for every width (the GLS runs on them for p <= 96);
    // the MFMA prep and split operands keep p <= kGlsMfmaMaxP (d.PG)
    d.PGA = seg ? 0 : 16 * ((p + 15) / 16);
    d.PG = p <= kGlsMfmaMaxP ? d.PGA : 0;
    if (d.PGA) {
        const int PG = d.PGA;
        std::vector<double> Xp((size_t)d.Tpad * PG, 0.0), Pp((size_t)d.Tpad * PG, 0.0);
        for (int t = 0; t < T; ++t)
            for (int j = 0; j < p; ++j) {
                Xp[(size_t)t * PG + j] = X[(size_t)t * p + j];
                Pp[(size_t)t * PG + j] = pinvT[(size_t)t * p + j];
            }
        if ((rc = P->alloc(&G64p, (size_t)K * d.Tpad * PG)) || (rc = upload(P, &dXp, Xp.data(), Xp.size(), s)) ||
            (rc = upload(P, &dPp, Pp.data(), Pp.size(), s)) || (rc = upload(P, &dXtX, XtX.data(), XtX.size(), s)))
            return rc;
        HIP_TRY(hipMemsetAsync(G64p, 0, sizeof(double) * K * d.Tpad * PG, s));
    }
    // GLS epilogue tables (f64 MFMA GLS widths, p <= 96): written after k_tables
    double *gls_epi = nullptr, *dXtXp = nullptr;
    if (d.PGA) {
        const int PG = d.PGA;
        std::vector<double> XtXp((size_t)PG * PG, 0.0);
        for (int a = 0; a < p; ++a)
            for (int b = 0; b < p; ++b) XtXp[(size_t)a * PG + b] = XtX[(size_t)a * p + b];
        if ((rc = P->alloc(&gls_epi, (size_t)K * (2 * PG * PG + PG))) ||
            (rc = upload(P, &dXtXp, XtXp.data(), XtXp.size(), s)))
            return rc;
    }
    d.gls_epi = gls_epi;
    d.XtXp = dXtXp;
    // recurrence tables for every p: K * Tpad * 32 bytes
    if ((rc = P->alloc(&coefR, (size_t)K * d.Tpad * 4))) return rc;
    HIP_TRY(hipMemsetAsync(coefR, 0, sizeof(double) * K * d.Tpad * 4, s));
    *to = TablesOut{dX,  soa, coef64, Xw64, G64, Lx,   Ainv,  Fm,  work, ld, ldR,
                    ldX, G32, Ghf,    Glf,  Ghb, Glb,  g16s,  G64p, coefR, ddep, dPi};
    *ka = KTablesArgs{T,  d.Tpad, p,   K,   dab, dwhiteflag, dgap,  d.n_task,
                      dtask, full_f, BN, npt, d.PGA, ridge, d.nseg};
    d.gap = dgap;
    d.ab = dab;
    d.ld = ld;
    d.coef64 = coef64;
    d.X = dX;
    d.pinvT = dpinv;
    d.Xw64 = Xw64;
    d.G64 = G64;
    d.Lx = Lx;
    d.Ainv = Ainv;
    d.Fm = Fm;
    d.task_idx = dtask;
    d.G32 = G32;
    d.ecoef = ecoef;
    d.ec = ec;
    d.es0 = es0;
    d.echunk = dchunk;
    d.white = dwhite;
    d.lut = dlut;
    d.a_axis = da_axis;
    d.b_axis = db_axis;
    d.white_flag = dwhiteflag;
    d.lp = lp;
    d.lc = lc;
    d.ls0 = ls0;
    d.Gh_f16 = Ghf;
    d.Gl_f16 = Glf;
    d.Gh_bf16 = Ghb;
    d.Gl_bf16 = Glb;
    d.g16_unscale = g16s;
    d.G64p = G64p;
    d.coefR = coefR;
    d.Xp = dXp;
    d.XtX = dXtX;
    d.pinvTp = dPp;
    d.proj_mode = 0;
    pend->P = P;
    pend->gls_only = gls_only;
    pend->soa = soa;
    pend->ldR = ldR;
    pend->ldX = ldX;
    pend->dpinv = dpinv;
    pend->ecoef = ecoef;
    pend->ec = ec;
    pend->es0 = es0;
    pend->lp = lp;
    pend->lc = lc;
    pend->ls0 = ls0;
    pend->dchunk = dchunk;
    pend->ridge = ridge;
    return 0;
}

// The per-pair tables of G plans of one timeline and pair list (designs may
// differ) in one k_tables launch over (pair, plan): a single plan's K one-wave
// blocks leave most of the chip idle.  Then each plan's search tables, one
// synchronisation, the status of every plan, the windows.
static int plans_build(std::vector<PlanPending>& pend, const std::vector<TablesOut>& touts, const KTablesArgs& ka,
                       hipStream_t s) {
    const int G = (int)pend.size();
    ffs_reml_plan* P0 = pend[0].P;
    TablesOut* douts;
    int32_t* dstatus;
    if (int rc = P0->alloc(&douts, G)) return rc;
    if (int rc = P0->alloc(&dstatus, G)) return rc;
    HIP_TRY(hipMemcpyAsync(douts, touts.data(), sizeof(TablesOut) * G, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemsetAsync(dstatus, 0, sizeof(int32_t) * G, s));
    // one block per (pair, design): 64 lanes, 256 for designs wider than 96
    // columns (the O(T p^2) loops; every table element is computed by one lane
    // in the same order either way)
    hipLaunchKernelGGL(k_tables, dim3(ka.K, G), dim3(ka.p > 96 ? 256 : 64), 0, s, ka, douts, dstatus);
    HIP_TRY(hipGetLastError());
    for (PlanPending& q : pend) {
        if (q.gls_only) continue;
        const PlanDev& d = q.P->d;
        hipLaunchKernelGGL(k_local_tables, dim3(d.K), dim3(128), 0, s, d.T, d.K, q.soa, q.lp, q.lc, q.ls0);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(k_energy_tables, dim3(d.n_chunks), dim3(128), 0, s, d.T, d.n_chunks, q.dchunk, q.soa,
                           q.ecoef, q.ec, q.es0);
        HIP_TRY(hipGetLastError());
    }
    for (PlanPending& q : pend) {
        const PlanDev& d = q.P->d;
        if (!d.gls_epi) continue;
        hipLaunchKernelGGL(k_gls_epi_tables, dim3(d.K), dim3(64), 0, s, d, const_cast<double*>(d.gls_epi));
        HIP_TRY(hipGetLastError());
    }
    std::vector<int32_t> status(G, 0);
    HIP_TRY(hipMemcpyAsync(status.data(), dstatus, sizeof(int32_t) * G, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (int g = 0; g < G; ++g) {
        const std::string who = G > 1 ? "plan_create_many (design " + std::to_string(g) + "): " : "plan_create: ";
        if (status[g] == 1) return fail(-2, who + "X'R^-1X is not positive definite (rank-deficient design)");
        if (status[g] == 2) return fail(-2, who + "task block of (X'R^-1X)^-1 is singular");
    }
    for (PlanPending& q : pend) {
        if (!q.gls_only && !q.P->d.nseg)  // window sets: set projection, p <= 96
            if (int rc = build_windows(q.P, s)) return rc;
        q.P->pinvT = q.dpinv;
        q.P->ridge = q.ridge;
        q.P->ldR = q.ldR;
        q.P->ldX = q.ldX;
    }
    return 0;
}

static int plan_create_impl(ffs_reml_plan** out, int32_t T, int32_t p, const double* X, const double* pinvT,
                            const int32_t* gap, int32_t K, const double* ab, const int32_t* white, int32_t n_a,
                            int32_t n_b, const int32_t* lut, int32_t zero_a, int32_t zero_b, int32_t ltop,
                            int32_t n_task, const int32_t* task_idx, double ridge, void* stream) {
    if (!out) return fail(-3, "plan_create: null output handle");
    *out = nullptr;
    hipStream_t s = (hipStream_t)stream;
    auto* P = new ffs_reml_plan();
    std::unique_ptr<ffs_reml_plan> guard(P);
    std::vector<PlanPending> pend(1);
    std::vector<TablesOut> touts(1);
    KTablesArgs ka{};
    if (int rc = plan_alloc(P, T, p, X, pinvT, gap, K, ab, white, n_a, n_b, lut, zero_a, zero_b, ltop, n_task,
                            task_idx, ridge, s, &touts[0], &ka, &pend[0]))
        return rc;
    if (int rc = plans_build(pend, touts, ka, s)) return rc;
    *out = guard.release();
    return 0;
}

extern "C" int ffs_reml_plan_create_many(ffs_reml_plan** out, int32_t G, int32_t T, int32_t p, const double* X,
                                         const double* pinvT, const int32_t* gap, int32_t K, const double* ab,
                                         const int32_t* white, int32_t n_a, int32_t n_b, const int32_t* lut,
                                         int32_t zero_a, int32_t zero_b, int32_t ltop, int32_t n_task,
                                         const int32_t* task_idx, void* stream) {
    if (!out) return fail(-3, "plan_create_many: null output array");
    if (G < 1) return fail(-3, "plan_create_many: need G >= 1");
    if (!X || !pinvT) return fail(-3, "plan_create_many: null design");
    for (int g = 0; g < G; ++g) out[g] = nullptr;
    hipStream_t s = (hipStream_t)stream;
    std::vector<std::unique_ptr<ffs_reml_plan>> owned;
    std::vector<PlanPending> pend(G);
    std::vector<TablesOut> touts(G);
    KTablesArgs ka{};
    for (int g = 0; g < G; ++g) {
        owned.emplace_back(new ffs_reml_plan());
        KTablesArgs kg{};
        if (int rc = plan_alloc(owned.back().get(), T, p, X + (size_t)g * T * p, pinvT + (size_t)g * T * p, gap, K,
                                ab, white, n_a, n_b, lut, zero_a, zero_b, ltop, n_task, task_idx, 0.0, s, &touts[g],
                                &kg, &pend[g]))
            return rc;
        if (g == 0) ka = kg;
    }
    if (int rc = plans_build(pend, touts, ka, s)) return rc;
    for (int g = 0; g < G; ++g) out[g] = owned[g].release();
    return 0;
}

extern "C" int ffs_reml_plan_destroy(ffs_reml_plan* plan) {
    delete plan;
    return 0;
}

extern "C" int ffs_reml_plan_logdets(const ffs_reml_plan* P, double* logdet_R, double* logdet_XRX) {
    if (!P) return fail(-3, "plan_logdets: null plan");
    if (logdet_R) HIP_TRY(hipMemcpy(logdet_R, P->ldR, sizeof(double) * P->d.K, hipMemcpyDeviceToHost));
    if (logdet_XRX) HIP_TRY(hipMemcpy(logdet_XRX, P->ldX, sizeof(double) * P->d.K, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int ffs_reml_plan_ainv(const ffs_reml_plan* P, double* ainv) {
    if (!P || !ainv) return fail(-3, "plan_ainv: null argument");
    HIP_TRY(hipMemcpy(ainv, P->d.Ainv, sizeof(double) * P->d.K * P->d.p * P->d.p, hipMemcpyDeviceToHost));
    return 0;
}

// Rows of the time-major buffer that ffs_reml_prep writes and ffs_reml_project
// reads: Tpad, or 2 Tpad when the projection runs on pre-split halves.
// prep writes the split halves of the residual (rows Tpad..2 Tpad-1 of Yt) in
// the split projection modes for every p <= 96 (f64 MFMA prep)
static bool split_halves(const ffs_reml_plan* P) { return (P->d.PGA != 0 || P->d.nseg) && P->d.proj_mode != 0; }
static int yt_rows(const ffs_reml_plan* P) { return split_halves(P) ? 2 * P->d.Tpad : P->d.Tpad; }

extern "C" int ffs_reml_plan_yt_rows(const ffs_reml_plan* P, int32_t* rows) {
    if (!P || !rows) return fail(-3, "plan_yt_rows: null argument");
    *rows = yt_rows(P);
    return 0;
}

extern "C" int ffs_reml_plan_dims(const ffs_reml_plan* P, int32_t* Tpad, int32_t* vox_pad, int32_t* K) {
    if (!P) return fail(-3, "plan_dims: null plan");
    if (Tpad) *Tpad = P->d.Tpad;
    if (vox_pad) *vox_pad = kVoxPad;
    if (K) *K = P->d.K;
    return 0;
}

static int64_t vpad_of(int64_t V) { return (V + kVoxPad - 1) / kVoxPad * kVoxPad; }

// Grouped layouts (ffs_reml_group_*): `plans` (device, one PlanDev per design)
// and `grp` (the design of each 64-column block, or of each projection voxel
// tile, or of each voxel for the GLS) select each block's tables; null for a
// single-design call.
template <int PM>
static void launch_prep(const ffs_reml_plan* P, const float* Y, int64_t ldY, int V, const int32_t* rows, int ncols,
                        int ncols_pad, float* Yt, int64_t ldYt, float* scale, int normalize, float* feats,
                        hipStream_t s, const PlanDev* plans, const int32_t* grp) {
    const dim3 grid((ncols_pad + 63) / 64), block(64);
    if (feats)
        hipLaunchKernelGGL((k_prep<PM, 10>), grid, block, 0, s, P->d.T, P->d.p, Y, ldY, V, rows, ncols, ncols_pad,
                           P->pinvT, P->d.X, normalize, Yt, ldYt, scale, feats, plans, grp);
    else
        hipLaunchKernelGGL((k_prep<PM, 0>), grid, block, 0, s, P->d.T, P->d.p, Y, ldY, V, rows, ncols, ncols_pad,
                           P->pinvT, P->d.X, normalize, Yt, ldYt, scale, feats, plans, grp);
}

static int prep_impl(const ffs_reml_plan* P, const float* Y, int64_t ldY, int64_t V, const int32_t* rows,
                     int64_t ncols, float* Yt, int64_t ldYt, float* scale, int normalize, float* feats,
                     hipStream_t s, const int32_t* ncols_live = nullptr, bool write_split = true,
                     bool record_split = true, const double* coef_in = nullptr, double* coef_out = nullptr,
                     const PlanDev* plans = nullptr, const int32_t* grp = nullptr) {
    const int64_t ncp = vpad_of(ncols);
    if (Yt && ldYt < ncp) return fail(-3, "prep: ldYt must be >= the padded column count");
    if (ldY < P->d.T) return fail(-3, "prep: ldY must be >= T");
    if (ncp > 0x7fffffff || V > 0x7fffffff) return fail(-3, "prep: batch too large; split it");
    const int p = P->d.p;
    const int v = (int)V, nc = (int)ncols, ncpi = (int)ncp;
    if (P->d.nseg) {
        // p > 96: FP64 prep, kSegPrepVox columns per block; writes rows
        // T..Tpad-1 as zeros, and the split halves in the split projection modes
        if (feats || coef_in || coef_out) return fail(-4, "prep: features and coefficients need p <= 96");
        if (plans) return fail(-4, "prep: grouped layouts need p <= 96");
        uint16_t *Ysh = nullptr, *Ysl = nullptr;
        if (write_split && Yt && P->d.proj_mode != 0) {
            Ysh = reinterpret_cast<uint16_t*>(Yt + (size_t)P->d.Tpad * ldYt);
            Ysl = Ysh + (size_t)P->d.Tpad * ldYt;
        }
        if (write_split && record_split) P->split_yt = Ysh ? Yt : nullptr;
        const int sbf = P->d.proj_mode == 1;
        const dim3 grid((unsigned)(ncp / kSegPrepVox)), block(256);
        const size_t lds = seg_prep_lds(p);
#define FFS_PREP_SEG(MJ_)                                                                                          \
    hipLaunchKernelGGL(k_prep_seg<MJ_>, grid, block, lds, s, P->d.T, P->d.Tpad, p, Y, ldY, v, rows, nc, ncpi,      \
                       P->pinvT, P->d.X, normalize, Yt, ldYt, scale, Ysh, Ysl, sbf)
        if (p <= 256) FFS_PREP_SEG(1);
        else if (p <= 512) FFS_PREP_SEG(2);
        else FFS_PREP_SEG(3);
#undef FFS_PREP_SEG
        HIP_TRY(hipGetLastError());
        return 0;
    }
    if (P->d.PG) {
        // f64 MFMA path: also writes the zero rows T..Tpad-1; in the split
        // projection modes also the hi / lo halves into rows Tpad..2 Tpad-1
        // of Yt (see split_rows_needed)
        uint16_t *Ysh = nullptr, *Ysl = nullptr;
        if (write_split && Yt && P->d.proj_mode != 0) {
            Ysh = reinterpret_cast<uint16_t*>(Yt + (size_t)P->d.Tpad * ldYt);
            Ysl = Ysh + (size_t)P->d.Tpad * ldYt;
        }
        if (write_split && record_split) P->split_yt = Ysh ? Yt : nullptr;
        const int sbf = P->d.proj_mode == 1;
        const dim3 grid((unsigned)(ncp / 64)), block(64 * kPrepWaves);
        if (coef_in && (feats || !scale)) return fail(-3, "prep: given coefficients need the scales and no features");
#define FFS_PREP(NT_, NF_, CI_)                                                                                     \
    hipLaunchKernelGGL((k_prep_mfma<NT_, NF_, CI_>), grid, block, 0, s, P->d, Y, ldY, v, rows, nc, ncpi, normalize, \
                       Yt, ldYt, scale, feats, ncols_live, Ysh, Ysl, sbf, coef_in, coef_out, plans, grp)
        if (P->d.PG == 16) {
            if (feats) FFS_PREP(1, 10, false);
            else if (coef_in) FFS_PREP(1, 0, true);
            else FFS_PREP(1, 0, false);
        } else {
            if (feats) FFS_PREP(2, 10, false);
            else if (coef_in) FFS_PREP(2, 0, true);
            else FFS_PREP(2, 0, false);
        }
#undef FFS_PREP
        HIP_TRY(hipGetLastError());
        return 0;
    }
    if (coef_in || coef_out) return fail(-4, "prep: residual coefficients are kept only for p <= 32");
    if (!feats && P->d.PGA >= 48 && P->d.PGA <= 96) {
        // 32 < p <= 96 on f64 MFMA (writes the zero rows T..Tpad-1 too, and
        // the split halves in the split projection modes)
        uint16_t *Ysh = nullptr, *Ysl = nullptr;
        if (write_split && Yt && P->d.proj_mode != 0) {
            Ysh = reinterpret_cast<uint16_t*>(Yt + (size_t)P->d.Tpad * ldYt);
            Ysl = Ysh + (size_t)P->d.Tpad * ldYt;
        }
        if (write_split && record_split) P->split_yt = Ysh ? Yt : nullptr;
        const int sbf = P->d.proj_mode == 1;
        const dim3 grid((unsigned)(ncp / 64)), block(64 * kPrepWaves);
#define FFS_PREPW(NT_)                                                                                             \
    hipLaunchKernelGGL((k_prep_mfma_w<NT_>), grid, block, 0, s, P->d, Y, ldY, v, rows, nc, ncpi, normalize, Yt, \
                       ldYt, scale, ncols_live, Ysh, Ysl, sbf, plans, grp)
        switch (P->d.PGA / 16) {
            case 3: FFS_PREPW(3); break;
            case 4: FFS_PREPW(4); break;
            case 5: FFS_PREPW(5); break;
            default: FFS_PREPW(6); break;
        }
#undef FFS_PREPW
        HIP_TRY(hipGetLastError());
        return 0;
    }
    if (p <= 8)
        launch_prep<8>(P, Y, ldY, v, rows, nc, ncpi, Yt, ldYt, scale, normalize, feats, s, plans, grp);
    else if (p <= 16)
        launch_prep<16>(P, Y, ldY, v, rows, nc, ncpi, Yt, ldYt, scale, normalize, feats, s, plans, grp);
    else if (p <= 32)
        launch_prep<32>(P, Y, ldY, v, rows, nc, ncpi, Yt, ldYt, scale, normalize, feats, s, plans, grp);
    else
        launch_prep<96>(P, Y, ldY, v, rows, nc, ncpi, Yt, ldYt, scale, normalize, feats, s, plans, grp);
    HIP_TRY(hipGetLastError());
    // rows T..Tpad-1 of the time-major buffer are zero (GEMM k padding)
    if (Yt && P->d.Tpad > P->d.T) {
        const int64_t n = (int64_t)(P->d.Tpad - P->d.T) * ncp;
        hipLaunchKernelGGL(k_zero_rows, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, Yt, ldYt, P->d.T,
                           P->d.Tpad, ncp);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

extern "C" int ffs_reml_prep(const ffs_reml_plan* P, const float* Y, int64_t ldY, int64_t V, const int32_t* rows,
                             int64_t ncols, float* Yt, int64_t ldYt, float* scale, int32_t normalize, float* feats,
                             void* stream) {
    if (!P) return fail(-3, "prep: null plan");
    if (!rows) ncols = V;
    if (ncols <= 0) return 0;
    return prep_impl(P, Y, ldY, V, rows, ncols, Yt, ldYt, scale, normalize, feats, (hipStream_t)stream);
}

extern "C" int ffs_reml_plan_coef_stride(const ffs_reml_plan* P, int32_t* stride) {
    if (!P || !stride) return fail(-3, "plan_coef_stride: null argument");
    *stride = P->d.PG;
    return 0;
}

extern "C" int ffs_reml_prep_features(const ffs_reml_plan* P, const float* Y, int64_t ldY, int64_t V, float* scale,
                                      int32_t normalize, float* feats, double* coef, void* stream) {
    if (!P) return fail(-3, "prep_features: null plan");
    FFS_NEED_NARROW(P, "prep_features");
    if (!scale || !feats) return fail(-3, "prep_features: scale and feats are required");
    if (V <= 0) return 0;
    return prep_impl(P, Y, ldY, V, nullptr, V, nullptr, 0, scale, normalize, feats, (hipStream_t)stream, nullptr,
                     false, false, nullptr, coef);
}

// n_live (nullable, device): voxels past *n_live are skipped (grid sized for V).
static int energy_impl(const ffs_reml_plan* P, const float* Yt, int64_t ldYt, int64_t V, float* E, int64_t ldE,
                       const int32_t* n_live, hipStream_t s) {
    const int64_t Vpad = vpad_of(V);
    if (ldYt < Vpad) return fail(-3, "energy: ldYt must be >= the padded voxel count");
    if (ldE < P->d.K) return fail(-3, "energy: ldE must be >= K");
    const int64_t n_xt = (V + kEnergyVox - 1) / kEnergyVox;
    const int64_t n_xo = (n_xt + kEnergyWaves - 1) / kEnergyWaves;
    const int64_t nblk = 8 * ((n_xo * P->d.n_chunks + 7) / 8);  // see k_energy's item map
    if (nblk > 0x7fffffff) return fail(-3, "energy: grid too large; split the voxel batch");
    hipLaunchKernelGGL(k_energy, dim3((unsigned)nblk), dim3(kEnergyVox * kEnergyWaves), 0, s, P->d, Yt, ldYt, (int)V,
                       E, ldE, n_live);
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int ffs_reml_energy(const ffs_reml_plan* P, const float* Yt, int64_t ldYt, int64_t V, float* E,
                               int64_t ldE, void* stream) {
    if (!P) return fail(-3, "energy: null plan");
    FFS_NEED_SEARCH(P, "energy");
    if (V <= 0) return 0;
    return energy_impl(P, Yt, ldYt, V, E, ldE, nullptr, (hipStream_t)stream);
}

static int project_impl(const ffs_reml_plan* P, const float* Yt, int64_t ldYt, int64_t V, float* S, int64_t ldS,
                        const int32_t* n_live, bool presplit, hipStream_t s, const PlanDev* plans = nullptr,
                        const int32_t* grp = nullptr) {
    const int64_t Vpad = vpad_of(V);
    if (ldYt < Vpad) return fail(-3, "project: ldYt must be >= the padded voxel count");
    if (ldS < P->d.K) return fail(-3, "project: ldS must be >= K");
    const int64_t nvt = (V + kProjBM - 1) / kProjBM;
    const int64_t grp8 = 8 * ((P->d.n_tiles + 7) / 8);  // see xcd_tile_map
    const int64_t nb = nvt * grp8;
    if (nb > 0x7fffffff) return fail(-3, "project: grid too large; split the voxel batch");
    const dim3 g((unsigned)nb), b(256);
    const int mode = P->d.proj_mode;
    if (P->d.nseg) {
        // p > 96: fp32 MFMA over each pair's segments, items (voxel tile, pair)
        if (plans) return fail(-4, "project: grouped layouts need p <= 96");
        if (mode != 0) {
            // split 16-bit modes: 256-voxel items on the halves prep wrote
            if (!presplit)
                return fail(-4, "project: the split modes for p > 96 read the halves that ffs_reml_prep wrote into "
                                "this buffer in the same mode");
            const int64_t nb16s = (V + kProj16BM - 1) / kProj16BM * 8 * ((P->d.K + 7) / 8);
            if (nb16s > 0x7fffffff) return fail(-3, "project: grid too large; split the voxel batch");
            if (mode == 1)
                hipLaunchKernelGGL((k_project16_seg<true>), dim3((unsigned)nb16s), b, 0, s, P->d, Yt, ldYt, (int)V, S,
                                   ldS, n_live);
            else
                hipLaunchKernelGGL((k_project16_seg<false>), dim3((unsigned)nb16s), b, 0, s, P->d, Yt, ldYt, (int)V, S,
                                   ldS, n_live);
            HIP_TRY(hipGetLastError());
            return 0;
        }
        const int64_t nbs = nvt * 8 * ((P->d.K + 7) / 8);
        if (nbs > 0x7fffffff) return fail(-3, "project: grid too large; split the voxel batch");
        hipLaunchKernelGGL(k_project_seg, dim3((unsigned)nbs), b, 0, s, P->d, Yt, ldYt, (int)V, S, ldS, n_live);
        HIP_TRY(hipGetLastError());
        return 0;
    }
    // split modes: 256-voxel blocks
    const int64_t nb16 = (V + kProj16BM - 1) / kProj16BM * grp8;
    const dim3 g16((unsigned)nb16);
    const int v = (int)V;
    if (P->d.BN == 64) {
        if (mode == 1 && presplit)
            hipLaunchKernelGGL((k_project16b<64, true>), g16, b, 0, s, P->d, Yt, ldYt, v, S, ldS, n_live, plans, grp);
        else if (mode == 2 && presplit)
            hipLaunchKernelGGL((k_project16b<64, false>), g16, b, 0, s, P->d, Yt, ldYt, v, S, ldS, n_live, plans, grp);
        else if (mode == 1)
            hipLaunchKernelGGL((k_project16<64, true>), g, b, 0, s, P->d, Yt, ldYt, v, S, ldS, n_live, plans, grp);
        else if (mode == 2)
            hipLaunchKernelGGL((k_project16<64, false>), g, b, 0, s, P->d, Yt, ldYt, v, S, ldS, n_live, plans, grp);
        else
            hipLaunchKernelGGL(k_project<64>, g, b, 0, s, P->d, Yt, ldYt, v, S, ldS, n_live, plans, grp);
    } else {
        if (mode == 1 && presplit)
            hipLaunchKernelGGL((k_project16b<96, true>), g16, b, 0, s, P->d, Yt, ldYt, v, S, ldS, n_live, plans, grp);
        else if (mode == 2 && presplit)
            hipLaunchKernelGGL((k_project16b<96, false>), g16, b, 0, s, P->d, Yt, ldYt, v, S, ldS, n_live, plans, grp);
        else if (mode == 1)
            hipLaunchKernelGGL((k_project16<96, true>), g, b, 0, s, P->d, Yt, ldYt, v, S, ldS, n_live, plans, grp);
        else if (mode == 2)
            hipLaunchKernelGGL((k_project16<96, false>), g, b, 0, s, P->d, Yt, ldYt, v, S, ldS, n_live, plans, grp);
        else
            hipLaunchKernelGGL(k_project<96>, g, b, 0, s, P->d, Yt, ldYt, v, S, ldS, n_live, plans, grp);
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int ffs_reml_project(const ffs_reml_plan* P, const float* Yt, int64_t ldYt, int64_t V, float* S,
                                int64_t ldS, void* stream) {
    if (!P) return fail(-3, "project: null plan");
    FFS_NEED_SEARCH(P, "project");
    if (V <= 0) return 0;
    // split modes read the halves that prep wrote into this buffer (f64 MFMA
    // prep, p <= 96); otherwise they split the residual in the kernel
    const bool presplit = split_halves(P) && Yt == P->split_yt;
    return project_impl(P, Yt, ldYt, V, S, ldS, nullptr, presplit, (hipStream_t)stream);
}

// Energy and projection of voxel chunks pipelined over two streams: the
// energy of chunk i+1 (VALU-bound, auxiliary stream) runs while the projection
// of chunk i (MFMA-bound, caller's stream) does.  The caller's stream is
// ordered after all of it on return.
extern "C" int ffs_reml_surface(const ffs_reml_plan* P, const float* Yt, int64_t ldYt, int64_t V, float* S,
                                int64_t ldS, int32_t n_chunks, void* stream) {
    if (!P) return fail(-3, "surface: null plan");
    FFS_NEED_SEARCH(P, "surface");
    if (V <= 0) return 0;
    if (n_chunks < 1 || n_chunks > ffs_reml_plan::kMaxSurfaceChunks)
        return fail(-3, "surface: n_chunks must be in [1, 16]");
    const int64_t chunk = (V + n_chunks - 1) / n_chunks;
    const int64_t step = (chunk + kVoxPad - 1) / kVoxPad * kVoxPad;  // column offsets stay tile-aligned
    if (n_chunks == 1 || step >= V) {
        if (int rc = ffs_reml_energy(P, Yt, ldYt, V, S, ldS, stream)) return rc;
        return ffs_reml_project(P, Yt, ldYt, V, S, ldS, stream);
    }
    hipStream_t s = (hipStream_t)stream;
    if (!P->aux) {
        int dev = 0;
        HIP_TRY(hipGetDevice(&dev));
        if (dev != P->device) return fail(-3, "surface: the plan belongs to another device");
        HIP_TRY(hipStreamCreateWithFlags(&P->aux, hipStreamNonBlocking));
        for (hipEvent_t& e : P->ev) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    // the auxiliary stream starts after the caller's preceding work (e.g. prep)
    HIP_TRY(hipEventRecord(P->ev[ffs_reml_plan::kMaxSurfaceChunks], s));
    HIP_TRY(hipStreamWaitEvent(P->aux, P->ev[ffs_reml_plan::kMaxSurfaceChunks], 0));
    int i = 0;
    for (int64_t v0 = 0; v0 < V; v0 += step, ++i) {
        const int64_t n = std::min(step, V - v0);
        if (int rc = ffs_reml_energy(P, Yt + v0, ldYt, n, S + v0 * ldS, ldS, P->aux)) return rc;
        HIP_TRY(hipEventRecord(P->ev[i], P->aux));
        HIP_TRY(hipStreamWaitEvent(s, P->ev[i], 0));
        if (int rc = ffs_reml_project(P, Yt + v0, ldYt, n, S + v0 * ldS, ldS, stream)) return rc;
    }
    return 0;
}

extern "C" int ffs_reml_select(const ffs_reml_plan* P, const float* S, int64_t ldS, int64_t V, int32_t mode,
                               int32_t* best, float* best_score, void* stream) {
    if (!P) return fail(-3, "select: null plan");
    FFS_NEED_SEARCH(P, "select");
    if (mode < 0 || mode > 2)
        return fail(-3, "select: mode must be 0 (exhaustive), 1 (hierarchical) or 2 (subset descent)");
    if (V <= 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    if (mode == 0)
        hipLaunchKernelGGL(k_select_first<float>, dim3((unsigned)((V + 3) / 4)), dim3(256), 0, s, P->d, S, ldS, (int)V,
                           nullptr, best, best_score);
    else
        hipLaunchKernelGGL(k_select<float>, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, s, P->d, S, ldS, (int)V, mode,
                           best, best_score);
    HIP_TRY(hipGetLastError());
    return 0;
}

template <int PM>
static void launch_gls(const ffs_reml_plan* P, const float* Y, int64_t ldY, const float* scale, const int32_t* perm,
                       const int32_t* best, int V, float* betas, float* tstats, float* fstats, float* sigma2,
                       float* r2, double* resid_w, int64_t ldR, float* resid, float* pred, hipStream_t s,
                       const PlanDev* plans = nullptr, const int32_t* vgroup = nullptr) {
    hipLaunchKernelGGL(k_gls<PM>, dim3((V + 63) / 64), dim3(64), 0, s, P->d, Y, ldY, scale,
                       perm, best, V, betas, tstats, fstats, sigma2, r2, resid_w, ldR, resid, pred, plans, vgroup);
}

// Voxels grouped by key (see k_bucket_count): cnt[nb], off[nb + 1], cursor[nb],
// perm[cap] with cap >= V + nb * (pad - 1).  Returns 0 or an error status.
static int group_by_key(const int32_t* keys, int64_t V, int nb, int pad, int32_t* cnt, int32_t* off,
                        int32_t* cursor, int32_t* perm, int64_t cap, hipStream_t s) {
    if (nb < 1 || pad < 1) return fail(-3, "group_by_key: bad bucket count or padding");
    // only non-empty groups are padded: at most min(V, nb) of them
    if (V + std::min<int64_t>(V, nb) * (pad - 1) > cap) return fail(-3, "group_by_key: permutation buffer too small");
    hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, s, cnt, (int64_t)nb, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((cap + 255) / 256)), dim3(256), 0, s, perm, cap, -1);
    if (V > 0) {
        if (nb <= kSortLdsMax) {
            const unsigned g = (unsigned)((V + kSortBlock * kSortItems - 1) / (kSortBlock * kSortItems));
            hipLaunchKernelGGL(k_bucket_count, dim3(g), dim3(kSortBlock), sizeof(int32_t) * nb, s, keys, (int)V, nb,
                               cnt);
            hipLaunchKernelGGL(k_bucket_scan, dim3(1), dim3(1024), 0, s, cnt, nb, pad, off, cursor);
            hipLaunchKernelGGL(k_bucket_place, dim3(g), dim3(kSortBlock), 2 * sizeof(int32_t) * nb, s, keys, (int)V,
                               nb, cursor, perm);
        } else {
            const unsigned g = (unsigned)((V + 255) / 256);
            hipLaunchKernelGGL(k_bucket_count_global, dim3(g), dim3(256), 0, s, keys, (int)V, nb, cnt);
            hipLaunchKernelGGL(k_bucket_scan, dim3(1), dim3(1024), 0, s, cnt, nb, pad, off, cursor);
            hipLaunchKernelGGL(k_bucket_place_global, dim3(g), dim3(256), 0, s, keys, (int)V, nb, cursor, perm);
        }
    } else {
        hipLaunchKernelGGL(k_bucket_scan, dim3(1), dim3(1024), 0, s, cnt, nb, pad, off, cursor);
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

// Wide f64 MFMA GLS (32 < p <= 96) over `cap` slots of 64-slot (pair) groups.
static int launch_gls_wide(const PlanDev& d, const float* Y, int64_t ldY, const float* scale, const int32_t* perm,
                           const int32_t* best, int64_t cap, float* betas, float* tstats, float* fstats, float* sigma2,
                           float* r2, const PlanDev* plans, const int32_t* vgroup, hipStream_t s) {
    const dim3 g((unsigned)(cap / 64)), b(256);
    switch (d.PGA) {
        case 48:
            hipLaunchKernelGGL(k_gls_wide<3>, g, b, 0, s, d, Y, ldY, scale, perm, best, betas, tstats, fstats, sigma2,
                               r2, plans, vgroup);
            break;
        case 64:
            hipLaunchKernelGGL(k_gls_wide<4>, g, b, 0, s, d, Y, ldY, scale, perm, best, betas, tstats, fstats, sigma2,
                               r2, plans, vgroup);
            break;
        case 80:
            hipLaunchKernelGGL(k_gls_wide<5>, g, b, 0, s, d, Y, ldY, scale, perm, best, betas, tstats, fstats, sigma2,
                               r2, plans, vgroup);
            break;
        case 96:
            hipLaunchKernelGGL(k_gls_wide<6>, g, b, 0, s, d, Y, ldY, scale, perm, best, betas, tstats, fstats, sigma2,
                               r2, plans, vgroup);
            break;
        default:
            return fail(-3, "gls: no wide GLS kernel for this design width");
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

// int32 workspace of the GLS call: cnt, off, cursor, perm (64-slot padded groups).
static int64_t gls_perm_cap(int64_t V, int K) { return (V + std::min<int64_t>(V, K) * 63 + 63) / 64 * 64; }
static int64_t gls_work_ints(int64_t V, int K) { return 3 * (int64_t)K + 1 + gls_perm_cap(V, K); }

extern "C" int ffs_reml_gls_workspace(const ffs_reml_plan* P, int64_t V, int64_t* n_int32) {
    if (!P || !n_int32) return fail(-3, "gls_workspace: null argument");
    *n_int32 = gls_work_ints(std::max<int64_t>(V, 0), P->d.K);
    return 0;
}

static int gls_impl(const ffs_reml_plan* P, const float* Y, int64_t ldY, const float* scale, const int32_t* best,
                    int64_t V, float* betas, float* tstats, float* fstats, float* sigma2, float* r2, double* resid_w,
                    int64_t ldR, float* resid, float* pred, int32_t* work, hipStream_t s);

extern "C" int ffs_reml_gls(const ffs_reml_plan* P, const float* Y, int64_t ldY, const float* scale,
                            const int32_t* best, int64_t V, float* betas, float* tstats, float* fstats,
                            float* sigma2, float* r2, double* resid_w, int64_t ldR, float* resid, float* pred,
                            int32_t* work, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (int rc = gls_impl(P, Y, ldY, scale, best, V, betas, tstats, fstats, sigma2, r2, resid_w, ldR, resid, pred, work,
                          s))
        return rc;
    if (V > 0 && P->d.Pi && !P->d.nseg) {
        // rank-deficient design: minimum-norm beta and its statistics (the
        // segmented GLS computes beta with the pseudo-inverse directly)
        hipLaunchKernelGGL(k_gls_rank, dim3((unsigned)V), dim3(64), 2 * sizeof(double) * P->d.p, s, P->d, (int)V, scale, best, betas, tstats,
                           fstats, sigma2, nullptr, nullptr);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

extern "C" int32_t ffs_reml_plan_rank(const ffs_reml_plan* P) { return P ? P->d.rank : -1; }

static int gls_impl(const ffs_reml_plan* P, const float* Y, int64_t ldY, const float* scale, const int32_t* best,
                    int64_t V, float* betas, float* tstats, float* fstats, float* sigma2, float* r2, double* resid_w,
                    int64_t ldR, float* resid, float* pred, int32_t* work, hipStream_t s) {
    if (!P) return fail(-3, "gls: null plan");
    if (P->ridge > 0.0) return fail(-3, "gls: a ridge plan is search-only");
    if (V <= 0) return 0;
    if (V > 0x7fffffff - 64 * (int64_t)P->d.K) return fail(-3, "gls: batch too large; split it");
    if (ldY < P->d.T) return fail(-3, "gls: ldY must be >= T");
    if ((resid_w || resid || pred) && ldR < P->d.T) return fail(-3, "gls: ldR must be >= T");
    const int K = P->d.K, p = P->d.p;
    hipLaunchKernelGGL(k_gls_invalid, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, s, best, (int)V, K, p, betas,
                       tstats, fstats, sigma2, r2);
    const bool series_out = resid_w || resid || pred;
    if (P->d.nseg) {
        // p > 96: nv voxels of one pair per wave (groups padded to nv slots;
        // 4 for p <= 256, 2 above, so the per-voxel columns stay in registers)
        // when a workspace is given, else one voxel per wave
        const int nvg = p <= 256 ? kSegGlsVox : kSegGlsVox / 2;
        int32_t* perm = nullptr;
        int64_t n = V;
        if (work) {
            int32_t* cnt = work;
            int32_t* off = cnt + K;
            int32_t* cursor = off + K + 1;
            perm = cursor + K;
            n = gls_perm_cap(V, K);
            if (int rc = group_by_key(best, V, K, nvg, cnt, off, cursor, perm, n, s)) return rc;
        }
        const int nv = perm ? nvg : 1;
        const int64_t waves = (n + nv - 1) / nv;
        const dim3 g((unsigned)((waves + kSegGlsWaves - 1) / kSegGlsWaves)), b(64 * kSegGlsWaves);
#define FFS_GLS_SEG(MJ_, NV_)                                                                                   \
    hipLaunchKernelGGL((k_gls_seg<MJ_, NV_>), g, b, 0, s, P->d, Y, ldY, scale, perm, best, (int)n, (int)V, betas, \
                       tstats, fstats, sigma2, r2, resid_w, ldR, resid, pred)
        if (perm) {
            if (p <= 256) FFS_GLS_SEG(4, kSegGlsVox);
            else if (p <= 512) FFS_GLS_SEG(8, kSegGlsVox / 2);
            else FFS_GLS_SEG(12, kSegGlsVox / 2);
        } else {
            if (p <= 256) FFS_GLS_SEG(4, 1);
            else if (p <= 512) FFS_GLS_SEG(8, 1);
            else FFS_GLS_SEG(12, 1);
        }
#undef FFS_GLS_SEG
        HIP_TRY(hipGetLastError());
        return 0;
    }
    if (work && P->d.PG && !series_out) {
        // pair-uniform waves: groups padded to 64 slots
        int32_t* cnt = work;
        int32_t* off = cnt + K;
        int32_t* cursor = off + K + 1;
        int32_t* perm = cursor + K;
        const int64_t cap = gls_perm_cap(V, K);
        if (int rc = group_by_key(best, V, K, 64, cnt, off, cursor, perm, cap, s)) return rc;
        const dim3 g((unsigned)(cap / 64)), b(64);
        if (P->d.PG == 16)
            hipLaunchKernelGGL(k_gls_mfma<1>, g, b, 0, s, P->d, Y, ldY, scale, perm, best, betas, tstats, fstats,
                               sigma2, r2, nullptr, nullptr);
        else
            hipLaunchKernelGGL(k_gls_mfma<2>, g, b, 0, s, P->d, Y, ldY, scale, perm, best, betas, tstats, fstats,
                               sigma2, r2, nullptr, nullptr);
        HIP_TRY(hipGetLastError());
        return 0;
    }
    if (work && !series_out) {
        // 32 < p <= 96: wide f64 MFMA GLS, four waves per 64-slot pair group
        int32_t* cnt = work;
        int32_t* off = cnt + K;
        int32_t* cursor = off + K + 1;
        int32_t* perm = cursor + K;
        const int64_t cap = gls_perm_cap(V, K);
        if (int rc = group_by_key(best, V, K, 64, cnt, off, cursor, perm, cap, s)) return rc;
        if (int rc = launch_gls_wide(P->d, Y, ldY, scale, perm, best, cap, betas, tstats, fstats, sigma2, r2, nullptr,
                                     nullptr, s))
            return rc;
        return 0;
    }
    int32_t* perm = nullptr;
    if (work) {
        // contiguous groups (no padding); invalid pairs are dropped from the order
        int32_t* cnt = work;
        int32_t* off = cnt + K;
        int32_t* cursor = off + K + 1;
        perm = cursor + K;
        if (int rc = group_by_key(best, V, K, 1, cnt, off, cursor, perm, gls_perm_cap(V, K), s)) return rc;
    }
    if (p <= 8)
        launch_gls<8>(P, Y, ldY, scale, perm, best, (int)V, betas, tstats, fstats, sigma2, r2, resid_w, ldR, resid, pred, s);
    else if (p <= 16)
        launch_gls<16>(P, Y, ldY, scale, perm, best, (int)V, betas, tstats, fstats, sigma2, r2, resid_w, ldR, resid, pred, s);
    else if (p <= 32)
        launch_gls<32>(P, Y, ldY, scale, perm, best, (int)V, betas, tstats, fstats, sigma2, r2, resid_w, ldR, resid, pred, s);
    else
        launch_gls<96>(P, Y, ldY, scale, perm, best, (int)V, betas, tstats, fstats, sigma2, r2, resid_w, ldR, resid, pred, s);
    HIP_TRY(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// Allocation-free GLS: the voxel grouping lives in a file-scope device array
// ---------------------------------------------------------------------------
namespace {
// int32 capacity of the grouping scratch of ffs_reml_gls_noalloc (8 MiB per GPU).
constexpr int64_t kNoAllocInts = FFS_REML_GLS_NOALLOC_MAX_INT32;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int32_t* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// The array's name is a host-side stand-in; kernels get the device address
// from hipGetSymbolAddress (one copy per GPU, each at its own address).
static __device__ __align__(256) int32_t scratch[kNoAllocInts];

static int scratch_for_device(int dev, int32_t** out) {
    if (dev < 0 || dev >= kMaxDevices) return fail(-3, "gls_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(-1, std::string("gls_noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<int32_t*>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

extern "C" int ffs_reml_gls_noalloc(const ffs_reml_plan* P, const float* Y, int64_t ldY, const float* scale,
                                    const int32_t* best, int64_t V, float* betas, float* tstats, float* fstats,
                                    float* sigma2, float* r2, double* resid_w, int64_t ldR, float* resid, float* pred,
                                    void* stream) {
    if (!P) return fail(-3, "gls_noalloc: null plan");
    if (V <= 0) return 0;
    const int64_t need = gls_work_ints(V, P->d.K);
    if (need > kNoAllocInts)
        return fail(-3, "gls_noalloc: V = " + std::to_string(V) + " with K = " + std::to_string(P->d.K) +
                            " needs " + std::to_string(need) + " int32 of grouping scratch; the maximum is " +
                            std::to_string(kNoAllocInts) + " (split the batch or use ffs_reml_gls)");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev != P->device) return fail(-3, "gls_noalloc: the plan belongs to another device");
    int32_t* work = nullptr;
    if (int rc = scratch_for_device(dev, &work)) return rc;
    return ffs_reml_gls(P, Y, ldY, scale, best, V, betas, tstats, fstats, sigma2, r2, resid_w, ldR, resid, pred, work,
                        stream);
}

// ---------------------------------------------------------------------------
// FP64 search (the reference's use_double search)
// ---------------------------------------------------------------------------
template <typename TY>
static void launch_surface_f64(const ffs_reml_plan* P, const TY* Y, int64_t ldY, const float* scale, int V, double* S,
                               int64_t ldS, hipStream_t s) {
    const dim3 g((unsigned)((int64_t)((V + 63) / 64) * P->d.K)), b(64);
    const int p = P->d.p;
    if (P->d.nseg)
        hipLaunchKernelGGL((k_surface_f64_seg<TY>), g, b, 0, s, P->d, Y, ldY, scale, V, S, ldS);
    else if (p <= 8)
        hipLaunchKernelGGL((k_surface_f64<8, TY>), g, b, 0, s, P->d, Y, ldY, scale, V, S, ldS);
    else if (p <= 16)
        hipLaunchKernelGGL((k_surface_f64<16, TY>), g, b, 0, s, P->d, Y, ldY, scale, V, S, ldS);
    else if (p <= 32)
        hipLaunchKernelGGL((k_surface_f64<32, TY>), g, b, 0, s, P->d, Y, ldY, scale, V, S, ldS);
    else
        hipLaunchKernelGGL((k_surface_f64<96, TY>), g, b, 0, s, P->d, Y, ldY, scale, V, S, ldS);
}

extern "C" int ffs_reml_surface_f64(const ffs_reml_plan* P, const void* Y, int32_t y_is_f64, int64_t ldY, int64_t V,
                                    const float* scale, double* S, int64_t ldS, void* stream) {
    if (!P) return fail(-3, "surface_f64: null plan");
    FFS_NEED_SEARCH(P, "surface_f64");
    if (V <= 0) return 0;
    if (!Y || !S) return fail(-3, "surface_f64: null argument");
    if (ldY < P->d.T) return fail(-3, "surface_f64: ldY must be >= T");
    if (ldS < P->d.K) return fail(-3, "surface_f64: ldS must be >= K");
    if ((int64_t)((V + 63) / 64) * P->d.K > 0x7fffffffLL) return fail(-3, "surface_f64: batch too large; split it");
    hipStream_t s = (hipStream_t)stream;
    if (y_is_f64)
        launch_surface_f64(P, static_cast<const double*>(Y), ldY, scale, (int)V, S, ldS, s);
    else
        launch_surface_f64(P, static_cast<const float*>(Y), ldY, scale, (int)V, S, ldS, s);
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int ffs_reml_select_f64(const ffs_reml_plan* P, const double* S, int64_t ldS, int64_t V, int32_t mode,
                                   int32_t* best, double* best_score, void* stream) {
    if (!P) return fail(-3, "select_f64: null plan");
    FFS_NEED_SEARCH(P, "select_f64");
    if (mode < 0 || mode > 2)
        return fail(-3, "select_f64: mode must be 0 (exhaustive), 1 (hierarchical) or 2 (subset descent)");
    if (V <= 0) return 0;
    if (ldS < P->d.K) return fail(-3, "select_f64: ldS must be >= K");
    hipStream_t s = (hipStream_t)stream;
    if (mode == 0)
        hipLaunchKernelGGL(k_select_first<double>, dim3((unsigned)((V + 3) / 4)), dim3(256), 0, s, P->d, S, ldS,
                           (int)V, nullptr, best, best_score);
    else
        hipLaunchKernelGGL(k_select<double>, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, s, P->d, S, ldS,
                           (int)V, mode, best, best_score);
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int ffs_reml_ljung_box(const double* resid_w, int64_t ldR, int64_t V, int32_t n, int32_t nbins,
                                  const int32_t* bin_off, const int32_t* pairs, const double* wbin, float* out,
                                  void* stream) {
    if (V <= 0) return 0;
    if (ldR < n) return fail(-3, "ljung_box: ldR must be >= n");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_ljung_box, dim3((unsigned)((V + 3) / 4)), dim3(256), 0, s, resid_w, ldR, (int)V, n, nbins,
                       bin_off, pairs, wbin, out);
    HIP_TRY(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// Pair-set search (learned path): 5 x 5 windows and lambda bands
// ---------------------------------------------------------------------------
namespace {
int host_clamp_center(int n_a, int n_b, int ai, int bi) {
    if (n_a >= kWin) ai = std::min(std::max(ai, kWinHalf), n_a - 1 - kWinHalf);
    if (n_b >= kWin) bi = std::min(std::max(bi, kWinHalf), n_b - 1 - kWinHalf);
    return ai * n_b + bi;
}

// Uploads a family of pair sets (pairs in pair order) with their edge / next
// flags and builds the energy chunk tables.  Energy chunks: the slots of a set
// are grouped by AR coefficient sequence (same a and lambda == 0 flag; groups
// in order of first slot), each group cut into pieces of gsize positions, and
// ngroups pieces make a chunk.
int build_set_table(ffs_reml_plan* P, const std::vector<std::vector<int32_t>>& sets,
                    const std::vector<std::vector<int32_t>>& edges, const std::vector<std::vector<int32_t>>* nexts,
                    SetTable* out, hipStream_t s, int gsize, int ngroups) {
    PlanDev& d = P->d;
    size_t maxn = 0;
    for (const auto& v : sets) maxn = std::max(maxn, v.size());
    const int unit = std::lcm(kEnergyNB, kSetWaves * d.npt);
    const int SB = (int)std::max<size_t>(unit, (maxn + unit - 1) / unit * unit);
    const int n = (int)sets.size();
    std::vector<int32_t> pairs((size_t)n * SB, -1), edge((size_t)n * SB, 0), next, cnt(n, 0);
    if (nexts) next.assign((size_t)n * SB, -1);
    const int NB = gsize * ngroups;
    std::vector<std::vector<int32_t>> chunks(n);  // positions of each set's chunks
    int MC = 1;
    for (int i = 0; i < n; ++i) {
        cnt[i] = (int32_t)sets[i].size();
        for (size_t q = 0; q < sets[i].size(); ++q) {
            pairs[(size_t)i * SB + q] = sets[i][q];
            edge[(size_t)i * SB + q] = edges[i][q];
            if (nexts) next[(size_t)i * SB + q] = (*nexts)[i][q];
        }
        std::vector<std::vector<int32_t>> groups;
        for (size_t q = 0; q < sets[i].size(); ++q) {
            const int k = sets[i][q];
            size_t g = 0;
            for (; g < groups.size(); ++g) {
                const int k0 = sets[i][groups[g][0]];
                if (P->h_ab[2 * k0] == P->h_ab[2 * k] && P->h_white[k0] == P->h_white[k]) break;
            }
            if (g == groups.size()) groups.emplace_back();
            groups[g].push_back((int32_t)q);
        }
        std::vector<int32_t>& c = chunks[i];
        for (const auto& g : groups)
            for (size_t q = 0; q < g.size(); q += gsize)
                for (int j = 0; j < gsize; ++j) c.push_back(q + j < g.size() ? g[q + j] : -1);
        while (c.size() % NB) c.push_back(-1);
        MC = std::max(MC, (int)(c.size() / NB));
    }
    std::vector<int32_t> eslot((size_t)n * MC * NB, -1), ecnt(n);
    for (int i = 0; i < n; ++i) {
        ecnt[i] = (int32_t)(chunks[i].size() / NB);
        std::copy(chunks[i].begin(), chunks[i].end(), eslot.begin() + (size_t)i * MC * NB);
    }
    int32_t *dpairs, *dedge, *dcnt, *dnext = nullptr, *deslot, *decnt;
    float *coef, *s0;
    const size_t nchunks = (size_t)n * MC;
    int rc = 0;
    if ((rc = upload(P, &dpairs, pairs.data(), pairs.size(), s)) ||
        (rc = upload(P, &dedge, edge.data(), edge.size(), s)) || (rc = upload(P, &dcnt, cnt.data(), cnt.size(), s)) ||
        (nexts && (rc = upload(P, &dnext, next.data(), next.size(), s))) ||
        (rc = upload(P, &deslot, eslot.data(), eslot.size(), s)) ||
        (rc = upload(P, &decnt, ecnt.data(), ecnt.size(), s)) ||
        (rc = P->alloc(&coef, nchunks * d.T * (ngroups + 2 * NB))) || (rc = P->alloc(&s0, nchunks * NB)))
        return rc;
    SetTable t{};
    t.n = n;
    t.sb = SB;
    t.cnt = dcnt;
    t.pairs = dpairs;
    t.edge = dedge;
    t.next = dnext;
    t.enb = NB;
    t.eng = ngroups;
    t.mc = MC;
    t.ecnt = decnt;
    t.eslot = deslot;
    t.coef = coef;
    t.s0 = s0;
    hipLaunchKernelGGL(k_set_tables, dim3((unsigned)nchunks), dim3(128), 0, s, d, t, coef, s0);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));  // the host vectors above are uploaded
    *out = t;
    return 0;
}

// Windows of every lattice cell: the kWin x kWin lattice cells around it.
// Slots are the valid pairs among them in pair order (a-major, b-minor);
// edge = the slot lies on a window side that the lattice continues past (a
// better pair may lie beyond it, also past an invalid neighbour); next = the
// window centred on that slot (kept inside the lattice).
int build_windows(ffs_reml_plan* P, hipStream_t s) {
    const int n_a = P->d.n_a, n_b = P->d.n_b;
    std::vector<std::vector<int32_t>> sets((size_t)n_a * n_b), edges(sets.size()), nexts(sets.size());
    for (int ca = 0; ca < n_a; ++ca)
        for (int cb = 0; cb < n_b; ++cb) {
            const size_t w = (size_t)ca * n_b + cb;
            for (int r = 0; r < kWin; ++r)
                for (int c = 0; c < kWin; ++c) {
                    const int ai = ca - kWinHalf + r, bi = cb - kWinHalf + c;
                    if (ai < 0 || ai >= n_a || bi < 0 || bi >= n_b) continue;
                    const int k = P->h_lut[(size_t)ai * n_b + bi];
                    if (k < 0) continue;
                    const int e = (r == 0 && ai > 0) || (r == kWin - 1 && ai < n_a - 1) || (c == 0 && bi > 0) ||
                                  (c == kWin - 1 && bi < n_b - 1);
                    sets[w].push_back(k);
                    edges[w].push_back(e);
                    nexts[w].push_back(e ? host_clamp_center(n_a, n_b, ai, bi) : -1);
                }
        }
    // a window row (one a) forms one chunk of kWin = 5 positions
    static_assert(kWin == 5, "window rows use the 5-position energy chunk");
    return build_set_table(P, sets, edges, &nexts, &P->d.win, s, kWin, 1);
}

struct SetLayout {
    int64_t n, max_cols, max_tiles, yt_rows;
    size_t off_key, off_cnt, off_fill, off_cursor, off_tset, off_tcol, off_nt, off_rows, off_yt, off_s, total;
};
SetLayout set_layout(const ffs_reml_plan* P, const SetTable& ST, int64_t V) {
    SetLayout L{};
    L.n = std::max(ST.n, 1);
    const int64_t groups = std::min<int64_t>(V, L.n);
    L.max_cols = vpad_of(V + (kSetTile - 1) * groups);
    L.max_tiles = L.max_cols / kSetTile;
    L.yt_rows = yt_rows(P);
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o = align256(o + bytes);
        return at;
    };
    L.off_key = take(4 * (size_t)V);
    L.off_cnt = take(4 * (size_t)L.n);
    L.off_fill = take(4 * (size_t)(L.n + 1));
    L.off_cursor = take(4 * (size_t)L.n);
    L.off_tset = take(4 * (size_t)L.max_tiles);
    L.off_tcol = take(4 * (size_t)L.max_tiles);
    L.off_nt = take(4);
    L.off_rows = take(4 * (size_t)L.max_cols);
    L.off_yt = take(4 * (size_t)L.yt_rows * L.max_cols);
    L.off_s = take(4 * (size_t)L.max_cols * std::max(ST.sb, 1));
    L.total = o;
    return L;
}
struct SetPtrs {
    int32_t *key, *cnt, *fill, *cursor, *tset, *tcol, *nt, *rows;
    float *Yt, *S;
};
SetPtrs set_ptrs(const SetLayout& L, void* workspace) {
    char* ws = static_cast<char*>(workspace);
    SetPtrs q;
    q.key = reinterpret_cast<int32_t*>(ws + L.off_key);
    q.cnt = reinterpret_cast<int32_t*>(ws + L.off_cnt);
    q.fill = reinterpret_cast<int32_t*>(ws + L.off_fill);
    q.cursor = reinterpret_cast<int32_t*>(ws + L.off_cursor);
    q.tset = reinterpret_cast<int32_t*>(ws + L.off_tset);
    q.tcol = reinterpret_cast<int32_t*>(ws + L.off_tcol);
    q.nt = reinterpret_cast<int32_t*>(ws + L.off_nt);
    q.rows = reinterpret_cast<int32_t*>(ws + L.off_rows);
    q.Yt = reinterpret_cast<float*>(ws + L.off_yt);
    q.S = reinterpret_cast<float*>(ws + L.off_s);
    return q;
}
int set_check(const ffs_reml_plan* P, const SetTable& ST, const char* what, int64_t V, void* workspace,
              int64_t ws_bytes, SetLayout* L) {
    if (!P) return fail(-3, std::string(what) + ": null plan");
    FFS_NEED_SEARCH(P, what);
    if (ST.n <= 0) return fail(-4, std::string(what) + ": no sets on this plan");
    if (P->d.proj_mode != 0 && !split_halves(P))
        return fail(-4, std::string(what) + ": the split projection modes need p <= 96 here (use fp32)");
    *L = set_layout(P, ST, V);
    if (!workspace || ws_bytes < (int64_t)L->total) return fail(-3, std::string(what) + ": workspace too small");
    if (L->max_cols > 0x7fffffff) return fail(-3, std::string(what) + ": batch too large; split it");
    return 0;
}
// voxels grouped by q.key, residualised (with the split halves in the split
// modes) into the set layout
int set_group_prep(const ffs_reml_plan* P, const SetTable& ST, const float* Y, int64_t ldY, int64_t V, float* scale,
                   const double* coef, const SetLayout& L, const SetPtrs& q, hipStream_t s,
                   double* coef_out = nullptr) {
    if (int rc = group_by_key(q.key, V, (int)L.n, kSetTile, q.cnt, q.fill, q.cursor, q.rows, L.max_cols, s)) return rc;
    hipLaunchKernelGGL(k_group_tiles, dim3((unsigned)((L.max_tiles + 255) / 256)), dim3(256), 0, s, q.fill, (int)L.n,
                       (int)L.max_tiles, kSetTile, q.tset, q.tcol, q.nt);
    HIP_TRY(hipGetLastError());
    const bool split = split_halves(P);
    return prep_impl(P, Y, ldY, V, q.rows, L.max_cols, q.Yt, L.max_cols, scale, 1, nullptr, s, q.fill + L.n, split,
                     false, coef, coef_out);
}
// the same layout gathered from a residual buffer prepared for all V voxels
// (ldres columns, with the split halves in the split modes)
int set_group_gather(const ffs_reml_plan* P, const SetTable& ST, const float* Yres, int64_t ldres, int64_t V,
                     const SetLayout& L, const SetPtrs& q, hipStream_t s) {
    if (int rc = group_by_key(q.key, V, (int)L.n, kSetTile, q.cnt, q.fill, q.cursor, q.rows, L.max_cols, s)) return rc;
    hipLaunchKernelGGL(k_group_tiles, dim3((unsigned)((L.max_tiles + 255) / 256)), dim3(256), 0, s, q.fill, (int)L.n,
                       (int)L.max_tiles, kSetTile, q.tset, q.tcol, q.nt);
    const int64_t n = L.max_cols * (P->d.Tpad / 8);
    if (n > 0)
        hipLaunchKernelGGL(k_gather_cols, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, Yres, ldres, q.Yt,
                           L.max_cols, q.rows, (int)L.max_cols, P->d.Tpad, (int)split_halves(P));
    HIP_TRY(hipGetLastError());
    return 0;
}
int set_energy(const ffs_reml_plan* P, const SetTable& ST, const SetLayout& L, const SetPtrs& q, hipStream_t s) {
    // one block per (tile octet, chunk) of the largest tile count, 8 * ceil(items / 8)
    const int64_t items = (L.max_tiles + kSetEnergyWaves - 1) / kSetEnergyWaves * ST.mc;
    const dim3 g((unsigned)(8 * ((items + 7) / 8))), b(64 * kSetEnergyWaves);
    const int64_t ld = L.max_cols;
    if (ST.enb == 5 && ST.eng == 1)
        hipLaunchKernelGGL((k_set_energy<5, 1>), g, b, 0, s, P->d, ST, q.Yt, ld, q.tset, q.nt, q.S);
    else if (ST.enb == 3 && ST.eng == 1)
        hipLaunchKernelGGL((k_set_energy<3, 1>), g, b, 0, s, P->d, ST, q.Yt, ld, q.tset, q.nt, q.S);
    else if (ST.enb == kEnergyNB && ST.eng == kEnergyNB)
        hipLaunchKernelGGL((k_set_energy<kEnergyNB, kEnergyNB>), g, b, 0, s, P->d, ST, q.Yt, ld, q.tset, q.nt,
                           q.S);
    else
        return fail(-3, "set_energy: no kernel for this chunk shape");
    HIP_TRY(hipGetLastError());
    return 0;
}
// scores of every column's set pairs into q.S (the full-grid projection
// arithmetic over the residual tile, score epilogue over the set energies)
int set_scores(const ffs_reml_plan* P, const SetTable& ST, const SetLayout& L, const SetPtrs& q, hipStream_t s) {
    const int64_t items = L.max_tiles * (ST.sb / (kSetWaves * P->d.npt));
    const dim3 g((unsigned)std::min<int64_t>(1024, std::max<int64_t>(8, (items + 7) / 8 * 8))), b(64 * kSetWaves);
    const int64_t ld = L.max_cols;
    const int mode = P->d.proj_mode;
#define FFS_SET_PROJ(BN_, MODE_, SEG_) \
    hipLaunchKernelGGL((k_set_project<BN_, MODE_, SEG_>), g, b, 0, s, P->d, ST, q.Yt, ld, q.tset, q.nt, q.S)
    if (P->d.nseg) {  // p > 96: pairs over 64-column segments
        if (mode == 1) FFS_SET_PROJ(64, 1, true);
        else if (mode == 2) FFS_SET_PROJ(64, 2, true);
        else FFS_SET_PROJ(64, 0, true);
    } else if (P->d.BN == 64) {
        if (mode == 1) FFS_SET_PROJ(64, 1, false);
        else if (mode == 2) FFS_SET_PROJ(64, 2, false);
        else FFS_SET_PROJ(64, 0, false);
    } else {
        if (mode == 1) FFS_SET_PROJ(96, 1, false);
        else if (mode == 2) FFS_SET_PROJ(96, 2, false);
        else FFS_SET_PROJ(96, 0, false);
    }
#undef FFS_SET_PROJ
    HIP_TRY(hipGetLastError());
    return 0;
}
int set_project(const ffs_reml_plan* P, const SetTable& ST, const SetLayout& L, const SetPtrs& q, int32_t* best,
                float* best_score, int32_t* need, int32_t* next_key, hipStream_t s) {
    if (int rc = set_scores(P, ST, L, q, s)) return rc;
    hipLaunchKernelGGL(k_set_select, dim3((unsigned)(L.max_cols / 256)), dim3(256), 0, s, P->d, ST, q.rows, q.tset,
                       q.nt, q.S, best, best_score, need, next_key);
    HIP_TRY(hipGetLastError());
    return 0;
}
}  // namespace

extern "C" int ffs_reml_local_workspace(const ffs_reml_plan* P, int64_t V, int64_t* bytes) {
    if (!P || !bytes) return fail(-3, "local_workspace: null argument");
    FFS_NEED_NARROW(P, "local_workspace");
    *bytes = (int64_t)set_layout(P, P->d.win, std::max<int64_t>(V, 0)).total;
    return 0;
}

extern "C" int ffs_reml_local_prepare(const ffs_reml_plan* P, const float* Y, int64_t ldY, int64_t V,
                                      const float* props, float conf_min, int32_t* need_full, float* scale,
                                      const double* coef, void* workspace, int64_t ws_bytes, void* stream) {
    if (V <= 0) return 0;
    SetLayout L;
    if (int rc = set_check(P, P->d.win, "local search", V, workspace, ws_bytes, &L)) return rc;
    SetPtrs q = set_ptrs(L, workspace);
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_win_key, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, s, P->d, props, (int)V, conf_min,
                       q.key, need_full);
    HIP_TRY(hipGetLastError());
    return set_group_prep(P, P->d.win, Y, ldY, V, scale, coef, L, q, s);
}

extern "C" int ffs_reml_local_recenter(const ffs_reml_plan* P, const float* Y, int64_t ldY, int64_t V, float* scale,
                                       const double* coef, void* workspace, int64_t ws_bytes, void* stream) {
    if (V <= 0) return 0;
    SetLayout L;
    if (int rc = set_check(P, P->d.win, "local search", V, workspace, ws_bytes, &L)) return rc;
    return set_group_prep(P, P->d.win, Y, ldY, V, scale, coef, L, set_ptrs(L, workspace), (hipStream_t)stream);
}

extern "C" int ffs_reml_local_energy(const ffs_reml_plan* P, int64_t V, void* workspace, int64_t ws_bytes,
                                     void* stream) {
    if (V <= 0) return 0;
    SetLayout L;
    if (int rc = set_check(P, P->d.win, "local search", V, workspace, ws_bytes, &L)) return rc;
    return set_energy(P, P->d.win, L, set_ptrs(L, workspace), (hipStream_t)stream);
}

extern "C" int ffs_reml_local_project(const ffs_reml_plan* P, int64_t V, int32_t* best, float* best_score,
                                      int32_t* need_full, void* workspace, int64_t ws_bytes, void* stream) {
    if (V <= 0) return 0;
    SetLayout L;
    if (int rc = set_check(P, P->d.win, "local search", V, workspace, ws_bytes, &L)) return rc;
    SetPtrs q = set_ptrs(L, workspace);
    // the window to continue with (walk) replaces the voxel's key
    return set_project(P, P->d.win, L, q, best, best_score, need_full, q.key, (hipStream_t)stream);
}

extern "C" int ffs_reml_local_search(const ffs_reml_plan* P, const float* Y, int64_t ldY, int64_t V,
                                     const float* props, float conf_min, int32_t* best, float* best_score,
                                     int32_t* need_full, float* scale, const double* coef, void* workspace,
                                     int64_t ws_bytes, void* stream) {
    if (int rc = ffs_reml_local_prepare(P, Y, ldY, V, props, conf_min, need_full, scale, coef, workspace, ws_bytes,
                                        stream))
        return rc;
    if (int rc = ffs_reml_local_energy(P, V, workspace, ws_bytes, stream)) return rc;
    return ffs_reml_local_project(P, V, best, best_score, need_full, workspace, ws_bytes, stream);
}

extern "C" int ffs_reml_compact(const int32_t* flags, int64_t V, int32_t* idx, int32_t* count, void* stream) {
    if (V <= 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, count, (int64_t)1, 0);
    hipLaunchKernelGGL(k_compact_flags, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, s, flags, (int)V, idx, count);
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int ffs_reml_scatter_best(const int32_t* idx, int64_t n, const int32_t* best_sub, const float* score_sub,
                                     int32_t* best, float* score, void* stream) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_scatter_best, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, idx,
                       (int)n, nullptr, best_sub, score_sub, best, score);
    HIP_TRY(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// Hierarchical descent on pair sets
// ---------------------------------------------------------------------------
// The lattice points each level of the descent can visit, enumerated from the
// start point: pass 0 is one set holding the top level's points, the start
// point and every point of the level below's neighbourhoods around the points
// the top level can end on; each later level has one set per centre the
// previous levels can end on (its clamped 3 x 3 neighbourhood at the level's
// step, distinct valid pairs in visit order) and a map from the 9 visits to
// the set's slots.  Visit orders are those of k_select mode 1.
extern "C" int ffs_reml_plan_build_hier(ffs_reml_plan* P, void* stream) {
    if (!P) return fail(-3, "plan_build_hier: null plan");
    FFS_NEED_SEARCH(P, "plan_build_hier");
    if (!P->hier.empty()) return 0;
    hipStream_t s = (hipStream_t)stream;
    const int n_a = P->d.n_a, n_b = P->d.n_b, ltop = P->d.ltop;
    const int za = P->d.zero_a, zb = P->d.zero_b;
    auto kof = [&](int x, int y) { return P->h_lut[(size_t)x * n_b + y]; };
    auto nbh = [&](int cx, int cy, int lev) {
        const int step = 1 << lev;
        std::vector<std::pair<int, int>> out;
        for (int oy = -1; oy <= 1; ++oy) {
            const int y = std::min(std::max(cy + oy * step, 0), n_b - 1);
            for (int ox = -1; ox <= 1; ++ox) out.push_back({std::min(std::max(cx + ox * step, 0), n_a - 1), y});
        }
        return out;
    };
    const size_t NL = (size_t)n_a * n_b;
    std::vector<char> W(NL, 0), U(NL, 0);  // points a voxel can be at; points of pass 0's set
    W[(size_t)za * n_b + zb] = 1;
    U[(size_t)za * n_b + zb] = 1;
    const int st = 1 << ltop;
    for (int y = 0; y < n_b; y += st)
        for (int x = 0; x < n_a; x += st) {
            U[(size_t)x * n_b + y] = 1;
            if (kof(x, y) >= 0) W[(size_t)x * n_b + y] = 1;
        }
    if (ltop >= 1) {
        std::vector<char> W2(NL, 0);
        for (size_t c = 0; c < NL; ++c) {
            if (!W[c]) continue;
            W2[c] = 1;
            for (auto [x, y] : nbh((int)(c / n_b), (int)(c % n_b), ltop - 1)) {
                U[(size_t)x * n_b + y] = 1;
                if (kof(x, y) >= 0) W2[(size_t)x * n_b + y] = 1;
            }
        }
        W.swap(W2);
    }
    std::vector<HierPass> passes;
    {
        std::vector<std::vector<int32_t>> sets(1), edges(1);
        std::vector<int32_t> pslot(NL, -1);
        for (size_t c = 0; c < NL; ++c) {
            const int k = U[c] ? kof((int)(c / n_b), (int)(c % n_b)) : -1;
            if (k < 0) continue;
            pslot[c] = (int32_t)sets[0].size();
            sets[0].push_back(k);
            edges[0].push_back(0);
        }
        HierPass h{};
        // lattice rows of the top levels, 5 points to a chunk
        if (int rc = build_set_table(P, sets, edges, nullptr, &h.ST, s, 5, 1)) return rc;
        int32_t* d;
        if (int rc = upload(P, &d, pslot.data(), pslot.size(), s)) return rc;
        h.point_slot = d;
        h.level = ltop;
        h.first = 1;
        passes.push_back(h);
    }
    for (int lev = ltop - 2; lev >= 0; --lev) {
        std::vector<std::vector<int32_t>> sets, edges;
        std::vector<int32_t> cset(NL, -1), vmap;
        std::vector<char> W2(NL, 0);
        for (size_t c = 0; c < NL; ++c) {
            if (!W[c]) continue;
            W2[c] = 1;
            cset[c] = (int32_t)sets.size();
            std::vector<int32_t> pairs, edge;
            for (auto [x, y] : nbh((int)(c / n_b), (int)(c % n_b), lev)) {
                const int k = kof(x, y);
                int slot = -1;
                if (k >= 0) {
                    W2[(size_t)x * n_b + y] = 1;
                    for (size_t q = 0; q < pairs.size(); ++q)
                        if (pairs[q] == k) slot = (int)q;
                    if (slot < 0) {
                        slot = (int)pairs.size();
                        pairs.push_back(k);
                        edge.push_back(0);
                    }
                }
                vmap.push_back(slot);
            }
            sets.push_back(pairs);
            edges.push_back(edge);
        }
        HierPass h{};
        // a 3 x 3 neighbourhood: each of its rows (one a) forms a chunk
        if (int rc = build_set_table(P, sets, edges, nullptr, &h.ST, s, 3, 1)) return rc;
        int32_t *dc, *dv;
        if (int rc = upload(P, &dc, cset.data(), cset.size(), s)) return rc;
        if (int rc = upload(P, &dv, vmap.data(), vmap.size(), s)) return rc;
        HIP_TRY(hipStreamSynchronize(s));  // the host vectors above are uploaded
        h.center_set = dc;
        h.vmap = dv;
        h.level = lev;
        passes.push_back(h);
        W.swap(W2);
    }
    HIP_TRY(hipStreamSynchronize(s));
    P->hier = passes;
    return 0;
}

namespace {
struct HierLayout {
    size_t off_ba, off_bb, off_cur, off_coef, off_res, off_sets, total;
    int stride;
};
HierLayout hier_layout(const ffs_reml_plan* P, int64_t V) {
    HierLayout H{};
    H.stride = P->d.PG;  // residualisation coefficients per voxel (0: no f64 MFMA prep, p > 32)
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o = align256(o + bytes);
        return at;
    };
    H.off_ba = take(4 * (size_t)V);
    H.off_bb = take(4 * (size_t)V);
    H.off_cur = take(4 * (size_t)V);
    H.off_coef = take(8 * (size_t)V * H.stride);
    // p > 96: the residual of every voxel, prepared once (passes gather from it)
    H.off_res = take(P->d.nseg ? 4 * (size_t)yt_rows(P) * vpad_of(V) : 0);
    H.off_sets = o;
    size_t mx = 0;
    for (const HierPass& h : P->hier) mx = std::max(mx, set_layout(P, h.ST, V).total);
    H.total = o + mx;
    return H;
}
}  // namespace

extern "C" int ffs_reml_hier_workspace(const ffs_reml_plan* P, int64_t V, int64_t* bytes) {
    if (!P || !bytes) return fail(-3, "hier_workspace: null argument");
    if (P->hier.empty()) return fail(-4, "hier_workspace: call ffs_reml_plan_build_hier first");
    *bytes = (int64_t)hier_layout(P, std::max<int64_t>(V, 0)).total;
    return 0;
}

extern "C" int ffs_reml_hier_search(const ffs_reml_plan* P, const float* Y, int64_t ldY, int64_t V, int32_t* best,
                                    float* best_score, float* scale, void* workspace, int64_t ws_bytes, void* stream) {
    if (!P) return fail(-3, "hier_search: null plan");
    FFS_NEED_SEARCH(P, "hier_search");
    if (V <= 0) return 0;
    if (P->hier.empty()) return fail(-4, "hier_search: call ffs_reml_plan_build_hier first");
    if (!Y || !best || !best_score || !scale) return fail(-3, "hier_search: null argument");
    if (P->d.proj_mode != 0 && !split_halves(P))
        return fail(-4, "hier_search: the split projection modes need p <= 96 here (use fp32)");
    const HierLayout H = hier_layout(P, V);
    if (!workspace || ws_bytes < (int64_t)H.total) return fail(-3, "hier_search: workspace too small");
    char* ws = static_cast<char*>(workspace);
    int32_t* hba = reinterpret_cast<int32_t*>(ws + H.off_ba);
    int32_t* hbb = reinterpret_cast<int32_t*>(ws + H.off_bb);
    float* hcur = reinterpret_cast<float*>(ws + H.off_cur);
    double* coef = H.stride ? reinterpret_cast<double*>(ws + H.off_coef) : nullptr;
    hipStream_t s = (hipStream_t)stream;
    const int np = (int)P->hier.size();
    // p > 96: one prep of every voxel (with the split halves in the split
    // modes); each pass gathers its set layout's columns from it
    float* Yres = P->d.nseg ? reinterpret_cast<float*>(ws + H.off_res) : nullptr;
    const int64_t ldres = vpad_of(V);
    if (Yres)
        if (int rc = prep_impl(P, Y, ldY, V, nullptr, V, Yres, ldres, scale, 1, nullptr, s, nullptr, split_halves(P),
                               false))
            return rc;
    for (int i = 0; i < np; ++i) {
        const HierPass& h = P->hier[i];
        const SetLayout L = set_layout(P, h.ST, V);
        if (L.max_cols > 0x7fffffff) return fail(-3, "hier_search: batch too large; split it");
        const SetPtrs q = set_ptrs(L, ws + H.off_sets);
        if (i == 0)
            hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, s, q.key, V, 0);
        else
            hipLaunchKernelGGL(k_hier_key, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, s, P->d, h.center_set,
                               hba, hbb, (int)V, q.key);
        HIP_TRY(hipGetLastError());
        if (Yres) {
            if (int rc = set_group_gather(P, h.ST, Yres, ldres, V, L, q, s)) return rc;
        } else if (int rc = set_group_prep(P, h.ST, Y, ldY, V, scale, i == 0 ? nullptr : coef, L, q, s,
                                           i == 0 ? coef : nullptr)) {
            // pass 0 residualises from Y and keeps each voxel's coefficients and
            // scale; the later passes run the second half of the prep from them
            return rc;
        }
        if (int rc = set_energy(P, h.ST, L, q, s)) return rc;
        if (int rc = set_scores(P, h.ST, L, q, s)) return rc;
        hipLaunchKernelGGL(k_hier_select, dim3((unsigned)(L.max_cols / 256)), dim3(256), 0, s, P->d, h, q.rows,
                           q.tset, q.nt, q.S, hba, hbb, hcur, (int)(i == np - 1), best, best_score);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Lambda bands
// ---------------------------------------------------------------------------
extern "C" int ffs_reml_plan_set_bands(ffs_reml_plan* P, double half_width, double bin_width, void* stream) {
    if (!P) return fail(-3, "plan_set_bands: null plan");
    FFS_NEED_NARROW(P, "plan_set_bands");
    FFS_NEED_SEARCH(P, "plan_set_bands");
    PlanDev& d = P->d;
    if (!(half_width > 0.0) || !(bin_width > 0.0)) {
        d.band = SetTable{};
        return 0;
    }
    if (bin_width < 1e-3 || bin_width > 2.0) return fail(-3, "plan_set_bands: bin_width must be in [0.001, 2]");
    const int K = d.K, n_a = d.n_a, n_b = d.n_b;
    const int nbins = (int)std::floor(2.0 / bin_width + 0.5) + 1;
    bin_width = (double)(float)bin_width;  // the centres the kernels use (fp32 step)
    std::vector<double> lam(K);
    for (int k = 0; k < K; ++k) {
        const double a = P->h_ab[2 * k], b = P->h_ab[2 * k + 1];
        const double den = 1.0 + 2.0 * a * b + b * b;
        lam[k] = den > 0.0 ? (1.0 + a * b) * (a + b) / den : 0.0;
    }
    std::vector<int> ai(K, -1), bi(K, -1);
    for (int x = 0; x < n_a; ++x)
        for (int y = 0; y < n_b; ++y) {
            const int q = P->h_lut[(size_t)x * n_b + y];
            if (q >= 0 && q < K) {
                ai[q] = x;
                bi[q] = y;
            }
        }
    std::vector<std::vector<int32_t>> bands(nbins), edges(nbins);
    size_t maxn = 0;
    for (int i = 0; i < nbins; ++i) {
        const double c = -1.0 + i * bin_width;
        auto inside = [&](int k) { return std::fabs(lam[k] - c) <= half_width; };
        for (int k = 0; k < K; ++k) {
            if (!inside(k)) continue;
            int e = 0;
            if (ai[k] >= 0) {
                const int nx[4] = {ai[k] - 1, ai[k] + 1, ai[k], ai[k]}, ny[4] = {bi[k], bi[k], bi[k] - 1, bi[k] + 1};
                for (int q = 0; q < 4; ++q) {
                    if (nx[q] < 0 || nx[q] >= n_a || ny[q] < 0 || ny[q] >= n_b) continue;
                    const int kn = P->h_lut[(size_t)nx[q] * n_b + ny[q]];
                    if (kn >= 0 && kn < K && !inside(kn)) e = 1;
                }
            }
            bands[i].push_back(k);
            edges[i].push_back(e);
        }
        maxn = std::max(maxn, bands[i].size());
    }
    if (maxn == 0) return fail(-3, "plan_set_bands: no pair lies in any band");
    SetTable t{};
    // a band holds few pairs of each a: one position per group, 8 to a chunk
    if (int rc = build_set_table(P, bands, edges, nullptr, &t, (hipStream_t)stream, 1, kEnergyNB)) return rc;
    d.band = t;
    d.band_lo = -1.f;
    d.band_h = (float)bin_width;
    return 0;
}

extern "C" int ffs_reml_plan_band_info(const ffs_reml_plan* P, int32_t* n_bands, int32_t* slots, int32_t* pairs,
                                       int32_t* counts, int32_t* edge) {
    if (!P) return fail(-3, "plan_band_info: null plan");
    const SetTable& t = P->d.band;
    if (n_bands) *n_bands = t.n;
    if (slots) *slots = t.sb;
    const size_t n = (size_t)t.n * t.sb;
    if (t.n > 0) {
        if (pairs) HIP_TRY(hipMemcpy(pairs, t.pairs, 4 * n, hipMemcpyDeviceToHost));
        if (edge) HIP_TRY(hipMemcpy(edge, t.edge, 4 * n, hipMemcpyDeviceToHost));
        if (counts) HIP_TRY(hipMemcpy(counts, t.cnt, 4 * (size_t)t.n, hipMemcpyDeviceToHost));
    }
    return 0;
}

extern "C" int ffs_reml_plan_window_info(const ffs_reml_plan* P, int32_t* n_windows, int32_t* slots,
                                         int32_t* pairs, int32_t* counts, int32_t* edge) {
    if (!P) return fail(-3, "plan_window_info: null plan");
    const SetTable& t = P->d.win;
    if (n_windows) *n_windows = t.n;
    if (slots) *slots = t.sb;
    const size_t n = (size_t)t.n * t.sb;
    if (t.n > 0) {
        if (pairs) HIP_TRY(hipMemcpy(pairs, t.pairs, 4 * n, hipMemcpyDeviceToHost));
        if (edge) HIP_TRY(hipMemcpy(edge, t.edge, 4 * n, hipMemcpyDeviceToHost));
        if (counts) HIP_TRY(hipMemcpy(counts, t.cnt, 4 * (size_t)t.n, hipMemcpyDeviceToHost));
    }
    return 0;
}

extern "C" int ffs_reml_band_workspace(const ffs_reml_plan* P, int64_t V, int64_t* bytes) {
    if (!P || !bytes) return fail(-3, "band_workspace: null argument");
    FFS_NEED_NARROW(P, "band_workspace");
    *bytes = (int64_t)set_layout(P, P->d.band, std::max<int64_t>(V, 0)).total;
    return 0;
}

extern "C" int ffs_reml_band_prepare(const ffs_reml_plan* P, const float* Y, int64_t ldY, int64_t V,
                                     const float* props, const int32_t* need_full, float* scale,
                                     const double* coef, void* workspace, int64_t ws_bytes, void* stream) {
    if (V <= 0) return 0;
    SetLayout L;
    if (int rc = set_check(P, P->d.band, "band search", V, workspace, ws_bytes, &L)) return rc;
    SetPtrs q = set_ptrs(L, workspace);
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_band_key, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, s, P->d, props, (int)V, need_full,
                       q.key);
    HIP_TRY(hipGetLastError());
    return set_group_prep(P, P->d.band, Y, ldY, V, scale, coef, L, q, s);
}

extern "C" int ffs_reml_band_energy(const ffs_reml_plan* P, int64_t V, void* workspace, int64_t ws_bytes,
                                    void* stream) {
    if (V <= 0) return 0;
    SetLayout L;
    if (int rc = set_check(P, P->d.band, "band search", V, workspace, ws_bytes, &L)) return rc;
    return set_energy(P, P->d.band, L, set_ptrs(L, workspace), (hipStream_t)stream);
}

extern "C" int ffs_reml_band_project(const ffs_reml_plan* P, int64_t V, int32_t* best, float* best_score,
                                     int32_t* need_full, void* workspace, int64_t ws_bytes, void* stream) {
    if (V <= 0) return 0;
    SetLayout L;
    if (int rc = set_check(P, P->d.band, "band search", V, workspace, ws_bytes, &L)) return rc;
    return set_project(P, P->d.band, L, set_ptrs(L, workspace), best, best_score, need_full, nullptr,
                       (hipStream_t)stream);
}

extern "C" int ffs_reml_band_search(const ffs_reml_plan* P, const float* Y, int64_t ldY, int64_t V,
                                    const float* props, int32_t* best, float* best_score, int32_t* need_full,
                                    float* scale, const double* coef, void* workspace, int64_t ws_bytes,
                                    void* stream) {
    if (int rc = ffs_reml_band_prepare(P, Y, ldY, V, props, need_full, scale, coef, workspace, ws_bytes, stream))
        return rc;
    if (int rc = ffs_reml_band_energy(P, V, workspace, ws_bytes, stream)) return rc;
    return ffs_reml_band_project(P, V, best, best_score, need_full, workspace, ws_bytes, stream);
}

// ---------------------------------------------------------------------------
// Full-grid search of a device-counted voxel subset
// ---------------------------------------------------------------------------
namespace {
struct SubsetLayout {
    int64_t cap_pad, rows;
    size_t off_yt, off_s, off_b, off_sc, total;
};
SubsetLayout subset_layout(const ffs_reml_plan* P, int64_t cap) {
    SubsetLayout L{};
    L.cap_pad = vpad_of(std::max<int64_t>(cap, 1));
    L.rows = yt_rows(P);
    size_t o = 0;
    L.off_yt = o;
    o = align256(o + 4 * (size_t)L.rows * L.cap_pad);
    L.off_s = o;
    o = align256(o + 4 * (size_t)L.cap_pad * P->d.K);
    L.off_b = o;
    o = align256(o + 4 * (size_t)L.cap_pad);
    L.off_sc = o;
    o = align256(o + 4 * (size_t)L.cap_pad);
    L.total = o;
    return L;
}
}  // namespace

extern "C" int ffs_reml_subset_workspace(const ffs_reml_plan* P, int64_t cap, int64_t* bytes) {
    if (!P || !bytes) return fail(-3, "subset_workspace: null argument");
    FFS_NEED_NARROW(P, "subset_workspace");
    *bytes = (int64_t)subset_layout(P, cap).total;
    return 0;
}

extern "C" int ffs_reml_subset_search(const ffs_reml_plan* P, const float* Y, int64_t ldY, int64_t V,
                                      const int32_t* idx, const int32_t* n_live, int64_t cap, int32_t* best,
                                      float* best_score, float* scale, const double* coef, void* workspace,
                                      int64_t ws_bytes, void* stream) {
    if (!P) return fail(-3, "subset_search: null plan");
    FFS_NEED_NARROW(P, "subset_search");
    FFS_NEED_SEARCH(P, "subset_search");
    if (V <= 0 || cap <= 0) return 0;
    if (!idx || !n_live) return fail(-3, "subset_search: idx and n_live are required");
    if (cap > V) cap = V;
    const SubsetLayout L = subset_layout(P, cap);
    if (!workspace || ws_bytes < (int64_t)L.total) return fail(-3, "subset_search: workspace too small");
    if (L.cap_pad > 0x7fffffff) return fail(-3, "subset_search: cap too large");
    char* ws = static_cast<char*>(workspace);
    float* Yt = reinterpret_cast<float*>(ws + L.off_yt);
    float* S = reinterpret_cast<float*>(ws + L.off_s);
    int32_t* bsub = reinterpret_cast<int32_t*>(ws + L.off_b);
    float* ssub = reinterpret_cast<float*>(ws + L.off_sc);
    hipStream_t s = (hipStream_t)stream;
    const int K = P->d.K;
    const bool split = split_halves(P);
    if (int rc = prep_impl(P, Y, ldY, V, idx, cap, Yt, L.cap_pad, scale, 1, nullptr, s, n_live, split, false, coef))
        return rc;
    if (int rc = energy_impl(P, Yt, L.cap_pad, cap, S, K, n_live, s)) return rc;
    if (int rc = project_impl(P, Yt, L.cap_pad, cap, S, K, n_live, split, s)) return rc;
    hipLaunchKernelGGL(k_select_first<float>, dim3((unsigned)((cap + 3) / 4)), dim3(256), 0, s, P->d, S, (int64_t)K,
                       (int)cap, n_live, bsub, ssub);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_scatter_best, dim3((unsigned)((cap + 255) / 256)), dim3(256), 0, s, idx, (int)cap, n_live,
                       bsub, ssub, best, best_score);
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int ffs_reml_whiten_cols(const ffs_reml_plan* P, const int32_t* pair, int64_t V, int32_t m,
                                    const double* cols, int64_t ld, double* out, void* stream) {
    if (!P) return fail(-3, "whiten_cols: null plan");
    if (V <= 0 || m <= 0) return 0;
    if (!pair || !cols || !out) return fail(-3, "whiten_cols: null argument");
    if (ld < P->d.T) return fail(-3, "whiten_cols: ld must be >= T");
    if (!P->d.coefR) return fail(-4, "whiten_cols: plan has no recurrence tables");
    const int64_t n = V * (int64_t)m;
    if (n > 0x7fffffffLL * 64 || V > 0x7fffffff) return fail(-3, "whiten_cols: batch too large; split it");
    hipLaunchKernelGGL(k_whiten_cols, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, P->d,
                       pair, (int)V, (int)m, cols, ld, out);
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int ffs_reml_dsort_workspace(const ffs_reml_plan* P, int64_t V, int32_t q, int32_t n_f, int64_t* n_double) {
    if (!P || !n_double) return fail(-3, "dsort_workspace: null argument");
    *n_double = std::max<int64_t>(V, 0) * dsort_work_doubles(P->d.p + q, n_f, P->d.T);
    return 0;
}

extern "C" int ffs_reml_dsort_gls(const ffs_reml_plan* P, const int32_t* pair, int64_t V, int32_t q, const double* D,
                                  const double* Y, const int32_t* fci, int32_t n_f, double* beta, double* tstat,
                                  double* sigma2, double* fstat, double* r2, double* Ainv, double* resid_w,
                                  double* resid, double* pred, int32_t* info, double* work, int64_t work_doubles,
                                  void* stream) {
    if (!P) return fail(-3, "dsort_gls: null plan");
    if (V <= 0) return 0;
    if (V > 0x7fffffff) return fail(-3, "dsort_gls: batch too large; split it");
    if (!P->d.coefR || !P->d.Xw64) return fail(-4, "dsort_gls: plan has no recurrence tables");
    const int pe = P->d.p + q;
    if (q < 0 || q > 63 || pe + 1 > kDsortMaxM)
        return fail(-3, "dsort_gls: need 0 <= q <= 63 and p + q <= " + std::to_string(kDsortMaxM - 1));
    if (n_f < 1 || n_f > pe || !fci) return fail(-3, "dsort_gls: need 1 <= n_f <= p + q F columns");
    if (!pair || !Y || (q > 0 && !D) || !beta || !tstat || !sigma2 || !fstat || !r2 || !info || !work)
        return fail(-3, "dsort_gls: null argument");
    const int64_t ws = dsort_work_doubles(pe, n_f, P->d.T);
    if (work_doubles < V * ws) return fail(-3, "dsort_gls: workspace too small (ffs_reml_dsort_workspace)");
    DsortArgs A{};
    A.pair = pair;
    A.V = (int)V;
    A.q = q;
    A.D = D;
    A.Y = Y;
    A.n_f = n_f;
    for (int i = 0; i < n_f; ++i) {
        if (fci[i] < 0 || fci[i] >= pe) return fail(-3, "dsort_gls: F column out of range");
        A.fci[i] = fci[i];
    }
    A.beta = beta;
    A.tstat = tstat;
    A.sigma2 = sigma2;
    A.fstat = fstat;
    A.r2 = r2;
    A.Ai_out = Ainv;
    A.rw = resid_w;
    A.resid = resid;
    A.pred = pred;
    A.info = info;
    A.work = work;
    A.ws = ws;
    const int m = pe + 1;
    const size_t shm = sizeof(double) * ((size_t)m * (m + 1) / 2 + (size_t)kDsortChunk * m);
    if (shm > 64 * 1024)
        HIP_TRY(hipFuncSetAttribute((const void*)k_dsort_gls, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    hipLaunchKernelGGL(k_dsort_gls, dim3((unsigned)V), dim3(64), shm, (hipStream_t)stream, P->d, A);
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" int ffs_reml_plan_set_projection(ffs_reml_plan* P, int32_t mode) {
    if (!P) return fail(-3, "plan_set_projection: null plan");
    if (mode < 0 || mode > 2) return fail(-3, "plan_set_projection: mode must be 0 (fp32), 1 (bf16x3) or 2 (fp16x3)");
    P->d.proj_mode = mode;
    return 0;
}

extern "C" int ffs_reml_dense_likelihood(const double* X, const double* Y, int32_t m, const double* R, int32_t T,
                                         int32_t p, double* work, int64_t work_doubles, double* out, void* stream) {
    if (T < 1 || p < 1 || m < 1 || p > T) return fail(-3, "dense_likelihood: need 1 <= p <= T and m >= 1");
    const int64_t need = (int64_t)T * T + (int64_t)T * (p + m) + (int64_t)p * p + (int64_t)p * m;
    if (!work || work_doubles < need) return fail(-3, "dense_likelihood: workspace too small");
    hipLaunchKernelGGL(k_dense_reml, dim3(1), dim3(kDenseThreads), 0, (hipStream_t)stream, X, Y, (int)m, R, (int)T,
                       (int)p, work, out);
    HIP_TRY(hipGetLastError());
    return 0;
}


// ---------------------------------------------------------------------------
// Grouped fits: one launch per stage over the voxels of every group
// ---------------------------------------------------------------------------
struct ffs_reml_group {
    int device = 0;
    int G = 0, K = 0, T = 0, p = 0;
    int64_t V = 0, ncols = 0;  // voxels; padded columns (each group padded to kVoxPad)
    std::vector<ffs_reml_plan*> plans;  // borrowed
    PlanDev* dplans = nullptr;
    int32_t *rows = nullptr, *blk_grp = nullptr, *tile_grp = nullptr, *vgroup = nullptr;
    float *Yt = nullptr, *S = nullptr, *score_col = nullptr;
    int32_t *best_col = nullptr, *key = nullptr, *work = nullptr;
    int64_t yt_rows = 0, work_ints = 0;
    std::vector<int64_t> group_cols;  // padded columns of each group
    // launch configuration (autotuned): surface in n_chunks column chunks on
    // two graph branches; split projection on pre-split halves (p <= 96) or not
    int n_chunks = 1;
    bool presplit = true;
    bool rank_deficient = false;  // some design has dependent columns
    std::string tuning;
    hipStream_t gs = nullptr, gs2 = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    std::vector<hipEvent_t> ev_chunk;
    hipGraphExec_t exec = nullptr;
    std::vector<const void*> graph_key;
    std::vector<void*> allocs;
    template <class Tp>
    int alloc(Tp** out, size_t n) {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, (n ? n : 1) * sizeof(Tp));
        if (e != hipSuccess) return fail(-1, std::string("hipMalloc: ") + hipGetErrorString(e));
        allocs.push_back(q);
        *out = static_cast<Tp*>(q);
        return 0;
    }
    ~ffs_reml_group() {
        if (exec) (void)hipGraphExecDestroy(exec);
        for (hipEvent_t e : ev_chunk) (void)hipEventDestroy(e);
        if (ev_in) (void)hipEventDestroy(ev_in);
        if (ev_out) (void)hipEventDestroy(ev_out);
        if (gs) (void)hipStreamDestroy(gs);
        if (gs2) (void)hipStreamDestroy(gs2);
        for (void* q : allocs) (void)hipFree(q);
    }
};

extern "C" int ffs_reml_group_create(ffs_reml_group** out, int32_t G, ffs_reml_plan* const* plans, int64_t V,
                                     const int32_t* group_of_voxel, void* stream) {
    if (!out) return fail(-3, "group_create: null output handle");
    *out = nullptr;
    if (G < 1 || !plans || V < 1 || !group_of_voxel) return fail(-3, "group_create: need G >= 1 plans and V >= 1");
    const ffs_reml_plan* P0 = plans[0];
    for (int g = 0; g < G; ++g) {
        const ffs_reml_plan* P = plans[g];
        if (!P) return fail(-3, "group_create: null plan");
        if (P->gls_only || P->ridge > 0.0) return fail(-3, "group_create: plans must be lattice (search) plans");
        FFS_NEED_NARROW(P, "group_create");
        if (P->d.T != P0->d.T || P->d.p != P0->d.p || P->d.K != P0->d.K || P->d.BN != P0->d.BN ||
            P->d.proj_mode != P0->d.proj_mode || P->d.n_task != P0->d.n_task || P->h_ab != P0->h_ab ||
            P->h_lut != P0->h_lut || P->device != P0->device)
            return fail(-3, "group_create: plans must share the timeline, pair list, lattice, width, projection "
                            "mode, F columns and device");
    }
    if (V > 0x7fffffff / 2) return fail(-3, "group_create: too many voxels; split them");
    hipStream_t s = (hipStream_t)stream;
    auto* R = new ffs_reml_group();
    std::unique_ptr<ffs_reml_group> guard(R);
    R->device = P0->device;
    R->G = G;
    R->K = P0->d.K;
    R->T = P0->d.T;
    R->p = P0->d.p;
    R->V = V;
    R->plans.assign(plans, plans + G);
    // padded layout: groups in id order, each padded to kVoxPad columns
    std::vector<std::vector<int32_t>> members(G);
    for (int64_t v = 0; v < V; ++v) {
        const int g = group_of_voxel[v];
        if (g < 0 || g >= G) return fail(-3, "group_create: group index out of range");
        members[g].push_back((int32_t)v);
    }
    std::vector<int32_t> rows, blk, tile, vg(group_of_voxel, group_of_voxel + V);
    for (int g = 0; g < G; ++g) {
        const int64_t n = (int64_t)members[g].size();
        const int64_t npad = (n + kVoxPad - 1) / kVoxPad * kVoxPad;
        R->group_cols.push_back(npad);
        for (int64_t i = 0; i < npad; ++i) rows.push_back(i < n ? members[g][i] : -1);
        for (int64_t i = 0; i < npad / 64; ++i) blk.push_back(g);
        for (int64_t i = 0; i < npad / kProjBM; ++i) tile.push_back(g);
    }
    R->ncols = (int64_t)rows.size();
    if (R->ncols > 0x7fffffff) return fail(-3, "group_create: padded layout too large");
    std::vector<PlanDev> hp(G);
    for (int g = 0; g < G; ++g) {
        hp[g] = plans[g]->d;
        if (hp[g].Pi) R->rank_deficient = true;
    }
    R->yt_rows = yt_rows(P0);
    R->work_ints = 3 * (int64_t)G * R->K + 1 + gls_perm_cap(V, G * R->K);
    int rc = 0;
    if ((rc = R->alloc(&R->dplans, G)) || (rc = R->alloc(&R->rows, rows.size())) ||
        (rc = R->alloc(&R->blk_grp, blk.size())) || (rc = R->alloc(&R->tile_grp, tile.size())) ||
        (rc = R->alloc(&R->vgroup, vg.size())) || (rc = R->alloc(&R->Yt, (size_t)R->yt_rows * R->ncols)) ||
        (rc = R->alloc(&R->S, (size_t)R->ncols * R->K)) || (rc = R->alloc(&R->score_col, R->ncols)) ||
        (rc = R->alloc(&R->best_col, R->ncols)) || (rc = R->alloc(&R->key, V)) ||
        (rc = R->alloc(&R->work, R->work_ints)))
        return rc;
    HIP_TRY(hipMemcpyAsync(R->dplans, hp.data(), sizeof(PlanDev) * G, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(R->rows, rows.data(), 4 * rows.size(), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(R->blk_grp, blk.data(), 4 * blk.size(), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(R->tile_grp, tile.data(), 4 * tile.size(), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(R->vgroup, vg.data(), 4 * vg.size(), hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));  // the host vectors above are uploaded
    HIP_TRY(hipStreamCreateWithFlags(&R->gs, hipStreamNonBlocking));
    HIP_TRY(hipStreamCreateWithFlags(&R->gs2, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&R->ev_in, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&R->ev_out, hipEventDisableTiming));
    R->presplit = split_halves(P0);
    *out = guard.release();
    return 0;
}

extern "C" int ffs_reml_group_destroy(ffs_reml_group* grp) {
    delete grp;
    return 0;
}

// The fit stages of a grouped layout, enqueued on s (and on the group's second
// stream for the odd surface chunks): prep, energy + projection per column
// chunk, selection, scatter to voxel order, GLS grouped by (design, pair).
struct GroupIO {
    const float* Y;
    int64_t ldY;
    int32_t mode;
    int32_t *best;
    float *score, *scale, *betas, *tstats, *fstats, *sigma2, *r2, *S_out;
    int64_t ldS_out;
};

static int group_enqueue(ffs_reml_group* R, const GroupIO& io, hipStream_t s) {
    const ffs_reml_plan* P0 = R->plans[0];
    const int K = R->K;
    const int64_t nc = R->ncols;
    const bool split_prep = split_halves(P0) && R->presplit;
    if (int rc = prep_impl(P0, io.Y, io.ldY, R->V, R->rows, nc, R->Yt, nc, io.scale, 1, nullptr, s, nullptr,
                           split_prep, false, nullptr, nullptr, R->dplans, R->blk_grp))
        return rc;
    // surface: energy of a chunk, then its projection; chunks alternate
    // between the two streams so that one chunk's VALU energy can overlap the
    // previous chunk's MFMA projection
    // the pre-split halves are addressed by absolute column: one chunk only
    const int nch = split_prep ? 1 : std::max(1, std::min<int>(R->n_chunks, (int)(nc / kVoxPad)));
    const int64_t step = (nc / kVoxPad + nch - 1) / nch * kVoxPad;
    while ((int)R->ev_chunk.size() < 2 * nch + 2) {
        hipEvent_t e;
        HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        R->ev_chunk.push_back(e);
    }
    HIP_TRY(hipEventRecord(R->ev_chunk[0], s));
    HIP_TRY(hipStreamWaitEvent(R->gs2, R->ev_chunk[0], 0));
    int i = 0;
    for (int64_t c0 = 0; c0 < nc; c0 += step, ++i) {
        const int64_t n = std::min(step, nc - c0);
        hipStream_t cs = (i & 1) ? R->gs2 : s;
        if (int rc = energy_impl(P0, R->Yt + c0, nc, n, R->S + c0 * K, K, nullptr, cs)) return rc;
        if (int rc = project_impl(P0, R->Yt + c0, nc, n, R->S + c0 * K, K, nullptr, split_prep, cs, R->dplans,
                                  R->tile_grp + c0 / kProjBM))
            return rc;
    }
    HIP_TRY(hipEventRecord(R->ev_chunk[1], R->gs2));
    HIP_TRY(hipStreamWaitEvent(s, R->ev_chunk[1], 0));
    if (io.mode == 0)
        hipLaunchKernelGGL(k_select_first<float>, dim3((unsigned)((nc + 3) / 4)), dim3(256), 0, s, P0->d, R->S,
                           (int64_t)K, (int)nc, nullptr, R->best_col, R->score_col);
    else
        hipLaunchKernelGGL(k_select<float>, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, s, P0->d, R->S,
                           (int64_t)K, (int)nc, (int)io.mode, R->best_col, R->score_col);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_group_scatter, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, s, R->rows, (int)nc,
                       R->best_col, R->score_col, R->vgroup, K, io.best, io.score, R->key);
    HIP_TRY(hipGetLastError());
    if (io.S_out) {
        hipLaunchKernelGGL(k_group_gather_rows, dim3((unsigned)nc), dim3(256), 0, s, R->rows, (int)nc, R->S,
                           (int64_t)K, K, io.S_out, io.ldS_out);
        HIP_TRY(hipGetLastError());
    }
    // GLS: voxels grouped by (design, pair) in 64-slot groups, so every wave
    // (or block) holds one design's pair
    const int nb = R->G * K;
    int32_t* cnt = R->work;
    int32_t* off = cnt + nb;
    int32_t* cursor = off + nb + 1;
    int32_t* perm = cursor + nb;
    const int64_t cap = gls_perm_cap(R->V, nb);
    hipLaunchKernelGGL(k_gls_invalid, dim3((unsigned)((R->V + 255) / 256)), dim3(256), 0, s, io.best, (int)R->V, K,
                       R->p, io.betas, io.tstats, io.fstats, io.sigma2, io.r2);
    if (int rc = group_by_key(R->key, R->V, nb, 64, cnt, off, cursor, perm, cap, s)) return rc;
    if (!P0->d.PG) {
        if (int rc = launch_gls_wide(P0->d, io.Y, io.ldY, io.scale, perm, io.best, cap, io.betas, io.tstats,
                                     io.fstats, io.sigma2, io.r2, R->dplans, R->vgroup, s))
            return rc;
    } else {
        const dim3 g((unsigned)(cap / 64)), b(64);
        if (P0->d.PG == 16)
            hipLaunchKernelGGL(k_gls_mfma<1>, g, b, 0, s, P0->d, io.Y, io.ldY, io.scale, perm, io.best, io.betas,
                               io.tstats, io.fstats, io.sigma2, io.r2, R->dplans, R->vgroup);
        else
            hipLaunchKernelGGL(k_gls_mfma<2>, g, b, 0, s, P0->d, io.Y, io.ldY, io.scale, perm, io.best, io.betas,
                               io.tstats, io.fstats, io.sigma2, io.r2, R->dplans, R->vgroup);
    }
    HIP_TRY(hipGetLastError());
    if (R->rank_deficient) {
        hipLaunchKernelGGL(k_gls_rank, dim3((unsigned)R->V), dim3(64), 2 * sizeof(double) * P0->d.p, s, P0->d, (int)R->V, io.scale, io.best,
                           io.betas, io.tstats, io.fstats, io.sigma2, R->dplans, R->vgroup);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

// Captures the stages into a graph on the group's own streams (pointers of
// the call are baked in: a call with other pointers or mode captures again).
static int group_capture(ffs_reml_group* R, const GroupIO& io) {
    if (R->exec) {
        (void)hipGraphExecDestroy(R->exec);
        R->exec = nullptr;
    }
    HIP_TRY(hipStreamBeginCapture(R->gs, hipStreamCaptureModeThreadLocal));
    const int rc = group_enqueue(R, io, R->gs);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(R->gs, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return fail(-1, std::string("group_fit: hipStreamEndCapture: ") + hipGetErrorString(e));
    const hipError_t e2 = hipGraphInstantiate(&R->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e2 != hipSuccess) return fail(-1, std::string("group_fit: hipGraphInstantiate: ") + hipGetErrorString(e2));
    return 0;
}

static int group_launch(ffs_reml_group* R, const GroupIO& io, hipStream_t s) {
    const std::vector<const void*> gk = {io.Y, (const void*)(intptr_t)io.ldY, (const void*)(intptr_t)io.mode,
                                         io.best, io.score, io.scale, io.betas, io.tstats, io.fstats, io.sigma2,
                                         io.r2, io.S_out, (const void*)(intptr_t)io.ldS_out,
                                         (const void*)(intptr_t)R->n_chunks, (const void*)(intptr_t)R->presplit};
    if (!R->exec || gk != R->graph_key) {
        R->graph_key.clear();
        if (int rc = group_capture(R, io)) return rc;
        R->graph_key = gk;
    }
    HIP_TRY(hipEventRecord(R->ev_in, s));
    HIP_TRY(hipStreamWaitEvent(R->gs, R->ev_in, 0));
    HIP_TRY(hipGraphLaunch(R->exec, R->gs));
    HIP_TRY(hipEventRecord(R->ev_out, R->gs));
    HIP_TRY(hipStreamWaitEvent(s, R->ev_out, 0));
    return 0;
}

extern "C" int ffs_reml_group_fit(ffs_reml_group* R, const float* Y, int64_t ldY, int32_t mode, int32_t* best,
                                  float* score, float* scale, float* betas, float* tstats, float* fstats,
                                  float* sigma2, float* r2, float* S_out, int64_t ldS_out, void* stream) {
    if (!R) return fail(-3, "group_fit: null group");
    if (mode < 0 || mode > 2) return fail(-3, "group_fit: mode must be 0, 1 or 2");
    if (!Y || !best || !score || !scale || !betas || !tstats || !fstats || !sigma2 || !r2)
        return fail(-3, "group_fit: null argument");
    if (ldY < R->T) return fail(-3, "group_fit: ldY must be >= T");
    if (S_out && ldS_out < R->K) return fail(-3, "group_fit: ldS_out must be >= K");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev != R->device) return fail(-3, "group_fit: the group belongs to another device");
    for (const ffs_reml_plan* P : R->plans)
        if (P->d.proj_mode != R->plans[0]->d.proj_mode)
            return fail(-3, "group_fit: the plans' projection modes changed after group_create");
    GroupIO io{Y, ldY, mode, best, score, scale, betas, tstats, fstats, sigma2, r2, S_out, ldS_out};
    return group_launch(R, io, (hipStream_t)stream);
}

// Autotuner: times `reps` fits of each candidate launch configuration on the
// caller's data (graph replays timed by events on the caller's stream, after
// one untimed capture + replay) and keeps the fastest.  Candidates: surface
// column chunks n in {1, 2, 4} (chunks alternate between two graph branches,
// so the VALU energy of one overlaps the MFMA projection of the other) and,
// in a split projection mode, the projection voxel tile (256
// voxels on pre-split halves, one chunk; or 128 voxels split in the kernel).
extern "C" int ffs_reml_group_tune(ffs_reml_group* R, const float* Y, int64_t ldY, int32_t mode, int32_t* best,
                                   float* score, float* scale, float* betas, float* tstats, float* fstats,
                                   float* sigma2, float* r2, int32_t reps, char* report, int64_t report_len,
                                   void* stream) {
    if (!R) return fail(-3, "group_tune: null group");
    if (reps < 1) reps = 1;
    hipStream_t s = (hipStream_t)stream;
    const ffs_reml_plan* P0 = R->plans[0];
    std::vector<std::pair<int, bool>> cands;
    const bool can_split = split_halves(P0);
    for (int n : {1, 2, 4})
        for (bool ps : {true, false}) {
            if (!can_split && !ps) continue;  // the tile choice exists only for pre-split operands
            if (can_split && ps && n > 1) continue;  // pre-split halves: one surface chunk
            cands.push_back({n, can_split && ps});
        }
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    std::string rep = "{\"candidates\": [";
    double best_ms = 1e300;
    std::pair<int, bool> best_c = cands[0];
    int rc = 0;
    for (size_t ci = 0; ci < cands.size() && !rc; ++ci) {
        R->n_chunks = cands[ci].first;
        R->presplit = cands[ci].second;
        GroupIO io{Y, ldY, mode, best, score, scale, betas, tstats, fstats, sigma2, r2, nullptr, 0};
        if ((rc = group_launch(R, io, s))) break;  // capture + first replay (discarded)
        hipError_t e = hipEventRecord(e0, s);
        for (int r = 0; r < reps && !rc && e == hipSuccess; ++r) rc = group_launch(R, io, s);
        if (!rc && e == hipSuccess) e = hipEventRecord(e1, s);
        if (!rc && e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (!rc && e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (!rc && e != hipSuccess) rc = fail(-1, std::string("group_tune: ") + hipGetErrorString(e));
        if (rc) break;
        const double per = ms / reps;
        char buf[160];
        snprintf(buf, sizeof buf, "%s{\"surface_chunks\": %d, \"projection_voxel_tile\": %d, \"ms\": %.4f}",
                 ci ? ", " : "", cands[ci].first, (can_split && cands[ci].second) ? 256 : 128, per);
        rep += buf;
        if (per < best_ms) {
            best_ms = per;
            best_c = cands[ci];
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    R->n_chunks = best_c.first;
    R->presplit = best_c.second;
    char buf[160];
    snprintf(buf, sizeof buf, "], \"chosen\": {\"surface_chunks\": %d, \"projection_voxel_tile\": %d, \"ms\": %.4f}}",
             best_c.first, (can_split && best_c.second) ? 256 : 128, best_ms);
    rep += buf;
    R->tuning = rep;
    if (report && report_len > 0) {
        const size_t n = std::min<size_t>(rep.size(), (size_t)report_len - 1);
        memcpy(report, rep.data(), n);
        report[n] = '\0';
    }
    return 0;
}
