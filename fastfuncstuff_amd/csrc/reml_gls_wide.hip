// (iv) FP64 GLS refit on f64 MFMA for designs of 33..96 columns
// (glm/arma.py:4784-5601), the sums of k_gls_mfma:
//   h = G_k' y, q = X' y                       (f64 MFMA, 16-column tiles)
//   E = |L^-1 y|^2, sum y, sum y^2             (FP64 innovations recurrence)
//   beta = Lx^-T h / s, ssw = E - |h|^2, ssres = y'y - 2 beta'q + beta'X'X beta,
//   sstot = y'y - T ybar^2, t, F, sigma^2, R^2 with the reference's epsilons.
// One block of four waves per 64 slots of one (design, pair) group; wave w
// takes slots 16 w .. 16 w + 15.
//
// Register layout.  MFMA 16x16x4 f64: A lane l = A[l % 16][l / 16], B lane l
// = B[l / 16][l % 16], D lane l element r = D[l / 16 + 4 r][l % 16].  With
// j = l % 16 (the voxel) and kq = l / 16, the four lanes of a voxel each hold
// every fourth coefficient: after h = G'y, lane kq holds h[4 s + kq] in
// accG[s / 4][s % 4] (s = 0 .. PG/4 - 1), and that is also the B operand
// layout of a product M * beta (k-step s, k-slot kq).  So beta is kept in the
// same distributed form (bb[s] = beta[4 s + kq]): beta = M h with the plan's
// M = Lx^-T and the two quadratic forms beta'X'X beta and beta'F beta are f64
// MFMA products over zero-padded tables whose output lands in that layout
// again.  In the main loop each 16-sample step feeds the
// MFMAs (lane samples t0 + 4 kq + m, m = 0..3) and, after a shuffle gather of
// the voxel's 16 samples, the recurrence (computed alike on the voxel's four
// lanes).  Included by reml_api.hip.
#include "reml_kernels.h"

namespace remlk {

// Sum over the four lanes of a voxel (lanes j, j + 16, j + 32, j + 48).
__device__ __forceinline__ double sum4(double x) {
    x += __shfl_xor(x, 16);
    x += __shfl_xor(x, 32);
    return x;
}

// M (PG x PG, row-major, zero past p) times a vector of 16 voxels held in the
// distributed layout: out[s] = (M x)[4 s + kq] for the lane's voxel.  The
// padded table needs no bounds tests, so a row's loads issue together.
template <int NT>
__device__ __forceinline__ void mat_vec(const double* __restrict__ M, int j, int kq, const double (&x)[4 * NT],
                                        double (&out)[4 * NT]) {
    constexpr int PG = 16 * NT;
    f64x4 acc[NT];
#pragma unroll
    for (int rt = 0; rt < NT; ++rt) acc[rt] = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 4 * NT; ++s) {
        const int col = 4 * s + kq;
#pragma unroll
        for (int rt = 0; rt < NT; ++rt)
            acc[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(M[(16 * rt + j) * PG + col], x[s], acc[rt], 0, 0, 0);
    }
#pragma unroll
    for (int s = 0; s < 4 * NT; ++s) out[s] = acc[s / 4][s % 4];
}

template <int NT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_gls_wide(PlanDev P, const float* __restrict__ Y, int64_t ldY,
                                                  const float* __restrict__ scale, const int32_t* __restrict__ perm,
                                                  const int32_t* __restrict__ best, float* __restrict__ betas,
                                                  float* __restrict__ tstats, float* __restrict__ fstats,
                                                  float* __restrict__ sigma2, float* __restrict__ r2,
                                                  const PlanDev* __restrict__ plans,
                                                  const int32_t* __restrict__ vgroup) {
    constexpr int PG = 16 * NT;
    constexpr int NS = 4 * NT;  // k-steps of a 16-row tile product; also beta values per lane
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int64_t slot0 = (int64_t)blockIdx.x * 64;
    const int v0 = __builtin_amdgcn_readfirstlane(perm[slot0]);
    if (v0 < 0) return;  // padding or beyond the last group
    if (plans) P = plans[vgroup[v0]];
    const int k = __builtin_amdgcn_readfirstlane(best[v0]);
    const int T = P.T, Tpad = P.Tpad, p = P.p;
    const int j = lane & 15, kq = lane >> 4;
    const int v = perm[slot0 + 16 * w + j];
    const float* __restrict__ row = Y + (size_t)(v >= 0 ? v : v0) * ldY;
    const double* __restrict__ Gk = P.G64p + (size_t)k * Tpad * PG;
    const double* __restrict__ Xp = P.Xp;
    const double* __restrict__ cR = P.coefR + (size_t)k * Tpad * 4;

    f64x4 accG[NT], accX[NT];
#pragma unroll
    for (int rt = 0; rt < NT; ++rt) {
        accG[rt] = f64x4{0.0, 0.0, 0.0, 0.0};
        accX[rt] = f64x4{0.0, 0.0, 0.0, 0.0};
    }
    double E = 0.0, sy = 0.0, syy = 0.0, e = 0.0, yprev = 0.0;
    const bool vec4 = (ldY & 3) == 0 && (reinterpret_cast<uintptr_t>(Y) & 15) == 0;
    for (int t0 = 0; t0 < Tpad; t0 += 16) {
        const int tq = t0 + 4 * kq;
        float y4[4];
        if (vec4 && tq + 4 <= T) {
            const f32x4 q4 = *reinterpret_cast<const f32x4*>(row + tq);
            y4[0] = q4.x;
            y4[1] = q4.y;
            y4[2] = q4.z;
            y4[3] = q4.w;
        } else {
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const float x = row[min(tq + m, T - 1)];
                y4[m] = (tq + m < T) ? x : 0.f;
            }
        }
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const double b = (double)y4[m];
#pragma unroll
            for (int rt = 0; rt < NT; ++rt) {
                const size_t o = (size_t)(tq + m) * PG + 16 * rt + j;
                accG[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(Gk[o], b, accG[rt], 0, 0, 0);
                accX[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(Xp[o], b, accX[rt], 0, 0, 0);
            }
        }
        // recurrence over samples t0 .. t0 + 15 of the voxel (rows t >= T have
        // zero coefficients and samples)
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const double yt = (double)__shfl(y4[i & 3], 16 * (i >> 2) + j);
            const double* cr = cR + 4 * (t0 + i);
            e = fma(cr[2], e, fma(-cr[1], yprev, yt));
            yprev = yt;
            const double wv = e * cr[0];
            E = fma(wv, wv, E);
            sy += yt;
            syy = fma(yt, yt, syy);
        }
    }
    const double s = (v >= 0 && scale) ? (double)scale[v] : 1.0;
    const double inv = 1.0 / s, inv2 = inv * inv;
    E *= inv2;
    sy *= inv;
    syy *= inv2;
    // h (scaled) in the distributed layout; its columns >= p are zero
    double bb[NS], hv[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q) hv[q] = accG[q / 4][q % 4] * inv;
    double hh = 0.0;
#pragma unroll
    for (int q = 0; q < NS; ++q) hh = fma(hv[q], hv[q], hh);
    hh = sum4(hh);
    // beta = M h with M = Lx^-T from the plan's epilogue tables (the back
    // substitution's zero-pivot rule included); then the quadratic forms
    const double* __restrict__ ep = P.gls_epi + (size_t)k * (2 * PG * PG + PG);
    mat_vec<NT>(ep, j, kq, hv, bb);
    double bq = 0.0, bxb = 0.0;
#pragma unroll
    for (int q = 0; q < NS; ++q) bq = fma(bb[q], accX[q / 4][q % 4] * inv, bq);  // q = X'y is dead after this
    bq = sum4(bq);
    double ra[NS];
    mat_vec<NT>(P.XtXp, j, kq, bb, ra);
#pragma unroll
    for (int q = 0; q < NS; ++q) bxb = fma(bb[q], ra[q], bxb);
    bxb = sum4(bxb);
    mat_vec<NT>(ep + PG * PG, j, kq, bb, ra);
    double quad = 0.0;
#pragma unroll
    for (int q = 0; q < NS; ++q) quad = fma(bb[q], ra[q], quad);
    quad = sum4(quad);
    if (v < 0) return;
    const double ssw = E - hh;
    const double ssres = syy - 2.0 * bq + bxb;
    const double sstot = syy - sy * (sy / T);
    const int dof = max(1, T - p);
    const double sig = ssw / dof;
    const double* __restrict__ Ad = ep + 2 * PG * PG;
#pragma unroll
    for (int q = 0; q < NS; ++q) {
        const int jj = 4 * q + kq;
        if (jj < p) {
            const double se = sqrt(sig * Ad[jj]);
            tstats[(size_t)v * p + jj] = (float)(bb[q] / (se + 1e-10));
            betas[(size_t)v * p + jj] = (float)(bb[q] * s);
        }
    }
    if (kq != 0) return;
    fstats[v] = (float)(quad / (sig * P.n_task + 1e-10));
    sigma2[v] = (float)(sig * s * s);
    r2[v] = (float)(sstot <= 1e-10 ? 0.0 : fmin(1.0 - ssres / sstot, 1.0));
}

}  // namespace remlk
