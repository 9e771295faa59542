// Band search: the pairs whose lag-1 correlation lambda lies near a proposed
// value.  Included by reml_api.hip after reml_proj16.hip (one translation unit).
//
// The REML surface of a weakly correlated series is a long valley along a
// contour of lambda = (1 + ab)(a + b) / (1 + 2ab + b^2), the lag-1
// correlation of the ARMA(1,1) process: the proposal network places such a
// voxel's lambda well but not its position along the valley.  Band i holds
// the pairs with |lambda_k - (band_lo + i band_h)| <= half-width, in pair
// order.  Voxels are grouped by band (groups padded to kBandTile columns) and
// every pair of their band is scored with the arithmetic of the full grid, so
// each band score equals the full-grid surface entry bitwise:
//   energy      k_band_energy  (the k_energy recurrence and accumulation
//               schedule; the AR-filtered sample is formed per pair, which
//               for pairs sharing a is the same value k_energy shares)
//   projection  proj16_item<BAND> (k_project16b with the tile's pairs
//               gathered column by column from the same split operand)
// The band optimum is the first minimum in pair order; it goes to the full
// grid when it has a lattice neighbour outside the band.
#pragma once

#include "reml_kernels.h"

namespace remlk {

constexpr int kBandWaves = 4;  // energy chunks per block (one wave each)

// Interleaved per-chunk energy coefficients of every band:
// coef[chunk][t][3 j + (0, 1, 2)] = (c_t, alpha_t, 1/v_{t+1}) of slot j, the
// same fp32 roundings as the full-grid chunk tables; s0 = 1/v_0.
__global__ void k_band_tables(PlanDev P, float* __restrict__ coef, float* __restrict__ s0) {
    const int MC = P.band_sb / kEnergyNB;
    const int cb = blockIdx.x;
    if (cb >= P.n_bands * MC) return;
    const int bin = cb / MC, chunk = cb % MC;
    const int T = P.T;
    const int32_t* slots = P.band_pairs + (size_t)bin * P.band_sb + chunk * kEnergyNB;
    for (int t = threadIdx.x; t < T; t += blockDim.x) {
        float* o = coef + ((size_t)cb * T + t) * 3 * kEnergyNB;
#pragma unroll
        for (int j = 0; j < kEnergyNB; ++j) {
            const int k = slots[j];
            float c = 0.f, al = 0.f, sn = 0.f;
            if (k >= 0) {
                const float2 q = P.lp[(size_t)k * T + t];
                c = P.lc[(size_t)k * T + t];
                al = q.x;
                sn = q.y;
            }
            o[3 * j] = c;
            o[3 * j + 1] = al;
            o[3 * j + 2] = sn;
        }
    }
    if (threadIdx.x < kEnergyNB) {
        const int k = slots[threadIdx.x];
        s0[(size_t)bin * P.band_sb + chunk * kEnergyNB + threadIdx.x] = k >= 0 ? P.ls0[k] : 0.f;
    }
}

__device__ inline float axis_clamp(const float* axis, int n, float x) {
    float lo = axis[0], hi = axis[0];
    for (int i = 1; i < n; ++i) {
        lo = fminf(lo, axis[i]);
        hi = fmaxf(hi, axis[i]);
    }
    return fminf(fmaxf(x, lo), hi);
}

// Band of each voxel still flagged after the window search (need[v] != 0)
// whose proposal is finite: lambda of the proposal clamped to the lattice
// box, nearest band centre; -1 for every other voxel and for empty bands.
__global__ void k_band_key(PlanDev P, const float* __restrict__ props, int V, const int32_t* __restrict__ need,
                           int32_t* __restrict__ key) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    int out = -1;
    if (need[v] != 0) {
        const float ah = props[3 * v], bh = props[3 * v + 1];
        if (isfinite(ah) && isfinite(bh)) {
            const double a = axis_clamp(P.a_axis, P.n_a, ah), b = axis_clamp(P.b_axis, P.n_b, bh);
            const double den = 1.0 + 2.0 * a * b + b * b;
            double lam = den > 0.0 ? (1.0 + a * b) * (a + b) / den : 0.0;
            lam = fmin(fmax(lam, -1.0), 1.0);
            int bin = (int)rint((lam - (double)P.band_lo) / (double)P.band_h);
            bin = min(max(bin, 0), P.n_bands - 1);
            if (P.band_n[bin] > 0) out = bin;
        }
    }
    key[v] = out;
}

// Band energies E[col][slot] for every live 64-column tile of the grouped
// layout and every 8-slot chunk of its band.  Work items (tile, chunk group)
// are split into 8 contiguous tile ranges, one per XCD (block id % 8), whose
// blocks stride over them, so the chunk groups of one tile share that XCD's L2.
__global__ void __launch_bounds__(64 * kBandWaves) k_band_energy(PlanDev P, const float* __restrict__ Yt, int64_t ldY,
                                                                 const int32_t* __restrict__ vb_bin,
                                                                 const int32_t* __restrict__ n_vb,
                                                                 const int32_t* __restrict__ off,
                                                                 const int32_t* __restrict__ cnt,
                                                                 float* __restrict__ Eb) {
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int SB = P.band_sb, MC = SB / kEnergyNB;
    const int n_cg = (MC + kBandWaves - 1) / kBandWaves;
    const int n_t = *n_vb * (kBandTile / 64);
    const int xcd = blockIdx.x & 7, lb = blockIdx.x >> 3, nlb = gridDim.x >> 3;
    const int per = (n_t + 7) >> 3;
    const int t_lo = xcd * per, t_hi = min(n_t, t_lo + per);
    const int n_items = max(0, t_hi - t_lo) * n_cg;
    const int T = P.T;
    constexpr int U = 8;  // steps per fp64 flush
    constexpr int G = 1;  // steps whose coefficients are fetched together (24 scalars each)
    const int nsteps = T - 1;
    const int nfull = nsteps / U * U;
    for (int it = lb; it < n_items; it += nlb) {
        const int tt = t_lo + it / n_cg;
        const int chunk = (it % n_cg) * kBandWaves + w;
        const int bin = vb_bin[tt / (kBandTile / 64)];
        const int col0 = tt * 64;
        if (col0 >= off[bin] + cnt[bin]) continue;     // padding tile of the group
        if (chunk * kEnergyNB >= P.band_n[bin]) continue;  // past the band
        const int col = col0 + lane;
        const float* __restrict__ y = Yt + col;
        const float* __restrict__ co = P.band_coef + ((size_t)bin * MC + chunk) * T * 3 * kEnergyNB;
        const float* __restrict__ s0 = P.band_s0 + (size_t)bin * SB + chunk * kEnergyNB;
        float e[kEnergyNB], part[kEnergyNB];
        double acc[kEnergyNB];
        const float y0 = y[0];
#pragma unroll
        for (int j = 0; j < kEnergyNB; ++j) {
            e[j] = y0;
            part[j] = y0 * y0 * s0[j];
            acc[j] = 0.0;
        }
        float yc = y0;
        auto step = [&](float yn, const float* c) __attribute__((always_inline)) {
#pragma unroll
            for (int j = 0; j < kEnergyNB; ++j) {
                const float z = fmaf(-c[3 * j], yc, yn);
                e[j] = fmaf(c[3 * j + 1], e[j], z);
                part[j] = fmaf(e[j] * c[3 * j + 2], e[j], part[j]);
            }
            yc = yn;
        };
        float ybuf[U];
#pragma unroll
        for (int i = 0; i < U; ++i) ybuf[i] = y[(size_t)min(1 + i, T - 1) * ldY];
        for (int t0 = 0; t0 < nfull; t0 += U) {
            float ycur[U];
#pragma unroll
            for (int i = 0; i < U; ++i) ycur[i] = ybuf[i];
#pragma unroll
            for (int i = 0; i < U; ++i) ybuf[i] = y[(size_t)min(t0 + U + 1 + i, T - 1) * ldY];
#pragma unroll
            for (int g0 = 0; g0 < U; g0 += G) {
                float cg[G][3 * kEnergyNB];
#pragma unroll
                for (int i = 0; i < G; ++i)
#pragma unroll
                    for (int q = 0; q < 3 * kEnergyNB; ++q) cg[i][q] = co[(size_t)(t0 + g0 + i) * 3 * kEnergyNB + q];
#pragma unroll
                for (int i = 0; i < G; ++i) step(ycur[g0 + i], cg[i]);
            }
#pragma unroll
            for (int j = 0; j < kEnergyNB; ++j) {
                acc[j] += part[j];
                part[j] = 0.f;
            }
        }
        for (int t = nfull; t < nsteps; ++t) step(y[(size_t)(t + 1) * ldY], co + (size_t)t * 3 * kEnergyNB);
        float* out = Eb + (size_t)col * SB + chunk * kEnergyNB;
#pragma unroll
        for (int j = 0; j < kEnergyNB; ++j) out[j] = (float)(acc[j] + part[j]);
    }
}

// Band projection + score over (256-column block, band tile) items, split into
// 8 contiguous block ranges (one per XCD) as in k_band_energy.
template <int BN, bool BF16>
__global__ void __launch_bounds__(256) k_band_project16(PlanDev P, const float* __restrict__ Yt, int64_t ldY,
                                                        int ncols, const int32_t* __restrict__ vb_bin,
                                                        const int32_t* __restrict__ n_vb, float* __restrict__ Sb) {
    __shared__ __attribute__((aligned(16))) unsigned char lds_raw[Proj16Lds<BN>::BYTES];
    const int SB = P.band_sb, npt = P.npt;
    const int MT = SB / npt;
    const int nvb = *n_vb;
    const int xcd = blockIdx.x & 7, lb = blockIdx.x >> 3, nlb = gridDim.x >> 3;
    const int per = (nvb + 7) >> 3;
    const int v_lo = xcd * per, v_hi = min(nvb, v_lo + per);
    const int n_items = max(0, v_hi - v_lo) * MT;
    for (int it = lb; it < n_items; it += nlb) {
        const int vb = v_lo + it / MT, tile = it % MT;
        const int bin = vb_bin[vb];
        if (tile * npt >= P.band_n[bin]) continue;  // block-uniform
        proj16_item<BN, BF16, true>(P, Yt, ldY, ncols, Sb, SB, vb, tile,
                                    P.band_pairs + (size_t)bin * SB + tile * npt, tile * npt, lds_raw);
    }
}

// First minimum over the band's slots (pair order), the lambda == 0 pairs
// mapped to the first of them as the global first-minimum rule does; need = 1
// when the optimum has a lattice neighbour outside the band (or no slot has a
// finite score), else 0.
__global__ void k_band_select(PlanDev P, const int32_t* __restrict__ rows, const int32_t* __restrict__ vb_bin,
                              const int32_t* __restrict__ n_vb, const float* __restrict__ Sb,
                              int32_t* __restrict__ best, float* __restrict__ best_score,
                              int32_t* __restrict__ need) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= *n_vb * kBandTile) return;
    const int v = rows[c];
    if (v < 0) return;
    const int bin = vb_bin[c / kBandTile];
    const int SB = P.band_sb, nb = P.band_n[bin];
    const float* sr = Sb + (size_t)c * SB;
    float bs = INFINITY;
    int bslot = -1;
    for (int s = 0; s < nb; ++s) {
        const float sc = sr[s];
        if (sc < bs) {
            bs = sc;
            bslot = s;
        }
    }
    if (bslot < 0) {
        need[v] = 1;
        return;
    }
    int k = P.band_pairs[(size_t)bin * SB + bslot];
    const int edge = P.band_edge[(size_t)bin * SB + bslot];
    if (P.white_flag[k]) k = P.white[0];
    best[v] = k;
    best_score[v] = bs;
    need[v] = edge ? 1 : 0;
}

}  // namespace remlk
