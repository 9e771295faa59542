// Split 16-bit projection GEMM with the REML score epilogue, 256-voxel blocks.
// Included by reml_api.hip after reml_kernels.hip (one translation unit).
#pragma once

#include "reml_kernels.h"

namespace remlk {

constexpr int kProj16BM = 256;  // voxels per block (4 waves x 64 rows)

// ---------------------------------------------------------------------------
// (ii) C = Ah Bh + Ah Bl + Al Bh on 32x32x16 MFMA (f16 or bf16 halves, fp32
// accumulation), then per pair k of the column tile:
//   P = sum_j C[v][slot*p + j]^2 (fp64, j ascending; x 2^-2e for fp16),
//   score = ld_k + (T - p) log(E - P + 1e-10) over E in the surface
// (glm/arma.py:1988-2004).  Per 16-sample k-step the three products are
// accumulated in the order al*bh, ah*bl, ah*bh, k-steps ascending in t: the
// window kernel k_local_project16 uses the same order and halves.
// Each wave owns 64 voxel rows (two 32-row MFMA blocks).  Its A operands are
// the hi / lo halves of the residual that prep wrote in rows Tpad..2 Tpad-1 of
// Yt as [t/8][column][8] (one 16-byte load per operand, each used by one lane
// of the block, so no LDS); the B halves of the column tile are staged in LDS
// and each B read serves both row blocks.  Consecutive MFMAs update different
// accumulators.
// ---------------------------------------------------------------------------
template <bool BF16>
__device__ __forceinline__ f32x16 mfma16u(u32x4 a, u32x4 b, f32x16 c) {
    if (BF16)
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                       0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c,
                                                      0, 0, 0);
}

template <int BN>
struct Proj16Lds {
    static constexpr int NTB = kProjBK / 8;
    static constexpr int RP = NTB * 8 + 8;  // B row pitch in halves (16-byte aligned rows)
    static constexpr int B_H = BN * RP;     // halves per plane
    static constexpr int C_LD = BN + 1;     // epilogue patch pitch (floats)
    static constexpr int STAGE_BYTES = 2 * B_H * 2;
    static constexpr int PATCH_BYTES = 4 * 32 * C_LD * 4;
    static constexpr int BYTES = STAGE_BYTES > PATCH_BYTES ? STAGE_BYTES : PATCH_BYTES;
};

// One (256-voxel block, column tile) of the split projection: S[v * ldS + k].
template <int BN, bool BF16>
__device__ __forceinline__ void proj16_item(const PlanDev& P, const float* __restrict__ Yt, int64_t ldY, int V,
                                            float* __restrict__ S, int64_t ldS, int vt, int tile,
                                            unsigned char* lds_raw) {
    using L = Proj16Lds<BN>;
    constexpr int BM = kProj16BM, NBLK = BN / 32, NTB = L::NTB, RP = L::RP, C_LD = L::C_LD;
    constexpr int B_N = 2 * NTB * BN;  // 16-byte B units per stage
    constexpr int B_U = B_N / 256;
    static_assert(B_N % 256 == 0, "B stage must split evenly over the block");
    uint16_t* Bh = reinterpret_cast<uint16_t*>(lds_raw);
    uint16_t* Bl = Bh + L::B_H;

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int v0 = vt * BM;
    const uint16_t* __restrict__ gh = BF16 ? P.Gh_bf16 : P.Gh_f16;
    const uint16_t* __restrict__ gl = BF16 ? P.Gl_bf16 : P.Gl_f16;
    // this lane's A column (voxel) of row block r: v0 + 64 w + 32 r + lane % 32;
    // k-slot half: samples 8 (lane / 32) .. + 8 of each 16-sample k-step
    const uint16_t* __restrict__ Ash = reinterpret_cast<const uint16_t*>(Yt + (size_t)P.Tpad * ldY);
    const uint16_t* __restrict__ Asl = Ash + (size_t)P.Tpad * ldY;
    const int a_col = v0 + 64 * w + (lane & 31);
    const int a_tb = lane >> 5;
    // source offset (halves, stage 0) of each B unit this thread stages
    size_t ub[B_U];
#pragma unroll
    for (int u = 0; u < B_U; ++u) {
        const int rr = (tid + u * 256) % (NTB * BN);
        ub[u] = (size_t)tile * P.Tpad * BN + ((size_t)(rr / BN) * BN + rr % BN) * 8;
    }

    u32x4 ah[2][2], al[2][2];  // [row block][k-step of the stage]
    u32x4 rb[B_U];
    auto load_a = [&](int kt) __attribute__((always_inline)) {
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) {
                const size_t o = ((size_t)(kt * NTB + 2 * ks + a_tb) * ldY + a_col + 32 * r) * 8;
                ah[r][ks] = *reinterpret_cast<const u32x4*>(Ash + o);
                al[r][ks] = *reinterpret_cast<const u32x4*>(Asl + o);
            }
    };
    auto load_b = [&](int kt) __attribute__((always_inline)) {
#pragma unroll
        for (int u = 0; u < B_U; ++u) {
            const int plane = (tid + u * 256) / (NTB * BN);
            rb[u] = *reinterpret_cast<const u32x4*>((plane ? gl : gh) + ub[u] + (size_t)kt * NTB * BN * 8);
        }
    };

    f32x16 acc[2][NBLK];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int i = 0; i < NBLK; ++i)
#pragma unroll
            for (int q = 0; q < 16; ++q) acc[r][i][q] = 0.f;
    const int nk = P.Tpad / kProjBK;
    load_a(0);
    load_b(0);
    for (int kt = 0; kt < nk; ++kt) {
#pragma unroll
        for (int u = 0; u < B_U; ++u) {
            const int q = tid + u * 256;
            const int plane = q / (NTB * BN), rr = q % (NTB * BN);
            const int tb = rr / BN, col = rr % BN;
            *reinterpret_cast<u32x4*>((plane ? Bl : Bh) + col * RP + tb * 8) = rb[u];
        }
        // this stage's A in registers; fetch the next stage behind the MFMAs
        u32x4 ahi[2][2], alo[2][2];
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) {
                ahi[r][ks] = ah[r][ks];
                alo[r][ks] = al[r][ks];
            }
        __syncthreads();
        if (kt + 1 < nk) {
            load_a(kt + 1);
            load_b(kt + 1);
        }
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            const int tb = 2 * ks + (lane >> 5);
            u32x4 bh[NBLK], bl[NBLK];
#pragma unroll
            for (int n = 0; n < NBLK; ++n) {
                const int col = n * 32 + (lane & 31);
                bh[n] = *reinterpret_cast<const u32x4*>(Bh + col * RP + tb * 8);
                bl[n] = *reinterpret_cast<const u32x4*>(Bl + col * RP + tb * 8);
            }
            // per accumulator: al*bh, then ah*bl, then ah*bh
#pragma unroll
            for (int n = 0; n < NBLK; ++n)
#pragma unroll
                for (int r = 0; r < 2; ++r) acc[r][n] = mfma16u<BF16>(alo[r][ks], bh[n], acc[r][n]);
#pragma unroll
            for (int n = 0; n < NBLK; ++n)
#pragma unroll
                for (int r = 0; r < 2; ++r) acc[r][n] = mfma16u<BF16>(ahi[r][ks], bl[n], acc[r][n]);
#pragma unroll
            for (int n = 0; n < NBLK; ++n)
#pragma unroll
                for (int r = 0; r < 2; ++r) acc[r][n] = mfma16u<BF16>(ahi[r][ks], bh[n], acc[r][n]);
        }
        __syncthreads();
    }
    // Epilogue per row block through a wave-private patch (32 rows x BN).
    float* Cw = reinterpret_cast<float*>(lds_raw) + w * 32 * C_LD;
    const int p = P.p, npt = P.npt, K = P.K;
    const double tp = (double)(P.T - p);
#pragma unroll
    for (int r = 0; r < 2; ++r) {
#pragma unroll
        for (int n = 0; n < NBLK; ++n)
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int row = 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
                Cw[row * C_LD + n * 32 + (lane & 31)] = acc[r][n][q];
            }
        __syncthreads();
        for (int it = lane; it < 32 * npt; it += 64) {
            const int row = it / npt, qq = it % npt;
            const int k = tile * npt + qq;
            const int v = v0 + 64 * w + 32 * r + row;
            if (k >= K || v >= V) continue;
            const float* cr = Cw + row * C_LD + qq * p;
            double pp = 0.0;
            for (int j = 0; j < p; ++j) {
                const double cv = (double)cr[j];
                pp = fma(cv, cv, pp);
            }
            if (!BF16) pp *= (double)P.g16_unscale[k];
            float* sp = S + (size_t)v * ldS + k;
            const double rss = (double)(*sp) - pp;
            *sp = (float)fma(tp, log(rss + 1e-10), P.ld[k]);
        }
        __syncthreads();
    }
}

template <int BN, bool BF16>
__global__ void __launch_bounds__(256) k_project16b(PlanDev P, const float* __restrict__ Yt, int64_t ldY, int V,
                                                    float* __restrict__ S, int64_t ldS,
                                                    const int32_t* __restrict__ n_live,
                                                    const PlanDev* __restrict__ plans,
                                                    const int32_t* __restrict__ grp) {
    __shared__ __attribute__((aligned(16))) unsigned char lds_raw[Proj16Lds<BN>::BYTES];
    int vt, tile;
    if (!xcd_tile_map(blockIdx.x, gridDim.x, P.n_tiles, &vt, &tile)) return;
    if (n_live && vt * kProj16BM >= *n_live) return;  // past the live voxels
    if (plans) {  // grouped layout: the design of this voxel tile (grp per kProjBM-voxel tile)
        const int g = grp[vt * (kProj16BM / kProjBM)];
        if (g < 0) return;
        P = plans[g];
    }
    proj16_item<BN, BF16>(P, Yt, ldY, V, S, ldS, vt, tile, lds_raw);
}

}  // namespace remlk
