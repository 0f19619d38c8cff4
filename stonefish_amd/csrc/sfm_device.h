// Device code of the surrogate network's fused inference, shared by the
// network library (sfh_mlp.hip: sfm_forward) and the hydrodynamics library
// (sfh_hydro.hip: the one-launch surrogate step).  gfx950 only.
//
// A workgroup of four wavefronts runs TR rows (samples) of one network
// through all five layers.  The hidden activations stay in LDS (two TR x 256
// BF16 buffers, row-major [sample][feature], row stride kActStride).  Each
// layer is computed transposed, H'^T = W H^T: the weights are the MFMA's A
// operand (16-byte loads of W[feature][k..k+7] from global memory; 405 KB of
// BF16 per network, L2-resident) and the activations its B operand (16-byte
// LDS reads of H[sample][k..k+7]), so an accumulator lane holds four
// consecutive output features of one sample and stores them with one 8-byte
// LDS write.  Wavefront w owns output features [64 w, 64 w + 64) of each
// hidden layer; the output layer is one 32-feature tile per 32-sample tile.
//
// v_mfma_f32_32x32x16_bf16 operand layout (checked on the device by
// sfm_probe_mfma): lane l holds A[l % 32][8 (l / 32) + 0..7] and
// B[8 (l / 32) + 0..7][l % 32]; accumulator register r of lane l is
// C[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/sfmlp.h"
#include "sfm_gelu.h"

namespace sfm {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

static constexpr int kActStride = SFM_HIDDEN + 8;   // LDS row stride (elements): spreads rows over banks

__device__ __forceinline__ bf16x8 ld8(const uint16_t* p) { return *reinterpret_cast<const bf16x8*>(p); }

// FP32 -> BF16 bits, round to nearest even (v_cvt_pk_bf16_f32).
__device__ __forceinline__ uint16_t f2bf(float f)
{
    const __bf16 h = (__bf16)f;
    return *reinterpret_cast<const uint16_t*>(&h);
}

// Two FP32 -> packed BF16 pair (low half = a).
__device__ __forceinline__ uint32_t f2bf2(float a, float b)
{
    bf16x2 v;
    v[0] = (__bf16)a;
    v[1] = (__bf16)b;
    return *reinterpret_cast<const uint32_t*>(&v);
}

// Row of the 32x32 accumulator tile held by register r of lane `lane`.
__device__ __forceinline__ int acc_row(int r, int lane) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

// One hidden layer of a TR-row tile: out = gelu(in W^T + b) in BF16, in rows
// at `in` with row stride in_stride (LDS, or global memory for the network
// input), K inputs.  The weights of the next k-step are loaded while the
// current one is multiplied.
template <int TR, int K>
__device__ __forceinline__ void mlp_layer(const uint16_t* __restrict__ w, const float* __restrict__ bias,
                                          const uint16_t* in, int in_stride, uint16_t* out)
{
    constexpr int RT = TR / 32;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 31, lk = (lane >> 5) * 8, lq = (lane >> 5) * 4;
    f32x16 acc[2][RT];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int nt = 0; nt < RT; ++nt) acc[mt][nt] = f32x16{};
    const uint16_t* wp0 = w + (size_t)(64 * wave + lr) * K + lk;
    const uint16_t* wp1 = wp0 + (size_t)32 * K;
    const uint16_t* bp = in + lr * in_stride + lk;
    bf16x8 a0 = ld8(wp0), a1 = ld8(wp1);
#pragma unroll
    for (int ks = 0; ks < K; ks += 16) {
        bf16x8 n0 = a0, n1 = a1;
        if (ks + 16 < K) {
            n0 = ld8(wp0 + ks + 16);
            n1 = ld8(wp1 + ks + 16);
        }
#pragma unroll
        for (int nt = 0; nt < RT; ++nt) {
            const bf16x8 b = ld8(bp + 32 * nt * in_stride + ks);
            acc[0][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b, acc[0][nt], 0, 0, 0);
            acc[1][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b, acc[1][nt], 0, 0, 0);
        }
        a0 = n0;
        a1 = n1;
    }
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int f0 = 64 * wave + 32 * mt + 8 * q + lq;   // four consecutive output features
            const float4 bq = *reinterpret_cast<const float4*>(bias + f0);
#pragma unroll
            for (int nt = 0; nt < RT; ++nt) {
                const f32x16& A = acc[mt][nt];
                uint2 pk;
                pk.x = f2bf2(gelu(A[4 * q] + bq.x), gelu(A[4 * q + 1] + bq.y));
                pk.y = f2bf2(gelu(A[4 * q + 2] + bq.z), gelu(A[4 * q + 3] + bq.w));
                *reinterpret_cast<uint2*>(out + (32 * nt + lr) * kActStride + f0) = pk;
            }
        }
    }
    __syncthreads();
}

// The four hidden layers of one TR-row tile.  Layer 0 reads its input rows
// from global memory (xg, row-major [row][32], rows xg .. xg + TR) when xg is
// not null, else from act[1] (row stride kActStride, columns 0..31).  On
// return act[1] holds the last hidden layer.  Each layer ends with a barrier.
template <int TR>
__device__ __forceinline__ void mlp_hidden(const uint16_t* const* W, const float* const* B,
                                           uint16_t (*act)[TR * kActStride], const uint16_t* __restrict__ xg)
{
    if (xg)
        mlp_layer<TR, SFM_IN_PAD>(W[0], B[0], xg, SFM_IN_PAD, act[0]);
    else
        mlp_layer<TR, SFM_IN_PAD>(W[0], B[0], act[1], kActStride, act[0]);
    mlp_layer<TR, SFM_HIDDEN>(W[1], B[1], act[0], kActStride, act[1]);
    mlp_layer<TR, SFM_HIDDEN>(W[2], B[2], act[1], kActStride, act[0]);
    mlp_layer<TR, SFM_HIDDEN>(W[3], B[3], act[0], kActStride, act[1]);
}

// Output layer of sample tile `wave` (< TR / 32) from act[1]: on return,
// registers 0..3 of lane l hold outputs 4 (l / 32) + 0..3 (bias included) of
// sample 32 wave + l % 32.
__device__ __forceinline__ void mlp_output(const uint16_t* __restrict__ W, const float* __restrict__ bias,
                                           const uint16_t* __restrict__ in, float o[4])
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 31, lk = (lane >> 5) * 8, lq = (lane >> 5) * 4;
    f32x16 acc = {};
    const uint16_t* wp = W + (size_t)lr * SFM_HIDDEN + lk;
    for (int ks = 0; ks < SFM_HIDDEN; ks += 16)
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ld8(wp + ks), ld8(in + (32 * wave + lr) * kActStride + ks + lk),
                                                      acc, 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; ++i) o[i] = acc[i] + (lq + i < SFM_OUT_PAD ? bias[lq + i] : 0.f);
}

}  // namespace sfm
