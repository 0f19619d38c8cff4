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

// Loads through a global-address-space pointer (global_load, not flat_load:
// a flat load also counts as an LDS access, and waiting for one waits for
// every memory access issued after it).
typedef const __attribute__((address_space(1))) bf16x8 gbf16x8;
__device__ __forceinline__ bf16x8 ld8g(const uint16_t* p) { return *(gbf16x8*)(p); }

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

// Fragment order of a layer's BF16 weights W [out x K] (row-major indices r,
// c) for the fused inference: the 8 consecutive columns 16 s + 8 h .. + 7 of
// row r = 64 w + 32 m + i are the 16 bytes that lane i + 32 h of wavefront w
// feeds to k-step s of its m-th 32-row tile, and each (w, m, s) owns 1 KB of
// contiguous memory, read by one fully coalesced load instruction.  (Row-major
// rows are 512 bytes apart, so the same fragments would touch 32 cache lines
// per instruction.)
__host__ __device__ __forceinline__ size_t frag_index(int r, int c, int K)
{
    const int w = r >> 6, m = (r >> 5) & 1, i = r & 31, s = c >> 4, h = (c >> 3) & 1, e = c & 7;
    return ((((size_t)(2 * w + m) * (K / 16) + s) * 64) + i + 32 * h) * 8 + e;
}

// Weight fragments of one wavefront for a hidden layer with K inputs: rows
// 64 w + l % 32 and 64 w + 32 + l % 32 of W, columns 8 (l / 32) + 16 s.
template <int K>
struct LayerW {
    bf16x8 a[2][K / 16];
};

// Requests a layer's weight fragments (fragment order, frag_index) (2 K / 16 16-byte loads per lane, all
// issued at once: one L2 round trip per layer instead of one per k-step).
// The weights do not depend on the activations, so a caller issues them one
// layer ahead, before the previous layer's epilogue.
template <int K>
__device__ __forceinline__ void load_layer(const uint16_t* __restrict__ w, LayerW<K>& L)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint16_t* wp0 = w + ((size_t)(2 * wave) * (K / 16) * 64 + lane) * 8;   // fragment order
    const uint16_t* wp1 = wp0 + (size_t)(K / 16) * 512;
#pragma unroll
    for (int s = 0; s < K / 16; ++s) {
        L.a[0][s] = ld8g(wp0 + 512 * s);
        L.a[1][s] = ld8g(wp1 + 512 * s);
    }
}

// acc = W^T-tile products of a TR-row tile: in rows at `in` with row stride
// in_stride (LDS, or global memory for the network input), K inputs.
template <int TR, int K>
__device__ __forceinline__ void layer_mma(const LayerW<K>& L, const uint16_t* in, int in_stride,
                                          f32x16 (&acc)[2][TR / 32])
{
    constexpr int RT = TR / 32;
    const int lane = threadIdx.x & 63;
    const uint16_t* bp = in + (lane & 31) * in_stride + (lane >> 5) * 8;
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int nt = 0; nt < RT; ++nt) acc[mt][nt] = f32x16{};
#pragma unroll
    for (int s = 0; s < K / 16; ++s) {
#pragma unroll
        for (int nt = 0; nt < RT; ++nt) {
            const bf16x8 b = ld8(bp + 32 * nt * in_stride + 16 * s);
            acc[0][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(L.a[0][s], b, acc[0][nt], 0, 0, 0);
            acc[1][nt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(L.a[1][s], b, acc[1][nt], 0, 0, 0);
        }
    }
}

// out = gelu(acc + b) in BF16 (LDS rows, stride kActStride), then a barrier.
// The bias is read from LDS: a global load here would have to wait for the
// next layer's weight loads issued before it (loads complete in order).
template <int TR>
__device__ __forceinline__ void layer_store(const f32x16 (&acc)[2][TR / 32], const float* bias, uint16_t* out)
{
    constexpr int RT = TR / 32;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 31, lq = (lane >> 5) * 4;
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

// Output-layer weight fragments of one lane: W[l % 32][8 (l / 32) + 16 s].
struct OutW {
    bf16x8 a[SFM_HIDDEN / 16];
};

__device__ __forceinline__ void load_output(const uint16_t* __restrict__ w, OutW& L)
{
    const uint16_t* wp = w + (size_t)(threadIdx.x & 63) * 8;   // fragment order
#pragma unroll
    for (int s = 0; s < SFM_HIDDEN / 16; ++s) L.a[s] = ld8g(wp + 512 * s);
}

// Hidden-layer biases of one network into LDS (bias[l][f], l < 4); the
// caller puts a barrier between this and mlp_hidden.
__device__ __forceinline__ void load_bias(const float* const* B, float (*bias)[SFM_HIDDEN])
{
    const int t = threadIdx.x;
    for (int i = t; i < (SFM_LAYERS - 1) * SFM_HIDDEN; i += blockDim.x)
        bias[i / SFM_HIDDEN][i % SFM_HIDDEN] = B[i / SFM_HIDDEN][i % SFM_HIDDEN];
}

// The four hidden layers of one TR-row tile, given layer 0's weight fragments
// (requested by the caller, possibly long before) and the hidden biases in
// LDS (load_bias).  Layer 0 reads its input
// rows from global memory (xg, row-major [row][32], rows xg .. xg + TR) when
// xg is not null, else from act[1] (row stride kActStride, columns 0..31).
// On return act[1] holds the last hidden layer and wavefronts w < TR / 32
// hold the output layer's fragments in ow.
template <int TR>
__device__ __forceinline__ void mlp_hidden(const uint16_t* const* W, const float (*B)[SFM_HIDDEN],
                                           uint16_t (*act)[TR * kActStride], const uint16_t* __restrict__ xg,
                                           const LayerW<SFM_IN_PAD>& w0, OutW& ow)
{
    f32x16 acc[2][TR / 32];
    LayerW<SFM_HIDDEN> wn;
    if (xg)
        layer_mma<TR, SFM_IN_PAD>(w0, xg, SFM_IN_PAD, acc);
    else
        layer_mma<TR, SFM_IN_PAD>(w0, act[1], kActStride, acc);
    load_layer<SFM_HIDDEN>(W[1], wn);
    layer_store<TR>(acc, B[0], act[0]);
    layer_mma<TR, SFM_HIDDEN>(wn, act[0], kActStride, acc);
    load_layer<SFM_HIDDEN>(W[2], wn);
    layer_store<TR>(acc, B[1], act[1]);
    layer_mma<TR, SFM_HIDDEN>(wn, act[1], kActStride, acc);
    load_layer<SFM_HIDDEN>(W[3], wn);
    layer_store<TR>(acc, B[2], act[0]);
    layer_mma<TR, SFM_HIDDEN>(wn, act[0], kActStride, acc);
    if ((threadIdx.x >> 6) < TR / 32) load_output(W[SFM_LAYERS - 1], ow);
    layer_store<TR>(acc, B[3], act[1]);
}

// Output layer of sample tile `wave` (< TR / 32) from act[1]: on return,
// registers 0..3 of lane l hold outputs 4 (l / 32) + 0..3 (bias included) of
// sample 32 wave + l % 32.
__device__ __forceinline__ void mlp_output(const OutW& L, const float* __restrict__ bias,
                                           const uint16_t* __restrict__ in, float o[4])
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 31, lk = (lane >> 5) * 8, lq = (lane >> 5) * 4;
    f32x16 acc = {};
#pragma unroll
    for (int s = 0; s < SFM_HIDDEN / 16; ++s)
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(L.a[s], ld8(in + (32 * wave + lr) * kActStride + 16 * s + lk),
                                                      acc, 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; ++i) o[i] = acc[i] + (lq + i < SFM_OUT_PAD ? bias[lq + i] : 0.f);
}

// Workgroup b of a grid of n in XCD-major order: the hardware deals
// workgroups to the eight XCDs round-robin (b % 8), so the workgroups that
// share an XCD (and its L2) get consecutive tiles, which mostly share one
// network's weights.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t n)
{
    const uint32_t q = n / 8, r = n % 8, x = b % 8, s = b / 8;
    return x * q + (x < r ? x : r) + s;
}

}  // namespace sfm
