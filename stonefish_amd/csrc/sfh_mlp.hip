// sfmlp: hand-written gfx950 BF16 MFMA kernels of the hydrodynamics surrogate
// (forward, backward, Adam).  No counterpart in the reference.
//
// Every matrix product is C = A * Bt^T with A [rows x K] and Bt [cols x K] both
// row-major and contiguous in K, so each lane's MFMA operands are one 16-byte
// load: the forward pass keeps transposed copies of its activations (H^T) and
// of the weights (W^T, refreshed by Adam) for the backward products.
//
//   forward   Z = H W^T + b,  H' = gelu(Z)                 (A = H,    Bt = W)
//   dH        dH = dZ W                                     (A = dZ,   Bt = W^T)
//   dW        dW = dZ^T H, split over the batch             (A = dZ^T, Bt = H^T)
//
// Tiles are 32x32 v_mfma_f32_32x32x16_bf16 accumulators: the products of the
// 256-wide layers run 64x64 per workgroup with the operands staged in LDS,
// the 32-wide ones one tile per wavefront.  The epilogues fuse bias, GELU
// (exact erf) and its derivative, the loss and its gradient, BF16 rounding
// and the transposed stores; the weight-gradient products also sum dZ's
// columns (the bias gradient).  Weight- and bias-gradient partials of the
// batch splits are summed in a fixed order by the Adam kernel.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/sfmlp.h"
#include "sfm_device.h"

using sfm::acc_row;
using sfm::bf16x8;
using sfm::f2bf;
using sfm::f2bf2;
using sfm::ld8g;
using sfm::f32x16;
using sfm::gelu;
using sfm::gelu_grad;
using sfm::kActStride;
using sfm::ld8;

static thread_local char g_err[512] = "";

static int fail(int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

#define HIPCHK(expr)                                                                    \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(SFM_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));    \
    } while (0)

extern "C" const char* sfm_last_error(void) { return g_err; }

static constexpr int kDims[SFM_LAYERS + 1] = {SFM_IN_PAD, SFM_HIDDEN, SFM_HIDDEN, SFM_HIDDEN, SFM_HIDDEN, SFM_OUT_PAD};
static constexpr int kReal[SFM_LAYERS + 1] = {SFM_IN, SFM_HIDDEN, SFM_HIDDEN, SFM_HIDDEN, SFM_HIDDEN, SFM_OUT};
static constexpr int kSplitK = 256;   // batch rows per weight-gradient split
static constexpr int kMaxSplits = 64; // batch splits of a weight-gradient product

// Batch splits of layer l's weight-gradient product at most: the 256 x 256
// products run 16 workgroups per split, the 32-wide ones 2, so the latter
// take more splits to fill the chip; every split's partial sum is read again
// by the Adam launch.
static constexpr int max_splits(int l) { return kDims[l] == SFM_HIDDEN && kDims[l + 1] == SFM_HIDDEN ? 32 : kMaxSplits; }

// ---------------------------------------------------------------------------
// numerics
// ---------------------------------------------------------------------------



// ---------------------------------------------------------------------------
// GEMM tile kernel
// ---------------------------------------------------------------------------

enum { EPI_HIDDEN = 0, EPI_OUT = 1, EPI_BWD = 2, EPI_WGRAD = 3 };

struct GemmArgs {
    const uint16_t* A;      // [rows x K], lda
    const uint16_t* Bt;     // [cols x K], ldb
    int lda, ldb;
    int rows, cols, K, ksplit;
    const float* bias;      // [cols]                      (HIDDEN, OUT)
    float* Z;               // [rows x cols] FP32          (HIDDEN: written; BWD: read)
    uint16_t* H;            // [rows x cols] BF16          (HIDDEN, OUT-train, BWD)
    uint16_t* Ht;           // [cols x rows] BF16          (optional)
    const float* target;    // [rows x SFM_OUT]            (OUT-train)
    float* y;               // [rows x SFM_OUT]            (OUT)
    float* loss;            //                              (OUT-train)
    float scale;            // 1 / (rows * SFM_OUT)         (OUT-train)
    float* part;            // [ksplit][rows x cols] FP32  (WGRAD)
    float* bpart;           // [ksplit][rows] FP32 row sums of A over each split (WGRAD, optional: bias gradient)
};

__device__ __forceinline__ float sum8(const bf16x8& v)
{
    return (((float)v[0] + (float)v[1]) + ((float)v[2] + (float)v[3])) +
           (((float)v[4] + (float)v[5]) + ((float)v[6] + (float)v[7]));
}


// The epilogue of a wavefront's TM x TN tiles of C: tile (ti, tj) covers rows
// (rt0 + ti) * 32 .. + 31 and columns (ct0 + tj) * 32 .. + 31.
template <int EPI, int TM, int TN>
__device__ __forceinline__ void gemm_store(const GemmArgs& g, const f32x16 (&acc)[TM][TN], int rt0, int ct0, int split,
                                           int lane)
{
    float sq = 0.f;
#pragma unroll
    for (int ti = 0; ti < TM; ++ti) {
#pragma unroll
        for (int tj = 0; tj < TN; ++tj) {
            const f32x16& A = acc[ti][tj];
            const int n = (ct0 + tj) * 32 + (lane & 31);
            const int mb = (rt0 + ti) * 32;
            if (EPI == EPI_WGRAD) {
                float* out = g.part + (size_t)split * g.rows * g.cols;
#pragma unroll
                for (int r = 0; r < 16; ++r) out[(size_t)(mb + acc_row(r, lane)) * g.cols + n] = A[r];
                continue;
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                uint16_t hv[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int r = 4 * q + i;
                    const int m = mb + acc_row(r, lane);
                    const size_t idx = (size_t)m * g.cols + n;
                    float out;
                    if (EPI == EPI_HIDDEN) {
                        const float z = A[r] + g.bias[n];
                        if (g.Z) g.Z[idx] = z;
                        out = gelu(z);
                    } else if (EPI == EPI_OUT) {
                        const float z = A[r] + g.bias[n];
                        out = 0.f;
                        if (n < SFM_OUT) {
                            if (g.y) g.y[(size_t)m * SFM_OUT + n] = z;
                            if (g.target) {
                                const float d = z - g.target[(size_t)m * SFM_OUT + n];
                                sq += d * d;
                                out = 2.0f * d * g.scale;   // dLoss/dz
                            }
                        }
                    } else {   // EPI_BWD: dZ = dH * gelu'(Z)
                        out = A[r] * gelu_grad(g.Z[idx]);
                    }
                    hv[i] = f2bf(out);
                    if (g.H) g.H[idx] = hv[i];
                }
                if (g.Ht) {   // four consecutive rows -> one 8-byte store of the transposed tile
                    const int m0 = mb + 8 * q + 4 * (lane >> 5);
                    uint2 pk;
                    pk.x = (uint32_t)hv[0] | ((uint32_t)hv[1] << 16);
                    pk.y = (uint32_t)hv[2] | ((uint32_t)hv[3] << 16);
                    *reinterpret_cast<uint2*>(g.Ht + (size_t)n * g.rows + m0) = pk;
                }
            }
        }
    }
    if (EPI == EPI_OUT && g.target) {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) sq += __shfl_xor(sq, off, 64);
        if (lane == 0) atomicAdd(g.loss, sq * g.scale);
    }
}

// One wavefront computes a (32 TM) x (32 TN) block of C as TM x TN MFMA
// tiles: per k-step TM + TN 16-byte operand loads feed TM * TN MFMAs.
template <int EPI, int TM, int TN>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g)
{
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int blocks_n = g.cols / (32 * TN);
    const int bm = wave / blocks_n, bn = wave - bm * blocks_n;
    if (bm * 32 * TM >= g.rows) return;
    const int split = blockIdx.y;
    const int kper = g.K / g.ksplit;
    const int k0 = split * kper;
    const int kk = (lane >> 5) * 8;
    const uint16_t* ap[TM];
    const uint16_t* bp[TN];
#pragma unroll
    for (int i = 0; i < TM; ++i) ap[i] = g.A + (size_t)((bm * TM + i) * 32 + (lane & 31)) * g.lda + k0 + kk;
#pragma unroll
    for (int j = 0; j < TN; ++j) bp[j] = g.Bt + (size_t)((bn * TN + j) * 32 + (lane & 31)) * g.ldb + k0 + kk;
    f32x16 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = f32x16{};
    // The operands of the next kPF k-steps are in flight while the current
    // one multiplies (a ring of kPF register stages, refilled right after use):
    // a k-step waits for a load issued kPF steps earlier, not for the last one.
    constexpr int kPF = 4;
    bf16x8 ra[kPF][TM], rb[kPF][TN];
#pragma unroll
    for (int s = 0; s < kPF; ++s)
        if (16 * s < kper) {
#pragma unroll
            for (int i = 0; i < TM; ++i) ra[s][i] = ld8g(ap[i] + 16 * s);
#pragma unroll
            for (int j = 0; j < TN; ++j) rb[s][j] = ld8g(bp[j] + 16 * s);
        }
    // the bias gradient rides along the weight gradient: the wavefronts of the
    // first column of tiles also sum their rows of A (each lane 8 of the 16 k
    // of a k-step, the two half-waves combined at the end)
    const bool rsum = EPI == EPI_WGRAD && g.bpart && bn == 0;
    float rs[TM] = {};
    for (int k0 = 0; k0 < kper; k0 += 16 * kPF) {
#pragma unroll
        for (int s = 0; s < kPF; ++s) {
            const int k = k0 + 16 * s;
            if (k < kper) {
                if (rsum) {
#pragma unroll
                    for (int i = 0; i < TM; ++i) rs[i] += sum8(ra[s][i]);
                }
#pragma unroll
                for (int i = 0; i < TM; ++i)
#pragma unroll
                    for (int j = 0; j < TN; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ra[s][i], rb[s][j], acc[i][j], 0, 0, 0);
                if (k + 16 * kPF < kper) {
#pragma unroll
                    for (int i = 0; i < TM; ++i) ra[s][i] = ld8g(ap[i] + k + 16 * kPF);
#pragma unroll
                    for (int j = 0; j < TN; ++j) rb[s][j] = ld8g(bp[j] + k + 16 * kPF);
                }
            }
        }
    }
    if (rsum) {
#pragma unroll
        for (int i = 0; i < TM; ++i) {
            const float v = rs[i] + __shfl_xor(rs[i], 32, 64);
            if (lane < 32) g.bpart[(size_t)split * g.rows + (bm * TM + i) * 32 + lane] = v;
        }
    }
    gemm_store<EPI, TM, TN>(g, acc, bm * TM, bn * TN, split, lane);
}

// Workgroup-tiled product for the 256-wide layers: the four wavefronts of a
// workgroup, in a 2 x 2 arrangement, cover (64 WM) x (64 WN) of C, each
// wavefront WM x WN 32x32 tiles.  K advances in chunks of kKC: the workgroup
// copies the chunk's rows of A and of Bt into LDS once (the next chunk's
// global loads in flight while the current one multiplies; two LDS buffers)
// and the wavefronts read their MFMA operands from there, so an operand row
// comes from L2 once per workgroup instead of once per wavefront that uses it.
static constexpr int kKC = 64;            // K per chunk
static constexpr int kLdsRow = kKC + 8;   // bf16 per LDS row: 16 bytes of padding against bank conflicts

template <int EPI, int WM, int WN>
__global__ __launch_bounds__(256) void k_gemm_lds(GemmArgs g)
{
    constexpr int BM = 64 * WM, BN = 64 * WN;
    constexpr int SEG = kKC / 8;                 // 16-byte segments per row and chunk
    constexpr int SA = BM * SEG / 256, SB = BN * SEG / 256;
    static_assert(SA * 256 == BM * SEG && SB * 256 == BN * SEG, "tile rows must fill the workgroup's copies");
    __shared__ __attribute__((aligned(16))) uint16_t As[2][BM * kLdsRow];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[2][BN * kLdsRow];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, wm = w >> 1, wn = w & 1;
    const int blocks_n = g.cols / BN;
    const int bm = blockIdx.x / blocks_n, bn = blockIdx.x - bm * blocks_n;
    const int split = blockIdx.y;
    const int kper = g.K / g.ksplit;
    const int kbase = split * kper;
    const int nch = kper / kKC;
    bf16x8 ra[SA], rb[SB];
    auto gload = [&](int c) {
        const int k0 = kbase + c * kKC;
#pragma unroll
        for (int i = 0; i < SA; ++i) {
            const int idx = t + 256 * i, row = idx / SEG, seg = idx % SEG;
            ra[i] = ld8g(g.A + (size_t)(bm * BM + row) * g.lda + k0 + seg * 8);
        }
#pragma unroll
        for (int i = 0; i < SB; ++i) {
            const int idx = t + 256 * i, row = idx / SEG, seg = idx % SEG;
            rb[i] = ld8g(g.Bt + (size_t)(bn * BN + row) * g.ldb + k0 + seg * 8);
        }
    };
    // bias gradient (see k_gemm): the workgroups of the first column of
    // tiles sum the rows of A they copy, each thread its 8-value segments
    const bool rsum = EPI == EPI_WGRAD && g.bpart && bn == 0;
    float rs[SA] = {};
    auto lstore = [&](int buf) {
#pragma unroll
        for (int i = 0; i < SA; ++i) {
            const int idx = t + 256 * i, row = idx / SEG, seg = idx % SEG;
            *reinterpret_cast<bf16x8*>(&As[buf][row * kLdsRow + seg * 8]) = ra[i];
            if (rsum) rs[i] += sum8(ra[i]);
        }
#pragma unroll
        for (int i = 0; i < SB; ++i) {
            const int idx = t + 256 * i, row = idx / SEG, seg = idx % SEG;
            *reinterpret_cast<bf16x8*>(&Bs[buf][row * kLdsRow + seg * 8]) = rb[i];
        }
    };
    f32x16 acc[WM][WN];
#pragma unroll
    for (int i = 0; i < WM; ++i)
#pragma unroll
        for (int j = 0; j < WN; ++j) acc[i][j] = f32x16{};
    gload(0);
    lstore(0);
    __syncthreads();
    for (int c = 0; c < nch; ++c) {
        const int buf = c & 1;
        if (c + 1 < nch) gload(c + 1);
#pragma unroll
        for (int s = 0; s < kKC / 16; ++s) {
            const int kk = 16 * s + (lane >> 5) * 8;
            bf16x8 fa[WM], fb[WN];
#pragma unroll
            for (int i = 0; i < WM; ++i)
                fa[i] = *reinterpret_cast<const bf16x8*>(&As[buf][((wm * WM + i) * 32 + (lane & 31)) * kLdsRow + kk]);
#pragma unroll
            for (int j = 0; j < WN; ++j)
                fb[j] = *reinterpret_cast<const bf16x8*>(&Bs[buf][((wn * WN + j) * 32 + (lane & 31)) * kLdsRow + kk]);
#pragma unroll
            for (int i = 0; i < WM; ++i)
#pragma unroll
                for (int j = 0; j < WN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
        }
        if (c + 1 < nch) lstore(buf ^ 1);
        __syncthreads();
    }
    if (rsum) {   // the SEG threads of a row are consecutive lanes
        static_assert(SEG == 8, "row segments per chunk");
#pragma unroll
        for (int i = 0; i < SA; ++i) {
            float v = rs[i];
            v += __shfl_xor(v, 1, 64);
            v += __shfl_xor(v, 2, 64);
            v += __shfl_xor(v, 4, 64);
            const int idx = t + 256 * i;
            if (idx % SEG == 0) g.bpart[(size_t)split * g.rows + bm * BM + idx / SEG] = v;
        }
    }
    gemm_store<EPI, WM, WN>(g, acc, (bm * 2 + wm) * WM, (bn * 2 + wn) * WN, split, lane);
}

// ---------------------------------------------------------------------------
// fused inference: all five layers of a tile of rows in one workgroup
// ---------------------------------------------------------------------------

struct NetSet {
    int nnets;
    uint32_t row_end[SFM_MAX_GROUP];   // exclusive end row of each network's rows
    const uint16_t* W[SFM_MAX_GROUP][SFM_LAYERS];
    const float* b[SFM_MAX_GROUP][SFM_LAYERS];
};

// Inference of TR consecutive rows of one network per workgroup (the fused
// network of sfm_device.h).
template <int TR>
__global__ __launch_bounds__(256) void k_mlp_fused(NetSet ns, const uint16_t* __restrict__ x, uint32_t rows,
                                                   float* __restrict__ y)
{
    __shared__ __attribute__((aligned(16))) uint16_t act[2][TR * kActStride];
    __shared__ __attribute__((aligned(16))) float bias[SFM_LAYERS - 1][SFM_HIDDEN];
    const uint32_t row0 = sfm::xcd_tile(blockIdx.x, gridDim.x) * TR;
    if (row0 >= rows) return;
    int k = 0;
    while (k + 1 < ns.nnets && row0 >= ns.row_end[k]) ++k;
    sfm::LayerW<SFM_IN_PAD> w0;
    sfm::load_layer<SFM_IN_PAD>(ns.W[k][0], w0);
    sfm::load_bias(ns.b[k], bias);
    __syncthreads();
    sfm::OutW ow;
    sfm::mlp_hidden<TR>(ns.W[k], bias, act, x + (size_t)row0 * SFM_IN_PAD, w0, ow);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (wave < TR / 32) {
        float o[4];
        sfm::mlp_output(ow, ns.b[k][SFM_LAYERS - 1], act[1], o);
        const int lq = (lane >> 5) * 4;
        float* yo = y + (size_t)(row0 + 32 * wave + (lane & 31)) * SFM_OUT;
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (lq + i < SFM_OUT) yo[lq + i] = o[i];
    }
}

// ---------------------------------------------------------------------------
// small kernels
// ---------------------------------------------------------------------------

// Transposes x [rows x cols] (BF16) into xt [cols x rows].
__global__ void k_transpose(const uint16_t* __restrict__ x, int rows, int cols, uint16_t* __restrict__ xt)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)rows * cols) return;
    int r = (int)(i / cols), c = (int)(i - (size_t)r * cols);
    xt[(size_t)c * rows + r] = x[i];
}

// Adam on n parameters of a [rows x cols] matrix (cols = 1 for biases): the
// gradient is the sum of nsplit partials (fixed order); keeps the FP32 master
// copy and refreshes the BF16 copy and its transpose.
// One parameter tensor of an Adam step: its FP32 weights and moments, the
// gradient partials of the batch splits (summed in a fixed order), where the
// summed gradient is kept, and (weights only) the BF16 copies to refresh.
struct AdamSeg {
    float *w, *mo, *ve;
    const float* part;
    float* gsave;
    uint16_t *wb, *wbt, *wp;
    int nsplit, rows, cols, block0;   // block0: first workgroup of this tensor in the launch
};

struct AdamSet {
    AdamSeg seg[2 * SFM_LAYERS];
    int nseg;
};

// Adam with bias correction for every weight and bias tensor of the network
// in one launch (each workgroup finds its tensor from the block offsets).
__global__ __launch_bounds__(256) void k_adam(AdamSet S, float lr, float b1, float b2, float eps, float bc1,
                                              float bc2)
{
    int si = 0;
    while (si + 1 < S.nseg && (int)blockIdx.x >= S.seg[si + 1].block0) ++si;
    const AdamSeg& g = S.seg[si];
    const int n = g.rows * g.cols;
    const int i = ((int)blockIdx.x - g.block0) * (int)blockDim.x + (int)threadIdx.x;
    if (i >= n) return;
    float gr = 0.f;
#pragma unroll 8
    for (int s = 0; s < g.nsplit; ++s) gr += g.part[(size_t)s * n + i];   // loads in flight 8 at a time
    g.gsave[i] = gr;
    const float m = b1 * g.mo[i] + (1.f - b1) * gr;
    const float v = b2 * g.ve[i] + (1.f - b2) * gr * gr;
    g.mo[i] = m;
    g.ve[i] = v;
    const float wn = g.w[i] - lr * (m / bc1) / (sqrtf(v / bc2) + eps);
    g.w[i] = wn;
    if (g.wb) {
        const uint16_t h = f2bf(wn);
        g.wb[i] = h;
        const int r = i / g.cols, c = i - r * g.cols;
        g.wbt[(size_t)c * g.rows + r] = h;
        g.wp[sfm::frag_index(r, c, g.cols)] = h;
    }
}

__global__ void k_to_bf16(const float* __restrict__ x, uint16_t* __restrict__ y, uint64_t n)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

// One 32x32x16 product with the operand mapping of k_gemm: a [32 x 16],
// b [16 x 32] row-major; c [32 x 32] row-major.
__global__ void k_probe(const uint16_t* __restrict__ a, const uint16_t* __restrict__ b, float* __restrict__ c)
{
    const int lane = threadIdx.x;
    const int kk = (lane >> 5) * 8;
    bf16x8 av, bv;
    for (int i = 0; i < 8; ++i) {
        uint16_t x = a[(lane & 31) * 16 + kk + i];
        uint16_t y = b[(kk + i) * 32 + (lane & 31)];
        av[i] = *reinterpret_cast<const __bf16*>(&x);
        bv[i] = *reinterpret_cast<const __bf16*>(&y);
    }
    f32x16 acc = {};
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
    for (int r = 0; r < 16; ++r) c[acc_row(r, lane) * 32 + (lane & 31)] = acc[r];
}

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------

struct Layer {
    float *W = nullptr, *b = nullptr, *mW = nullptr, *vW = nullptr, *mb = nullptr, *vb = nullptr;
    float *gW = nullptr, *gb = nullptr;
    uint16_t *Wb = nullptr, *Wbt = nullptr;
    uint16_t* Wp = nullptr;   // BF16 in fragment order (sfm::frag_index) for the fused inference
};

struct sfm_net_t {
    uint32_t maxb = 0;
    Layer L[SFM_LAYERS];
    uint16_t* H[SFM_LAYERS] = {};    // H[l] = input of layer l (H[0] unused: the caller's x)
    uint16_t* Ht[SFM_LAYERS] = {};   // transposed inputs
    float* Z[SFM_LAYERS] = {};       // pre-activations of layers 0..3 outputs (Z[l] = output of layer l)
    uint16_t *dZ[2] = {}, *dZt[2] = {};
    float* part[SFM_LAYERS] = {};    // per layer: weight-gradient partials of the batch splits (<= max_splits)
    float* dbias[SFM_LAYERS] = {};   // per layer: bias-gradient partials of the same splits
    uint64_t step = 0;
    bool timing = false;
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool pending = false;
    double pending_flops = 0, ms_sum = 0, flops_sum = 0;
    uint32_t timed = 0;
};

static uint64_t splitmix(uint64_t& s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static void net_free(sfm_net_t* n)
{
    for (Layer& l : n->L) {
        for (float* p : {l.W, l.b, l.mW, l.vW, l.mb, l.vb, l.gW, l.gb}) hipFree(p);
        hipFree(l.Wb);
        hipFree(l.Wbt);
        hipFree(l.Wp);
    }
    for (int i = 0; i < SFM_LAYERS; ++i) { hipFree(n->H[i]); hipFree(n->Ht[i]); hipFree(n->Z[i]); }
    for (int i = 0; i < 2; ++i) { hipFree(n->dZ[i]); hipFree(n->dZt[i]); }
    for (int i = 0; i < SFM_LAYERS; ++i) { hipFree(n->part[i]); hipFree(n->dbias[i]); }
    for (int i = 0; i < 2; ++i)
        if (n->ev[i]) hipEventDestroy(n->ev[i]);
}

static int upload_layer(sfm_net_t* n, int l, const float* w, const float* b)
{
    const int rows = kDims[l + 1], cols = kDims[l];
    std::vector<uint16_t> wb((size_t)rows * cols), wbt((size_t)rows * cols), wp((size_t)rows * cols);
    for (int r = 0; r < rows; ++r)
        for (int c = 0; c < cols; ++c) {
            float f = w[(size_t)r * cols + c];
            uint32_t u;
            memcpy(&u, &f, 4);
            u += 0x7fffu + ((u >> 16) & 1u);
            uint16_t h = (uint16_t)(u >> 16);
            wb[(size_t)r * cols + c] = h;
            wbt[(size_t)c * rows + r] = h;
            wp[sfm::frag_index(r, c, cols)] = h;
        }
    Layer& L = n->L[l];
    HIPCHK(hipMemcpy(L.W, w, sizeof(float) * rows * cols, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(L.b, b, sizeof(float) * rows, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(L.Wb, wb.data(), 2 * wb.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(L.Wbt, wbt.data(), 2 * wbt.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(L.Wp, wp.data(), 2 * wp.size(), hipMemcpyHostToDevice));
    return SFM_OK;
}

extern "C" int sfm_create(uint32_t max_batch, uint64_t seed, sfm_net* out)
{
    if (!out) return fail(SFM_ERR_INVALID, "null argument");
    if (max_batch == 0 || max_batch % kSplitK || max_batch > (1u << 22))
        return fail(SFM_ERR_SIZE, "max_batch must be a positive multiple of %d", kSplitK);
    sfm_net_t* n = new sfm_net_t();
    n->maxb = max_batch;
    auto alloc = [&](void** p, size_t bytes) -> bool {
        if (hipMalloc(p, bytes) != hipSuccess) return false;
        return hipMemset(*p, 0, bytes) == hipSuccess;
    };
    bool ok = true;
    for (int l = 0; l < SFM_LAYERS && ok; ++l) {
        const size_t nw = (size_t)kDims[l + 1] * kDims[l], nb = kDims[l + 1];
        Layer& L = n->L[l];
        ok = alloc((void**)&L.W, 4 * nw) && alloc((void**)&L.b, 4 * nb) && alloc((void**)&L.mW, 4 * nw) &&
             alloc((void**)&L.vW, 4 * nw) && alloc((void**)&L.mb, 4 * nb) && alloc((void**)&L.vb, 4 * nb) &&
             alloc((void**)&L.gW, 4 * nw) && alloc((void**)&L.gb, 4 * nb) && alloc((void**)&L.Wb, 2 * nw) &&
             alloc((void**)&L.Wbt, 2 * nw) && alloc((void**)&L.Wp, 2 * nw);
        if (ok && l > 0) ok = alloc((void**)&n->H[l], 2 * (size_t)max_batch * kDims[l]) &&
                               alloc((void**)&n->Z[l], 4 * (size_t)max_batch * kDims[l]);
        if (ok) ok = alloc((void**)&n->Ht[l], 2 * (size_t)max_batch * kDims[l]);
    }
    for (int i = 0; i < 2 && ok; ++i)
        ok = alloc((void**)&n->dZ[i], 2 * (size_t)max_batch * SFM_HIDDEN) &&
             alloc((void**)&n->dZt[i], 2 * (size_t)max_batch * SFM_HIDDEN);
    for (int l = 0; l < SFM_LAYERS && ok; ++l)
        ok = alloc((void**)&n->part[l], 4 * (size_t)max_splits(l) * kDims[l + 1] * kDims[l]) &&
             alloc((void**)&n->dbias[l], 4 * (size_t)max_splits(l) * kDims[l + 1]);
    if (!ok) { net_free(n); delete n; return fail(SFM_ERR_HIP, "allocation failed"); }
    // Kaiming-uniform weights on the real fan-in, zero biases, padding zero
    uint64_t s = seed;
    for (int l = 0; l < SFM_LAYERS; ++l) {
        const int rows = kDims[l + 1], cols = kDims[l];
        std::vector<float> w((size_t)rows * cols, 0.f), b(rows, 0.f);
        const float bound = 1.0f / std::sqrt((float)kReal[l]);
        for (int r = 0; r < kReal[l + 1]; ++r)
            for (int c = 0; c < kReal[l]; ++c) {
                double u = (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
                w[(size_t)r * cols + c] = (float)((2.0 * u - 1.0) * bound);
            }
        int rc = upload_layer(n, l, w.data(), b.data());
        if (rc != SFM_OK) { net_free(n); delete n; return rc; }
    }
    *out = n;
    return SFM_OK;
}

extern "C" int sfm_destroy(sfm_net n)
{
    if (!n) return fail(SFM_ERR_INVALID, "null network");
    hipDeviceSynchronize();
    net_free(n);
    delete n;
    return SFM_OK;
}

extern "C" int sfm_param_count(sfm_net n, uint64_t* real_params)
{
    if (!n || !real_params) return fail(SFM_ERR_INVALID, "null argument");
    uint64_t c = 0;
    for (int l = 0; l < SFM_LAYERS; ++l) c += (uint64_t)kReal[l + 1] * kReal[l] + kReal[l + 1];
    *real_params = c;
    return SFM_OK;
}

extern "C" int sfm_layer_io(sfm_net n, int l, float* w, float* b, int set)
{
    if (!n || l < 0 || l >= SFM_LAYERS || !w || !b) return fail(SFM_ERR_INVALID, "bad argument");
    HIPCHK(hipDeviceSynchronize());
    if (set) return upload_layer(n, l, w, b);
    const int rows = kDims[l + 1], cols = kDims[l];
    HIPCHK(hipMemcpy(w, n->L[l].W, sizeof(float) * rows * cols, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(b, n->L[l].b, sizeof(float) * rows, hipMemcpyDeviceToHost));
    return SFM_OK;
}

// The products whose rows and columns are multiples of 64 and whose K (per
// batch split) is a multiple of kKC run workgroup-tiled through LDS, 64 x 64
// per workgroup (the 256-wide layers' forward, backward and weight-gradient
// products: 1,024 workgroups for 16,384 x 256, where 128 x 128 tiles leave
// 256 and the operand latency exposed; measured 97 vs 85 TFLOP/s training);
// the 32-wide ones (the first layer's K, the output layer) one 32 x 32 tile
// per wavefront (k_gemm).
template <int EPI>
static void launch_gemm(const GemmArgs& g, hipStream_t s)
{
    if ((g.K / g.ksplit) % kKC == 0 && g.rows % 64 == 0 && g.cols % 64 == 0) {
        hipLaunchKernelGGL((k_gemm_lds<EPI, 1, 1>), dim3((g.rows / 64) * (g.cols / 64), g.ksplit), dim3(256), 0, s, g);
        return;
    }
    const bool m2 = false, n2 = false;
    const int blocks = (g.rows / (m2 ? 64 : 32)) * (g.cols / (n2 ? 64 : 32));
    const dim3 grid((blocks + 3) / 4, g.ksplit);
    if (m2 && n2) hipLaunchKernelGGL((k_gemm<EPI, 2, 2>), grid, dim3(256), 0, s, g);
    else if (m2) hipLaunchKernelGGL((k_gemm<EPI, 2, 1>), grid, dim3(256), 0, s, g);
    else if (n2) hipLaunchKernelGGL((k_gemm<EPI, 1, 2>), grid, dim3(256), 0, s, g);
    else hipLaunchKernelGGL((k_gemm<EPI, 1, 1>), grid, dim3(256), 0, s, g);
}

static GemmArgs gemm(const uint16_t* A, int lda, const uint16_t* Bt, int ldb, int rows, int cols, int K)
{
    GemmArgs g;
    memset(&g, 0, sizeof(g));
    g.A = A; g.lda = lda; g.Bt = Bt; g.ldb = ldb;
    g.rows = rows; g.cols = cols; g.K = K; g.ksplit = 1;
    return g;
}

static double forward_flops(uint32_t batch)
{
    double f = 0;
    for (int l = 0; l < SFM_LAYERS; ++l) f += 2.0 * kReal[l] * kReal[l + 1];
    return f * batch;
}

static void time_begin(sfm_net_t* n, hipStream_t s)
{
    if (!n->timing) return;
    if (n->pending) {
        float ms = 0.f;
        if (hipEventSynchronize(n->ev[1]) == hipSuccess && hipEventElapsedTime(&ms, n->ev[0], n->ev[1]) == hipSuccess) {
            n->ms_sum += ms;
            n->flops_sum += n->pending_flops;
            n->timed++;
        }
        n->pending = false;
    }
    hipEventRecord(n->ev[0], s);
}

static void time_end(sfm_net_t* n, hipStream_t s, double flops)
{
    if (!n->timing) return;
    hipEventRecord(n->ev[1], s);
    n->pending = true;
    n->pending_flops = flops;
}

// Forward through the hidden layers: x -> H[1..4] (and Z, H^T when training).
static void forward_hidden(sfm_net_t* n, const uint16_t* x, uint32_t M, bool train, hipStream_t s)
{
    for (int l = 0; l < SFM_LAYERS - 1; ++l) {
        const uint16_t* A = (l == 0) ? x : n->H[l];
        GemmArgs g = gemm(A, kDims[l], n->L[l].Wb, kDims[l], (int)M, kDims[l + 1], kDims[l]);
        g.bias = n->L[l].b;
        g.H = n->H[l + 1];
        if (train) { g.Z = n->Z[l + 1]; g.Ht = n->Ht[l + 1]; }
        launch_gemm<EPI_HIDDEN>(g, s);
    }
}

// Inference through the fused kernel: 64-row tiles when every network's row
// count allows it, else 32-row tiles.
static int forward_group(const sfm_net* nets, uint32_t nnets, const uint32_t* counts, const uint16_t* x, float* y,
                         hipStream_t s)
{
    if (!nets || !counts || !x || !y || nnets == 0) return fail(SFM_ERR_INVALID, "null argument");
    if (nnets > SFM_MAX_GROUP) return fail(SFM_ERR_SIZE, "at most %d networks per call", SFM_MAX_GROUP);
    if ((uintptr_t)x % 16) return fail(SFM_ERR_INVALID, "x must be 16-byte aligned");
    NetSet ns;
    memset(&ns, 0, sizeof(ns));
    ns.nnets = (int)nnets;
    uint64_t total = 0;
    bool even64 = true;
    for (uint32_t i = 0; i < nnets; ++i) {
        if (!nets[i]) return fail(SFM_ERR_INVALID, "null network");
        if (counts[i] % 32) return fail(SFM_ERR_SIZE, "row counts must be multiples of 32");
        even64 = even64 && counts[i] % 64 == 0;
        total += counts[i];
        if (total > (1ull << 31)) return fail(SFM_ERR_SIZE, "too many rows");
        ns.row_end[i] = (uint32_t)total;
        for (int l = 0; l < SFM_LAYERS; ++l) {
            ns.W[i][l] = nets[i]->L[l].Wp;
            ns.b[i][l] = nets[i]->L[l].b;
        }
    }
    if (total == 0) return SFM_OK;
    sfm_net_t* t = nets[0];
    time_begin(t, s);
    if (even64)
        hipLaunchKernelGGL(k_mlp_fused<64>, dim3((unsigned)(total / 64)), dim3(256), 0, s, ns, x, (uint32_t)total, y);
    else
        hipLaunchKernelGGL(k_mlp_fused<32>, dim3((unsigned)(total / 32)), dim3(256), 0, s, ns, x, (uint32_t)total, y);
    time_end(t, s, forward_flops((uint32_t)total));
    HIPCHK(hipGetLastError());
    return SFM_OK;
}

extern "C" int sfm_forward(sfm_net n, const uint16_t* x, uint32_t batch, float* y, void* stream)
{
    if (!n || !x || !y) return fail(SFM_ERR_INVALID, "null argument");
    if (batch == 0 || batch % 32) return fail(SFM_ERR_SIZE, "batch must be a positive multiple of 32");
    return forward_group(&n, 1, &batch, x, y, (hipStream_t)stream);
}

extern "C" int sfm_forward_group(const sfm_net* nets, uint32_t nnets, const uint32_t* counts, const uint16_t* x,
                                 float* y, void* stream)
{
    return forward_group(nets, nnets, counts, x, y, (hipStream_t)stream);
}

// Inference through the per-layer kernels of the training path (kept for
// checking the fused kernel against them).
extern "C" int sfm_forward_layers(sfm_net n, const uint16_t* x, uint32_t batch, float* y, void* stream)
{
    if (!n || !x || !y) return fail(SFM_ERR_INVALID, "null argument");
    if (batch == 0 || batch % 32 || batch > n->maxb) return fail(SFM_ERR_SIZE, "batch must be a multiple of 32 <= %u", n->maxb);
    if ((uintptr_t)x % 16) return fail(SFM_ERR_INVALID, "x must be 16-byte aligned");
    hipStream_t s = (hipStream_t)stream;
    time_begin(n, s);
    forward_hidden(n, x, batch, false, s);
    GemmArgs g = gemm(n->H[SFM_LAYERS - 1], SFM_HIDDEN, n->L[SFM_LAYERS - 1].Wb, SFM_HIDDEN, (int)batch, SFM_OUT_PAD,
                      SFM_HIDDEN);
    g.bias = n->L[SFM_LAYERS - 1].b;
    g.y = y;
    launch_gemm<EPI_OUT>(g, s);
    time_end(n, s, forward_flops(batch));
    HIPCHK(hipGetLastError());
    return SFM_OK;
}

extern "C" int sfm_train_step(sfm_net n, const uint16_t* x, const float* target, uint32_t batch, float lr, float beta1,
                              float beta2, float eps, float* loss, void* stream)
{
    if (!n || !x || !target || !loss) return fail(SFM_ERR_INVALID, "null argument");
    if (batch == 0 || batch % kSplitK || batch > n->maxb)
        return fail(SFM_ERR_SIZE, "batch must be a multiple of %d <= %u", kSplitK, n->maxb);
    if ((uintptr_t)x % 16) return fail(SFM_ERR_INVALID, "x must be 16-byte aligned");
    hipStream_t s = (hipStream_t)stream;
    const int M = (int)batch;
    time_begin(n, s);
    HIPCHK(hipMemsetAsync(loss, 0, sizeof(float), s));
    // x^T for the first layer's weight gradient
    hipLaunchKernelGGL(k_transpose, dim3((M * SFM_IN_PAD + 255) / 256), dim3(256), 0, s, x, M, SFM_IN_PAD, n->Ht[0]);
    forward_hidden(n, x, batch, true, s);
    // output layer: loss and dLoss/dZ (row-major and transposed)
    {
        GemmArgs g = gemm(n->H[SFM_LAYERS - 1], SFM_HIDDEN, n->L[SFM_LAYERS - 1].Wb, SFM_HIDDEN, M, SFM_OUT_PAD,
                          SFM_HIDDEN);
        g.bias = n->L[SFM_LAYERS - 1].b;
        g.target = target;
        g.loss = loss;
        g.scale = 1.0f / ((float)M * SFM_OUT);
        g.H = n->dZ[0];
        g.Ht = n->dZt[0];
        launch_gemm<EPI_OUT>(g, s);
    }
    n->step++;
    const float bc1 = 1.0f - std::pow(beta1, (float)n->step), bc2 = 1.0f - std::pow(beta2, (float)n->step);
    int cur = 0;
    AdamSet adam;
    adam.nseg = 0;
    int nblk = 0;
    for (int l = SFM_LAYERS - 1; l >= 0; --l) {
        const int dout = kDims[l + 1], din = kDims[l];
        Layer& L = n->L[l];
        // dW = dZ^T H, split over the batch
        {
            // batch splits of the weight-gradient product: at most
            // max_splits(l), each a multiple of 256 rows; their partial sums
            // (weights and biases) are summed in a fixed order by k_adam
            int splits = 1;
            for (int d = max_splits(l); d >= 1; --d)
                if ((M / kSplitK) % d == 0) { splits = d; break; }
            GemmArgs g = gemm(n->dZt[cur], M, n->Ht[l], M, dout, din, M);
            g.ksplit = splits;
            g.part = n->part[l];
            g.bpart = n->dbias[l];
            launch_gemm<EPI_WGRAD>(g, s);
            // dH = dZ W  (with the pre-update weights), then dZ of the layer below
            if (l > 0) {
                GemmArgs h = gemm(n->dZ[cur], dout, L.Wbt, dout, M, din, dout);
                h.Z = n->Z[l];
                h.H = n->dZ[cur ^ 1];
                h.Ht = n->dZt[cur ^ 1];
                launch_gemm<EPI_BWD>(h, s);
            }
            // every Adam update waits for the end of the step: the products
            // above read the pre-update weights, and each layer keeps its own
            // gradient partials, so one launch serves all tensors
            AdamSeg& sw = adam.seg[adam.nseg++];
            sw = AdamSeg{L.W, L.mW, L.vW, n->part[l], L.gW, L.Wb, L.Wbt, L.Wp, splits, dout, din, nblk};
            nblk += (dout * din + 255) / 256;
            AdamSeg& sb = adam.seg[adam.nseg++];
            sb = AdamSeg{L.b, L.mb, L.vb, n->dbias[l], L.gb, nullptr, nullptr, nullptr, splits, dout, 1, nblk};
            nblk += (dout + 255) / 256;
        }
        cur ^= 1;
    }
    hipLaunchKernelGGL(k_adam, dim3(nblk), dim3(256), 0, s, adam, lr, beta1, beta2, eps, bc1, bc2);
    time_end(n, s, 3.0 * forward_flops(batch));
    HIPCHK(hipGetLastError());
    return SFM_OK;
}

extern "C" int sfm_gradients(sfm_net n, int l, float* dw, float* db, void* stream)
{
    if (!n || l < 0 || l >= SFM_LAYERS || !dw || !db) return fail(SFM_ERR_INVALID, "bad argument");
    hipStream_t s = (hipStream_t)stream;
    HIPCHK(hipMemcpyAsync(dw, n->L[l].gW, sizeof(float) * kDims[l + 1] * kDims[l], hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(db, n->L[l].gb, sizeof(float) * kDims[l + 1], hipMemcpyDeviceToDevice, s));
    return SFM_OK;
}

extern "C" int sfm_to_bf16(const float* x, uint16_t* y, uint64_t n, void* stream)
{
    if ((!x || !y) && n) return fail(SFM_ERR_INVALID, "null argument");
    if (n == 0) return SFM_OK;
    hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, x, y, n);
    HIPCHK(hipGetLastError());
    return SFM_OK;
}

extern "C" int sfm_probe_mfma(const uint16_t* a, const uint16_t* b, float* c, void* stream)
{
    if (!a || !b || !c) return fail(SFM_ERR_INVALID, "null argument");
    hipLaunchKernelGGL(k_probe, dim3(1), dim3(64), 0, (hipStream_t)stream, a, b, c);
    HIPCHK(hipGetLastError());
    return SFM_OK;
}

extern "C" int sfm_weights(sfm_net n, const uint16_t** w, const float** b)
{
    if (!n || !w || !b) return fail(SFM_ERR_INVALID, "null argument");
    for (int l = 0; l < SFM_LAYERS; ++l) {
        w[l] = n->L[l].Wp;
        b[l] = n->L[l].b;
    }
    return SFM_OK;
}

extern "C" int sfm_timing(sfm_net n, int enable)
{
    if (!n) return fail(SFM_ERR_INVALID, "null network");
    if (enable && !n->ev[0])
        for (int i = 0; i < 2; ++i) HIPCHK(hipEventCreate(&n->ev[i]));
    n->timing = enable != 0;
    n->pending = false;
    n->ms_sum = n->flops_sum = 0;
    n->timed = 0;
    return SFM_OK;
}

extern "C" int sfm_kernel_ms(sfm_net n, double* ms, double* flops, uint32_t* count)
{
    if (!n) return fail(SFM_ERR_INVALID, "null network");
    if (n->pending) {
        float t = 0.f;
        if (hipEventSynchronize(n->ev[1]) == hipSuccess && hipEventElapsedTime(&t, n->ev[0], n->ev[1]) == hipSuccess) {
            n->ms_sum += t;
            n->flops_sum += n->pending_flops;
            n->timed++;
        }
        n->pending = false;
    }
    if (count) *count = n->timed;
    if (ms) *ms = n->timed ? n->ms_sum / n->timed : 0.0;
    if (flops) *flops = n->timed ? n->flops_sum / n->timed : 0.0;
    return SFM_OK;
}
