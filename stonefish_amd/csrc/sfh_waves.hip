// Headless wave producer: the reference's FFT ocean without OpenGL.
//
//   sfh_wave_spectrum    OpenGLOcean::GenerateWavesSpectrum (OpenGLOcean.cpp:1025-1046),
//                        GetSpectrumSample (1003-1022), spectrum (940-1001), omega (934-937),
//                        frandom / lrandom (SystemUtil.hpp:207-217); the sea-state
//                        parameters of OpenGLRealOcean (OpenGLRealOcean.cpp:43-45) and
//                        OpenGLOcean (OpenGLOcean.cpp:63-71).  Host code.
//   sfh_wave_fft_*       OpenGLOcean::Simulate (OpenGLOcean.cpp:455-530): oceanInit.frag
//                        (h(k, t) of the four grids, packed two per complex value), then
//                        the butterfly passes of oceanFFTX.frag / oceanFFTY.frag over the
//                        table of ComputeButterflyLookupTable (879-925), whose result's
//                        height layer is what OpenGLRealOcean::Simulate copies into fftData
//                        (OpenGLRealOcean.cpp:333-356) for Ocean::GetDepth.
//
// The reference stores the spectra and the butterfly table in RGBA16F
// textures: the spectra and twiddle factors are rounded to binary16 here the
// same way, the butterflies run in FP32 like its RGBA32F passes, and the
// phase w t is formed in FP64 (the reference's shader forms it in FP32).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/sfhydro.h"

int sfh_internal_fail(int code, const char* msg);   // sfh_hydro.hip: sets sfh_last_error

namespace {

constexpr uint32_t kMaxPasses = 11;   // 2048 x 2048 texels: two lines of float4 in LDS = 64 KB

float to_half(float x) { return (float)(_Float16)x; }

// ---- spectrum (host) --------------------------------------------------------

struct SeaState {
    float wind, omega, A, km, cm;
    bool propagate;
};

long lcg_next(long* seed)
{
    *seed = (*seed * 1103515245 + 12345) & 0x7FFFFFFF;
    return *seed;
}

float lcg_uniform(long* seed) { return (float)(lcg_next(seed) >> (31 - 24)) / (float)(1 << 24); }

float sq(float x) { return x * x; }

float dispersion(const SeaState& s, float k) { return std::sqrt(9.81 * k * (1.0 + sq(k / s.km))); }

// Directional Elfouhaily spectrum (Eq 67 of Elfouhaily et al. 1997, with the
// propagate option keeping the waves that run along +x at twice the energy).
float elfouhaily(const SeaState& s, float kx, float ky)
{
    const float U10 = s.wind, Om = s.omega;
    const float k = std::sqrt(kx * kx + ky * ky);
    const float c = dispersion(s, k) / k;
    const float kp = 9.81 * sq(Om / U10);
    const float cp = dispersion(s, kp) / kp;
    const float z0 = 3.7e-5 * sq(U10) / 9.81 * std::pow(U10 / cp, 0.9f);
    const float u_star = 0.41 * U10 / std::log(10.0 / z0);
    const float Lpm = std::exp(-5.0 / 4.0 * sq(kp / k));
    const float gamma = Om < 1.0 ? 1.7 : 1.7 + 6.0 * std::log(Om);
    const float sigma = 0.08 * (1.0 + 4.0 / std::pow(Om, 3.0f));
    const float Gamma = std::exp(-1.0 / (2.0 * sq(sigma)) * sq(std::sqrt(k / kp) - 1.0));
    const float Jp = std::pow(gamma, Gamma);
    const float Fp = Lpm * Jp * std::exp(-Om / std::sqrt(10.0) * (std::sqrt(k / kp) - 1.0));
    const float alphap = 0.006 * std::sqrt(Om);
    float Bl = 0.5 * alphap * cp / c * Fp;
    const float alpham = 0.01 * (u_star < s.cm ? 1.0 + std::log(u_star / s.cm) : 1.0 + 3.0 * std::log(u_star / s.cm));
    const float Fm = std::exp(-0.25 * sq(k / s.km - 1.0));
    float Bh = 0.5 * alpham * s.cm / c * Fm;
    Bh *= Lpm;
    const float a0 = std::log(2.0) / 4.0;
    const float am = 0.13 * u_star / s.cm;
    const float Delta = std::tanh(a0 + 4.0 * std::pow(c / cp, 2.5f) + am * std::pow(s.cm / c, 2.5f));
    const float phi = std::atan2(ky, kx);
    if (s.propagate) {
        if (kx < 0.0) return 0.0f;
        Bl *= 2.0;
        Bh *= 2.0;
    }
    return s.A * (Bl + Bh) * (1.0 + Delta * std::cos(2.0 * phi)) / (2.0 * M_PI * sq(sq(k)));
}

void spectrum_sample(const SeaState& s, int i, int j, float L, float kmin, long* seed, float* out)
{
    const float dk = 2.0 * M_PI / L;
    const float kx = i * dk, ky = j * dk;
    if (std::fabs(kx) < kmin && std::fabs(ky) < kmin) {
        out[0] = out[1] = 0.0f;
        return;
    }
    const float S = elfouhaily(s, kx, ky);
    const float h = std::sqrt(S / 2.0) * dk;
    const float phi = lcg_uniform(seed) * 2.0 * M_PI;
    out[0] = h * std::cos(phi);
    out[1] = h * std::sin(phi);
}

// ---- device passes ------------------------------------------------------------

struct Grids {
    double inv[4];   // 2 pi n / L_g as the float uniform of the init pass
};

// h(k, t) of the four grids at every texel, packed (h1 + i h2, h3 + i h4).
__global__ __launch_bounds__(256) void k_wave_init(const float4* __restrict__ s12, const float4* __restrict__ s34,
                                                   uint32_t n, uint32_t lg, Grids G, double t,
                                                   float4* __restrict__ out)
{
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * n) return;
    const uint32_t x = idx & (n - 1), y = idx >> lg;
    const uint32_t xc = (n - x) & (n - 1), yc = (n - y) & (n - 1);   // -k, the texture repeats
    const double fx = (double)((x >= n / 2) ? (int)x - (int)n : (int)x) / (double)n;
    const double fy = (double)((y >= n / 2) ? (int)y - (int)n : (int)y) / (double)n;
    const float r2 = 1.414213562f;
    const float4 a = s12[idx], ac = s12[(size_t)yc * n + xc];
    const float4 b = s34[idx], bc = s34[(size_t)yc * n + xc];
    const float s0[8] = {a.x * r2, a.y * r2, a.z * r2, a.w * r2, b.x * r2, b.y * r2, b.z * r2, b.w * r2};
    const float sc[8] = {ac.x * r2, ac.y * r2, ac.z * r2, ac.w * r2, bc.x * r2, bc.y * r2, bc.z * r2, bc.w * r2};
    double h[8];
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        const double kx = fx * G.inv[g], ky = fy * G.inv[g];
        const double K = sqrt(kx * kx + ky * ky);
        const double w = sqrt(9.81 * K * (1.0 + K * K / (370.0 * 370.0)));
        double s, c;
        sincos(w * t, &s, &c);
        const double px = s0[2 * g], py = s0[2 * g + 1], qx = sc[2 * g], qy = sc[2 * g + 1];
        h[2 * g] = (px + qx) * c - (py + qy) * s;
        h[2 * g + 1] = (px - qx) * s + (py - qy) * c;
    }
    out[idx] = make_float4((float)(h[0] - h[3]), (float)(h[1] + h[2]), (float)(h[4] - h[7]), (float)(h[5] + h[6]));
}

// The butterfly passes along one direction, one workgroup per line, the line
// held in LDS (ping-pong between two copies): out[i] = in[j1] + w * in[j2]
// for both packed complex values, with (j1, j2, w) the table's entry for
// output i of the pass.  dir 0: lines are rows (x varies), dir 1: columns.
__global__ __launch_bounds__(256) void k_wave_fft(float4* __restrict__ data, const int2* __restrict__ idx,
                                                  const float2* __restrict__ tw, uint32_t n, uint32_t passes,
                                                  uint32_t dir)
{
    extern __shared__ float4 lds[];
    float4* src = lds;
    float4* dst = lds + n;
    const uint32_t line = blockIdx.x;
    const size_t step = dir ? (size_t)n : 1;
    float4* base = data + (dir ? (size_t)line : (size_t)line * n);
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) src[i] = base[i * step];
    __syncthreads();
    for (uint32_t p = 0; p < passes; ++p) {
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
            const int2 j = idx[(size_t)p * n + i];
            const float2 w = tw[(size_t)p * n + i];
            const float4 u = src[j.x], v = src[j.y];
            float4 r;
            r.x = u.x + (w.x * v.x - w.y * v.y);
            r.y = u.y + (w.y * v.x + w.x * v.y);
            r.z = u.z + (w.x * v.z - w.y * v.w);
            r.w = u.w + (w.y * v.z + w.x * v.w);
            dst[i] = r;
        }
        __syncthreads();
        float4* t = src; src = dst; dst = t;
    }
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) base[i * step] = src[i];
}

uint32_t bit_reverse(uint32_t i, uint32_t n)
{
    uint32_t r = 0;
    for (uint32_t m = n >> 1, w = 1; m; m >>= 1, w <<= 1)
        if (i & m) r += w;
    return r;
}

}  // namespace

struct sfh_wave_fft_t {
    uint32_t n = 0, passes = 0;
    Grids grids;
    float4* d_s12 = nullptr;
    float4* d_s34 = nullptr;
    int2* d_idx = nullptr;
    float2* d_tw = nullptr;
};

static void wave_fft_free(sfh_wave_fft_t* w)
{
    hipFree(w->d_s12);
    hipFree(w->d_s34);
    hipFree(w->d_idx);
    hipFree(w->d_tw);
    delete w;
}

extern "C" int sfh_wave_spectrum(float state, uint32_t passes, const float grid_sizes[4], float* spectrum12,
                                 float* spectrum34)
{
    if (!spectrum12 || !spectrum34) return sfh_internal_fail(SFH_ERR_INVALID, "null argument");
    if (passes < 1 || passes > kMaxPasses) return sfh_internal_fail(SFH_ERR_SIZE, "passes must be 1 to 11");
    const float L[4] = {grid_sizes ? grid_sizes[0] : 893.f, grid_sizes ? grid_sizes[1] : 101.f,
                        grid_sizes ? grid_sizes[2] : 21.f, grid_sizes ? grid_sizes[3] : 11.f};
    for (int g = 0; g < 4; ++g)
        if (!(L[g] > 0.f)) return sfh_internal_fail(SFH_ERR_INVALID, "grid sizes must be positive");
    SeaState s;
    s.wind = state * 5.f + 2.f;
    s.A = 1.f;
    s.omega = 5.f * std::exp(-state) + 0.2f;
    s.km = 370.f;
    s.cm = 0.23f;
    s.propagate = true;
    const int n = 1 << passes;
    long seed = 1234;
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) {
            const size_t o = 4 * ((size_t)x + (size_t)y * n);
            const int i = x >= n / 2 ? x - n : x;
            const int j = y >= n / 2 ? y - n : y;
            spectrum_sample(s, i, j, L[0], M_PI / L[0], &seed, spectrum12 + o);
            spectrum_sample(s, i, j, L[1], M_PI * n / L[0], &seed, spectrum12 + o + 2);
            spectrum_sample(s, i, j, L[2], M_PI * n / L[1], &seed, spectrum34 + o);
            spectrum_sample(s, i, j, L[3], M_PI * n / L[2], &seed, spectrum34 + o + 2);
        }
    return SFH_OK;
}

extern "C" int sfh_wave_fft_create(uint32_t passes, const float grid_sizes[4], const float* spectrum12,
                                   const float* spectrum34, sfh_wave_fft* out)
{
    if (!out || !spectrum12 || !spectrum34 || !grid_sizes) return sfh_internal_fail(SFH_ERR_INVALID, "null argument");
    if (passes < 1 || passes > kMaxPasses) return sfh_internal_fail(SFH_ERR_SIZE, "passes must be 1 to 11");
    for (int g = 0; g < 4; ++g)
        if (!(grid_sizes[g] > 0.f)) return sfh_internal_fail(SFH_ERR_INVALID, "grid sizes must be positive");
    const uint32_t n = 1u << passes;
    const size_t texels = (size_t)n * n;
    sfh_wave_fft_t* w = new sfh_wave_fft_t();
    w->n = n;
    w->passes = passes;
    for (int g = 0; g < 4; ++g) w->grids.inv[g] = (double)(float)(2.0 * M_PI * (float)n / grid_sizes[g]);
    // spectra as RGBA16F texels
    std::vector<float> h12(4 * texels), h34(4 * texels);
    for (size_t i = 0; i < 4 * texels; ++i) {
        h12[i] = to_half(spectrum12[i]);
        h34[i] = to_half(spectrum34[i]);
    }
    // butterfly table (j1, j2, w) per output of each pass, twiddles as binary16
    std::vector<int2> idx((size_t)passes * n);
    std::vector<float2> tw((size_t)passes * n);
    for (uint32_t p = 0; p < passes; ++p) {
        const uint32_t nblocks = 1u << (passes - 1 - p), nh = 1u << p;
        for (uint32_t j = 0; j < nblocks; ++j)
            for (uint32_t k = 0; k < nh; ++k) {
                const uint32_t i1 = j * nh * 2 + k, i2 = i1 + nh;
                const uint32_t j1 = p == 0 ? bit_reverse(i1, n) : i1;
                const uint32_t j2 = p == 0 ? bit_reverse(i2, n) : i2;
                const double arg = 2.0 * M_PI * (double)(k * nblocks) / (double)(float)n;
                const float wr = to_half((float)std::cos(arg)), wi = to_half((float)std::sin(arg));
                idx[(size_t)p * n + i1] = make_int2((int)j1, (int)j2);
                idx[(size_t)p * n + i2] = make_int2((int)j1, (int)j2);
                tw[(size_t)p * n + i1] = make_float2(wr, wi);
                tw[(size_t)p * n + i2] = make_float2(-wr, -wi);
            }
    }
    hipError_t e = hipMalloc(&w->d_s12, sizeof(float4) * texels);
    if (e == hipSuccess) e = hipMalloc(&w->d_s34, sizeof(float4) * texels);
    if (e == hipSuccess) e = hipMalloc(&w->d_idx, sizeof(int2) * idx.size());
    if (e == hipSuccess) e = hipMalloc(&w->d_tw, sizeof(float2) * tw.size());
    if (e == hipSuccess) e = hipMemcpy(w->d_s12, h12.data(), sizeof(float) * h12.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(w->d_s34, h34.data(), sizeof(float) * h34.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(w->d_idx, idx.data(), sizeof(int2) * idx.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(w->d_tw, tw.data(), sizeof(float2) * tw.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        wave_fft_free(w);
        return sfh_internal_fail(SFH_ERR_HIP, "wave producer upload failed");
    }
    *out = w;
    return SFH_OK;
}

extern "C" int sfh_wave_fft_step(sfh_wave_fft w, float t, float* d_heights, void* stream)
{
    if (!w || !d_heights) return sfh_internal_fail(SFH_ERR_INVALID, "null argument");
    hipStream_t s = (hipStream_t)stream;
    const uint32_t n = w->n;
    float4* out = reinterpret_cast<float4*>(d_heights);
    hipLaunchKernelGGL(k_wave_init, dim3((n * n + 255) / 256), dim3(256), 0, s, w->d_s12, w->d_s34, n, w->passes,
                       w->grids, (double)t, out);
    const size_t lds = 2 * sizeof(float4) * n;
    hipLaunchKernelGGL(k_wave_fft, dim3(n), dim3(256), lds, s, out, w->d_idx, w->d_tw, n, w->passes, 0u);
    hipLaunchKernelGGL(k_wave_fft, dim3(n), dim3(256), lds, s, out, w->d_idx, w->d_tw, n, w->passes, 1u);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return sfh_internal_fail(SFH_ERR_HIP, hipGetErrorString(e));
    return SFH_OK;
}

extern "C" int sfh_wave_fft_info(sfh_wave_fft w, uint32_t* n)
{
    if (!w || !n) return sfh_internal_fail(SFH_ERR_INVALID, "null argument");
    *n = w->n;
    return SFH_OK;
}

extern "C" int sfh_wave_fft_destroy(sfh_wave_fft w)
{
    if (!w) return sfh_internal_fail(SFH_ERR_INVALID, "null argument");
    hipDeviceSynchronize();
    wave_fft_free(w);
    return SFH_OK;
}
