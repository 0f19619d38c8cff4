// GELU of the surrogate network, shared by the gfx950 kernels and the host
// harness that checks it.
//
// GELU(z) = z Phi(z) = 0.5 z (1 + erf(z / sqrt 2)), with erf from Abramowitz
// and Stegun, Handbook of Mathematical Functions (1964), formula 7.1.26:
//   erf(x) = 1 - (a1 t + a2 t^2 + a3 t^3 + a4 t^4 + a5 t^5) exp(-x^2),
//   t = 1 / (1 + p x),  x >= 0,  |error| <= 1.5e-7,
// odd in x.  The network stores activations in BF16 (8 significant bits), so
// this error is far below what its outputs resolve; it costs one reciprocal,
// one exponential and a handful of FMAs instead of a correctly rounded erf.
// The derivative GELU'(z) = Phi(z) + z phi(z), phi(z) = exp(-z^2/2) / sqrt(2 pi),
// reuses the same exponential.
#pragma once

#ifdef __HIPCC__
#define SFM_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define SFM_HD inline
#endif

namespace sfm {

#ifdef __HIP_DEVICE_COMPILE__
SFM_HD float g_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
SFM_HD float g_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
#else
SFM_HD float g_rcp(float x) { return 1.0f / x; }
SFM_HD float g_exp2(float x) { return exp2f(x); }
#endif

// Phi(z) and exp(-z^2 / 2).
SFM_HD void phi_parts(float z, float& Phi, float& e)
{
    const float x = fabsf(z) * 0.70710678118654752f;
    const float t = g_rcp(fmaf(0.3275911f, x, 1.0f));
    float p = fmaf(1.061405429f, t, -1.453152027f);
    p = fmaf(p, t, 1.421413741f);
    p = fmaf(p, t, -0.284496736f);
    p = fmaf(p, t, 0.254829592f);
    p *= t;
    e = g_exp2(-(x * x) * 1.4426950408889634f);   // exp(-x^2)
    const float erf_abs = fmaf(-p, e, 1.0f);
    const float erf = copysignf(erf_abs, z);
    Phi = fmaf(0.5f, erf, 0.5f);
}

SFM_HD float gelu(float z)
{
    float Phi, e;
    phi_parts(z, Phi, e);
    return z * Phi;
}

SFM_HD float gelu_grad(float z)
{
    float Phi, e;
    phi_parts(z, Phi, e);
    return fmaf(z * 0.3989422804014327f, e, Phi);
}

}  // namespace sfm
