// sfhydro: gfx950 kernels and C-ABI host code for Stonefish's geometry-based
// hydrodynamics (SolidEntity.cpp:1239-1850, Ocean.cpp:149-201,
// OpenGLRealOcean.cpp:280-331, Compound-free body path).
//
// One evaluation pass over a batch of bodies is three launches on one stream
// (four for compounds):
//   k_prologue  eight lanes per body, FP64: C transform, body/fluid
//               classification (8 AABB corners against the surface), per-body
//               FP32 constants of the face loop in the mesh frame, FP64
//               constants of the epilogue; the rows of bodies without face
//               work (outside, not hydrodynamic, no faces);
//   k_faces     one wavefront per chunk of up to 64*K faces of one body, the
//               submerged or the crossing loop as the body's position says:
//               FP32 face arithmetic with FP32 per-lane partials, FP64
//               wavefront reduction into one 24-double partial per chunk (a
//               crossing chunk first evaluates the FP64 depths of its vertex
//               window into LDS);
//   k_epilogue  one wavefront per body with face work, FP64: sums the body's
//               chunk partials (fixed order), rotates to world, moment about
//               the CoG, buoyancy, CorrectHydrodynamicForces, output write mask;
//   k_compound  compound batches: one lane per compound composes its parts.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/sfhydro.h"
#include "sfh_face.h"
#include "sfm_device.h"

using namespace sfh;

// ---------------------------------------------------------------------------
// error reporting
// ---------------------------------------------------------------------------

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
            return fail(SFH_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));    \
    } while (0)

extern "C" const char* sfh_last_error(void) { return g_err; }

// The error path of the host-only translation units of the library (sfh_hull.cpp).
int sfh_internal_fail(int code, const char* msg) { return fail(code, "%s", msg); }
int sfh_internal_na_init(void);

// ---------------------------------------------------------------------------
// device-side data
// ---------------------------------------------------------------------------

enum : uint32_t {
    BF_HYDRO = 1u,       // physics mode FLOATING or SUBMERGED
    BF_BUOYANT = 2u,     // isBuoyant()
    BF_DAMPING = 4u,     // dampingForces (per-face calls); body level: always on
    BF_MODE_SURF = 8u,   // static Surface call: no classification, no Correct
    BF_MODE_SUBM = 16u,  // static Submerged call: no classification, no Correct
    BF_EXPLICIT_TC = 32u,// state carries T_C explicitly (per-face calls)
    BF_PART = 64u,       // part of a compound (Compound::ComputeHydrodynamicForces)
    BF_EXTERNAL = 128u,  // part: CompoundPart::isExternal
    BF_PHYDRO = 256u,    // part: the part solid's mode is FLOATING or SUBMERGED
    BF_MODE_AERO = 512u, // aerodynamic loop: every face, normal-velocity drag
    BF_AERO_CORRECT = 1024u,  // body level: CorrectAerodynamicForces with the proxy in bits 11-12
    BF_APPROX_SHIFT = 11u
};

static constexpr int kMaxLayers = 4;
static constexpr int kChunkDoubles = 24;
static constexpr int kMaxCur = 4;   // jets and pipes evaluated per face

// The jets and pipes of the ocean in one body's mesh frame: anchor (jet
// outlet c, pipe inlet p1) and unit axis, with the scalar parameters: jet r,
// vout; pipe l, r1, r2, vin, gamma.
struct CurBody {
    int32_t n, pad_;
    int32_t type[kMaxCur];     // 1 jet, 2 pipe
    double c[kMaxCur][3], a[kMaxCur][3];
    double p[kMaxCur][5];
};

struct BodyStatic {
    const float* verts;
    const uint32_t* faces;
    uint32_t nf;
    uint32_t flags;
    uint32_t chunk0;
    uint32_t nchunks;
    uint32_t nv;
    uint32_t sidx;      // row of the state array (and active mask) the body reads: its compound for a part
    double cg2c[12];   // basis row-major, origin
    double cg2o[9];    // basis row-major
    double aabb_c[3];  // CG-frame AABB centre and half extents
    double aabb_h[3];
    double Cd[3], Cf[3];
    double volume, surface;
    double P_CB[3];
};

// Depth queries run in FP64: near the waterline the depth is a small
// difference of body-scale quantities (z of the vertex, height of the wave),
// and the clip points of thin wet slivers depend on it.
struct LayerK {
    double fu, fv;      // fractional texel coordinate of the mesh origin
    double ru[3], rv[3];// rows 0/1 of R scaled by n / length: texels per mesh unit
    int ku, kv;         // integer texel coordinate of the mesh origin, mod n
};

struct BodyDyn {
    int32_t pos;        // 0 outside, 1 inside (submerged loop), 2 crossing (surface loop)
    uint32_t flags;     // effective settings of the face loop
    int32_t cpos;       // classification before the part rules (the compound's position for a part)
    int32_t pad_;
    FaceBody fb;
    double rz[3];       // row 2 of R: world z of a mesh-frame vector
    double tz;          // world z of the mesh origin
    LayerK layer[kMaxLayers];
    CurBody cur;        // varying currents (cur.n == 0: none)
    // FP64 epilogue constants
    double R[9];        // basis of T_C, row-major
    double dt[3];       // R^T (t_C - t_CG)
    double Rcg[9];      // basis of T_CG, row-major
    double tcg[3];      // origin of T_CG
};

struct OceanK {
    double rho;
    double g[3];
    double len[kMaxLayers];
    double scale[kMaxLayers];         // n / len, texels per metre
    double u[3];                      // sum of enabled uniform currents
    const float4* quad[kMaxLayers];   // per layer: (h00, h10, h01, h11) per texel
    int32_t nlayers;                  // 0 = flat surface
    int32_t n;
    int32_t nmask;                    // n - 1 when n is a power of two, else -1
    int32_t ncur;                     // enabled jets and pipes
    int32_t cur_type[kMaxCur];
    double cur_c[kMaxCur][3], cur_a[kMaxCur][3];   // world anchor and unit axis
    double cur_p[kMaxCur][5];
};

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------

// Global-memory views of pointers the kernels read out of structures in
// memory.  Through a generic pointer the compiler issues flat loads, which
// count against the scalar-load counter as well as the vector-load counter,
// so waiting for a scalar load also waits for every flat load in flight —
// including the next face's vertices the face passes fetch ahead of use.
// Global loads are counted by the vector counter alone.
#define SFH_GLOBAL __attribute__((address_space(1)))
typedef float f32x4 __attribute__((ext_vector_type(4)));
template <typename T>
__device__ __forceinline__ const SFH_GLOBAL T* gptr(const T* p)
{
    return (const SFH_GLOBAL T*)p;
}
template <typename T>
__device__ __forceinline__ SFH_GLOBAL T* gptr_w(T* p)
{
    return (SFH_GLOBAL T*)p;
}

__device__ __forceinline__ int wrap_n(int i, int n)
{
    // i = k + floor(u) with k in [0, n) and |floor(u)| below n for any body
    // smaller than the wave tile: one conditional step wraps it.  The clamp
    // keeps the texel read in bounds for anything else (non-finite states).
    if (i < 0) i += n;
    if (i >= n) i -= n;
    return min(max(i, 0), n - 1);
}

// Bilinear mix of a texel's four corner heights at fractions (a, b).
__device__ __forceinline__ double texel_mix(double a, double b, float4 t)
{
    return (1.0 - a) * (1.0 - b) * t.x + a * (1.0 - b) * t.y + (1.0 - a) * b * t.z + a * b * t.w;
}

// Depth of the mesh-frame point (x, y, z), FP64: z - H(x, y) with
// H = -sum of layer heights (Ocean::GetDepth, OpenGLRealOcean.cpp:280-331).
// The texel coordinate is u = ku + fu + ru.p, the integer part of the body
// position carried exactly in ku; in exact arithmetic this is the reference's
// i0 = floor(u) mod n, i1 = i0 + 1 mod n, alpha = frac(u).
__device__ __forceinline__ double depth64(const BodyDyn& D, const OceanK& oc, double x, double y, double z)
{
    double d = D.tz + D.rz[0] * x + D.rz[1] * y + D.rz[2] * z;
    for (int l = 0; l < oc.nlayers; ++l) {
        const LayerK& L = D.layer[l];
        double u = L.fu + L.ru[0] * x + L.ru[1] * y + L.ru[2] * z;
        double v = L.fv + L.rv[0] * x + L.rv[1] * y + L.rv[2] * z;
        double fi = floor(u), fj = floor(v);
        double a = u - fi, b = v - fj;
        int i, j;
        if (oc.nmask >= 0) {   // power-of-two grid: the wrap is a mask (two's complement)
            i = (L.ku + (int)fi) & oc.nmask;
            j = (L.kv + (int)fj) & oc.nmask;
        } else {
            i = wrap_n(L.ku + (int)fi, oc.n);
            j = wrap_n(L.kv + (int)fj, oc.n);
        }
        float4 t = oc.quad[l][(size_t)j * oc.n + i];
        d += (1.0 - a) * (1.0 - b) * t.x + a * (1.0 - b) * t.y + (1.0 - a) * b * t.z + a * b * t.w;
    }
    return d;
}

__device__ __forceinline__ float depth_vertex(const BodyDyn& D, const OceanK& oc, f3 p)
{
    return (float)depth64(D, oc, p.x, p.y, p.z);
}

// Depth of a wet polygon's centroid for the pressure buoyancy
// (SolidEntity.cpp:1650-1658), with the body's wave constants hoisted out of
// the face loop for a chunk: NL layers known at compile time (0, 1, 2) on a
// power-of-two grid (the texel wrap is a mask), their constants read once
// into registers; NL < 0 reads the layer count at run time and wraps any
// grid size.  The centroid is formed in FP64 from its FP32 wet base vertex
// and offset (exact: both are FP32 numbers at most a face apart), and its
// depth is the FP64 query of depth64.  Near the waterline a depth is a
// small difference of body-scale values, so a centroid rounded to FP32 at
// body scale (~1e-7 m) would carry that error into the millimetre depths of
// thin wet slivers.
template <int NL>
struct WaveQ {
    static constexpr int M = NL > 0 ? NL : 1;
    double tz, rz[3];
    double fu[M], fv[M], ru[M][3], rv[M][3];
    int ku[M], kv[M];
    const SFH_GLOBAL f32x4* quad[M];
    int n, nmask;
    const BodyDyn* D;
    const OceanK* oc;

    __device__ __forceinline__ void load(const BodyDyn& Dy, const OceanK& o)
    {
        D = &Dy;
        oc = &o;
        if (NL < 0) return;
        tz = Dy.tz;
        for (int k = 0; k < 3; ++k) rz[k] = Dy.rz[k];
        n = o.n;
        nmask = o.nmask;
#pragma unroll
        for (int l = 0; l < (NL > 0 ? NL : 0); ++l) {
            const LayerK& L = Dy.layer[l];
            fu[l] = L.fu;
            fv[l] = L.fv;
            for (int k = 0; k < 3; ++k) { ru[l][k] = L.ru[k]; rv[l][k] = L.rv[k]; }
            ku[l] = L.ku;
            kv[l] = L.kv;
            quad[l] = (const SFH_GLOBAL f32x4*)o.quad[l];
        }
    }

    // depth() in two halves (NL >= 0): issue() forms the vertical part and the
    // texel fractions and requests the texels, finish() mixes them in; the
    // same operations in the same order as depth(), so a caller can put
    // other work between the texel requests and their use.
    struct Pending {
        double d, a[M], b[M];
        f32x4 t[M];
    };

    __device__ __forceinline__ Pending issue(f3 base, f3 off) const
    {
        Pending q;
        const double x = (double)base.x + (double)off.x, y = (double)base.y + (double)off.y,
                     z = (double)base.z + (double)off.z;
        q.d = tz + rz[0] * x + rz[1] * y + rz[2] * z;
#pragma unroll
        for (int l = 0; l < (NL > 0 ? NL : 0); ++l) {
            const double u = fu[l] + ru[l][0] * x + ru[l][1] * y + ru[l][2] * z;
            const double v = fv[l] + rv[l][0] * x + rv[l][1] * y + rv[l][2] * z;
            const double fi = floor(u), fj = floor(v);
            const int i = (ku[l] + (int)fi) & nmask, j = (kv[l] + (int)fj) & nmask;
            q.t[l] = quad[l][(uint32_t)j * (uint32_t)n + (uint32_t)i];
            q.a[l] = u - fi;
            q.b[l] = v - fj;
        }
        return q;
    }

    __device__ __forceinline__ double finish(const Pending& q) const
    {
        double d = q.d;
#pragma unroll
        for (int l = 0; l < (NL > 0 ? NL : 0); ++l)
            d += texel_mix(q.a[l], q.b[l], make_float4(q.t[l].x, q.t[l].y, q.t[l].z, q.t[l].w));
        return d;
    }

    // depth64 at the mesh-frame point base + off (FP32 parts, summed in FP64)
    __device__ __forceinline__ double depth(f3 base, f3 off) const
    {
        const double x = (double)base.x + (double)off.x, y = (double)base.y + (double)off.y,
                     z = (double)base.z + (double)off.z;
        if (NL < 0) return depth64(*D, *oc, x, y, z);
        double d = tz + rz[0] * x + rz[1] * y + rz[2] * z;
#pragma unroll
        for (int l = 0; l < (NL > 0 ? NL : 0); ++l) {
            const double u = fu[l] + ru[l][0] * x + ru[l][1] * y + ru[l][2] * z;
            const double v = fv[l] + rv[l][0] * x + rv[l][1] * y + rv[l][2] * z;
            const double fi = floor(u), fj = floor(v);
            const int i = (ku[l] + (int)fi) & nmask, j = (kv[l] + (int)fj) & nmask;
            const f32x4 t = quad[l][(uint32_t)j * (uint32_t)n + (uint32_t)i];
            d += texel_mix(u - fi, v - fj, make_float4(t.x, t.y, t.z, t.w));
        }
        return d;
    }
};

// floor(u) mod n in [0, n) for a texel coordinate floor fu.
__device__ __forceinline__ int wrap_texel(double fu, const OceanK& oc)
{
    if (oc.nmask >= 0) return (int)((long long)fu & oc.nmask);   // power-of-two grid (two's complement)
    long long k = (long long)fu % oc.n;
    if (k < 0) k += oc.n;
    return (int)k;
}

// FP64 wave height sum at world (x, y), for the classification corners.

__device__ double wave_sum64(const OceanK& oc, double x, double y)
{
    double h = 0.0;
    for (int l = 0; l < oc.nlayers; ++l) {
        const double s = oc.scale[l];   // the same IEEE quotient n / len the host formed
        double u = x * s - 0.5, v = y * s - 0.5;
        double fi = floor(u), fj = floor(v);
        double a = u - fi, b = v - fj;
        const int ii = wrap_texel(fi, oc), jj = wrap_texel(fj, oc);
        h += texel_mix(a, b, oc.quad[l][(size_t)jj * oc.n + (size_t)ii]);
    }
    return h;
}

__device__ __forceinline__ void gl_basis(const double* m, double* B)  // OpenGL -> row-major
{
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) B[3 * r + c] = m[4 * c + r];
}

__device__ __forceinline__ void mat_t_vec(const double* B, const double* x, double* y)  // B^T x
{
#pragma clang fp contract(on)   // FMAs only within an expression: the same rounding wherever this is inlined
    for (int i = 0; i < 3; ++i) y[i] = B[i] * x[0] + B[3 + i] * x[1] + B[6 + i] * x[2];
}

__device__ __forceinline__ void mat_vec(const double* B, const double* x, double* y)  // B x
{
#pragma clang fp contract(on)   // FMAs only within an expression: the same rounding wherever this is inlined
    for (int i = 0; i < 3; ++i) y[i] = B[3 * i] * x[0] + B[3 * i + 1] * x[1] + B[3 * i + 2] * x[2];
}

__device__ __forceinline__ void cross64(const double* a, const double* b, double* c)
{
#pragma clang fp contract(on)   // FMAs only within an expression: the same rounding wherever this is inlined
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

__device__ __forceinline__ f3 to_f3(const double* x) { return mk3((float)x[0], (float)x[1], (float)x[2]); }

// Corner k of a world AABB (min mn, max mx, extents d = mx - mn) in the
// order of CheckBodyFluidPosition (SolidEntity.cpp:1246-1253): min, max,
// min + (dx, 0, 0), (0, dy, 0), (dx, dy, 0), (0, 0, dz), (dx, 0, dz), (0, dy, dz).
__device__ __forceinline__ void aabb_corner(int k, const double* mn, const double* mx, const double* d, double* q)
{
    if (k == 1) {
        q[0] = mx[0]; q[1] = mx[1]; q[2] = mx[2];
        return;
    }
    const bool ax = k == 2 || k == 4 || k == 6, ay = k == 3 || k == 4 || k == 7, az = k >= 5;
    q[0] = mn[0] + (ax ? d[0] : 0.0);
    q[1] = mn[1] + (ay ? d[1] : 0.0);
    q[2] = mn[2] + (az ? d[2] : 0.0);
}

// btCompoundShape::getAabb(T_CG) of a CG-frame AABB (centre ac, half extents
// ah): min, max and extents.
__device__ __forceinline__ void world_aabb(const double* Rcg, const double* tcg, const double* ac, const double* ah,
                                           double* mn, double* mx, double* d)
{
    for (int i = 0; i < 3; ++i) {
        const double c = (Rcg[3 * i] * ac[0] + Rcg[3 * i + 1] * ac[1] + Rcg[3 * i + 2] * ac[2]) + tcg[i];
        const double e = fabs(Rcg[3 * i]) * ah[0] + fabs(Rcg[3 * i + 1]) * ah[1] + fabs(Rcg[3 * i + 2]) * ah[2];
        mn[i] = c - e;
        mx[i] = c + e;
        d[i] = mx[i] - mn[i];
    }
}

// Sum over aligned groups of G lanes.
template <int G, typename T>
__device__ __forceinline__ T gsum(T x)
{
#pragma unroll
    for (int off = G / 2; off >= 1; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

// CheckBodyFluidPosition (SolidEntity.cpp:1239-1261) on the world AABB
// btCompoundShape::getAabb(T_CG) of the CG-frame AABB (centre c, half extents
// h): 0 outside, 1 inside, 2 crossing.  The eight wave queries are spread
// over a group of G lanes (lane `sub` of the group takes corners sub,
// sub + G, ...); every lane of the group returns the result.
template <int G>
__device__ int classify64_g(const double* Rcg, const double* tcg, const double* ac, const double* ah, const OceanK& oc,
                            int sub)
{
    double mn[3], mx[3], d[3];
    world_aabb(Rcg, tcg, ac, ah, mn, mx, d);
    int under = 0;
    for (int k = sub; k < 8; k += G) {
        double q[3];
        aabb_corner(k, mn, mx, d, q);
        double depth = q[2];
        if (oc.nlayers) depth += wave_sum64(oc, q[0], q[1]);
        if (depth > 0.0) ++under;
    }
    under = gsum<G>(under);
    return under == 0 ? 0 : (under == 8 ? 1 : 2);
}

// ---------------------------------------------------------------------------
// k_prologue: one lane per body
// ---------------------------------------------------------------------------

// states: kind 0 = body state (T_CG[16], v, w; T_C = T_CG * T_CG2C),
//         kind 1 = explicit (T_CG[16], T_C[16], v, w).
// active (optional): bodies with active[b] == 0 get position -2 and are left
// out of the step entirely.
// kPrologueLanes lanes per body: the eight wave queries of the classification
// are spread over them (one memory round trip instead of eight in turn, the
// body's critical path); they compute the rest redundantly and the first
// writes the results.
// The prologue also writes the rows of the bodies without face work (not
// hydrodynamic, outside, or without faces); the epilogue writes the others'.
static constexpr int kPrologueLanes = 8;

__device__ void epilogue_finish(const BodyStatic& S, const BodyDyn& D, const OceanK& oc, const double* P,
                                double* __restrict__ o);

// Body b on its group of kPrologueLanes lanes (`sub` is the lane's index in
// the group); returns the body's position for this step (-2: left out).
__device__ __forceinline__ int prologue_body(const BodyStatic* __restrict__ st, const double* __restrict__ states,
                                             int state_kind, const OceanK& oc, const uint8_t* __restrict__ active,
                                             const uint32_t* __restrict__ sidx, BodyDyn* __restrict__ dyn,
                                             int8_t* __restrict__ bpos, double* __restrict__ out, uint32_t b, int sub)
{
    const BodyStatic& S = st[b];
    // A body the step skips is marked in bpos only: its static record and its
    // BodyDyn row are not touched (the later passes test bpos, and the
    // compound pass returns first for an inactive compound, whose parts share
    // its activity).
    if (active && !active[sidx[b]]) {
        if (sub == 0) bpos[b] = -2;
        return -2;
    }
    // The body's BodyDyn record is written by the group's lane 0 as its parts
    // are computed (a record held whole until one store at the end keeps ~1 KB
    // live in registers); the wave layers below nlayers, and the currents'
    // table only when there are varying currents (cur.n says how many), so a
    // body on one layer without jets or pipes writes 360 of its 944 bytes.
    // D keeps what the finish of a body without faces reads.
    BodyDyn* __restrict__ dd = dyn + b;
    const bool w0 = sub == 0;
    BodyDyn D;
    const double* s = states + (size_t)S.sidx * (state_kind ? 38 : 22);
    double Rcg[9], tcg[3], R[9], tc[3];
    gl_basis(s, Rcg);
    tcg[0] = s[12]; tcg[1] = s[13]; tcg[2] = s[14];
    const double* v = s + (state_kind ? 32 : 16);
    const double* w = v + 3;
    if (state_kind) {
        gl_basis(s + 16, R);
        tc[0] = s[28]; tc[1] = s[29]; tc[2] = s[30];
    } else {
        // btTransform operator*: basis Rcg * B, origin Rcg * o + tcg
        const double* B = S.cg2c;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                R[3 * i + j] = Rcg[3 * i] * B[j] + Rcg[3 * i + 1] * B[3 + j] + Rcg[3 * i + 2] * B[6 + j];
        for (int i = 0; i < 3; ++i)
            tc[i] = (Rcg[3 * i] * B[9] + Rcg[3 * i + 1] * B[10] + Rcg[3 * i + 2] * B[11]) + tcg[i];
    }

    uint32_t flags = S.flags;
    int pos;
    if (flags & BF_MODE_SURF) pos = 2;
    else if (flags & (BF_MODE_SUBM | BF_MODE_AERO)) pos = 1;
    else if (!(flags & BF_HYDRO)) pos = -1;
    else {
        pos = classify64_g<kPrologueLanes>(Rcg, tcg, S.aabb_c, S.aabb_h, oc, sub);
        // body level: damping always on (Ocean::ApplyFluidForces, Ocean.cpp:248-252)
        flags |= BF_DAMPING;
    }
    D.cpos = pos;
    if (flags & BF_PART) {
        // Compound::ComputeHydrodynamicForces (Compound.cpp:291-437): the parts
        // follow the compound's classification.  Inside: drag of the external
        // parts whose solid is hydrodynamic.  Crossing: every hydrodynamic
        // part; an internal one only for its buoyancy, without drag.
        const bool ext = flags & BF_EXTERNAL, ph = flags & BF_PHYDRO, pb = flags & BF_BUOYANT;
        if (pos == 0) pos = -2;
        else if (pos == 1 && !(ext && ph)) pos = -2;
        else if (pos == 2) {
            if (!ph || (!ext && !pb)) pos = -2;
            else if (!ext) flags &= ~BF_DAMPING;
        }
    }
    D.pos = pos;
    D.flags = flags;
    if (w0) {
        bpos[b] = (int8_t)pos;
        dd->pos = pos;
        dd->flags = flags;
        dd->cpos = D.cpos;
        dd->pad_ = 0;
    }

    // mesh-frame FP32 constants
    double x[3], y[3], vl[3], wl[3], ul[3], wd[3];
    mat_t_vec(R, v, vl);
    mat_t_vec(R, w, wl);
    mat_t_vec(R, oc.u, ul);
    FaceBody fb;
    fb.w = to_f3(wl);
    for (int i = 0; i < 3; ++i) x[i] = tc[i] - tcg[i];
    mat_t_vec(R, x, y);
    D.dt[0] = y[0]; D.dt[1] = y[1]; D.dt[2] = y[2];
    cross64(wl, D.dt, wd);
    for (int i = 0; i < 3; ++i) y[i] = vl[i] + wd[i] - ul[i];
    fb.vb = to_f3(y);
    x[0] = tcg[0] - tc[0]; x[1] = tcg[1] - tc[1]; x[2] = -tc[2];
    mat_t_vec(R, x, y);
    fb.o = to_f3(y);
    if (w0) {
        dd->fb = fb;
        dd->rz[0] = R[6]; dd->rz[1] = R[7]; dd->rz[2] = R[8];
        dd->tz = tc[2];
        for (int i = 0; i < 3; ++i) dd->dt[i] = D.dt[i];
    }

    for (int l = 0; l < kMaxLayers; ++l) {
        if (l >= oc.nlayers) break;
        LayerK L;
        const double s = oc.scale[l];
        double bu = tc[0] * s - 0.5, bv = tc[1] * s - 0.5;
        double fu = floor(bu), fv = floor(bv);
        L.fu = bu - fu;
        L.fv = bv - fv;
        L.ku = wrap_texel(fu, oc);
        L.kv = wrap_texel(fv, oc);
        for (int k = 0; k < 3; ++k) { L.ru[k] = R[k] * s; L.rv[k] = R[3 + k] * s; }
        if (w0) dd->layer[l] = L;
    }
    if (w0) {
        dd->cur.n = oc.ncur;
        dd->cur.pad_ = 0;
        for (int i = 0; i < kMaxCur && oc.ncur > 0; ++i) {
            CurBody& C = dd->cur;
            C.type[i] = 0;
            for (int k = 0; k < 3; ++k) C.c[i][k] = C.a[i][k] = 0.0;
            for (int k = 0; k < 5; ++k) C.p[i][k] = 0.0;
            if (i >= oc.ncur) continue;
            double d[3], cl[3], al[3];
            for (int k = 0; k < 3; ++k) d[k] = oc.cur_c[i][k] - tc[k];
            mat_t_vec(R, d, cl);
            mat_t_vec(R, oc.cur_a[i], al);
            for (int k = 0; k < 3; ++k) { C.c[i][k] = cl[k]; C.a[i][k] = al[k]; }
            C.type[i] = oc.cur_type[i];
            for (int k = 0; k < 5; ++k) C.p[i][k] = oc.cur_p[i][k];
        }
    }
    for (int i = 0; i < 9; ++i) { D.R[i] = R[i]; D.Rcg[i] = Rcg[i]; }
    for (int i = 0; i < 3; ++i) D.tcg[i] = tcg[i];
    if (w0) {
        for (int i = 0; i < 9; ++i) { dd->R[i] = R[i]; dd->Rcg[i] = Rcg[i]; }
        for (int i = 0; i < 3; ++i) dd->tcg[i] = tcg[i];
    }

    double* o = out + (size_t)b * SFH_OUT_DOUBLES;
    if (pos == -1) {          // not a hydrodynamic body: nothing assigned
        if (sub == 0) { o[20] = 0.0; o[21] = -1.0; }
    } else if (pos == 0) {    // OUTSIDE: everything zero (SolidEntity.cpp:1810-1821)
        for (int k = sub; k < 20; k += kPrologueLanes) o[k] = 0.0;
        if (sub == 0) { o[20] = 255.0; o[21] = 0.0; }
    } else if (pos > 0 && sub == 0 && S.nchunks == 0) {   // a mesh without faces: empty sums
        double P[23];
        for (int k = 0; k < 23; ++k) P[k] = 0.0;
        epilogue_finish(S, D, oc, P, o);
    }
    return pos;
}

// The step's work list: one array of nc entries (nc = the batch's chunks),
// the chunks of bodies in position 1 (submerged loop) filled from the front
// and those of bodies in position 2 (crossing loop) from the back, with the
// two counts in counts[0..1] (zero before the prologue; the epilogue zeroes
// them for the next step).  Each workgroup of the prologue sums its bodies'
// chunk counts (a scan over each wavefront's lanes, then over the
// wavefronts) and reserves its ranges with one atomic add per list; a body's
// lanes then write its chunk ids.  The order of the entries depends on the
// order of the atomics, the results do not: a chunk's partial has its own
// slot and the epilogue sums a body's partials in chunk order.
static constexpr int kPrologueBlock = 256;
// Whole unmasked steps of large batches use workgroups of 1,024 (128 bodies):
// every workgroup reserves its list ranges with one atomic add per list, and
// the atomics of all workgroups on the two counters serialise (the compound
// batch's 81,920 part bodies: 1,280 reservations per list instead of 5,120,
// its step 8-14 us faster); with few bodies, or a mask leaving most
// workgroups without work, the smaller workgroups hide the prologue's latency
// chain better (the fleet 2.5 us, the rollout's masked step 2 us faster).
static constexpr int kPrologueBlockLarge = 1024;
static constexpr uint32_t kPrologueLargeBodies = 32768;

template <int BLOCK>
__device__ __forceinline__ void list_append(uint32_t* __restrict__ list, uint32_t* __restrict__ counts,
                                            uint32_t nc, uint32_t c_sub, uint32_t c_cross, uint32_t chunk0,
                                            int sub)
{
    __shared__ uint32_t wsum[2][BLOCK / 64];
    __shared__ uint32_t wbase[2][BLOCK / 64];
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const uint32_t v0 = sub == 0 ? c_sub : 0u, v1 = sub == 0 ? c_cross : 0u;
    uint32_t i0 = v0, i1 = v1;   // inclusive scan over the wavefront's lanes
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y0 = __shfl_up(i0, off, 64), y1 = __shfl_up(i1, off, 64);
        if (lane >= off) { i0 += y0; i1 += y1; }
    }
    if (lane == 63) { wsum[0][w] = i0; wsum[1][w] = i1; }
    __syncthreads();
    if (threadIdx.x < 2) {
        const int k = threadIdx.x;
        uint32_t run = 0;
        for (int j = 0; j < BLOCK / 64; ++j) { wbase[k][j] = run; run += wsum[k][j]; }
        const uint32_t base = run ? atomicAdd(counts + k, run) : 0u;
        for (int j = 0; j < BLOCK / 64; ++j) wbase[k][j] += base;
    }
    __syncthreads();
    const int lead = lane & ~(kPrologueLanes - 1);
    const uint32_t o0 = wbase[0][w] + __shfl(i0 - v0, lead, 64);
    const uint32_t o1 = wbase[1][w] + __shfl(i1 - v1, lead, 64);
    // the two parts hold at most the batch's nc chunks together; the bounds
    // tests only guard the array against counts that were not zeroed
    for (uint32_t i = (uint32_t)sub; i < c_sub; i += kPrologueLanes)
        if (o0 + i < nc) list[o0 + i] = chunk0 + i;
    for (uint32_t i = (uint32_t)sub; i < c_cross; i += kPrologueLanes)
        if (o1 + i < nc) list[nc - 1u - (o1 + i)] = chunk0 + i;
}

// Every lane of the grid reaches the list build (lanes past the last body
// contribute nothing), so the scans see whole wavefronts and workgroups.
// The prologue of the bodies of one workgroup of BLOCK threads (gt: the
// thread's index in the grid); every thread of the workgroup calls it (the
// list build's barriers).
template <int BLOCK>
__device__ __forceinline__ void prologue_group(const BodyStatic* __restrict__ st, const double* __restrict__ states,
                                               int state_kind, uint32_t n, const OceanK& oc,
                                               const uint8_t* __restrict__ active, const uint32_t* __restrict__ sidx,
                                               BodyDyn* __restrict__ dyn, int8_t* __restrict__ bpos,
                                               double* __restrict__ out, uint32_t* __restrict__ list, uint32_t nc,
                                               uint32_t* __restrict__ counts, uint32_t gt)
{
    const uint32_t b = gt / kPrologueLanes;
    const int sub = (int)(gt % kPrologueLanes);
    int pos = -3;
    if (b < n) pos = prologue_body(st, states, state_kind, oc, active, sidx, dyn, bpos, out, b, sub);
    uint32_t c_sub = 0, c_cross = 0, chunk0 = 0;
    if (pos == 1 || pos == 2) {
        chunk0 = st[b].chunk0;
        if (pos == 1) c_sub = st[b].nchunks;
        else c_cross = st[b].nchunks;
    }
    list_append<BLOCK>(list, counts, nc, c_sub, c_cross, chunk0, sub);
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_prologue(const BodyStatic* __restrict__ st, const double* __restrict__ states, int state_kind,
                           uint32_t n, OceanK oc, const uint8_t* __restrict__ active,
                           const uint32_t* __restrict__ sidx, BodyDyn* __restrict__ dyn,
                           int8_t* __restrict__ bpos, double* __restrict__ out, uint32_t* __restrict__ list,
                           uint32_t nc, uint32_t* __restrict__ counts)
{
    prologue_group<BLOCK>(st, states, state_kind, n, oc, active, sidx, dyn, bpos, out, list, nc, counts,
                          blockIdx.x * blockDim.x + threadIdx.x);
}

// Jets and pipes at a mesh-frame point (Jet.cpp:54-75, Pipe.cpp:57-78): the
// reference's formulas, in FP64 as there (Scalar), on the body-relative
// frame of CurBody; the result is the face loop's FP32.
struct CurFlow {
    static constexpr bool on = true;
    const CurBody* C;
    __device__ f3 operator()(double x, double y, double z) const
    {
        double u[3] = {0.0, 0.0, 0.0};
        for (int i = 0; i < C->n; ++i) {
            const double* n = C->a[i];
            const double cp[3] = {x - C->c[i][0], y - C->c[i][1], z - C->c[i][2]};
            double xc[3];
            cross64(cp, n, xc);
            const double d = sqrt(xc[0] * xc[0] + xc[1] * xc[1] + xc[2] * xc[2]);
            const double t = cp[0] * n[0] + cp[1] * n[1] + cp[2] * n[2];
            const double* p = C->p[i];
            double f, vm;
            if (C->type[i] == 1) {   // jet: r = p0, vout = p1
                if (t < 0.0) continue;
                const double rr = 1.0 / 5.0 * (t + 5.0 * p[0]);
                if (d >= rr) continue;
                vm = 10.0 * p[0] / (t + 5.0 * p[0]) * p[1];
                f = exp(-50.0 * d * d / (t * t));
            } else {                 // pipe: l, r1, r2, vin, gamma = p0..p4
                if (t < 0.0 || t > p[0]) continue;
                const double r = p[1] + (p[2] - p[1]) * t / p[0];
                if (d >= r) continue;
                vm = p[1] / r * p[3];
                f = pow(1.0 - d / r, p[4]);
            }
            for (int k = 0; k < 3; ++k) u[k] += f * (vm * n[k]);
        }
        return mk3((float)u[0], (float)u[1], (float)u[2]);
    }
};

// ---------------------------------------------------------------------------
// chunks
// ---------------------------------------------------------------------------

// A work item of the face pass: a run of consecutive faces of one body and
// the window of vertex ids they reference (vertices are numbered in order of
// first use, so a run of faces of a coherent mesh references a compact
// window).  vn = 0: the window is larger than the crossing pass keeps in LDS.
struct Chunk {
    uint32_t body, first, count;
    uint32_t vlo, vn;
    uint32_t pad[3];
};

// Vertex-window capacity (vertices) of a crossing-pass wavefront for chunks
// of 64 K faces; the window's FP64 depths are kept in LDS as floats.
__host__ __device__ constexpr uint32_t win_cap(uint32_t K) { return K <= 16 ? 64u * K : 1536u; }

// ---------------------------------------------------------------------------
// k_faces: one wavefront per chunk of up to 64*K faces of one body
// ---------------------------------------------------------------------------

// Transposed FP64 reduction of 24 per-lane values across a wavefront.  Each
// step exchanges half of the remaining values with the partner lane, so the
// wave moves 12+6+3+2+1+1 doubles instead of 6 per value.  On return v[0] of
// lane l holds the wave total of quantity q = 12 b5 + 6 b4 + 3 b3 + 2 b2 + b1
// (b = bits of l); q == 23 and (2 b2 + b1) == 3 are padding.
__device__ __forceinline__ void xreduce24(double (&v)[24], int lane)
{
    bool hi = lane & 32;
#pragma unroll
    for (int i = 0; i < 12; ++i) {
        double send = hi ? v[i] : v[i + 12], keep = hi ? v[i + 12] : v[i];
        v[i] = keep + __shfl_xor(send, 32, 64);
    }
    hi = lane & 16;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        double send = hi ? v[i] : v[i + 6], keep = hi ? v[i + 6] : v[i];
        v[i] = keep + __shfl_xor(send, 16, 64);
    }
    hi = lane & 8;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double send = hi ? v[i] : v[i + 3], keep = hi ? v[i + 3] : v[i];
        v[i] = keep + __shfl_xor(send, 8, 64);
    }
    v[3] = 0.0;
    hi = lane & 4;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        double send = hi ? v[i] : v[i + 2], keep = hi ? v[i + 2] : v[i];
        v[i] = keep + __shfl_xor(send, 4, 64);
    }
    hi = lane & 2;
    {
        double send = hi ? v[0] : v[1], keep = hi ? v[1] : v[0];
        v[0] = keep + __shfl_xor(send, 2, 64);
    }
    v[0] += __shfl_xor(v[0], 1, 64);
}

__device__ __forceinline__ int xreduce24_slot(int lane)
{
    int g = ((lane >> 1) & 1) + 2 * ((lane >> 2) & 1);
    if (g == 3) return -1;
    int q = 12 * ((lane >> 5) & 1) + 6 * ((lane >> 4) & 1) + 3 * ((lane >> 3) & 1) + g;
    return (q < 23 && !(lane & 1)) ? q : -1;
}

__device__ __forceinline__ f3 vert(const SFH_GLOBAL float* __restrict__ verts, uint32_t i)
{
    const SFH_GLOBAL float* v = verts + 3 * (size_t)i;
    return mk3(v[0], v[1], v[2]);
}

// Transposed FP64 reduction of 12 per-lane values: on return v[0] of lane l
// holds the total of quantity q = 6 b5 + 3 b4 + 2 b3 + b2 ((2 b3 + b2) == 3 is
// padding); lanes with b1 = b0 = 0 hold it.
__device__ __forceinline__ void xreduce12(double (&v)[12], int lane)
{
    bool hi = lane & 32;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        double send = hi ? v[i] : v[i + 6], keep = hi ? v[i + 6] : v[i];
        v[i] = keep + __shfl_xor(send, 32, 64);
    }
    hi = lane & 16;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double send = hi ? v[i] : v[i + 3], keep = hi ? v[i + 3] : v[i];
        v[i] = keep + __shfl_xor(send, 16, 64);
    }
    v[3] = 0.0;
    hi = lane & 8;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        double send = hi ? v[i] : v[i + 2], keep = hi ? v[i + 2] : v[i];
        v[i] = keep + __shfl_xor(send, 8, 64);
    }
    hi = lane & 4;
    {
        double send = hi ? v[0] : v[1], keep = hi ? v[1] : v[0];
        v[0] = keep + __shfl_xor(send, 4, 64);
    }
    v[0] += __shfl_xor(v[0], 2, 64);
    v[0] += __shfl_xor(v[0], 1, 64);
}

__device__ __forceinline__ int xreduce12_slot(int lane)
{
    int g = ((lane >> 2) & 1) + 2 * ((lane >> 3) & 1);
    if (g == 3 || (lane & 3)) return -1;
    return 6 * ((lane >> 5) & 1) + 3 * ((lane >> 4) & 1) + g;
}

// Submerged loop (bodies fully below the surface), one chunk (faces
// [f0, f0 + nfc) of body S): per-face arithmetic in FP32 with FP32 per-lane
// partial sums over the lane's K faces, reduced across the wavefront in FP64
// into the chunk's partial (12 doubles: Fdq Mdq Fdf Mdf).
template <int K, bool NU, bool AERO>
__device__ __forceinline__ void sub_chunk(uint32_t c, uint32_t f0, uint32_t nfc, const BodyStatic& S,
                                          const BodyDyn& D, double* __restrict__ partials, int lane)
{
    const SFH_GLOBAL float* __restrict__ verts = gptr(S.verts);
    const SFH_GLOBAL uint32_t* __restrict__ faces = gptr(S.faces) + 3 * (size_t)f0;
    const FaceBody fb = D.fb;
    const CurFlow flow{&D.cur};
    AccDrag<float> acc;
    acc_zero(acc);
    // Two-stage software pipeline over the lane's K faces: the indices of
    // face j+2 and the vertices of face j+1 are in flight while face j is
    // computed.  Face numbers past the chunk are clamped to its last face
    // so every load stays in bounds; their results are not used.
    const uint32_t flast = nfc - 1;
    uint32_t fc0 = min((uint32_t)lane, flast), fc1 = min(64u + lane, flast);
    uint32_t a0 = faces[3 * fc0], a1 = faces[3 * fc0 + 1], a2 = faces[3 * fc0 + 2];
    uint32_t b0 = faces[3 * fc1], b1 = faces[3 * fc1 + 1], b2 = faces[3 * fc1 + 2];
    f3 p1 = vert(verts, a0), p2 = vert(verts, a1), p3 = vert(verts, a2);
#pragma unroll 2
    for (int j = 0; j < K; ++j) {
        const uint32_t fn2 = min(64u * (j + 2) + lane, flast);
        f3 q1 = vert(verts, b0), q2 = vert(verts, b1), q3 = vert(verts, b2);
        uint32_t c0 = faces[3 * fn2], c1 = faces[3 * fn2 + 1], c2 = faces[3 * fn2 + 2];
        if (64u * j + lane < nfc) {
            if (AERO) {
                if (NU)
                    face_aero(fb, p1, p2, p3, flow, acc);
                else
                    face_aero(fb, p1, p2, p3, NoFlow(), acc);
            } else if (NU) {
                face_submerged(fb, p1, p2, p3, flow, acc);
            } else {
                face_submerged(fb, p1, p2, p3, acc);
            }
        }
        p1 = q1; p2 = q2; p3 = q3;
        b0 = c0; b1 = c1; b2 = c2;
    }
    double vals[12];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        vals[k] = acc.Fdq[k]; vals[3 + k] = acc.Mdq[k];
        vals[6 + k] = acc.Fdf[k]; vals[9 + k] = acc.Mdf[k];
    }
    xreduce12(vals, lane);
    const int q = xreduce12_slot(lane);
    if (q >= 0) gptr_w(partials)[(size_t)c * kChunkDoubles + q] = vals[0];
}

// Contributions of a wet (possibly clipped) face: pressure buoyancy at the
// wet centroid when the surface has waves (1649-1658), drag, wetted area.
template <bool NU, int NL>
__device__ __forceinline__ void wet_contrib(const Wet& wt, const FaceBody& fb, const BodyDyn& D, const WaveQ<NL>& wq,
                                            bool wave_buoy, bool damping, Acc<float>& acc)
{
    if (!wt.ok) return;
    f3 cpt = wt.base + wt.off;
    if (NL > 0 && !NU) {
        // the centroid's texels are requested first and mixed in after the
        // drag, whose arithmetic covers their latency
        typename WaveQ<NL>::Pending q;
        if (wave_buoy) q = wq.issue(wt.base, wt.off);
        if (damping) face_drag(fb, cpt, wt.n1, wt.A, acc);
        if (wave_buoy) {
            const float dc = (float)wq.finish(q);
            f3 fbi = wt.n1 * (-wt.A * dc);
            add3(acc.Fb, fbi);
            add3(acc.Mb, cross(cpt, fbi));
        }
        acc.Swet += wt.A;
        return;
    }
    if (NL != 0 && wave_buoy) {
        const float dc = (float)wq.depth(wt.base, wt.off);
        f3 fbi = wt.n1 * (-wt.A * dc);
        add3(acc.Fb, fbi);
        add3(acc.Mb, cross(cpt, fbi));
    }
    if (damping) {
        if (NU)
            face_drag<true>(fb, cpt, wt.n1, wt.A,
                            CurFlow{&D.cur}((double)wt.base.x + (double)wt.off.x, (double)wt.base.y + (double)wt.off.y,
                                            (double)wt.base.z + (double)wt.off.z),
                            acc);
        else
            face_drag(fb, cpt, wt.n1, wt.A, acc);
    }
    acc.Swet += wt.A;
}

// Surface loop (bodies crossing the surface), one chunk: the FP64 depths of
// the chunk's vertex window (WIN: into the wavefront's LDS slice first;
// otherwise per face vertex as the faces come), then the faces: all-wet
// faces inline; faces the waterline cuts (clip cases A-F) queued in LDS and
// clipped together afterwards, 64 at a time, so the clip code runs once per
// 64 cut faces rather than on every pass over the chunk.  Tetra volumes,
// pressure buoyancy with waves, drag.
template <int K, bool NU, int NL, bool WIN>
__device__ __forceinline__ void cross_chunk(uint32_t c, uint32_t f0, uint32_t nfc, uint32_t vlo, uint32_t vn,
                                            const BodyStatic& S, const BodyDyn& D, const OceanK& oc,
                                            const WaveQ<NL>& wq, float* __restrict__ ldep,
                                            uint16_t* __restrict__ cutq, double* __restrict__ partials, int lane)
{
    const uint64_t below = (1ull << lane) - 1ull;
    const SFH_GLOBAL float* __restrict__ verts = gptr(S.verts);
    const SFH_GLOBAL uint32_t* __restrict__ faces = gptr(S.faces) + 3 * (size_t)f0;
    const FaceBody fb = D.fb;
    const uint32_t flags = D.flags;
    const bool damping = flags & BF_DAMPING;
    const bool wave_buoy = (flags & BF_BUOYANT) && oc.nlayers > 0;
    const f3 zero = mk3(0.f, 0.f, 0.f);
    if (WIN) {
        // four vertices per lane in flight: their loads, then their texel
        // gathers, issue together (indices past the window are clamped to
        // its last vertex; those results are not stored)
        bool wet = false;
        for (uint32_t v0 = 0; v0 < vn; v0 += 256u) {
            f3 p[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) p[k] = vert(verts, vlo + min(v0 + 64u * k + lane, vn - 1));
            float d[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) d[k] = (float)wq.depth(p[k], zero);
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (v0 + 64u * k + lane < vn) {
                    ldep[v0 + 64u * k + lane] = d[k];
                    wet |= !(d[k] < 0.f);
                }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (__ballot(wet) == 0ull) {
            // every vertex of the window is dry: every face of the chunk is
            // skipped (SolidEntity.cpp:1353), the chunk contributes nothing
            if (lane < kChunkDoubles) gptr_w(partials)[(size_t)c * kChunkDoubles + lane] = 0.0;
            return;
        }
    }
    auto dep = [&](uint32_t i, f3 p) -> float {
        if (WIN) return ldep[i - vlo];
        return (float)wq.depth(p, zero);   // = depth64 at p
    };
    // the surface's centroid terms are wanted on a flat surface only (with
    // waves the buoyancy comes from the pressure integral)
    constexpr bool CBON = NL <= 0;
    Acc<float> acc;
    acc_zero(acc);
    const uint32_t flast = nfc - 1;
    uint32_t fc0 = min((uint32_t)lane, flast), fc1 = min(64u + lane, flast);
    uint32_t a0 = faces[3 * fc0], a1 = faces[3 * fc0 + 1], a2 = faces[3 * fc0 + 2];
    uint32_t b0 = faces[3 * fc1], b1 = faces[3 * fc1 + 1], b2 = faces[3 * fc1 + 2];
    f3 p1 = vert(verts, a0), p2 = vert(verts, a1), p3 = vert(verts, a2);
    uint32_t ncut = 0;
    // unrolled by two: the pipeline's registers alternate instead of being
    // copied at the end of every face
#pragma unroll 2
    for (int j = 0; j < K; ++j) {
        const uint32_t fn2 = min(64u * (j + 2) + lane, flast);
        f3 q1 = vert(verts, b0), q2 = vert(verts, b1), q3 = vert(verts, b2);
        uint32_t c0 = faces[3 * fn2], c1 = faces[3 * fn2 + 1], c2 = faces[3 * fn2 + 2];
        const float d0 = dep(a0, p1), d1 = dep(a1, p2), d2 = dep(a2, p3);
        const bool valid = 64u * j + lane < nfc;
        const bool allwet = !(d0 < 0.f) && !(d1 < 0.f) && !(d2 < 0.f);
        const bool cut = valid && !allwet && !(d0 < 0.f && d1 < 0.f && d2 < 0.f);
        if (valid && allwet)
            wet_contrib<NU, NL>(face_allwet<float, CBON>(fb, p1, p2, p3, acc), fb, D, wq, wave_buoy, damping, acc);
        const uint64_t m = __ballot(cut);
        if (cut) cutq[ncut + (uint32_t)__popcll(m & below)] = (uint16_t)(64u * j + lane);
        ncut += (uint32_t)__popcll(m);
        p1 = q1; p2 = q2; p3 = q3;
        a0 = b0; a1 = b1; a2 = b2;
        b0 = c0; b1 = c1; b2 = c2;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (uint32_t qb = 0; qb < ncut; qb += 64) {
        if (qb + lane < ncut) {
            const uint32_t fo = cutq[qb + lane];
            const uint32_t i0 = faces[3 * fo], i1 = faces[3 * fo + 1], i2 = faces[3 * fo + 2];
            const f3 v0 = vert(verts, i0), v1 = vert(verts, i1), v2 = vert(verts, i2);
            const Wet wt = face_clip<float, CBON>(fb, v0, v1, v2, dep(i0, v0), dep(i1, v1), dep(i2, v2), acc);
            wet_contrib<NU, NL>(wt, fb, D, wq, wave_buoy, damping, acc);
        }
    }
    // the queue and the depth window are rewritten by the next chunk
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // chunk partial layout: Fdq Mdq Fdf Mdf Fb Mb CB (3 each), V6, Swet
    double vals[24];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        vals[k] = acc.Fdq[k]; vals[3 + k] = acc.Mdq[k];
        vals[6 + k] = acc.Fdf[k]; vals[9 + k] = acc.Mdf[k];
        vals[12 + k] = acc.Fb[k]; vals[15 + k] = acc.Mb[k];
        vals[18 + k] = acc.CB[k];
    }
    vals[21] = acc.V6;
    vals[22] = acc.Swet;
    vals[23] = 0.0;
    xreduce24(vals, lane);
    const int q = xreduce24_slot(lane);
    if (q >= 0) gptr_w(partials)[(size_t)c * kChunkDoubles + q] = vals[0];
}

// ---------------------------------------------------------------------------
// epilogue: one wavefront per body
// ---------------------------------------------------------------------------

// CorrectHydrodynamicForces (SolidEntity.cpp:1263-1287): M = inverse(basis of
// T_O) by cofactors (btMatrix3x3.h:1093-1103); safeNormalize (btVector3.h:286-299).
__device__ void correct64(double rho, const double* BO, const double* Cd, const double* Cf,
                          double* Fdq, double* Tdq, double* Fdf, double* Tdf)
{
#pragma clang fp contract(on)   // FMAs only within an expression: the same rounding wherever this is inlined
#define CF(r1, c1, r2, c2) (BO[3 * (r1) + (c1)] * BO[3 * (r2) + (c2)] - BO[3 * (r1) + (c2)] * BO[3 * (r2) + (c1)])
    double co0 = CF(1, 1, 2, 2), co1 = CF(1, 2, 2, 0), co2 = CF(1, 0, 2, 1);
    double det = BO[0] * co0 + BO[1] * co1 + BO[2] * co2;
    double s = 1.0 / det;
    double M[9] = {co0 * s, CF(0, 2, 2, 1) * s, CF(0, 1, 1, 2) * s,
                   co1 * s, CF(0, 0, 2, 2) * s, CF(0, 2, 1, 0) * s,
                   co2 * s, CF(0, 1, 2, 0) * s, CF(0, 0, 1, 1) * s};
#undef CF
    double* vecs[4] = {Fdq, Tdq, Fdf, Tdf};
    for (int q = 0; q < 4; ++q) {
        double* x = vecs[q];
        double d[3];
        mat_vec(M, x, d);
        double l2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        if (l2 >= 2.220446049250313e-16 * 2.220446049250313e-16) {
            double inv = 1.0 / sqrt(l2);   // btVector3::operator/= multiplies by the reciprocal
            d[0] *= inv; d[1] *= inv; d[2] *= inv;
        } else {
            d[0] = 1.0; d[1] = 0.0; d[2] = 0.0;
        }
        const double* C = (q < 2) ? Cd : Cf;
        double c = fabs(d[0]) * C[0] + fabs(d[1]) * C[1] + fabs(d[2]) * C[2];
        double k = (q < 2) ? 0.5 * rho * c : rho * c;
        x[0] *= k; x[1] *= k; x[2] *= k;
    }
}

// CorrectAerodynamicForces (SolidEntity.cpp:1925-1948): d = safeNormalize(
// inverse(basis of T_CG2H) * Fda); F and T scaled by |Cd . d| with the proxy's
// Cd: sphere 1, cylinder (0.5, 0.5, 1), ellipsoid (1/a, 1/b, 1/c) / max.
// Minv: the inverse basis, rows, from the host; params: fdApproxParams.
__device__ void correct_aero64(uint32_t type, const double* params, const double* Minv, double* F, double* T)
{
#pragma clang fp contract(on)   // FMAs only within an expression: the same rounding wherever this is inlined
    double d[3];
    mat_vec(Minv, F, d);
    const double l2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (l2 >= 2.220446049250313e-16 * 2.220446049250313e-16) {
        const double inv = 1.0 / sqrt(l2);
        d[0] *= inv; d[1] *= inv; d[2] *= inv;
    } else {
        d[0] = 1.0; d[1] = 0.0; d[2] = 0.0;
    }
    double cor = 1.0;
    if (type == 2u) {
        cor = 0.5 * d[0] + 0.5 * d[1] + d[2];
    } else if (type == 3u) {
        double c[3] = {1.0 / params[0], 1.0 / params[1], 1.0 / params[2]};
        const double mx = fmax(fmax(c[0], c[1]), c[2]);
        cor = c[0] / mx * d[0] + c[1] / mx * d[1] + c[2] / mx * d[2];
    }
    cor = fabs(cor);
    for (int k = 0; k < 3; ++k) { F[k] *= cor; T[k] *= cor; }
}

// Finishes the row o of a body in position 1 or 2 from the sums P of its
// chunk partials (one lane): drag in world axes with moments about the CoG,
// buoyancy, CorrectHydrodynamicForces, the write mask.
__device__ void epilogue_finish(const BodyStatic& S, const BodyDyn& D, const OceanK& oc, const double* P,
                                double* __restrict__ o)
{
#pragma clang fp contract(on)   // FMAs only within an expression: the same rounding wherever this is inlined
    const int pos = D.pos;
    const uint32_t flags = D.flags;
    const bool body_level = !(flags & (BF_MODE_SURF | BF_MODE_SUBM | BF_MODE_AERO));
    const bool damping = flags & BF_DAMPING;
    const bool buoyant = flags & BF_BUOYANT;
    uint32_t written = 0;
    // drag: world = R * F_local; torque = R * (M_local + dt x F_local)
    double drag[12];
    for (int q = 0; q < 2; ++q) {
        const double* F = P + 6 * q;
        const double* M = P + 6 * q + 3;
        double t[3], m[3];
        cross64(D.dt, F, t);
        for (int k = 0; k < 3; ++k) m[k] = M[k] + t[k];
        mat_vec(D.R, F, drag + 6 * q);
        mat_vec(D.R, m, drag + 6 * q + 3);
    }
    double* Fdq = drag;
    double* Tdq = drag + 3;
    double* Fdf = drag + 6;
    double* Tdf = drag + 9;
    if (flags & BF_MODE_AERO) {   // Fda, Tda = 0.5 rho_gas times the sums (SolidEntity.cpp:1915-1917)
        double Fa[3], Ta[3];
        for (int k = 0; k < 3; ++k) {
            Fa[k] = 0.5 * oc.rho * Fdq[k];
            Ta[k] = 0.5 * oc.rho * Tdq[k];
        }
        if (flags & BF_AERO_CORRECT) correct_aero64((flags >> BF_APPROX_SHIFT) & 3u, S.Cd, S.cg2o, Fa, Ta);
        for (int k = 0; k < 3; ++k) {
            o[6 + k] = Fa[k];
            o[9 + k] = Ta[k];
        }
        o[20] = (double)(4u | 8u);   // written: the quadratic-drag slots
        o[21] = (double)pos;
        return;
    }
    double Fb[3] = {0, 0, 0}, Tb[3] = {0, 0, 0};
    double Swet = 0.0, Vsub = 0.0;
    if (pos == 1) {
        if (body_level && !(flags & BF_PART)) {
            if (buoyant) {   // analytic buoyancy of a submerged body (1828-1836)
                double cb[3];
                for (int k = 0; k < 3; ++k) Fb[k] = -S.volume * oc.rho * oc.g[k];
                mat_vec(D.Rcg, S.P_CB, cb);
                cross64(cb, Fb, Tb);
                written |= 3u;
            }
            Swet = S.surface;
            written |= 64u;
        }
        written |= 4u | 8u | 16u | 32u;
    } else {
        const double V6 = P[21];
        if (buoyant && V6 > (double)1e-9f) {
            Vsub = V6 / 6.0;
            written |= 1u | 2u | 128u;
            if (oc.nlayers) {   // pressure integration (1694-1700)
                double s = oc.rho * oc.g[2];
                double t[3], m[3];
                cross64(D.dt, P + 12, t);
                for (int k = 0; k < 3; ++k) m[k] = P[15 + k] + t[k];
                mat_vec(D.R, P + 12, Fb);
                mat_vec(D.R, m, Tb);
                for (int k = 0; k < 3; ++k) { Fb[k] *= s; Tb[k] *= s; }
            } else {            // volume and centre of buoyancy (1701-1707)
                double cbl[3], cb[3], r[3];
                for (int k = 0; k < 3; ++k) cbl[k] = P[18 + k] / V6;
                mat_vec(D.R, cbl, cb);
                // CB - T_CG.origin = (p0 - t_CG) + R cbl, p0 - t_CG = (0, 0, -t_CG.z)
                r[0] = cb[0]; r[1] = cb[1]; r[2] = cb[2] - D.tcg[2];
                for (int k = 0; k < 3; ++k) Fb[k] = -Vsub * oc.rho * oc.g[k];
                cross64(r, Fb, Tb);
            }
        }
        if (damping) written |= 4u | 8u | 16u | 32u;
        Swet = P[22];
        written |= 64u;
    }
    if (body_level && damping) {
        double BO[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                BO[3 * i + j] = D.Rcg[3 * i] * S.cg2o[j] + D.Rcg[3 * i + 1] * S.cg2o[3 + j] + D.Rcg[3 * i + 2] * S.cg2o[6 + j];
        correct64(oc.rho, BO, S.Cd, S.Cf, Fdq, Tdq, Fdf, Tdf);
    }
    if (written & 1u) for (int k = 0; k < 3; ++k) o[k] = Fb[k];
    if (written & 2u) for (int k = 0; k < 3; ++k) o[3 + k] = Tb[k];
    if (written & 4u) for (int k = 0; k < 3; ++k) o[6 + k] = Fdq[k];
    if (written & 8u) for (int k = 0; k < 3; ++k) o[9 + k] = Tdq[k];
    if (written & 16u) for (int k = 0; k < 3; ++k) o[12 + k] = Fdf[k];
    if (written & 32u) for (int k = 0; k < 3; ++k) o[15 + k] = Tdf[k];
    if (written & 64u) o[18] = Swet;
    if (written & 128u) o[19] = Vsub;
    o[20] = (double)written;
    o[21] = (double)pos;
}

// Sums of the chunk partials of body b on one wavefront: lanes stride over
// the body's chunks, a transposed reduction across the wavefront (a fixed
// order, so the sums do not depend on which wavefront runs this) leaves the
// 23 totals on 23 lanes, which store them to `ps` (24 doubles of LDS).
__device__ void epilogue_sums(const BodyStatic& S, const double* partials, double* ps, int lane)
{
    double v[24];
#pragma unroll
    for (int k = 0; k < 24; ++k) v[k] = 0.0;
    for (uint32_t c = lane; c < S.nchunks; c += 64) {
        const double2* p = (const double2*)(partials + (size_t)(S.chunk0 + c) * kChunkDoubles);
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            const double2 x = p[k];
            v[2 * k] += x.x;
            v[2 * k + 1] += x.y;
        }
    }
    v[23] = 0.0;
    xreduce24(v, lane);
    const int q = xreduce24_slot(lane);
    if (q >= 0) ps[q] = v[0];
}

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// kEpilogueWaves wavefronts per workgroup; wavefront w of W takes the bodies
// w, w + W, w + 2W, ...: lane k reads the position of the k-th of the
// wavefront's next 64 bodies in one load.  Of those with face work (positions
// 1 and 2; the prologue wrote the others' rows, and those of bodies without
// faces), a body of one chunk is finished by its own lane from the chunk's
// partial; the larger ones are summed by the whole wavefront, one after the
// other, into LDS slots, and every kEpilogueBatch bodies (and at the end
// of the 64) lane j finishes the row of the body in slot j — world rotation,
// moment about the CoG, buoyancy, CorrectHydrodynamicForces, the write mask —
// so the finishing arithmetic runs once per batch, not once per body.
// A masked step with few active bodies costs two loads per 64 bodies.
// It also zeroes the work list's counts for the next step's prologue.
static constexpr int kEpilogueWaves = 4;
static constexpr int kEpilogueBatch = 16;

__global__ __launch_bounds__(64 * kEpilogueWaves, 4) void k_epilogue(
    const BodyStatic* __restrict__ st, const BodyDyn* __restrict__ dyn, const int8_t* __restrict__ bpos, uint32_t n,
    OceanK oc, const double* __restrict__ partials, double* __restrict__ out, uint32_t* __restrict__ counts)
{
    __shared__ double ps[kEpilogueWaves][kEpilogueBatch][24];
    const int lane = threadIdx.x & 63;
    const int wl = threadIdx.x >> 6;
    const uint32_t W = gridDim.x * kEpilogueWaves;
    const uint32_t w = blockIdx.x * kEpilogueWaves + wl;
    if (blockIdx.x == 0 && threadIdx.x < 2) counts[threadIdx.x] = 0u;
    for (uint32_t base = w; base < n; base += 64u * W) {
        const uint32_t bl = base + (uint32_t)lane * W;
        int p = -3;
        if (bl < n) p = bpos[bl];
        uint32_t nck = 0;
        if (p == 1 || p == 2) nck = st[bl].nchunks;
        // a body of one chunk: the chunk's partial is its sum, and its lane
        // finishes the row, every such body of the 64 at once
        if (nck == 1) {
            const double2* p2 = (const double2*)(partials + (size_t)st[bl].chunk0 * kChunkDoubles);
            double P[23];
#pragma unroll
            for (int k = 0; k < 11; ++k) {
                const double2 x = p2[k];
                P[2 * k] = x.x;
                P[2 * k + 1] = x.y;
            }
            P[22] = ((const double*)p2)[22];
            epilogue_finish(st[bl], dyn[bl], oc, P, out + (size_t)bl * SFH_OUT_DOUBLES);
        }
        uint64_t m = __ballot(nck > 1);
        uint32_t mine = 0;    // lane j < nb: the body in slot j
        uint32_t first = 0;   // the body in slot 0 (wave-uniform)
        int nb = 0;
        while (m) {
            const uint32_t k = (uint32_t)__builtin_ctzll(m);
            m &= m - 1ull;
            const uint32_t b = base + k * W;
            epilogue_sums(st[b], partials, ps[wl][nb], lane);
            if (nb == 0) first = b;
            if (lane == nb) mine = b;
            if (++nb == kEpilogueBatch || !m) {
                wave_lds_sync();
                // a batch of one keeps the body's index wave-uniform (its
                // constants come in through scalar loads); the finishing
                // arithmetic rounds the same either way
                if (nb == 1) {
                    if (lane == 0) {
                        double P[23];
#pragma unroll
                        for (int q = 0; q < 23; ++q) P[q] = ps[wl][0][q];
                        epilogue_finish(st[first], dyn[first], oc, P, out + (size_t)first * SFH_OUT_DOUBLES);
                    }
                } else if (lane < nb) {
                    double P[23];
#pragma unroll
                    for (int q = 0; q < 23; ++q) P[q] = ps[wl][lane][q];
                    epilogue_finish(st[mine], dyn[mine], oc, P, out + (size_t)mine * SFH_OUT_DOUBLES);
                }
                wave_lds_sync();   // the slots are rewritten by the next batch
                nb = 0;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// k_faces: one wavefront per chunk
// ---------------------------------------------------------------------------

// Wavefront w of W takes the entries w, w + W, w + 2W, ... of its PASS's
// part of the step's work list: kPassSub, the submerged chunks (the list's
// front, counts[0] entries); kPassCross, the crossing chunks (its back,
// counts[1] entries); kPassBoth, the crossing chunks then the submerged ones
// (the longer work first), in one launch; kPassAero, the front with the
// aerodynamic loop.  kPassSub and kPassCross are two launches, each at its
// own occupancy.
static constexpr int kPassSub = 1, kPassCross = 2, kPassBoth = 3, kPassAero = 4;

template <int K, bool NU, int NL, int PASS>
__global__ __launch_bounds__(256, NU ? 1 : ((PASS & kPassCross) ? 5 : 8)) void k_faces(
    const uint32_t* __restrict__ list, const uint32_t* __restrict__ counts, uint32_t nc,
    const Chunk* __restrict__ chunks, const BodyStatic* __restrict__ st, const BodyDyn* __restrict__ dyn, OceanK oc,
    double* __restrict__ partials)
{
    constexpr bool CROSS = (PASS & kPassCross) != 0;
    __shared__ uint16_t cutq[4][CROSS ? 64 * K : 1];
    __shared__ float ldep[4][CROSS ? win_cap(K) : 1];
    const int lane = threadIdx.x & 63;
    const int wl = threadIdx.x >> 6;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    const uint32_t n_cross = CROSS ? min(counts[1], nc) : 0u;
    const uint32_t n_sub = PASS == kPassCross ? 0u : min(counts[0], nc - min(counts[1], nc));
    const uint32_t* __restrict__ cross_list = list + (nc - n_cross);
    for (uint32_t i = wave; i < n_cross + n_sub; i += nwaves) {
        const bool is_cross = i < n_cross;
        const uint32_t c = __builtin_amdgcn_readfirstlane(is_cross ? cross_list[i] : list[i - n_cross]);
        const Chunk ch = chunks[c];
        const uint32_t b = __builtin_amdgcn_readfirstlane(ch.body);
        const uint32_t f0 = __builtin_amdgcn_readfirstlane(ch.first);
        const uint32_t nfc = __builtin_amdgcn_readfirstlane(ch.count);
        const BodyStatic& S = st[b];
        const BodyDyn& D = dyn[b];
        if (!CROSS || !is_cross) {
            sub_chunk<K, NU, PASS == kPassAero>(c, f0, nfc, S, D, partials, lane);
        } else {
            const uint32_t vlo = __builtin_amdgcn_readfirstlane(ch.vlo);
            const uint32_t vn = __builtin_amdgcn_readfirstlane(ch.vn);
            WaveQ<NL> wq;
            wq.load(D, oc);
            if (vn > 0 && vn <= win_cap(K))
                cross_chunk<K, NU, NL, true>(c, f0, nfc, vlo, vn, S, D, oc, wq, ldep[wl], cutq[wl], partials, lane);
            else
                cross_chunk<K, NU, NL, false>(c, f0, nfc, vlo, vn, S, D, oc, wq, ldep[wl], cutq[wl], partials, lane);
        }
    }
}

// Per row of the state array (a body, or a compound): what the surrogate's
// feature build needs.
struct UnitStatic {
    double P_CB[3];
    double aabb_c[3], aabb_h[3];
    double volume, surface;
    uint32_t hydro, buoyant;
    const float* hst;            // hydrostatic table of the unit's hull (null: none)
    uint32_t hs, hth, hph, pad_; // its dimensions: immersion, polar angle, azimuth
};

// Per compound: static data of Compound::ComputeHydrodynamicForces.
struct CompStatic {
    uint32_t part0, nparts;   // the compound's parts are bodies part0 .. part0+nparts-1 of the batch
    uint32_t flags;           // BF_BUOYANT of the compound
    uint32_t pad_;
    double volume, surface;
    double P_CB[3];
};

// Composes the part rows into the compound's row (Compound.cpp:291-437), one
// lane per compound, parts in their order.  Inside: the compound's analytic
// buoyancy and the sum of the parts' corrected drag.  Crossing: the sums of
// the part results, where a part's output the reference leaves untouched
// keeps the previous part's value (the per-part temporaries are declared once
// before the loop, Compound.cpp:392-399); internal parts add buoyancy only.
__global__ void k_compound(const CompStatic* __restrict__ cs, const BodyStatic* __restrict__ st,
                           const BodyDyn* __restrict__ dyn, uint32_t ncomp, OceanK oc,
                           const uint8_t* __restrict__ active, const double* __restrict__ part_out,
                           double* __restrict__ out)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncomp) return;
    if (active && !active[c]) return;
    const CompStatic& C = cs[c];
    double* o = out + (size_t)c * SFH_OUT_DOUBLES;
    const BodyDyn& D0 = dyn[C.part0];
    const int cpos = D0.cpos;
    if (cpos < 0) { o[20] = 0.0; o[21] = -1.0; return; }
    if (cpos == 0) {
        for (int k = 0; k < 20; ++k) o[k] = 0.0;
        o[20] = 255.0; o[21] = 0.0;
        return;
    }
    if (C.flags & BF_MODE_AERO) {   // Compound::ComputeAerodynamicForces (Compound.cpp:439-464)
        double acc[6] = {0, 0, 0, 0, 0, 0};
        for (uint32_t p = C.part0; p < C.part0 + C.nparts; ++p) {
            if (dyn[p].pos != 1) continue;
            const double* row = part_out + (size_t)p * SFH_OUT_DOUBLES;
            for (int k = 0; k < 6; ++k) acc[k] += row[6 + k];
        }
        for (int k = 0; k < 6; ++k) o[6 + k] = acc[k];
        o[20] = (double)(4u | 8u);
        o[21] = 1.0;
        return;
    }
    if (cpos == 1) {
        double acc[12];
        for (int k = 0; k < 12; ++k) acc[k] = 0.0;
        for (uint32_t p = C.part0; p < C.part0 + C.nparts; ++p) {
            if (dyn[p].pos != 1) continue;
            const double* row = part_out + (size_t)p * SFH_OUT_DOUBLES;
            for (int k = 0; k < 12; ++k) acc[k] += row[6 + k];
        }
        uint32_t written = 4u | 8u | 16u | 32u | 64u | 128u;
        if (C.flags & BF_BUOYANT) {
            double Fb[3], Tb[3], cb[3];
            for (int k = 0; k < 3; ++k) Fb[k] = -C.volume * oc.rho * oc.g[k];
            mat_vec(D0.Rcg, C.P_CB, cb);
            cross64(cb, Fb, Tb);
            for (int k = 0; k < 3; ++k) { o[k] = Fb[k]; o[3 + k] = Tb[k]; }
            written |= 3u;
        }
        for (int k = 0; k < 12; ++k) o[6 + k] = acc[k];
        o[18] = C.surface;
        o[19] = C.volume;
        o[20] = (double)written;
        o[21] = 1.0;
        return;
    }
    double carry[20], acc[20];
    for (int k = 0; k < 20; ++k) carry[k] = acc[k] = 0.0;
    for (uint32_t p = C.part0; p < C.part0 + C.nparts; ++p) {
        if (dyn[p].pos != 2) continue;
        const double* row = part_out + (size_t)p * SFH_OUT_DOUBLES;
        const uint32_t wm = (uint32_t)row[20];
        for (int q = 0; q < 6; ++q)
            if (wm & (1u << q))
                for (int k = 0; k < 3; ++k) carry[3 * q + k] = row[3 * q + k];
        if (wm & 64u) carry[18] = row[18];
        if (wm & 128u) carry[19] = row[19];
        if (st[p].flags & BF_EXTERNAL) {
            for (int k = 0; k < 20; ++k) acc[k] += carry[k];
        } else {
            for (int k = 0; k < 6; ++k) acc[k] += carry[k];
            acc[19] += carry[19];
        }
    }
    for (int k = 0; k < 20; ++k) o[k] = acc[k];
    o[20] = 255.0;
    o[21] = 2.0;
}

// Standalone CorrectHydrodynamicForces for the per-body entry point.
__global__ void k_correct(double rho, const double* __restrict__ in, double* __restrict__ io)
{
    // in: BO[9] row-major, Cd[3], Cf[3]; io: Fdq Tdq Fdf Tdf
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    correct64(rho, in, in + 9, in + 12, io, io + 3, io + 6, io + 9);
}

// ---------------------------------------------------------------------------
// wave grid builder
// ---------------------------------------------------------------------------

__global__ void k_wave_grid(const sfh_wave_component* __restrict__ comps, uint32_t ncomp, uint32_t n,
                            float* __restrict__ grid)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t j = blockIdx.y;
    if (i >= n) return;
    const float two_pi_n = 6.283185307179586f / (float)n;
    float h = 0.f;
    for (uint32_t k = 0; k < ncomp; ++k) {
        long long m = (long long)comps[k].mx * i + (long long)comps[k].my * j;
        m %= (long long)n;
        if (m < 0) m += n;
        h += comps[k].amplitude * cosf(two_pi_n * (float)m + comps[k].phase);
    }
    grid[(size_t)j * n + i] = h;
}

// quad texture: per texel (h[j][i], h[j][i+1], h[j+1][i], h[j+1][i+1]), wrapping
__global__ void k_wave_quad(const float* __restrict__ data, uint32_t n, uint32_t channels, uint32_t channel,
                            float4* __restrict__ quad)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t j = blockIdx.y;
    if (i >= n) return;
    uint32_t i1 = (i + 1 == n) ? 0 : i + 1;
    uint32_t j1 = (j + 1 == n) ? 0 : j + 1;
    float4 q;
    q.x = data[((size_t)j * n + i) * channels + channel];
    q.y = data[((size_t)j * n + i1) * channels + channel];
    q.z = data[((size_t)j1 * n + i) * channels + channel];
    q.w = data[((size_t)j1 * n + i1) * channels + channel];
    quad[(size_t)j * n + i] = q;
}

// ---------------------------------------------------------------------------
// host objects
// ---------------------------------------------------------------------------

// Meshes are reference counted: the creator holds one reference and every
// batch (and the per-body workspace) that keeps the mesh's device pointers holds
// another, so sfh_mesh_destroy never frees memory a batch still reads.
struct sfh_mesh_t {
    float* d_verts = nullptr;
    uint32_t* d_faces = nullptr;
    uint2* d_blkwin = nullptr;          // per block of 64 faces: lowest and highest vertex id referenced
    std::vector<uint2> blkwin;          // the same on the host
    uint32_t nv = 0, nf = 0;
    float aabb_min[3], aabb_max[3];
    uint64_t id = 0;
    std::atomic<int> refs{1};
};

static std::atomic<uint64_t> g_mesh_ids{1};

static void mesh_release(sfh_mesh_t* m)
{
    if (m && m->refs.fetch_sub(1) == 1) {
        hipFree(m->d_verts);
        hipFree(m->d_faces);
        hipFree(m->d_blkwin);
        delete m;
    }
}

struct Current {
    int type, enabled;
    double p[12];
};

struct sfh_ocean_t {
    double rho = 1000.0;
    double g[3] = {0, 0, 9.81};
    int nlayers = 0;
    uint32_t n = 0;
    double len[kMaxLayers] = {0, 0, 0, 0};
    float4* d_quad[kMaxLayers] = {nullptr, nullptr, nullptr, nullptr};
    std::vector<Current> currents;
    bool currents_enabled = false;
};

struct sfh_batch_t {
    std::vector<sfh_mesh_t*> meshes;   // one reference per body with a mesh
    std::vector<sfh_mesh_t*> meshes_by_body;
    uint32_t n = 0;         // bodies of the face pass (the parts, for a compound batch)
    uint32_t nunits = 0;    // rows of the state and output arrays: bodies, or compounds
    bool compound = false;
    bool aero = false;      // every body runs the aerodynamic face loop (sfh_forces_aero)
    CompStatic* d_comp = nullptr;
    double* d_part_out = nullptr;
    UnitStatic* d_units = nullptr;  // per unit: P_CB and AABB (CG frame), for the surrogate features
    uint32_t nchunks = 0;
    uint64_t nfaces = 0;
    uint32_t K = 8;
    uint32_t waves = 0;
    uint32_t split = 0;     // 1: the submerged and the crossing loops in launches of their own
    std::vector<BodyStatic> h_static;
    BodyStatic* d_static = nullptr;
    uint32_t* d_sidx = nullptr;     // per body: BodyStatic::sidx, compact (the masked prologue's first read)
    double* d_hstates = nullptr;    // device copies of sfh_batch_compute_host's host buffers (first call)
    double* d_hout = nullptr;
    Chunk* d_chunks = nullptr;
    int8_t* d_bpos = nullptr;       // per body: this step's position (-2 inactive, -1 not hydrodynamic, 0, 1, 2)
    uint32_t* d_lists = nullptr;    // the step's work list [nchunks], then its two counts
    BodyDyn* d_dyn = nullptr;
    double* d_partials = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    bool timing = false;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    bool ev_pending = false;
    double face_ms_sum = 0.0, total_ms_sum = 0.0;
    uint32_t timed = 0;
    int device = 0;
};

static OceanK make_oceank(const sfh_ocean_t* o)
{
    OceanK k;
    memset(&k, 0, sizeof(k));
    k.rho = o->rho;
    for (int i = 0; i < 3; ++i) k.g[i] = o->g[i];
    k.nlayers = o->nlayers;
    k.n = (int32_t)o->n;
    k.nmask = (o->n && (o->n & (o->n - 1)) == 0) ? (int32_t)o->n - 1 : -1;
    for (int l = 0; l < kMaxLayers; ++l) {
        k.len[l] = o->len[l];
        k.scale[l] = o->len[l] > 0.0 ? (double)o->n / o->len[l] : 0.0;
        k.quad[l] = o->d_quad[l];
    }
    if (o->currents_enabled)
        for (const Current& c : o->currents) {
            if (!c.enabled) continue;
            if (c.type == 0) {
                for (int i = 0; i < 3; ++i) k.u[i] += c.p[i];
            } else if (k.ncur < kMaxCur) {   // sfh_ocean_add_current keeps at most kMaxCur
                const int j = k.ncur++;
                k.cur_type[j] = c.type;
                for (int i = 0; i < 3; ++i) {
                    k.cur_c[j][i] = c.p[i];
                    k.cur_a[j][i] = c.p[3 + i];
                }
                for (int i = 0; i < 5; ++i) k.cur_p[j][i] = c.p[6 + i];
            }
        }
    return k;
}

extern "C" int sfh_init(int device)
{
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) return fail(SFH_ERR_NODEVICE, "no HIP device");
    if (device < 0 || device >= count) return fail(SFH_ERR_INVALID, "device %d out of range", device);
    HIPCHK(hipSetDevice(device));
    return sfh_internal_na_init();   // this device's no-alloc scratch address
}

// ---- meshes -----------------------------------------------------------------

extern "C" int sfh_mesh_create(const void* vertex_pos, size_t vertex_stride, uint32_t nv,
                               const uint32_t* face_ids, uint32_t nf, sfh_mesh* out)
{
    if (!out || (!vertex_pos && nv) || (!face_ids && nf)) return fail(SFH_ERR_INVALID, "null argument");
    if (vertex_stride < 12) return fail(SFH_ERR_INVALID, "vertex stride %zu < 12", vertex_stride);
    if (nf > (1u << 30)) return fail(SFH_ERR_SIZE, "too many faces");
    for (size_t i = 0; i < 3 * (size_t)nf; ++i)
        if (face_ids[i] >= nv) return fail(SFH_ERR_INVALID, "face %zu references vertex %u >= %u", i / 3, face_ids[i], nv);
    // Renumber vertices in order of first use by the faces, so each run of
    // consecutive faces references a compact vertex window.  Positions are
    // unchanged, so every per-face result is unchanged.
    std::vector<uint32_t> remap(nv, UINT32_MAX);
    std::vector<uint32_t> order;
    order.reserve(nv);
    std::vector<uint32_t> faces(3 * (size_t)nf);
    for (size_t i = 0; i < 3 * (size_t)nf; ++i) {
        uint32_t v = face_ids[i];
        if (remap[v] == UINT32_MAX) {
            remap[v] = (uint32_t)order.size();
            order.push_back(v);
        }
        faces[i] = remap[v];
    }
    for (uint32_t v = 0; v < nv; ++v)
        if (remap[v] == UINT32_MAX) {
            remap[v] = (uint32_t)order.size();
            order.push_back(v);
        }
    std::vector<float> verts(3 * (size_t)nv);
    const char* src = (const char*)vertex_pos;
    float mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
    for (uint32_t i = 0; i < nv; ++i) {
        float p[3];
        memcpy(p, src + (size_t)order[i] * vertex_stride, 12);
        for (int k = 0; k < 3; ++k) {
            verts[3 * (size_t)i + k] = p[k];
            if (i == 0 || p[k] < mn[k]) mn[k] = p[k];
            if (i == 0 || p[k] > mx[k]) mx[k] = p[k];
        }
    }
    sfh_mesh_t* m = new sfh_mesh_t();
    m->id = g_mesh_ids.fetch_add(1);
    m->nv = nv;
    m->nf = nf;
    m->blkwin.resize((nf + 63) / 64);
    for (size_t k = 0; k < m->blkwin.size(); ++k) {
        uint32_t lo = UINT32_MAX, hi = 0;
        for (size_t i = 3 * 64 * k; i < std::min(3 * 64 * (k + 1), 3 * (size_t)nf); ++i) {
            lo = std::min(lo, faces[i]);
            hi = std::max(hi, faces[i]);
        }
        m->blkwin[k] = make_uint2(lo, hi);
    }
    for (int k = 0; k < 3; ++k) { m->aabb_min[k] = mn[k]; m->aabb_max[k] = mx[k]; }
    if (nv) {
        hipError_t e = hipMalloc(&m->d_verts, verts.size() * sizeof(float));
        if (e == hipSuccess) e = hipMemcpy(m->d_verts, verts.data(), verts.size() * sizeof(float), hipMemcpyHostToDevice);
        if (e != hipSuccess) { hipFree(m->d_verts); delete m; return fail(SFH_ERR_HIP, "mesh upload: %s", hipGetErrorString(e)); }
    }
    if (nf) {
        hipError_t e = hipMalloc(&m->d_faces, faces.size() * sizeof(uint32_t));
        if (e == hipSuccess) e = hipMemcpy(m->d_faces, faces.data(), faces.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMalloc(&m->d_blkwin, m->blkwin.size() * sizeof(uint2));
        if (e == hipSuccess)
            e = hipMemcpy(m->d_blkwin, m->blkwin.data(), m->blkwin.size() * sizeof(uint2), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            hipFree(m->d_verts); hipFree(m->d_faces); hipFree(m->d_blkwin); delete m;
            return fail(SFH_ERR_HIP, "mesh upload: %s", hipGetErrorString(e));
        }
    }
    *out = m;
    return SFH_OK;
}

extern "C" int sfh_mesh_destroy(sfh_mesh m)
{
    if (!m) return fail(SFH_ERR_INVALID, "null mesh");
    mesh_release(m);
    return SFH_OK;
}

extern "C" int sfh_mesh_info(sfh_mesh m, uint32_t* nv, uint32_t* nf, float aabb_min[3], float aabb_max[3])
{
    if (!m) return fail(SFH_ERR_INVALID, "null mesh");
    if (nv) *nv = m->nv;
    if (nf) *nf = m->nf;
    for (int k = 0; k < 3; ++k) {
        if (aabb_min) aabb_min[k] = m->aabb_min[k];
        if (aabb_max) aabb_max[k] = m->aabb_max[k];
    }
    return SFH_OK;
}

// ---- ocean ------------------------------------------------------------------

extern "C" int sfh_ocean_create(double density, const double gravity[3], sfh_ocean* out)
{
    if (!out || !gravity) return fail(SFH_ERR_INVALID, "null argument");
    sfh_ocean_t* o = new sfh_ocean_t();
    o->rho = density;
    for (int k = 0; k < 3; ++k) o->g[k] = gravity[k];
    *out = o;
    return SFH_OK;
}

static void ocean_free_waves(sfh_ocean_t* o)
{
    for (int l = 0; l < kMaxLayers; ++l) {
        if (o->d_quad[l]) hipFree(o->d_quad[l]);
        o->d_quad[l] = nullptr;
    }
    o->nlayers = 0;
    o->n = 0;
}

extern "C" int sfh_ocean_destroy(sfh_ocean o)
{
    if (!o) return fail(SFH_ERR_INVALID, "null ocean");
    ocean_free_waves(o);
    delete o;
    return SFH_OK;
}

extern "C" int sfh_ocean_set_waves(sfh_ocean o, const float* data, int host_data, uint32_t n,
                                   uint32_t channels, uint32_t layers, const float* lengths)
{
    if (!o) return fail(SFH_ERR_INVALID, "null ocean");
    HIPCHK(hipDeviceSynchronize());
    ocean_free_waves(o);
    if (!data || layers == 0) return SFH_OK;
    if (layers > (uint32_t)kMaxLayers || channels < layers || n < 2 || n > 16384 || !lengths)
        return fail(SFH_ERR_INVALID, "bad wave layout (n %u, channels %u, layers %u)", n, channels, layers);
    for (uint32_t l = 0; l < layers; ++l)
        if (!(lengths[l] > 0.f)) return fail(SFH_ERR_INVALID, "layer %u length must be positive", l);
    size_t count = (size_t)n * n * channels;
    float* d_data = nullptr;
    const float* src = data;
    if (host_data) {
        HIPCHK(hipMalloc(&d_data, count * sizeof(float)));
        HIPCHK(hipMemcpy(d_data, data, count * sizeof(float), hipMemcpyHostToDevice));
        src = d_data;
    }
    for (uint32_t l = 0; l < layers; ++l) {
        hipError_t e = hipMalloc(&o->d_quad[l], (size_t)n * n * sizeof(float4));
        if (e != hipSuccess) { hipFree(d_data); ocean_free_waves(o); return fail(SFH_ERR_HIP, "wave alloc: %s", hipGetErrorString(e)); }
        dim3 grid((n + 255) / 256, n);
        hipLaunchKernelGGL(k_wave_quad, grid, dim3(256), 0, 0, src, n, channels, l, o->d_quad[l]);
        o->len[l] = (double)lengths[l];
    }
    hipError_t e = hipDeviceSynchronize();
    if (d_data) hipFree(d_data);
    if (e != hipSuccess) { ocean_free_waves(o); return fail(SFH_ERR_HIP, "wave setup: %s", hipGetErrorString(e)); }
    o->nlayers = (int)layers;
    o->n = n;
    return SFH_OK;
}

extern "C" int sfh_ocean_add_current(sfh_ocean o, int type, int enabled, const double params[12])
{
    if (!o || !params) return fail(SFH_ERR_INVALID, "null argument");
    if (type < 0 || type > 2) return fail(SFH_ERR_UNSUPPORTED, "current type %d (0 uniform, 1 jet, 2 pipe)", type);
    if (type != 0) {
        int nvar = 0;
        for (const Current& x : o->currents) nvar += x.type != 0;
        if (nvar >= kMaxCur) return fail(SFH_ERR_SIZE, "at most %d jets and pipes", kMaxCur);
        const double an = params[3] * params[3] + params[4] * params[4] + params[5] * params[5];
        if (!(fabs(an - 1.0) < 1e-9)) return fail(SFH_ERR_INVALID, "the axis params[3..5] must be a unit vector");
        if (type == 2 && !(params[6] > 0.0)) return fail(SFH_ERR_INVALID, "pipe length must be positive");
    }
    Current c;
    c.type = type;
    c.enabled = enabled;
    memcpy(c.p, params, sizeof(c.p));
    o->currents.push_back(c);
    return SFH_OK;
}

extern "C" int sfh_ocean_enable_currents(sfh_ocean o, int enabled)
{
    if (!o) return fail(SFH_ERR_INVALID, "null ocean");
    o->currents_enabled = enabled != 0;
    return SFH_OK;
}

// ---- batches ----------------------------------------------------------------

static void gl_to_rows(const double* m, double* B, double* t)
{
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) B[3 * r + c] = m[4 * c + r];
    if (t) { t[0] = m[12]; t[1] = m[13]; t[2] = m[14]; }
}

static void batch_free(sfh_batch_t* b)
{
    if (b->graph_exec) hipGraphExecDestroy(b->graph_exec);
    if (b->graph) hipGraphDestroy(b->graph);
    for (int i = 0; i < 4; ++i)
        if (b->ev[i]) hipEventDestroy(b->ev[i]);
    hipFree(b->d_static);
    hipFree(b->d_sidx);
    hipFree(b->d_hstates);
    hipFree(b->d_hout);
    b->d_hstates = b->d_hout = nullptr;
    hipFree(b->d_chunks);
    hipFree(b->d_bpos);
    hipFree(b->d_lists);
    hipFree(b->d_dyn);
    hipFree(b->d_partials);
    hipFree(b->d_comp);
    hipFree(b->d_part_out);
    hipFree(b->d_units);
    b->d_comp = nullptr; b->d_part_out = nullptr; b->d_units = nullptr;
    for (sfh_mesh_t* m : b->meshes) mesh_release(m);
    b->meshes.clear();
}

static uint32_t default_waves()
{
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, dev) == hipSuccess && p.multiProcessorCount > 0) cus = p.multiProcessorCount;
    }
    return (uint32_t)cus * 64u;
}

// Vertex window of the faces [first, first + count) of a mesh from its
// per-64-face block windows: (lowest id, number of ids), or (0, 0) when the
// window is larger than `cap` (the crossing pass then evaluates depths per
// face vertex).
static void chunk_window(const std::vector<uint2>& blk, uint32_t first, uint32_t count, uint32_t cap, uint32_t& vlo,
                         uint32_t& vn)
{
    uint32_t lo = UINT32_MAX, hi = 0;
    for (uint32_t k = first / 64; k < (first + count + 63) / 64 && k < blk.size(); ++k) {
        lo = std::min(lo, blk[k].x);
        hi = std::max(hi, blk[k].y);
    }
    vlo = 0;
    vn = 0;
    if (lo <= hi && hi - lo + 1 <= cap) {
        vlo = lo;
        vn = hi - lo + 1;
    }
}

// (Re)builds the chunks: every body's faces cut into chunks of 64*K
// consecutive faces, each with its vertex window.
static int batch_build_chunks(sfh_batch_t* b)
{
    std::vector<Chunk> chunks;
    const uint32_t csize = 64u * b->K;
    for (uint32_t i = 0; i < b->n; ++i) {
        BodyStatic& s = b->h_static[i];
        const sfh_mesh_t* m = i < b->meshes_by_body.size() ? b->meshes_by_body[i] : nullptr;
        s.chunk0 = (uint32_t)chunks.size();
        s.nchunks = (s.nf + csize - 1) / csize;
        for (uint32_t c = 0; c < s.nchunks; ++c) {
            Chunk ch;
            memset(&ch, 0, sizeof(ch));
            ch.body = i;
            ch.first = c * csize;
            ch.count = std::min(csize, s.nf - c * csize);
            if (m) chunk_window(m->blkwin, ch.first, ch.count, win_cap(b->K), ch.vlo, ch.vn);
            chunks.push_back(ch);
        }
    }
    hipFree(b->d_static);
    hipFree(b->d_sidx);
    hipFree(b->d_chunks);
    hipFree(b->d_partials);
    hipFree(b->d_bpos);
    hipFree(b->d_lists);
    b->d_static = nullptr; b->d_sidx = nullptr; b->d_chunks = nullptr;
    b->d_partials = nullptr; b->d_bpos = nullptr; b->d_lists = nullptr;
    b->nchunks = (uint32_t)chunks.size();
    HIPCHK(hipMalloc(&b->d_static, sizeof(BodyStatic) * std::max<size_t>(1, b->n)));
    HIPCHK(hipMemcpy(b->d_static, b->h_static.data(), sizeof(BodyStatic) * b->n, hipMemcpyHostToDevice));
    {
        std::vector<uint32_t> sx(b->n);
        for (uint32_t i = 0; i < b->n; ++i) sx[i] = b->h_static[i].sidx;
        HIPCHK(hipMalloc(&b->d_sidx, sizeof(uint32_t) * std::max<size_t>(1, b->n)));
        if (b->n) HIPCHK(hipMemcpy(b->d_sidx, sx.data(), sizeof(uint32_t) * b->n, hipMemcpyHostToDevice));
    }
    HIPCHK(hipMalloc(&b->d_chunks, sizeof(Chunk) * std::max<size_t>(1, chunks.size())));
    if (!chunks.empty())
        HIPCHK(hipMemcpy(b->d_chunks, chunks.data(), sizeof(Chunk) * chunks.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&b->d_partials, sizeof(double) * kChunkDoubles * std::max<size_t>(1, chunks.size())));
    HIPCHK(hipMalloc(&b->d_bpos, std::max<size_t>(1, b->n)));
    HIPCHK(hipMalloc(&b->d_lists, sizeof(uint32_t) * ((size_t)b->nchunks + 2)));
    HIPCHK(hipMemset(b->d_lists + b->nchunks, 0, 2 * sizeof(uint32_t)));
    if (b->graph_exec) { hipGraphExecDestroy(b->graph_exec); b->graph_exec = nullptr; }
    if (b->graph) { hipGraphDestroy(b->graph); b->graph = nullptr; }
    return SFH_OK;
}

static int apply_opts(sfh_batch_t* b, const sfh_batch_opts* opts)
{
    uint32_t K = opts && opts->faces_per_lane ? opts->faces_per_lane : 16u;
    if (K != 1 && K != 2 && K != 4 && K != 8 && K != 16 && K != 32)
        return fail(SFH_ERR_INVALID, "faces_per_lane must be 1, 2, 4, 8, 16 or 32");
    if (opts && opts->split_passes > 1) return fail(SFH_ERR_INVALID, "split_passes must be 0 or 1");
    b->waves = opts && opts->waves ? opts->waves : default_waves();
    b->split = opts ? opts->split_passes : 0u;
    if (K != b->K || !b->d_chunks) {
        b->K = K;
        return batch_build_chunks(b);
    }
    return SFH_OK;
}

static BodyStatic make_static(const sfh_mesh_t* m, uint32_t flags)
{
    BodyStatic s;
    memset(&s, 0, sizeof(s));
    s.verts = m ? m->d_verts : nullptr;
    s.faces = m ? m->d_faces : nullptr;
    s.nf = m ? m->nf : 0;
    s.nv = m ? m->nv : 0;
    s.flags = flags;
    for (int i = 0; i < 3; ++i) s.cg2c[4 * i] = s.cg2o[4 * i] = 1.0;
    return s;
}

static UnitStatic make_unit(const sfh_body_desc& d)
{
    UnitStatic u;
    memset(&u, 0, sizeof(u));
    for (int k = 0; k < 3; ++k) {
        u.P_CB[k] = d.P_CB[k];
        u.aabb_c[k] = 0.5 * (d.aabb_max[k] + d.aabb_min[k]);
        u.aabb_h[k] = 0.5 * (d.aabb_max[k] - d.aabb_min[k]);
    }
    u.volume = d.volume;
    u.surface = d.surface;
    u.hydro = d.hydro ? 1u : 0u;
    u.buoyant = d.buoyant ? 1u : 0u;
    return u;
}

static int upload_units(sfh_batch_t* b, const std::vector<UnitStatic>& units)
{
    HIPCHK(hipMalloc(&b->d_units, sizeof(UnitStatic) * std::max<size_t>(1, units.size())));
    if (!units.empty())
        HIPCHK(hipMemcpy(b->d_units, units.data(), sizeof(UnitStatic) * units.size(), hipMemcpyHostToDevice));
    return SFH_OK;
}

extern "C" int sfh_batch_create(const sfh_body_desc* bodies, uint32_t n, const sfh_batch_opts* opts, sfh_batch* out)
{
    if (!out || (!bodies && n)) return fail(SFH_ERR_INVALID, "null argument");
    sfh_batch_t* b = new sfh_batch_t();
    b->n = n;
    hipGetDevice(&b->device);
    b->h_static.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        const sfh_body_desc& d = bodies[i];
        uint32_t flags = 0;
        if (d.hydro) flags |= BF_HYDRO;
        if (d.buoyant) flags |= BF_BUOYANT;
        BodyStatic s = make_static(d.mesh, flags);
        s.sidx = i;
        if (d.mesh) {
            d.mesh->refs.fetch_add(1);
            b->meshes.push_back(d.mesh);
        }
        b->meshes_by_body.push_back(d.mesh);
        gl_to_rows(d.T_CG2C, s.cg2c, s.cg2c + 9);
        gl_to_rows(d.T_CG2O, s.cg2o, nullptr);
        for (int k = 0; k < 3; ++k) {
            s.aabb_c[k] = 0.5 * (d.aabb_max[k] + d.aabb_min[k]);
            s.aabb_h[k] = 0.5 * (d.aabb_max[k] - d.aabb_min[k]);
            s.Cd[k] = d.Cd[k];
            s.Cf[k] = d.Cf[k];
            s.P_CB[k] = d.P_CB[k];
        }
        s.volume = d.volume;
        s.surface = d.surface;
        b->h_static[i] = s;
        b->nfaces += s.nf;
    }
    b->K = 0;
    b->nunits = n;
    std::vector<UnitStatic> units(n);
    for (uint32_t i = 0; i < n; ++i) units[i] = make_unit(bodies[i]);
    int rc = hipMalloc(&b->d_dyn, sizeof(BodyDyn) * std::max<size_t>(1, n)) == hipSuccess ? SFH_OK
             : fail(SFH_ERR_HIP, "batch alloc failed");
    if (rc == SFH_OK) rc = upload_units(b, units);
    if (rc == SFH_OK) rc = apply_opts(b, opts);
    if (rc != SFH_OK) { batch_free(b); delete b; return rc; }
    *out = b;
    return SFH_OK;
}

static void gl_mul_host(const double* a, const double* c, double* r)   // OpenGL column-major 4x4: r = a c
{
    double t[16];
    for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 4; ++i)
            t[4 * j + i] = a[i] * c[4 * j] + a[4 + i] * c[4 * j + 1] + a[8 + i] * c[4 * j + 2] + a[12 + i] * c[4 * j + 3];
    memcpy(r, t, sizeof(t));
}

extern "C" int sfh_compound_batch_create(const sfh_compound_desc* comps, uint32_t n, const sfh_batch_opts* opts,
                                         sfh_batch* out)
{
    if (!out || (!comps && n)) return fail(SFH_ERR_INVALID, "null argument");
    uint64_t nparts = 0;
    for (uint32_t c = 0; c < n; ++c) {
        if (comps[c].nparts == 0 || !comps[c].parts) return fail(SFH_ERR_INVALID, "compound %u has no parts", c);
        for (uint32_t p = 0; p < comps[c].nparts; ++p)
            if (!comps[c].parts[p].mesh) return fail(SFH_ERR_INVALID, "compound %u part %u has no mesh", c, p);
        nparts += comps[c].nparts;
    }
    if (nparts > (1u << 26)) return fail(SFH_ERR_SIZE, "too many parts");
    sfh_batch_t* b = new sfh_batch_t();
    b->n = (uint32_t)nparts;
    b->nunits = n;
    b->compound = true;
    hipGetDevice(&b->device);
    std::vector<CompStatic> cs(n);
    std::vector<UnitStatic> units(n);
    for (uint32_t c = 0; c < n; ++c) {
        const sfh_compound_desc& C = comps[c];
        CompStatic k;
        memset(&k, 0, sizeof(k));
        k.part0 = (uint32_t)b->h_static.size();
        k.nparts = C.nparts;
        k.flags = C.body.buoyant ? BF_BUOYANT : 0u;
        k.volume = C.body.volume;
        k.surface = C.body.surface;
        for (int i = 0; i < 3; ++i) k.P_CB[i] = C.body.P_CB[i];
        cs[c] = k;
        units[c] = make_unit(C.body);
        for (uint32_t p = 0; p < C.nparts; ++p) {
            const sfh_part_desc& P = C.parts[p];
            uint32_t flags = BF_PART;
            if (C.body.hydro) flags |= BF_HYDRO;
            if (P.external) flags |= BF_EXTERNAL;
            if (P.buoyant) flags |= BF_BUOYANT;
            if (P.hydro) flags |= BF_PHYDRO;
            BodyStatic s = make_static(P.mesh, flags);
            s.sidx = c;
            P.mesh->refs.fetch_add(1);
            b->meshes.push_back(P.mesh);
            b->meshes_by_body.push_back(P.mesh);
            // T_C of the part = T_O * origin * O2C, T_O of the part = T_O * origin,
            // with T_O = T_CG * T_CG2O (Compound.cpp:348-349, 408-409)
            double o2p[16], cg2c[16];
            gl_mul_host(C.body.T_CG2O, P.origin, o2p);
            gl_mul_host(o2p, P.O2C, cg2c);
            gl_to_rows(cg2c, s.cg2c, s.cg2c + 9);
            gl_to_rows(o2p, s.cg2o, nullptr);
            for (int i = 0; i < 3; ++i) {
                s.aabb_c[i] = 0.5 * (C.body.aabb_max[i] + C.body.aabb_min[i]);
                s.aabb_h[i] = 0.5 * (C.body.aabb_max[i] - C.body.aabb_min[i]);
                s.Cd[i] = P.Cd[i];
                s.Cf[i] = P.Cf[i];
            }
            b->h_static.push_back(s);
            b->nfaces += s.nf;
        }
    }
    b->K = 0;
    int rc = SFH_OK;
    if (hipMalloc(&b->d_dyn, sizeof(BodyDyn) * std::max<uint64_t>(1, nparts)) != hipSuccess ||
        hipMalloc(&b->d_comp, sizeof(CompStatic) * std::max<uint32_t>(1, n)) != hipSuccess ||
        hipMalloc(&b->d_part_out, sizeof(double) * SFH_OUT_DOUBLES * std::max<uint64_t>(1, nparts)) != hipSuccess)
        rc = fail(SFH_ERR_HIP, "compound batch alloc failed");
    if (rc == SFH_OK && n &&
        hipMemcpy(b->d_comp, cs.data(), sizeof(CompStatic) * n, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(SFH_ERR_HIP, "compound upload failed");
    if (rc == SFH_OK) rc = upload_units(b, units);
    if (rc == SFH_OK) rc = apply_opts(b, opts);
    if (rc != SFH_OK) { batch_free(b); delete b; return rc; }
    *out = b;
    return SFH_OK;
}

// Rows of the inverse of a GL matrix's basis, by cofactors (btMatrix3x3::inverse,
// btMatrix3x3.h:1093-1103).
static bool basis_inverse_rows(const double* gl, double* M)
{
    double B[9];
    gl_to_rows(gl, B, nullptr);
#define CF(r1, c1, r2, c2) (B[3 * (r1) + (c1)] * B[3 * (r2) + (c2)] - B[3 * (r1) + (c2)] * B[3 * (r2) + (c1)])
    const double co0 = CF(1, 1, 2, 2), co1 = CF(1, 2, 2, 0), co2 = CF(1, 0, 2, 1);
    const double det = B[0] * co0 + B[1] * co1 + B[2] * co2;
    if (!(det != 0.0) || !std::isfinite(det)) return false;
    const double s = 1.0 / det;
    const double R[9] = {co0 * s, CF(0, 2, 2, 1) * s, CF(0, 1, 1, 2) * s,
                         co1 * s, CF(0, 0, 2, 2) * s, CF(0, 2, 1, 0) * s,
                         co2 * s, CF(0, 1, 2, 0) * s, CF(0, 0, 1, 1) * s};
#undef CF
    memcpy(M, R, sizeof(R));
    return true;
}

// Flags and correction data of an aerodynamic body or part: the proxy type in
// the flags, fdApproxParams in Cd, the inverse proxy basis in cg2o.
static int aero_static(BodyStatic& s, int32_t type, const double* params, const double* T_CG2H)
{
    if (type < SFH_APPROX_AUTO || type > SFH_APPROX_ELLIPSOID) return fail(SFH_ERR_INVALID, "bad approx_type");
    if (!basis_inverse_rows(T_CG2H, s.cg2o)) return fail(SFH_ERR_INVALID, "T_CG2H basis is singular");
    for (int k = 0; k < 3; ++k) s.Cd[k] = params[k];
    s.flags |= BF_MODE_AERO | BF_AERO_CORRECT | ((uint32_t)type << BF_APPROX_SHIFT);
    return SFH_OK;
}

extern "C" int sfh_aero_batch_create(const sfh_aero_desc* bodies, uint32_t n, const sfh_batch_opts* opts,
                                     sfh_batch* out)
{
    if (!out || (!bodies && n)) return fail(SFH_ERR_INVALID, "null argument");
    sfh_batch_t* b = new sfh_batch_t();
    b->n = n;
    b->nunits = n;
    b->aero = true;
    hipGetDevice(&b->device);
    b->h_static.resize(n);
    std::vector<UnitStatic> units(n);
    int rc = SFH_OK;
    for (uint32_t i = 0; i < n && rc == SFH_OK; ++i) {
        const sfh_aero_desc& d = bodies[i];
        BodyStatic s = make_static(d.mesh, 0u);
        s.sidx = i;
        if (d.mesh) {
            d.mesh->refs.fetch_add(1);
            b->meshes.push_back(d.mesh);
        }
        b->meshes_by_body.push_back(d.mesh);
        gl_to_rows(d.T_CG2C, s.cg2c, s.cg2c + 9);
        if (d.aero) rc = aero_static(s, d.approx_type, d.approx_params, d.T_CG2H);
        b->h_static[i] = s;
        b->nfaces += s.nf;
    }
    if (rc == SFH_OK && hipMalloc(&b->d_dyn, sizeof(BodyDyn) * std::max<size_t>(1, n)) != hipSuccess)
        rc = fail(SFH_ERR_HIP, "batch alloc failed");
    if (rc == SFH_OK) rc = upload_units(b, units);
    if (rc == SFH_OK) rc = apply_opts(b, opts);
    if (rc != SFH_OK) { batch_free(b); delete b; return rc; }
    *out = b;
    return SFH_OK;
}

extern "C" int sfh_aero_compound_batch_create(const sfh_aero_compound_desc* comps, uint32_t n,
                                              const sfh_batch_opts* opts, sfh_batch* out)
{
    if (!out || (!comps && n)) return fail(SFH_ERR_INVALID, "null argument");
    uint64_t nparts = 0;
    for (uint32_t c = 0; c < n; ++c) {
        if (comps[c].nparts == 0 || !comps[c].parts) return fail(SFH_ERR_INVALID, "compound %u has no parts", c);
        for (uint32_t p = 0; p < comps[c].nparts; ++p)
            if (!comps[c].parts[p].mesh) return fail(SFH_ERR_INVALID, "compound %u part %u has no mesh", c, p);
        nparts += comps[c].nparts;
    }
    if (nparts > (1u << 26)) return fail(SFH_ERR_SIZE, "too many parts");
    sfh_batch_t* b = new sfh_batch_t();
    b->n = (uint32_t)nparts;
    b->nunits = n;
    b->compound = true;
    b->aero = true;
    hipGetDevice(&b->device);
    std::vector<CompStatic> cs(n);
    std::vector<UnitStatic> units(n);
    int rc = SFH_OK;
    for (uint32_t c = 0; c < n; ++c) {
        const sfh_aero_compound_desc& C = comps[c];
        CompStatic k;
        memset(&k, 0, sizeof(k));
        k.part0 = (uint32_t)b->h_static.size();
        k.nparts = C.nparts;
        k.flags = C.aero ? BF_MODE_AERO : 0u;
        cs[c] = k;
        for (uint32_t p = 0; p < C.nparts; ++p) {
            const sfh_aero_part_desc& P = C.parts[p];
            // part rule of the device prologue: a part runs when inside and
            // external with BF_PHYDRO, so BF_PHYDRO marks every aerodynamic part
            BodyStatic s = make_static(P.mesh, BF_PART | (P.external ? BF_EXTERNAL : 0u));
            s.sidx = c;
            P.mesh->refs.fetch_add(1);
            b->meshes.push_back(P.mesh);
            b->meshes_by_body.push_back(P.mesh);
            // T_C of the part = T_CG * T_CG2O * origin * O2C (Compound.cpp:458)
            double o2p[16], cg2c[16];
            gl_mul_host(C.T_CG2O, P.origin, o2p);
            gl_mul_host(o2p, P.O2C, cg2c);
            gl_to_rows(cg2c, s.cg2c, s.cg2c + 9);
            if (C.aero && rc == SFH_OK) {
                s.flags |= BF_PHYDRO;
                rc = aero_static(s, P.approx_type, P.approx_params, P.T_CG2H);
            }
            b->h_static.push_back(s);
            b->nfaces += s.nf;
        }
    }
    if (rc == SFH_OK &&
        (hipMalloc(&b->d_dyn, sizeof(BodyDyn) * std::max<uint64_t>(1, nparts)) != hipSuccess ||
         hipMalloc(&b->d_comp, sizeof(CompStatic) * std::max<uint32_t>(1, n)) != hipSuccess ||
         hipMalloc(&b->d_part_out, sizeof(double) * SFH_OUT_DOUBLES * std::max<uint64_t>(1, nparts)) != hipSuccess))
        rc = fail(SFH_ERR_HIP, "compound batch alloc failed");
    if (rc == SFH_OK && n &&
        hipMemcpy(b->d_comp, cs.data(), sizeof(CompStatic) * n, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(SFH_ERR_HIP, "compound upload failed");
    if (rc == SFH_OK) rc = upload_units(b, units);
    if (rc == SFH_OK) rc = apply_opts(b, opts);
    if (rc != SFH_OK) { batch_free(b); delete b; return rc; }
    *out = b;
    return SFH_OK;
}

extern "C" int sfh_batch_destroy(sfh_batch b)
{
    if (!b) return fail(SFH_ERR_INVALID, "null batch");
    hipDeviceSynchronize();
    batch_free(b);
    delete b;
    return SFH_OK;
}

extern "C" int sfh_batch_set_opts(sfh_batch b, const sfh_batch_opts* opts)
{
    if (!b) return fail(SFH_ERR_INVALID, "null batch");
    return apply_opts(b, opts);
}

extern "C" int sfh_batch_info(sfh_batch b, uint32_t* nbodies, uint32_t* nchunks, uint64_t* nfaces)
{
    if (!b) return fail(SFH_ERR_INVALID, "null batch");
    if (nbodies) *nbodies = b->nunits;
    if (nchunks) *nchunks = b->nchunks;
    if (nfaces) *nfaces = b->nfaces;
    return SFH_OK;
}

// Launches the face pass over every chunk of the batch.
static int launch_faces(sfh_batch_t* b, const OceanK& oc, hipStream_t s)
{
    if (b->nchunks == 0) return SFH_OK;
    const uint32_t waves = std::min<uint32_t>(std::max<uint32_t>(b->waves, 1u), b->nchunks);   // grid-strided
    const uint32_t blocks = (waves + 3) / 4;
#define LAUNCH_F(KK, NU, NL, PASS)                                                                              \
    {                                                                                                           \
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_faces<KK, NU, NL, PASS>), dim3(blocks), dim3(256), 0, s,          \
                           b->d_lists, b->d_lists + b->nchunks, b->nchunks, b->d_chunks, b->d_static, b->d_dyn, \
                           oc, b->d_partials);                                                                  \
    }
#define LAUNCH_X(KK, NU, NL)                                                                                    \
    {                                                                                                           \
        if (b->split) {                                                                                         \
            LAUNCH_F(KK, NU, 0, kPassSub)                                                                       \
            LAUNCH_F(KK, NU, NL, kPassCross)                                                                    \
        } else {                                                                                                \
            LAUNCH_F(KK, NU, NL, kPassBoth)                                                                     \
        }                                                                                                       \
    }
#define LAUNCH_K(KK)                                                                                            \
    case KK:                                                                                                    \
        if (b->aero) {                                                                                          \
            if (oc.ncur) LAUNCH_F(KK, true, 0, kPassAero) else LAUNCH_F(KK, false, 0, kPassAero)                \
        } else if (oc.ncur) LAUNCH_X(KK, true, -1)                                                              \
        else if (oc.nlayers > 2 || (oc.nlayers > 0 && oc.nmask < 0)) LAUNCH_X(KK, false, -1)                   \
        else if (oc.nlayers == 0) LAUNCH_X(KK, false, 0)                                                        \
        else if (oc.nlayers == 1) LAUNCH_X(KK, false, 1)                                                        \
        else LAUNCH_X(KK, false, 2)                                                                             \
        break;
    switch (b->K) {
        LAUNCH_K(1) LAUNCH_K(2) LAUNCH_K(4) LAUNCH_K(8) LAUNCH_K(16) LAUNCH_K(32)
        default: return fail(SFH_ERR_INVALID, "bad K");
    }
#undef LAUNCH_K
#undef LAUNCH_X
#undef LAUNCH_F
    return SFH_OK;
}

// One evaluation: the prologue, the face pass, the epilogue and, for a
// compound batch, the composition of the parts' rows.  timed: events before
// the prologue (ev[0]), around the face launches (ev[1], ev[2]) and after the
// last launch (ev[3]).
static int batch_launch(sfh_batch_t* b, const sfh_ocean_t* o, const double* d_states, int kind, double* d_out,
                        hipStream_t s, bool timed, const uint8_t* active = nullptr, bool skip_prologue = false)
{
    if (b->n == 0) return SFH_OK;
    // every launch of the step runs or none does: the epilogue zeroes the
    // counts the prologue's list build adds to
    if (b->K != 1 && b->K != 2 && b->K != 4 && b->K != 8 && b->K != 16 && b->K != 32)
        return fail(SFH_ERR_INVALID, "bad K");
    OceanK oc = make_oceank(o);
    double* rows = b->compound ? b->d_part_out : d_out;
    if (timed) HIPCHK(hipEventRecord(b->ev[0], s));
    if (!skip_prologue) {
        const bool large = !active && b->n >= kPrologueLargeBodies;
        const uint32_t blk = large ? kPrologueBlockLarge : kPrologueBlock;
        const uint32_t blocks = (uint32_t)(((uint64_t)b->n * kPrologueLanes + blk - 1) / blk);
        if (large)
            hipLaunchKernelGGL(k_prologue<kPrologueBlockLarge>, dim3(blocks), dim3(blk), 0, s, b->d_static, d_states,
                               kind, b->n, oc, active, b->d_sidx, b->d_dyn, b->d_bpos, rows, b->d_lists, b->nchunks,
                               b->d_lists + b->nchunks);
        else
            hipLaunchKernelGGL(k_prologue<kPrologueBlock>, dim3(blocks), dim3(blk), 0, s, b->d_static, d_states, kind,
                               b->n, oc, active, b->d_sidx, b->d_dyn, b->d_bpos, rows, b->d_lists, b->nchunks,
                               b->d_lists + b->nchunks);
    }
    if (timed) HIPCHK(hipEventRecord(b->ev[1], s));
    int rc = launch_faces(b, oc, s);
    if (rc != SFH_OK) return rc;
    if (timed) HIPCHK(hipEventRecord(b->ev[2], s));
    if (b->nchunks) {
        const uint32_t ewaves = std::min<uint32_t>(b->n, std::max<uint32_t>(1, b->waves / 2));
        const uint32_t eblocks = (ewaves + kEpilogueWaves - 1) / kEpilogueWaves;
        hipLaunchKernelGGL(k_epilogue, dim3(eblocks), dim3(64 * kEpilogueWaves), 0, s, b->d_static, b->d_dyn, b->d_bpos, b->n, oc,
                           b->d_partials, rows, b->d_lists + b->nchunks);
    }
    if (b->compound && b->nunits)
        hipLaunchKernelGGL(k_compound, dim3((b->nunits + 127) / 128), dim3(128), 0, s, b->d_comp, b->d_static,
                           b->d_dyn, b->nunits, oc, active, b->d_part_out, d_out);
    if (timed) HIPCHK(hipEventRecord(b->ev[3], s));
    HIPCHK(hipGetLastError());
    return SFH_OK;
}

static int check_ocean(const sfh_ocean_t* o)
{
    if (!o) return fail(SFH_ERR_INVALID, "null ocean");
    return SFH_OK;
}

static void harvest_timing(sfh_batch_t* b)
{
    if (!b->ev_pending) return;
    float t_face = 0.f, t_tot = 0.f;
    if (hipEventSynchronize(b->ev[3]) == hipSuccess &&
        hipEventElapsedTime(&t_face, b->ev[1], b->ev[2]) == hipSuccess &&
        hipEventElapsedTime(&t_tot, b->ev[0], b->ev[3]) == hipSuccess) {
        b->face_ms_sum += t_face;
        b->total_ms_sum += t_tot;
        b->timed++;
    }
    b->ev_pending = false;
}

extern "C" int sfh_batch_compute(sfh_batch b, sfh_ocean o, const double* d_states, double* d_out, void* stream)
{
    if (!b || !d_states || !d_out) return fail(SFH_ERR_INVALID, "null argument");
    int rc = check_ocean(o);
    if (rc != SFH_OK) return rc;
    hipStream_t s = (hipStream_t)stream;
    bool timed = b->timing;
    if (timed) harvest_timing(b);
    rc = batch_launch(b, o, d_states, 0, d_out, s, timed);
    if (rc == SFH_OK && timed) b->ev_pending = true;
    return rc;
}

extern "C" int sfh_batch_compute_masked(sfh_batch b, sfh_ocean o, const double* d_states, const uint8_t* d_active,
                                        double* d_out, void* stream)
{
    if (!b || !d_states || !d_out) return fail(SFH_ERR_INVALID, "null argument");
    int rc = check_ocean(o);
    if (rc != SFH_OK) return rc;
    hipStream_t s = (hipStream_t)stream;
    bool timed = b->timing;
    if (timed) harvest_timing(b);
    rc = batch_launch(b, o, d_states, 0, d_out, s, timed, d_active);
    if (rc == SFH_OK && timed) b->ev_pending = true;
    return rc;
}

// ---- autotuner ------------------------------------------------------------------

// Times the batch's whole step on its own inputs for each (faces per lane,
// resident wavefronts per CU, one or two face launches) candidate and keeps
// the fastest.  The step is timed with HIP events on the caller's stream
// around `reps` eager launches after one untimed launch; the time of the face
// launches of the last launch comes from the batch's own event pair.  A batch's chunks depend on the faces per
// lane, so a candidate with another value rebuilds them (and any graph the
// caller captured from the batch before must be captured again).
extern "C" int sfh_batch_autotune(sfh_batch b, sfh_ocean o, const double* d_states, const uint8_t* d_active,
                                  double* d_out, void* stream, uint32_t reps, sfh_tune_point* points,
                                  uint32_t max_points, uint32_t* npoints)
{
    if (!b || !d_states || !d_out) return fail(SFH_ERR_INVALID, "null argument");
    int rc = check_ocean(o);
    if (rc != SFH_OK) return rc;
    if (reps == 0) reps = 5;
    hipStream_t s = (hipStream_t)stream;
    int dev = 0, cus = 256;
    HIPCHK(hipGetDevice(&dev));
    {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, dev) == hipSuccess && p.multiProcessorCount > 0) cus = p.multiProcessorCount;
    }
    static const uint32_t kK[] = {4, 8, 16, 32};
    static const uint32_t kWpc[] = {8, 16, 24, 32, 48, 64, 96, 128, 192, 256};
    const uint32_t nsplit = b->aero ? 1u : 2u;   // an aerodynamic batch has one loop
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    const bool timing = b->timing;
    if (!b->ev[0])
        for (int i = 0; i < 4; ++i) HIPCHK(hipEventCreate(&b->ev[i]));
    double best = 1e30;
    sfh_batch_opts best_opts = {b->K, b->waves, b->split};
    uint32_t np = 0;
    for (uint32_t K : kK) {
        for (uint32_t wpc : kWpc) {
          for (uint32_t split = 0; split < nsplit && rc == SFH_OK; ++split) {
            sfh_batch_opts opts = {K, wpc * (uint32_t)cus, split};
            rc = apply_opts(b, &opts);
            if (rc != SFH_OK) break;
            rc = batch_launch(b, o, d_states, 0, d_out, s, false, d_active);   // warm-up
            if (rc != SFH_OK) break;
            hipEventRecord(e0, s);
            for (uint32_t r = 0; r < reps && rc == SFH_OK; ++r)
                rc = batch_launch(b, o, d_states, 0, d_out, s, r + 1 == reps, d_active);
            hipEventRecord(e1, s);
            if (rc != SFH_OK) break;
            hipError_t e = hipEventSynchronize(e1);
            float ms = 0.f, face = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (e == hipSuccess) e = hipEventElapsedTime(&face, b->ev[1], b->ev[2]);
            if (e != hipSuccess) { rc = fail(SFH_ERR_HIP, "autotune timing: %s", hipGetErrorString(e)); break; }
            const double step = (double)ms / reps;
            if (points && np < max_points) {
                points[np].faces_per_lane = K;
                points[np].waves_per_cu = wpc;
                points[np].split_passes = split;
                points[np].pad_ = 0;
                points[np].step_ms = step;
                points[np].face_ms = face;
            }
            ++np;
            if (step < best) {
                best = step;
                best_opts = opts;
            }
          }
          if (rc != SFH_OK) break;
        }
        if (rc != SFH_OK) break;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    b->ev_pending = false;
    b->timing = timing;
    const int rc2 = apply_opts(b, &best_opts);
    if (npoints) *npoints = std::min(np, max_points);
    return rc != SFH_OK ? rc : rc2;
}

extern "C" int sfh_batch_get_opts(sfh_batch b, sfh_batch_opts* opts)
{
    if (!b || !opts) return fail(SFH_ERR_INVALID, "null argument");
    opts->faces_per_lane = b->K;
    opts->waves = b->waves;
    opts->split_passes = b->split;
    return SFH_OK;
}

// ---- tick driver ---------------------------------------------------------------

// ApplyHydrodynamicForces (SolidEntity.cpp:1958-1962) / ApplyAerodynamicForces
// (1964-1968) of every row the fluid overlaps this tick: the force and torque
// the entity hands to Bullet from the members the last recompute left.
__global__ __launch_bounds__(128) void k_tick_apply(const double* __restrict__ out, const uint8_t* __restrict__ overlap,
                                                    uint32_t n, int aero, double* __restrict__ applied)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* o = out + (size_t)i * SFH_OUT_DOUBLES;
    double* a = applied + (size_t)i * 6;
    if (overlap && !overlap[i]) {
        for (int k = 0; k < 6; ++k) a[k] = 0.0;
        return;
    }
    for (int k = 0; k < 3; ++k) {
        a[k] = aero ? o[6 + k] : (o[k] + o[6 + k]) + o[12 + k];
        a[3 + k] = aero ? o[9 + k] : (o[3 + k] + o[9 + k]) + o[15 + k];
    }
}

struct sfh_tick_t {
    sfh_batch batch = nullptr;
    uint32_t prescaler = 1;
    uint64_t counter = 0;
};

extern "C" int sfh_tick_create(sfh_batch batch, uint32_t prescaler, sfh_tick* out)
{
    if (!batch || !out) return fail(SFH_ERR_INVALID, "null argument");
    if (prescaler == 0) return fail(SFH_ERR_INVALID, "prescaler must be at least 1");
    sfh_tick_t* t = new sfh_tick_t();
    t->batch = batch;
    t->prescaler = prescaler;
    *out = t;
    return SFH_OK;
}

extern "C" int sfh_tick_step(sfh_tick t, sfh_ocean fluid, const double* d_states, const uint8_t* d_overlap,
                             double* d_out, double* d_applied, int* recomputed, void* stream)
{
    if (!t || !d_states || !d_out || !d_applied) return fail(SFH_ERR_INVALID, "null argument");
    int rc = check_ocean(fluid);
    if (rc != SFH_OK) return rc;
    // recompute = fdCounter % fdPrescaler == 0; ++fdCounter (SimulationManager.cpp:1584-1585)
    const bool recompute = t->counter % t->prescaler == 0;
    ++t->counter;
    hipStream_t s = (hipStream_t)stream;
    if (recompute) {
        rc = sfh_batch_compute_masked(t->batch, fluid, d_states, d_overlap, d_out, stream);
        if (rc != SFH_OK) return rc;
    }
    const uint32_t n = t->batch->nunits;
    if (n) hipLaunchKernelGGL(k_tick_apply, dim3((n + 127) / 128), dim3(128), 0, s, d_out, d_overlap, n,
                              t->batch->aero ? 1 : 0, d_applied);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(SFH_ERR_HIP, "tick apply: %s", hipGetErrorString(e));
    if (recomputed) *recomputed = recompute ? 1 : 0;
    return SFH_OK;
}

extern "C" int sfh_tick_destroy(sfh_tick t)
{
    if (!t) return fail(SFH_ERR_INVALID, "null argument");
    delete t;
    return SFH_OK;
}

extern "C" int sfh_batch_compute_host(sfh_batch b, sfh_ocean o, const double* states, double* out)
{
    if (!b || !states || !out) return fail(SFH_ERR_INVALID, "null argument");
    size_t ns = (size_t)b->nunits * SFH_STATE_DOUBLES, no = (size_t)b->nunits * SFH_OUT_DOUBLES;
    if (!b->d_hstates) {   // the device copies live with the batch (sizes fixed per batch)
        HIPCHK(hipMalloc(&b->d_hstates, sizeof(double) * std::max<size_t>(1, ns)));
        hipError_t ea = hipMalloc(&b->d_hout, sizeof(double) * std::max<size_t>(1, no));
        if (ea != hipSuccess) {
            hipFree(b->d_hstates);
            b->d_hstates = nullptr;
            return fail(SFH_ERR_HIP, "alloc: %s", hipGetErrorString(ea));
        }
    }
    int rc = SFH_OK;
    hipError_t e = hipMemcpy(b->d_hstates, states, sizeof(double) * ns, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b->d_hout, out, sizeof(double) * no, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        rc = sfh_batch_compute(b, o, b->d_hstates, b->d_hout, nullptr);
        if (rc == SFH_OK) e = hipMemcpy(out, b->d_hout, sizeof(double) * no, hipMemcpyDeviceToHost);
    }
    if (e != hipSuccess) return fail(SFH_ERR_HIP, "batch host compute: %s", hipGetErrorString(e));
    return rc;
}

extern "C" int sfh_batch_capture(sfh_batch b, sfh_ocean o, const double* d_states, double* d_out, void* stream)
{
    if (!b || !d_states || !d_out) return fail(SFH_ERR_INVALID, "null argument");
    int rc = check_ocean(o);
    if (rc != SFH_OK) return rc;
    hipStream_t s = (hipStream_t)stream;
    if (!s) return fail(SFH_ERR_INVALID, "graph capture needs a non-default stream");
    if (b->graph_exec) { hipGraphExecDestroy(b->graph_exec); b->graph_exec = nullptr; }
    if (b->graph) { hipGraphDestroy(b->graph); b->graph = nullptr; }
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    rc = batch_launch(b, o, d_states, 0, d_out, s, false);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc != SFH_OK) { if (g) hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return fail(SFH_ERR_HIP, "end capture: %s", hipGetErrorString(e));
    b->graph = g;
    HIPCHK(hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0));
    return SFH_OK;
}

extern "C" int sfh_batch_replay(sfh_batch b, void* stream)
{
    if (!b || !b->graph_exec) return fail(SFH_ERR_INVALID, "no captured graph");
    HIPCHK(hipGraphLaunch(b->graph_exec, (hipStream_t)stream));
    return SFH_OK;
}

extern "C" int sfh_batch_timing(sfh_batch b, int enable)
{
    if (!b) return fail(SFH_ERR_INVALID, "null batch");
    if (enable && !b->ev[0])
        for (int i = 0; i < 4; ++i) HIPCHK(hipEventCreate(&b->ev[i]));
    if (enable) harvest_timing(b);
    b->timing = enable != 0;
    b->face_ms_sum = b->total_ms_sum = 0.0;
    b->timed = 0;
    b->ev_pending = false;
    return SFH_OK;
}

extern "C" int sfh_batch_kernel_ms(sfh_batch b, double* face_ms, double* total_ms, uint32_t* count)
{
    if (!b) return fail(SFH_ERR_INVALID, "null batch");
    harvest_timing(b);
    if (count) *count = b->timed;
    if (face_ms) *face_ms = b->timed ? b->face_ms_sum / b->timed : 0.0;
    if (total_ms) *total_ms = b->timed ? b->total_ms_sum / b->timed : 0.0;
    return SFH_OK;
}

// ---- per-body evaluations ------------------------------------------------------

// One per-body evaluation (a single-body batch) needs, beyond the mesh: the
// body's static record and per-step constants, its chunks with their vertex
// windows, the work list and the chunk partials.  They are
// carved out of one device buffer sized for meshes of up to F faces: a
// workspace of the per-body statics (device memory from hipMalloc, grown on
// demand) or the no-alloc entry point's file-scope array.  One kernel writes
// the static record and the chunks from its arguments and the mesh's block
// windows and, in its first workgroup, runs the body's prologue (which gives
// the body the position its call names: the Surface call the crossing loop,
// the Submerged and aerodynamic calls the submerged one); the step is that
// kernel, the face pass and the epilogue.
static constexpr uint32_t kSingleK = 8;   // faces per lane of per-body calls

struct SingleHead {
    BodyStatic st;
    BodyDyn dyn;
    uint32_t sidx;
    int8_t bpos;
    double state[38];   // T_CG, T_C, v, omega of a call whose state came by value
};

// A per-body call's state passed by value to the setup kernel (no copy
// command of its own).
struct StateArg {
    double v[38];
};

__host__ __device__ constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
__host__ __device__ constexpr uint32_t single_nchunks(uint32_t F) { return (F + 64u * kSingleK - 1) / (64u * kSingleK); }

// Byte offsets of the parts in a buffer for F faces: head, work list (with
// its counts), chunks, partials, end.
struct SingleLayout {
    size_t lists, chunks, partials, end;
};

__host__ __device__ constexpr SingleLayout single_layout(uint32_t F)
{
    const size_t nc = single_nchunks(F);
    const size_t lists = align256(sizeof(SingleHead));
    const size_t chunks = align256(lists + sizeof(uint32_t) * (nc + 2));
    const size_t partials = align256(chunks + sizeof(Chunk) * (nc ? nc : 1));
    const size_t end = align256(partials + sizeof(double) * kChunkDoubles * (nc ? nc : 1));
    return SingleLayout{lists, chunks, partials, end};
}

// Writes the body's static record and its chunks with their vertex windows
// (from the mesh's per-64-face block windows), and zeroes the work list's
// counts (the buffer may hold another call's).
// The setup of a per-body call and, in its first workgroup once the record,
// the state and the counts are written, the body's prologue (one launch
// instead of two; the batch's own prologue launch is skipped).
__global__ __launch_bounds__(kPrologueBlock) void k_single_setup(SingleHead* __restrict__ H, Chunk* __restrict__ chunks,
                                                      uint32_t* __restrict__ counts, const uint2* __restrict__ blkwin,
                                                      BodyStatic st, StateArg sa, int with_state,
                                                      const double* __restrict__ states, OceanK oc,
                                                      double* __restrict__ out, uint32_t* __restrict__ list)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t csize = 64u * kSingleK;
    const uint32_t nc = st.nchunks;
    if (with_state && t < 38) H->state[t] = sa.v[t];
    if (t == 0) {
        H->st = st;
        H->sidx = 0;
        counts[0] = 0u;
        counts[1] = 0u;
    }
    if (t < nc) {
        Chunk c;
        c.body = 0;
        c.first = t * csize;
        c.count = min(csize, st.nf - t * csize);
        uint32_t lo = 0xffffffffu, hi = 0;
        const uint32_t nblk = (st.nf + 63) / 64;
        for (uint32_t k = t * kSingleK; k < min((t + 1) * kSingleK, nblk); ++k) {
            const uint2 w = blkwin[k];
            lo = min(lo, w.x);
            hi = max(hi, w.y);
        }
        const bool fits = lo <= hi && hi - lo + 1 <= win_cap(kSingleK);
        c.vlo = fits ? lo : 0u;
        c.vn = fits ? hi - lo + 1 : 0u;
        c.pad[0] = c.pad[1] = c.pad[2] = 0;
        chunks[t] = c;
    }
    if (blockIdx.x == 0) {
        __syncthreads();   // the record, the state and the zeroed counts are this workgroup's writes
        prologue_group<kPrologueBlock>(&H->st, states, 1, 1u, oc, nullptr, &H->sidx, &H->dyn, &H->bpos, out, list,
                                       nc, counts, threadIdx.x);
    }
}

// One evaluation of `m` with `flags` from the 38-double state d_state into
// the output row d_out, on stream s, in the buffer `buf` laid out for the
// mesh's face count.  No allocation.
// d_state: the state in device memory, or nullptr with h_state: the state's
// 38 doubles on the host, passed to the setup kernel by value.
static int single_launch(unsigned char* buf, const sfh_mesh_t* m, uint32_t flags, const sfh_ocean_t* o,
                         const double* d_state, double* d_out, hipStream_t s, uint32_t waves,
                         const double* h_state = nullptr)
{
    const SingleLayout L = single_layout(m->nf);
    SingleHead* H = (SingleHead*)buf;
    Chunk* chunks = (Chunk*)(buf + L.chunks);
    uint32_t* lists = (uint32_t*)(buf + L.lists);
    BodyStatic st = make_static(m, flags);
    st.chunk0 = 0;
    st.nchunks = single_nchunks(m->nf);
    st.sidx = 0;
    const uint32_t nt = std::max(st.nchunks, 38u);
    StateArg sa;
    const int with_state = d_state == nullptr;
    if (with_state) {
        memcpy(sa.v, h_state, sizeof(sa.v));
        d_state = H->state;
    } else {
        memset(sa.v, 0, sizeof(sa.v));
    }
    hipLaunchKernelGGL(k_single_setup, dim3((nt + kPrologueBlock - 1) / kPrologueBlock), dim3(kPrologueBlock), 0, s,
                       H, chunks, lists + st.nchunks, m->d_blkwin, st, sa, with_state, d_state, make_oceank(o), d_out,
                       lists);
    // a batch view of the buffer: no allocation (empty vectors, borrowed pointers)
    sfh_batch_t b;
    b.n = 1;
    b.nunits = 1;
    b.aero = (flags & BF_MODE_AERO) != 0;
    b.nchunks = st.nchunks;
    b.nfaces = m->nf;
    b.K = kSingleK;
    b.waves = waves;
    b.d_static = &H->st;
    b.d_sidx = &H->sidx;
    b.d_chunks = chunks;
    b.d_bpos = &H->bpos;
    b.d_lists = lists;
    b.d_dyn = &H->dyn;
    b.d_partials = (double*)(buf + L.partials);
    return batch_launch(&b, o, d_state, 1, d_out, s, false, nullptr, true);   // the prologue ran in the setup
}

// ---- per-body reference-shaped calls ------------------------------------------

// The reference calls its statics from pool threads, one body per task, with
// any mesh.  Each call takes a workspace from a per-device pool (one per
// concurrently calling thread): its own stream, pinned host staging for the
// state and outputs, and a device buffer grown to the largest mesh it has
// evaluated.  Calls on different workspaces run concurrently; nothing is
// rebuilt when consecutive calls use different meshes.
struct Workspace {
    int device = -1;
    hipStream_t stream = nullptr;
    double* h_io = nullptr;       // pinned: state (38) | outputs (24) | aux (64)
    double* d_io = nullptr;       // device, same layout
    unsigned char* d_buf = nullptr;
    size_t cap = 0;               // bytes of d_buf
    uint32_t waves = 0;
};
static constexpr size_t kIoDoubles = 128;

static std::mutex g_ws_mu;
static std::vector<Workspace*> g_ws_free[64];

static void ws_destroy(Workspace* w)
{
    if (!w) return;
    if (w->stream) hipStreamDestroy(w->stream);
    if (w->h_io) hipHostFree(w->h_io);
    hipFree(w->d_io);
    hipFree(w->d_buf);
    delete w;
}

static int ws_acquire(Workspace** out)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(SFH_ERR_UNSUPPORTED, "device %d beyond the 64 the per-body pool tracks", dev);
    {
        std::lock_guard<std::mutex> lk(g_ws_mu);
        if (!g_ws_free[dev].empty()) {
            *out = g_ws_free[dev].back();
            g_ws_free[dev].pop_back();
            return SFH_OK;
        }
    }
    Workspace* w = new Workspace();
    w->device = dev;
    w->waves = default_waves();
    hipError_t e = hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipHostMalloc((void**)&w->h_io, sizeof(double) * kIoDoubles, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc(&w->d_io, sizeof(double) * kIoDoubles);
    if (e != hipSuccess) {
        ws_destroy(w);
        return fail(SFH_ERR_HIP, "per-body workspace: %s", hipGetErrorString(e));
    }
    *out = w;
    return SFH_OK;
}

static void ws_release(Workspace* w)
{
    std::lock_guard<std::mutex> lk(g_ws_mu);
    g_ws_free[w->device].push_back(w);
}

// Grows the workspace's buffer to the layout of an F-face mesh.
static int ws_reserve(Workspace* w, uint32_t F)
{
    const size_t need = single_layout(F).end;
    if (need <= w->cap) return SFH_OK;
    const size_t cap = std::max(need, 2 * w->cap);
    HIPCHK(hipStreamSynchronize(w->stream));
    hipFree(w->d_buf);
    w->d_buf = nullptr;
    w->cap = 0;
    HIPCHK(hipMalloc(&w->d_buf, cap));
    w->cap = cap;
    return SFH_OK;
}

// Runs one per-body evaluation with host inputs and returns the 24-double
// output row in out.
static int single_call(sfh_mesh_t* m, uint32_t flags, const sfh_ocean_t* o, const double T_CG[16],
                       const double T_C[16], const double v[3], const double w[3], double* out)
{
    Workspace* ws = nullptr;
    int rc = ws_acquire(&ws);
    if (rc != SFH_OK) return rc;
    rc = ws_reserve(ws, m->nf);
    if (rc == SFH_OK) {
        double* h = ws->h_io;
        memcpy(h, T_CG, 16 * sizeof(double));
        memcpy(h + 16, T_C, 16 * sizeof(double));
        memcpy(h + 32, v, 3 * sizeof(double));
        memcpy(h + 35, w, 3 * sizeof(double));
        // the state goes to the device as the setup kernel's argument
        rc = single_launch(ws->d_buf, m, flags, o, nullptr, ws->d_io + 38, ws->stream, ws->waves, h);
        if (rc == SFH_OK) {
            hipError_t e = hipMemcpyAsync(h + 38, ws->d_io + 38, SFH_OUT_DOUBLES * sizeof(double),
                                          hipMemcpyDeviceToHost, ws->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(ws->stream);
            if (e != hipSuccess) rc = fail(SFH_ERR_HIP, "per-body call: %s", hipGetErrorString(e));
            else memcpy(out, h + 38, SFH_OUT_DOUBLES * sizeof(double));
        }
    }
    ws_release(ws);
    return rc;
}

extern "C" int sfh_forces_surface(const sfh_settings* settings, sfh_mesh mesh, sfh_ocean ocean,
                                  const double T_CG[16], const double T_C[16], const double v[3], const double omega[3],
                                  double Fb[3], double Tb[3], double Fdq[3], double Tdq[3], double Fdf[3],
                                  double Tdf[3], double* Swet, double* Vsub)
{
    if (!settings || !T_CG || !T_C || !v || !omega || !Fb || !Tb || !Fdq || !Tdq || !Fdf || !Tdf || !Swet || !Vsub)
        return fail(SFH_ERR_INVALID, "null argument");
    int rc = check_ocean(ocean);
    if (rc != SFH_OK) return rc;
    if (!mesh) {   // SolidEntity.cpp:1293-1312
        if (settings->realistic_buoyancy)
            for (int k = 0; k < 3; ++k) Fb[k] = Tb[k] = 0.0;
        if (settings->damping_forces)
            for (int k = 0; k < 3; ++k) Fdq[k] = Tdq[k] = Fdf[k] = Tdf[k] = 0.0;
        *Swet = 0.0;
        *Vsub = 0.0;
        return SFH_OK;
    }
    uint32_t flags = BF_HYDRO | BF_MODE_SURF | BF_EXPLICIT_TC;
    if (settings->damping_forces) flags |= BF_DAMPING;
    if (settings->realistic_buoyancy) flags |= BF_BUOYANT;
    double out[SFH_OUT_DOUBLES];
    rc = single_call(mesh, flags, ocean, T_CG, T_C, v, omega, out);
    if (rc != SFH_OK) return rc;
    uint32_t wm = (uint32_t)out[20];
    if (wm & 1u) memcpy(Fb, out, 24);
    if (wm & 2u) memcpy(Tb, out + 3, 24);
    if (wm & 4u) memcpy(Fdq, out + 6, 24);
    if (wm & 8u) memcpy(Tdq, out + 9, 24);
    if (wm & 16u) memcpy(Fdf, out + 12, 24);
    if (wm & 32u) memcpy(Tdf, out + 15, 24);
    if (wm & 64u) *Swet = out[18];
    if (wm & 128u) *Vsub = out[19];
    return SFH_OK;
}

extern "C" int sfh_forces_submerged(sfh_mesh mesh, sfh_ocean ocean, const double T_CG[16], const double T_C[16],
                                    const double v[3], const double omega[3], double Fdq[3], double Tdq[3],
                                    double Fdf[3], double Tdf[3])
{
    if (!T_CG || !T_C || !v || !omega || !Fdq || !Tdq || !Fdf || !Tdf) return fail(SFH_ERR_INVALID, "null argument");
    int rc = check_ocean(ocean);
    if (rc != SFH_OK) return rc;
    if (!mesh) {
        for (int k = 0; k < 3; ++k) Fdq[k] = Tdq[k] = Fdf[k] = Tdf[k] = 0.0;
        return SFH_OK;
    }
    double out[SFH_OUT_DOUBLES];
    rc = single_call(mesh, BF_HYDRO | BF_MODE_SUBM | BF_EXPLICIT_TC | BF_DAMPING, ocean, T_CG, T_C, v, omega, out);
    if (rc != SFH_OK) return rc;
    memcpy(Fdq, out + 6, 24);
    memcpy(Tdq, out + 9, 24);
    memcpy(Fdf, out + 12, 24);
    memcpy(Tdf, out + 15, 24);
    return SFH_OK;
}

extern "C" int sfh_forces_aero(sfh_mesh mesh, sfh_ocean atmosphere, const double T_CG[16], const double T_C[16],
                               const double v[3], const double omega[3], double Fda[3], double Tda[3])
{
    if (!T_CG || !T_C || !v || !omega || !Fda || !Tda) return fail(SFH_ERR_INVALID, "null argument");
    int rc = check_ocean(atmosphere);
    if (rc != SFH_OK) return rc;
    if (!mesh) {
        for (int k = 0; k < 3; ++k) Fda[k] = Tda[k] = 0.0;
        return SFH_OK;
    }
    double out[SFH_OUT_DOUBLES];
    rc = single_call(mesh, BF_MODE_AERO | BF_EXPLICIT_TC | BF_DAMPING, atmosphere, T_CG, T_C, v, omega, out);
    if (rc != SFH_OK) return rc;
    memcpy(Fda, out + 6, 24);
    memcpy(Tda, out + 9, 24);
    return SFH_OK;
}

extern "C" int sfh_correct(sfh_ocean ocean, double Fdq[3], double Tdq[3], double Fdf[3], double Tdf[3],
                           const double Cd[3], const double Cf[3], const double T_O[16])
{
    if (!ocean || !Fdq || !Tdq || !Fdf || !Tdf || !Cd || !Cf || !T_O) return fail(SFH_ERR_INVALID, "null argument");
    Workspace* ws = nullptr;
    int rc = ws_acquire(&ws);
    if (rc != SFH_OK) return rc;
    double* h = ws->h_io + 64;   // aux: in[15] at 0, io[12] at 16
    gl_to_rows(T_O, h, nullptr);
    memcpy(h + 9, Cd, 24);
    memcpy(h + 12, Cf, 24);
    memcpy(h + 16, Fdq, 24); memcpy(h + 19, Tdq, 24); memcpy(h + 22, Fdf, 24); memcpy(h + 25, Tdf, 24);
    double* d = ws->d_io + 64;
    hipError_t e = hipMemcpyAsync(d, h, 28 * sizeof(double), hipMemcpyHostToDevice, ws->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_correct, dim3(1), dim3(64), 0, ws->stream, ocean->rho, d, d + 16);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(h + 16, d + 16, 12 * sizeof(double), hipMemcpyDeviceToHost, ws->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ws->stream);
    if (e != hipSuccess) rc = fail(SFH_ERR_HIP, "correct: %s", hipGetErrorString(e));
    else {
        memcpy(Fdq, h + 16, 24); memcpy(Tdq, h + 19, 24); memcpy(Fdf, h + 22, 24); memcpy(Tdf, h + 25, 24);
    }
    ws_release(ws);
    return rc;
}

// ---- per-body call that allocates nothing ----------------------------------------

// The buffer of the no-alloc entry point: the per-body layout for meshes of
// up to kNaMaxFaces faces, in one file-scope device array.  Each GPU the library runs on holds its own copy until the process
// ends; every call on a GPU shares it, so calls that overlap on one GPU (two
// streams, or two host threads) race.
static constexpr uint32_t kNaMaxFaces = 1u << 19;   // 524,288
static constexpr size_t kNaBytes = single_layout(kNaMaxFaces).end;
static_assert(kNaBytes <= (256u << 10), "no-alloc scratch above 256 KiB");

static __device__ __align__(256) unsigned char scratch[kNaBytes];

static std::mutex g_na_mu;
static void* g_na_ptr[64];       // per device: the address of `scratch` (hipGetSymbolAddress)
static uint32_t g_na_waves[64];  // per device: resident wavefronts of the face pass

// Looks up this device's copy of `scratch` once (with the device current):
// from sfh_init, or from the first no-alloc call on a device sfh_init did not see.
static int na_scratch(unsigned char** out, uint32_t* waves)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(SFH_ERR_UNSUPPORTED, "device %d beyond the 64 the no-alloc path tracks", dev);
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return fail(SFH_ERR_HIP, "scratch symbol address: %s", hipGetErrorString(e));
        g_na_waves[dev] = default_waves();
        g_na_ptr[dev] = p;
    }
    if (out) *out = (unsigned char*)g_na_ptr[dev];
    if (waves) *waves = g_na_waves[dev];
    return SFH_OK;
}

int sfh_internal_na_init(void) { return na_scratch(nullptr, nullptr); }

extern "C" int sfh_forces_noalloc(const sfh_settings* settings, int mode, sfh_mesh mesh, sfh_ocean ocean,
                                  const double* d_state, double* d_out, void* stream)
{
    if (!settings || !mesh || !d_state || !d_out) return fail(SFH_ERR_INVALID, "null argument");
    if (mode != SFH_NOALLOC_SURFACE && mode != SFH_NOALLOC_SUBMERGED) return fail(SFH_ERR_INVALID, "mode %d", mode);
    int rc = check_ocean(ocean);
    if (rc != SFH_OK) return rc;
    if (mesh->nf > kNaMaxFaces)
        return fail(SFH_ERR_SIZE, "mesh of %u faces exceeds the no-alloc limit of %u faces", mesh->nf, kNaMaxFaces);
    unsigned char* buf = nullptr;
    uint32_t waves = 0;
    rc = na_scratch(&buf, &waves);
    if (rc != SFH_OK) return rc;
    uint32_t flags = BF_HYDRO | BF_EXPLICIT_TC;
    if (mode == SFH_NOALLOC_SURFACE) {
        flags |= BF_MODE_SURF;
        if (settings->damping_forces) flags |= BF_DAMPING;
        if (settings->realistic_buoyancy) flags |= BF_BUOYANT;
    } else {
        flags |= BF_MODE_SUBM | BF_DAMPING;
    }
    // the layout for this mesh fits in the array's: every part's offset and
    // size grows with F
    return single_launch(buf, mesh, flags, ocean, d_state, d_out, (hipStream_t)stream, waves);
}

// ---- wave grid ----------------------------------------------------------------

extern "C" int sfh_wave_grid_build(const sfh_wave_component* comps, uint32_t ncomp, uint32_t n, float* d_grid,
                                   void* stream)
{
    if ((!comps && ncomp) || !d_grid) return fail(SFH_ERR_INVALID, "null argument");
    if (n < 2 || n > 16384 || ncomp > 4096) return fail(SFH_ERR_SIZE, "bad grid size or component count");
    sfh_wave_component* d_c = nullptr;
    HIPCHK(hipMalloc(&d_c, sizeof(sfh_wave_component) * std::max<uint32_t>(1, ncomp)));
    hipStream_t s = (hipStream_t)stream;
    if (ncomp) HIPCHK(hipMemcpyAsync(d_c, comps, sizeof(sfh_wave_component) * ncomp, hipMemcpyHostToDevice, s));
    dim3 grid((n + 255) / 256, n);
    hipLaunchKernelGGL(k_wave_grid, grid, dim3(256), 0, s, d_c, ncomp, n, d_grid);
    hipError_t e = hipStreamSynchronize(s);
    hipFree(d_c);
    if (e != hipSuccess) return fail(SFH_ERR_HIP, "wave grid: %s", hipGetErrorString(e));
    return SFH_OK;
}

// ---------------------------------------------------------------------------
// surrogate glue: feature build, out-of-distribution mask, decode and audit
// ---------------------------------------------------------------------------

struct SurNorm {
    float in_shift[SFH_SUR_FEATURES], in_scale[SFH_SUR_FEATURES];
    float in_lo[SFH_SUR_FEATURES], in_hi[SFH_SUR_FEATURES];
    float out_shift[6], out_scale[6];
    float gate_band;
};

__device__ __forceinline__ uint16_t bf16_bits(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);   // round to nearest even
    return (uint16_t)(u >> 16);
}

// Total F and T (world frame) of one output row, fields with a clear written
// bit counted as zero.
__device__ __forceinline__ void exact_totals(const double* e, double* F, double* T)
{
    const uint32_t wm = (uint32_t)e[20];
    for (int k = 0; k < 3; ++k) { F[k] = 0.0; T[k] = 0.0; }
    for (int q = 0; q < 3; ++q) {
        if (wm & (1u << (2 * q)))
            for (int k = 0; k < 3; ++k) F[k] += e[6 * q + k];
        if (wm & (1u << (2 * q + 1)))
            for (int k = 0; k < 3; ++k) T[k] += e[6 * q + 3 + k];
    }
}

// Everything the surrogate derives from one body's state: the rotation, the
// 13 features, the body/fluid position, and the physical baseline and scales
// the network's outputs are relative to.  Features, all in the body (CG)
// frame: relative linear velocity R^T (v - u) (0-2), angular velocity (3-5),
// the direction of gravity (6-8), the depth of the centre of buoyancy below
// the local surface (9) and that surface's unit normal (10-12; the plane
// fitted to the wave heights over the body's footprint).  The forces depend
// on the orientation only through these directions, so the network does not
// have to learn that a yaw about gravity changes nothing on a flat sea.  The
// network predicts
//   y_F = (F - Fb0) / sF,  y_T = (T - Tb0) / sT   (CG frame)
// where Fb0, Tb0 is a hydrostatic estimate (the analytic buoyancy inside the
// water; when crossing, the hull's hydrostatic table, or without one the
// analytic buoyancy scaled by the immersed fraction of the body's vertical
// extent) and
//   sF = rho g V + rho (A/4) (u^3 + u),  sT = sF Lc,
// u = |v_rel| + |omega| Lc, Lc the largest half extent of the body's AABB, A
// its surface.  Outputs stay O(1) whatever the speed and immersion.  (Scales
// that follow the drag law more closely, rho (A/4) (u^2 + u), or that shrink
// with the submerged volume of a crossing body, gave larger held-out errors:
// the latter lets the crossing bodies dominate the loss.)
struct SurCtx {
    double R[9], t[3];
    double f[SFH_SUR_FEATURES];
    double base[6];
    double sF, sT;
    int pos;
};

// Trilinear lookup of a hydrostatic table: s in [-1, 1] (immersion of the
// AABB: -1 top at the surface, 1 bottom at the surface), gh the unit gravity
// direction in the CG frame, as polar angle acos(gh.z) on hth nodes over
// [0, pi] and azimuth atan2(gh.y, gh.x) on hph periodic nodes over [-pi, pi).
// Entries: 6 floats, F / (rho g V) and T / (rho g V Lc), CG frame.  FP32: the
// result is a baseline the network corrects.
__device__ void hst_lookup(const UnitStatic& U, float sv, const float* gh, float* out)
{
    const float th = acosf(fminf(1.f, fmaxf(-1.f, gh[2])));
    const float ph = atan2f(gh[1], gh[0]);
    const float fs = (fminf(1.f, fmaxf(-1.f, sv)) + 1.f) * 0.5f * (float)(U.hs - 1);
    const float ft = th * (float)(1.0 / 3.141592653589793) * (float)(U.hth - 1);
    float fp = (ph + 3.14159265f) * (float)(0.5 / 3.141592653589793) * (float)U.hph;
    if (!(fp >= 0.f && fp <= (float)U.hph)) fp = 0.f;   // non-finite direction
    const int is = min((int)fs, (int)U.hs - 2), it = min((int)ft, (int)U.hth - 2);
    int ip = (int)floorf(fp);
    const float as = fs - is, at = ft - it, ap = fp - ip;
    ip %= (int)U.hph;
    if (ip < 0) ip += (int)U.hph;
    const int ip1 = (ip + 1 == (int)U.hph) ? 0 : ip + 1;
    for (int k = 0; k < 6; ++k) out[k] = 0.f;
    for (int c = 0; c < 8; ++c) {
        const int ds = c & 1, dt = (c >> 1) & 1, dp = (c >> 2) & 1;
        const float wgt = (ds ? as : 1.f - as) * (dt ? at : 1.f - at) * (dp ? ap : 1.f - ap);
        const float* e = U.hst + (((size_t)(is + ds) * U.hth + (size_t)(it + dt)) * U.hph + (size_t)(dp ? ip1 : ip)) * 6;
        for (int k = 0; k < 6; ++k) out[k] += wgt * e[k];
    }
}

// The context of one unit computed by a group of G lanes (G = 1: one lane);
// every lane of the group ends with the full context.  The 17 wave queries
// (the classification's 8 AABB corners and the footprint plane's 3 x 3
// stencil) are spread over the group and issued together, one memory round
// trip.  Positions, wave sums and the classification are FP64 (the
// classification is the exact path's, bit for bit); the rest is FP32, far
// finer than the BF16 network input.
template <int G>
__device__ void sur_context_g(const UnitStatic& U, const double* __restrict__ s, const OceanK& oc, SurCtx& C, int sub)
{
    gl_basis(s, C.R);
    for (int k = 0; k < 3; ++k) C.t[k] = s[12 + k];
    double vr[3];
    for (int k = 0; k < 3; ++k) vr[k] = s[16 + k] - oc.u[k];
    double* f = C.f;
    mat_t_vec(C.R, vr, f);
    mat_t_vec(C.R, s + 19, f + 3);
    {   // the direction of gravity in the body frame
        const double gnn = sqrt(oc.g[0] * oc.g[0] + oc.g[1] * oc.g[1] + oc.g[2] * oc.g[2]);
        double gh[3] = {0.0, 0.0, 1.0};
        if (gnn > 0.0)
            for (int k = 0; k < 3; ++k) gh[k] = oc.g[k] / gnn;
        mat_t_vec(C.R, gh, f + 6);
    }
    double mn[3], mx[3], dd[3], cb[3];
    world_aabb(C.R, C.t, U.aabb_c, U.aabb_h, mn, mx, dd);
    mat_vec(C.R, U.P_CB, cb);
    for (int k = 0; k < 3; ++k) cb[k] += C.t[k];
    // the local surface is the plane fitted to the wave field over the body's
    // horizontal footprint (half extents ex, ey of its world AABB, at least
    // one texel): weights (1/4, 1/2, 1/4) on a 3 x 3 stencil give its height
    // at the CoB and least squares its slope
    const double tex = oc.nlayers ? oc.len[0] / (double)oc.n : 0.0;
    const double ex = fmax(0.5 * dd[0], tex), ey = fmax(0.5 * dd[1], tex);
    constexpr int NQ = (17 + G - 1) / G;
    double qx[NQ], qy[NQ], qz[NQ], h[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        const int k = sub + i * G;
        double q[3];
        aabb_corner(k < 8 ? k : 0, mn, mx, dd, q);
        const int j = k < 8 ? 4 : (k < 17 ? k - 8 : 4), a = j / 3, c = j - 3 * (j / 3);
        qx[i] = k < 8 ? q[0] : cb[0] + (a - 1) * ex;
        qy[i] = k < 8 ? q[1] : cb[1] + (c - 1) * ey;
        qz[i] = q[2];
        h[i] = 0.0;
    }
    for (int l = 0; l < oc.nlayers; ++l) {
        const double sc = oc.scale[l];
        float4 tq[NQ];
        double fa[NQ], fb[NQ];
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            const double u = qx[i] * sc - 0.5, v = qy[i] * sc - 0.5;
            const double fi = floor(u), fj = floor(v);
            fa[i] = u - fi;
            fb[i] = v - fj;
            tq[i] = oc.quad[l][(size_t)wrap_texel(fj, oc) * oc.n + (size_t)wrap_texel(fi, oc)];
        }
#pragma unroll
        for (int i = 0; i < NQ; ++i) h[i] += texel_mix(fa[i], fb[i], tq[i]);
    }
    int under = 0;
    double hm = 0.0, sx = 0.0, sy = 0.0;
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
        const int k = sub + i * G;
        if (k < 8) {
            if (qz[i] + h[i] > 0.0) ++under;   // wave_sum64 order: depth = z, then += the layer sum
        } else if (k < 17) {
            const int a = (k - 8) / 3, c = (k - 8) - 3 * ((k - 8) / 3);
            const double w = (a == 1 ? 0.5 : 0.25) * (c == 1 ? 0.5 : 0.25);
            hm += w * h[i];
            sx += w * (a - 1) * h[i];
            sy += w * (c - 1) * h[i];
        }
    }
    under = gsum<G>(under);
    C.pos = U.hydro ? (under == 0 ? 0 : (under == 8 ? 1 : 2)) : -1;
    double dep = cb[2], gx = 0.0, gy = 0.0;
    if (oc.nlayers) {
        hm = gsum<G>(hm);
        sx = gsum<G>(sx);
        sy = gsum<G>(sy);
        dep += hm;
        gx = (double)((float)sx * (2.f / (float)ex));   // sum of w (a-1)^2 = 1/2
        gy = (double)((float)sy * (2.f / (float)ey));
    }
    f[9] = dep;
    {   // the local plane's unit normal (towards increasing depth) in the body frame
        const double in = 1.0 / sqrt(1.0 + gx * gx + gy * gy);
        const double nw[3] = {gx * in, gy * in, in};
        mat_t_vec(C.R, nw, f + 10);
    }
    // hydrostatic estimate
    const float Lc = (float)fmax(fmax(U.aabb_h[0], U.aabb_h[1]), U.aabb_h[2]);
    const float g0 = (float)oc.g[0], g1 = (float)oc.g[1], g2 = (float)oc.g[2];
    const float gn = sqrtf(g0 * g0 + g1 * g1 + g2 * g2);
    const float rho = (float)oc.rho, vol = (float)U.volume;
    float gb[3];
    {
        double g64[3];
        mat_t_vec(C.R, oc.g, g64);
        for (int k = 0; k < 3; ++k) gb[k] = (float)g64[k];
    }
    const float ah0 = (float)U.aabb_h[0], ah1 = (float)U.aabb_h[1], ah2 = (float)U.aabb_h[2];
    // immersed fraction of the AABB's vertical extent (1 inside, 0 outside)
    float frac = C.pos == 1 ? 1.f : 0.f;
    if (C.pos == 2 && U.buoyant && !(U.hst && gn > 0.f)) {
        const float hz = gn > 0.f ? (fabsf(gb[0]) * ah0 + fabsf(gb[1]) * ah1 + fabsf(gb[2]) * ah2) / gn : 0.f;
        frac = hz > 0.f ? fminf(1.f, fmaxf(0.f, ((float)dep + hz) / (2.f * hz))) : 0.f;
    }
    if (U.buoyant && C.pos == 2 && U.hst && gn > 0.f) {
        // the hull's table: static buoyancy on a flat surface against the
        // immersion of its AABB and the direction of gravity (CG frame)
        const float ig = 1.f / gn;
        const float gh[3] = {gb[0] * ig, gb[1] * ig, gb[2] * ig};
        const float hz = fabsf(gh[0]) * ah0 + fabsf(gh[1]) * ah1 + fabsf(gh[2]) * ah2;
        float sv = 0.f;
        for (int k = 0; k < 3; ++k) sv += gh[k] * (float)(U.aabb_c[k] - U.P_CB[k]);
        sv = hz > 0.f ? ((float)dep + sv) / hz : 0.f;
        float tab[6];
        hst_lookup(U, sv, gh, tab);
        const float sb = rho * gn * vol;
        for (int k = 0; k < 3; ++k) {
            C.base[k] = (double)(sb * tab[k]);
            C.base[3 + k] = (double)(sb * Lc * tab[3 + k]);
        }
    } else {
        const float fb = U.buoyant ? frac : 0.f;   // analytic buoyancy times the immersed fraction
        const float pc0 = (float)U.P_CB[0], pc1 = (float)U.P_CB[1], pc2 = (float)U.P_CB[2];
        const float b0 = -rho * vol * fb * gb[0], b1 = -rho * vol * fb * gb[1], b2 = -rho * vol * fb * gb[2];
        C.base[0] = b0;
        C.base[1] = b1;
        C.base[2] = b2;
        C.base[3] = pc1 * b2 - pc2 * b1;
        C.base[4] = pc2 * b0 - pc0 * b2;
        C.base[5] = pc0 * b1 - pc1 * b0;
    }
    // scales: the sizes of the buoyancy and drag terms
    const float v0 = (float)f[0], v1 = (float)f[1], v2 = (float)f[2], w0 = (float)f[3], w1 = (float)f[4],
                w2 = (float)f[5];
    const float u = sqrtf(v0 * v0 + v1 * v1 + v2 * v2) + sqrtf(w0 * w0 + w1 * w1 + w2 * w2) * Lc;
    float sF = rho * gn * vol + rho * 0.25f * (float)U.surface * (u * u * u + u);
    if (!(sF > 0.f)) sF = 1.f;
    C.sF = sF;
    C.sT = (double)(sF * fmaxf(Lc, 1e-6f));
}

__device__ __forceinline__ void sur_context(const UnitStatic& U, const double* __restrict__ s, const OceanK& oc,
                                            SurCtx& C)
{
    sur_context_g<1>(U, s, oc, C, 0);
}

// Network input row (BF16 bits, 32 per row) of a body's features.
__device__ __forceinline__ void sur_input_row(const double* f, const SurNorm& nm, uint32_t w[SFH_SUR_ROW / 2])
{
    for (int k = 0; k < SFH_SUR_ROW / 2; ++k) w[k] = 0;
    for (int k = 0; k < SFH_SUR_FEATURES; ++k) {
        const uint32_t h = bf16_bits(((float)f[k] - nm.in_shift[k]) * nm.in_scale[k]);
        w[k >> 1] |= h << (16 * (k & 1));
    }
}

// Out of distribution: a feature outside the training range, or jets and
// pipes in the ocean (the features carry only the uniform current); and the
// error-aware gate: a body crossing the surface whose depth feature is in the
// band above the training minimum where the network's training error was
// large.
__device__ __forceinline__ uint8_t sur_ood(const double* f, int pos, const SurNorm& nm, const OceanK& oc)
{
    uint8_t o = oc.ncur > 0 ? 1 : 0;
    for (int k = 0; k < SFH_SUR_FEATURES; ++k) {
        const float v = (float)f[k];
        if (!(v >= nm.in_lo[k] && v <= nm.in_hi[k])) o = 1;   // also catches NaN
    }
    if (pos == 2 && nm.gate_band > 0.f && (float)f[9] < nm.in_lo[9] + nm.gate_band) o = 1;
    return pos > 0 ? o : (uint8_t)0;   // out of the water (or not hydrodynamic): the result is known
}

__global__ __launch_bounds__(128) void k_sur_features(const UnitStatic* __restrict__ units, const double* __restrict__ states,
                               uint32_t first, uint32_t count, uint32_t rows, OceanK oc, SurNorm nm,
                               float* __restrict__ raw, uint16_t* __restrict__ x, uint8_t* __restrict__ ood,
                               int8_t* __restrict__ posout, const double* __restrict__ exact,
                               float* __restrict__ raw_target, float* __restrict__ target)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    if (i >= count) {
        if (x) {
            uint4* row = reinterpret_cast<uint4*>(x + (size_t)i * SFH_SUR_ROW);
            for (int k = 0; k < SFH_SUR_ROW / 8; ++k) row[k] = make_uint4(0, 0, 0, 0);
        }
        return;
    }
    const uint32_t b = first + i;
    SurCtx C;
    sur_context(units[b], states + (size_t)b * SFH_STATE_DOUBLES, oc, C);
    const double* f = C.f;
    const int pos = C.pos;
    if (posout) posout[i] = (int8_t)pos;
    if (raw)
        for (int k = 0; k < SFH_SUR_FEATURES; ++k) raw[(size_t)i * SFH_SUR_FEATURES + k] = (float)f[k];
    if (ood) ood[i] = sur_ood(f, pos, nm, oc);
    if (x) {
        uint32_t w[SFH_SUR_ROW / 2];
        sur_input_row(f, nm, w);
        uint4* row = reinterpret_cast<uint4*>(x + (size_t)i * SFH_SUR_ROW);
        for (int k = 0; k < SFH_SUR_ROW / 8; ++k) row[k] = make_uint4(w[4 * k], w[4 * k + 1], w[4 * k + 2], w[4 * k + 3]);
    }
    if (exact && (raw_target || target)) {
        double F[3], T[3], y[6];
        exact_totals(exact + (size_t)b * SFH_OUT_DOUBLES, F, T);
        mat_t_vec(C.R, F, y);
        mat_t_vec(C.R, T, y + 3);
        for (int k = 0; k < 6; ++k) {
            const float r = (float)((y[k] - C.base[k]) / (k < 3 ? C.sF : C.sT));
            if (raw_target) raw_target[(size_t)i * 6 + k] = r;
            if (target) target[(size_t)i * 6 + k] = (r - nm.out_shift[k]) / nm.out_scale[k];
        }
    }
}

// CG-frame network output -> world-frame F, T of one body.
__device__ __forceinline__ void sur_decode(const double* R, const double* base, double sF, double sT, int pos,
                                           const SurNorm& nm, const float* y, double* F, double* T)
{
    double yb[6];
    for (int k = 0; k < 6; ++k)
        yb[k] = base[k] + (k < 3 ? sF : sT) * (double)(y[k] * nm.out_scale[k] + nm.out_shift[k]);
    mat_vec(R, yb, F);
    mat_vec(R, yb + 3, T);
    if (pos <= 0)   // outside the water or not hydrodynamic: no force
        for (int k = 0; k < 3; ++k) F[k] = T[k] = 0.0;
}

__device__ __forceinline__ void sur_decode(const SurCtx& C, const SurNorm& nm, const float* y, double* F, double* T)
{
    sur_decode(C.R, C.base, C.sF, C.sT, C.pos, nm, y, F, T);
}



__global__ __launch_bounds__(128) void k_sur_apply(const UnitStatic* __restrict__ units, const double* __restrict__ states, uint32_t first,
                            uint32_t count, OceanK oc, SurNorm nm, const float* __restrict__ y,
                            const uint8_t* __restrict__ ood, int8_t* __restrict__ posout,
                            const double* __restrict__ exact, double* __restrict__ ft, float* __restrict__ err)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint32_t b = first + i;
    SurCtx C;
    sur_context(units[b], states + (size_t)b * SFH_STATE_DOUBLES, oc, C);
    if (posout) posout[i] = (int8_t)C.pos;
    double Fs[3], Ts[3];
    sur_decode(C, nm, y + (size_t)i * 6, Fs, Ts);
    double Fe[3], Te[3];
    if (exact) exact_totals(exact + (size_t)b * SFH_OUT_DOUBLES, Fe, Te);
    const bool fallback = exact && ood && ood[i];
    double* o = ft + (size_t)i * 6;
    for (int k = 0; k < 3; ++k) {
        o[k] = fallback ? Fe[k] : Fs[k];
        o[3 + k] = fallback ? Te[k] : Ts[k];
    }
    if (err && exact) {
        double dF = 0, nF = 0, dT = 0, nT = 0;
        for (int k = 0; k < 3; ++k) {
            dF += (Fs[k] - Fe[k]) * (Fs[k] - Fe[k]);
            nF += Fe[k] * Fe[k];
            dT += (Ts[k] - Te[k]) * (Ts[k] - Te[k]);
            nT += Te[k] * Te[k];
        }
        err[2 * (size_t)i] = (float)(dF == 0.0 ? 0.0 : sqrt(dF / nF));
        err[2 * (size_t)i + 1] = (float)(dT == 0.0 ? 0.0 : sqrt(dT / nT));
    }
}

static SurNorm sur_norm(const sfh_surrogate_norm* n)
{
    SurNorm k;
    memset(&k, 0, sizeof(k));
    if (!n) {
        for (int i = 0; i < SFH_SUR_FEATURES; ++i) k.in_scale[i] = 1.f;
        for (int i = 0; i < 6; ++i) k.out_scale[i] = 1.f;
        return k;
    }
    memcpy(k.in_shift, n->in_shift, sizeof(k.in_shift));
    memcpy(k.in_scale, n->in_scale, sizeof(k.in_scale));
    memcpy(k.in_lo, n->in_lo, sizeof(k.in_lo));
    memcpy(k.in_hi, n->in_hi, sizeof(k.in_hi));
    memcpy(k.out_shift, n->out_shift, sizeof(k.out_shift));
    memcpy(k.out_scale, n->out_scale, sizeof(k.out_scale));
    k.gate_band = n->gate_band;
    return k;
}

extern "C" int sfh_surrogate_features(sfh_batch b, sfh_ocean o, const double* d_states, uint32_t first,
                                      uint32_t count, uint32_t rows, const sfh_surrogate_norm* norm, float* d_raw,
                                      uint16_t* d_x, uint8_t* d_ood, int8_t* d_pos, const double* d_exact,
                                      float* d_raw_target, float* d_target, void* stream)
{
    if (!b || !d_states) return fail(SFH_ERR_INVALID, "null argument");
    if (b->aero) return fail(SFH_ERR_INVALID, "the surrogate models hydrodynamic batches, not aerodynamic ones");
    int rc = check_ocean(o);
    if (rc != SFH_OK) return rc;
    if ((uint64_t)first + count > b->nunits) return fail(SFH_ERR_SIZE, "body range out of the batch");
    if ((d_x || d_ood || d_target) && !norm) return fail(SFH_ERR_INVALID, "normalised outputs need a norm");
    if ((d_raw_target || d_target) && !d_exact) return fail(SFH_ERR_INVALID, "targets need the exact outputs");
    if (d_x && (rows < count || rows % 32)) return fail(SFH_ERR_SIZE, "rows must be a multiple of 32 >= count");
    if (d_x && ((uintptr_t)d_x % 16)) return fail(SFH_ERR_INVALID, "d_x must be 16-byte aligned");
    const uint32_t r = d_x ? rows : count;
    if (r == 0) return SFH_OK;
    hipLaunchKernelGGL(k_sur_features, dim3((r + 127) / 128), dim3(128), 0, (hipStream_t)stream, b->d_units,
                       d_states, first, count, r, make_oceank(o), sur_norm(norm), d_raw, d_x, d_ood, d_pos, d_exact,
                       d_raw_target, d_target);
    HIPCHK(hipGetLastError());
    return SFH_OK;
}

extern "C" int sfh_surrogate_apply(sfh_batch b, sfh_ocean o, const double* d_states, uint32_t first, uint32_t count,
                                   const sfh_surrogate_norm* norm, const float* d_y, const uint8_t* d_ood,
                                   int8_t* d_pos, const double* d_exact, double* d_ft, float* d_err,
                                   void* stream)
{
    if (!b || !d_states || !norm || !d_y || !d_ft) return fail(SFH_ERR_INVALID, "null argument");
    if (b->aero) return fail(SFH_ERR_INVALID, "the surrogate models hydrodynamic batches, not aerodynamic ones");
    int rc = check_ocean(o);
    if (rc != SFH_OK) return rc;
    if ((uint64_t)first + count > b->nunits) return fail(SFH_ERR_SIZE, "body range out of the batch");
    if (count == 0) return SFH_OK;
    hipLaunchKernelGGL(k_sur_apply, dim3((count + 127) / 128), dim3(128), 0, (hipStream_t)stream, b->d_units,
                       d_states, first, count, make_oceank(o), sur_norm(norm), d_y, d_ood, d_pos, d_exact, d_ft,
                       d_err);
    HIPCHK(hipGetLastError());
    return SFH_OK;
}

// ---------------------------------------------------------------------------
// rollout harness: rigid-body step of the bodies from their total F/T
// ---------------------------------------------------------------------------

// Semi-implicit Euler of the Newton-Euler equations about the CG: linear
// velocity from F + m g, angular velocity in the CG frame from
// T - w x (I w) with principal moments I, then position from the new
// velocity and orientation by the exact rotation of w dt (Rodrigues).
__global__ void k_rollout(double* __restrict__ states, uint32_t count, double* __restrict__ ft,
                          const uint8_t* __restrict__ ood, const double* __restrict__ exact, double mass,
                          double I0, double I1, double I2, double g0, double g1, double g2, double dt)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    double* s = states + (size_t)i * SFH_STATE_DOUBLES;
    double* f = ft + (size_t)i * 6;
    if (ood && ood[i]) {   // the exact totals of a flagged body (merged here rather than in a launch of its own)
        double Fe[3], Te[3];
        exact_totals(exact + (size_t)i * SFH_OUT_DOUBLES, Fe, Te);
        for (int k = 0; k < 3; ++k) { f[k] = Fe[k]; f[3 + k] = Te[k]; }
    }
    double R[9];
    gl_basis(s, R);
    const double g[3] = {g0, g1, g2};
    for (int k = 0; k < 3; ++k) s[16 + k] += dt * (f[k] / mass + g[k]);
    double wb[3], tb[3];
    mat_t_vec(R, s + 19, wb);
    mat_t_vec(R, f + 3, tb);
    const double I[3] = {I0, I1, I2};
    double Iw[3] = {I0 * wb[0], I1 * wb[1], I2 * wb[2]}, gyro[3];
    cross64(wb, Iw, gyro);
    for (int k = 0; k < 3; ++k) wb[k] += dt * (tb[k] - gyro[k]) / I[k];
    mat_vec(R, wb, s + 19);
    for (int k = 0; k < 3; ++k) s[12 + k] += dt * s[16 + k];
    // R <- R exp([wb dt]x)
    const double ax = wb[0] * dt, ay = wb[1] * dt, az = wb[2] * dt;
    const double th = sqrt(ax * ax + ay * ay + az * az);
    double E[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (th > 0.0) {
        const double kx = ax / th, ky = ay / th, kz = az / th, c = cos(th), sn = sin(th), cc = 1.0 - c;
        E[0] = c + kx * kx * cc;      E[1] = kx * ky * cc - kz * sn; E[2] = kx * kz * cc + ky * sn;
        E[3] = ky * kx * cc + kz * sn; E[4] = c + ky * ky * cc;      E[5] = ky * kz * cc - kx * sn;
        E[6] = kz * kx * cc - ky * sn; E[7] = kz * ky * cc + kx * sn; E[8] = c + kz * kz * cc;
    }
    double Rn[9];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) Rn[3 * r + c] = R[3 * r] * E[c] + R[3 * r + 1] * E[3 + c] + R[3 * r + 2] * E[6 + c];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) s[4 * c + r] = Rn[3 * r + c];
}

extern "C" int sfh_rollout_integrate(sfh_ocean o, double* d_states, uint32_t count, double* d_ft,
                                     const uint8_t* d_ood, const double* d_exact, const sfh_rigid* rigid, double dt,
                                     void* stream)
{
    if (!o || !d_states || !d_ft || !rigid) return fail(SFH_ERR_INVALID, "null argument");
    if (d_ood && !d_exact) return fail(SFH_ERR_INVALID, "d_ood needs d_exact");
    if (!(rigid->mass > 0.0) || !(rigid->inertia[0] > 0.0) || !(rigid->inertia[1] > 0.0) || !(rigid->inertia[2] > 0.0))
        return fail(SFH_ERR_INVALID, "mass and inertia must be positive");
    if (count == 0) return SFH_OK;
    hipLaunchKernelGGL(k_rollout, dim3((count + 127) / 128), dim3(128), 0, (hipStream_t)stream, d_states, count, d_ft,
                       d_ood, d_exact, rigid->mass, rigid->inertia[0], rigid->inertia[1], rigid->inertia[2], o->g[0], o->g[1], o->g[2],
                       dt);
    HIPCHK(hipGetLastError());
    return SFH_OK;
}

// ---------------------------------------------------------------------------
// one-launch surrogate step: features, network, decode
// ---------------------------------------------------------------------------

struct SurPlanDev {
    const uint16_t* W[SFM_MAX_GROUP][SFM_LAYERS];
    const float* B[SFM_MAX_GROUP][SFM_LAYERS];
    SurNorm norm[SFM_MAX_GROUP];
};

struct SurTile {
    uint32_t unit0, nvalid, group, pad_;
};

struct SurRowCtx {
    double R[9], base[6], sF, sT;
    int32_t pos, pad_;
};

// One workgroup per tile of TR units of one group (one network), tiles in
// XCD-major order: groups of 256 / TR lanes build the features of one unit
// each (the FP64 context of k_sur_features, wave queries spread over the
// group) into the network's input rows in LDS while the first layer's
// weights are in flight, the four wavefronts run the fused network
// (sfm_device.h), and lanes 0..TR-1 decode their unit's outputs to
// world-frame F, T.
template <int TR>
__global__ __launch_bounds__(256) void k_sur_step(const SurPlanDev* __restrict__ P, const SurTile* __restrict__ tiles,
                                                  const UnitStatic* __restrict__ units,
                                                  const double* __restrict__ states, OceanK oc,
                                                  uint8_t* __restrict__ ood, int8_t* __restrict__ posout,
                                                  double* __restrict__ ft)
{
    __shared__ __attribute__((aligned(16))) uint16_t act[2][TR * sfm::kActStride];
    __shared__ __attribute__((aligned(16))) float bias[SFM_LAYERS - 1][SFM_HIDDEN];
    __shared__ SurRowCtx ctx[TR];
    __shared__ float yo[TR][SFM_OUT];
    const SurTile tl = tiles[sfm::xcd_tile(blockIdx.x, gridDim.x)];
    const int g = (int)tl.group;
    const int t = threadIdx.x;
    sfm::LayerW<SFM_IN_PAD> w0;   // requested now, used after the features
    sfm::load_layer<SFM_IN_PAD>(P->W[g][0], w0);
    sfm::load_bias(P->B[g], bias);
    {
        // G lanes per unit build its context together
        constexpr int G = 256 / TR;
        const int r = t / G, sub = t % G;
        const bool valid = (uint32_t)r < tl.nvalid;
        const uint32_t u = tl.unit0 + (valid ? (uint32_t)r : 0u);
        SurCtx C;
        sur_context_g<G>(units[u], states + (size_t)u * SFH_STATE_DOUBLES, oc, C, sub);
        if (sub == 0) {
            uint4* xr = reinterpret_cast<uint4*>(act[1] + r * sfm::kActStride);
            if (valid) {
                const SurNorm& nm = P->norm[g];
                uint32_t w[SFH_SUR_ROW / 2];
                sur_input_row(C.f, nm, w);
                for (int k = 0; k < SFH_SUR_ROW / 8; ++k)
                    xr[k] = make_uint4(w[4 * k], w[4 * k + 1], w[4 * k + 2], w[4 * k + 3]);
                if (ood) ood[u] = sur_ood(C.f, C.pos, nm, oc);
                if (posout) posout[u] = (int8_t)C.pos;
                SurRowCtx& X = ctx[r];
                for (int k = 0; k < 9; ++k) X.R[k] = C.R[k];
                for (int k = 0; k < 6; ++k) X.base[k] = C.base[k];
                X.sF = C.sF;
                X.sT = C.sT;
                X.pos = C.pos;
            } else {
                for (int k = 0; k < SFH_SUR_ROW / 8; ++k) xr[k] = make_uint4(0, 0, 0, 0);
            }
        }
    }
    __syncthreads();
    sfm::OutW ow;
    sfm::mlp_hidden<TR>(P->W[g], bias, act, nullptr, w0, ow);
    const int lane = t & 63, wave = t >> 6;
    if (wave < TR / 32) {
        float o[4];
        sfm::mlp_output(ow, P->B[g][SFM_LAYERS - 1], act[1], o);
        const int lq = (lane >> 5) * 4, row = 32 * wave + (lane & 31);
        for (int i = 0; i < 4; ++i)
            if (lq + i < SFM_OUT) yo[row][lq + i] = o[i];
    }
    __syncthreads();
    if (t < TR && (uint32_t)t < tl.nvalid) {
        const uint32_t u = tl.unit0 + (uint32_t)t;
        const SurRowCtx& X = ctx[t];
        double F[3], T[3];
        sur_decode(X.R, X.base, X.sF, X.sT, X.pos, P->norm[g], yo[t], F, T);
        double* o = ft + (size_t)u * 6;
        for (int k = 0; k < 3; ++k) { o[k] = F[k]; o[3 + k] = T[k]; }
    }
}

// Exact totals for the out-of-distribution units; with err, first the
// surrogate's error against the exact totals for every unit.
__global__ void k_sur_merge(uint32_t n, const uint8_t* __restrict__ ood, const double* __restrict__ exact,
                            double* __restrict__ ft, float* __restrict__ err)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double* o = ft + (size_t)i * 6;
    double Fe[3], Te[3];
    exact_totals(exact + (size_t)i * SFH_OUT_DOUBLES, Fe, Te);
    if (err) {
        double dF = 0, nF = 0, dT = 0, nT = 0;
        for (int k = 0; k < 3; ++k) {
            dF += (o[k] - Fe[k]) * (o[k] - Fe[k]);
            nF += Fe[k] * Fe[k];
            dT += (o[3 + k] - Te[k]) * (o[3 + k] - Te[k]);
            nT += Te[k] * Te[k];
        }
        err[2 * (size_t)i] = (float)(dF == 0.0 ? 0.0 : sqrt(dF / nF));
        err[2 * (size_t)i + 1] = (float)(dT == 0.0 ? 0.0 : sqrt(dT / nT));
    }
    if (ood && ood[i])
        for (int k = 0; k < 3; ++k) { o[k] = Fe[k]; o[3 + k] = Te[k]; }
}

struct sfh_sur_plan_t {
    sfh_batch_t* batch = nullptr;
    SurPlanDev* d_plan = nullptr;
    SurTile* d_tiles = nullptr;
    uint32_t ntiles = 0;
    int TR = 32;
};

extern "C" int sfh_surrogate_plan_create(sfh_batch b, uint32_t ngroups, const uint32_t* counts,
                                         const sfh_surrogate_norm* norms, const sfh_net_ptrs* nets,
                                         sfh_sur_plan* out)
{
    if (!b || !counts || !norms || !nets || !out) return fail(SFH_ERR_INVALID, "null argument");
    if (b->aero) return fail(SFH_ERR_INVALID, "the surrogate models hydrodynamic batches, not aerodynamic ones");
    if (ngroups == 0 || ngroups > SFM_MAX_GROUP) return fail(SFH_ERR_SIZE, "1 to %d groups", SFM_MAX_GROUP);
    uint64_t total = 0;
    for (uint32_t g = 0; g < ngroups; ++g) {
        total += counts[g];
        for (int l = 0; l < SFM_LAYERS; ++l)
            if (!nets[g].weights[l] || !nets[g].bias[l]) return fail(SFH_ERR_INVALID, "group %u: null weights", g);
    }
    if (total != b->nunits) return fail(SFH_ERR_SIZE, "group counts must add up to the batch's %u units", b->nunits);
    SurPlanDev h;
    memset(&h, 0, sizeof(h));
    for (uint32_t g = 0; g < ngroups; ++g) {
        for (int l = 0; l < SFM_LAYERS; ++l) {
            h.W[g][l] = nets[g].weights[l];
            h.B[g][l] = nets[g].bias[l];
        }
        h.norm[g] = sur_norm(&norms[g]);
    }
    int cus = 256, dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
        cus = prop.multiProcessorCount;
    const int TR = (total / 64 >= 2 * (uint64_t)cus) ? 64 : 32;
    std::vector<SurTile> tiles;
    uint64_t u0 = 0;
    for (uint32_t g = 0; g < ngroups; ++g) {
        for (uint32_t r = 0; r < counts[g]; r += TR) {
            SurTile t;
            t.unit0 = (uint32_t)(u0 + r);
            t.nvalid = std::min<uint32_t>(TR, counts[g] - r);
            t.group = g;
            t.pad_ = 0;
            tiles.push_back(t);
        }
        u0 += counts[g];
    }
    sfh_sur_plan_t* p = new sfh_sur_plan_t();
    p->batch = b;
    p->TR = TR;
    p->ntiles = (uint32_t)tiles.size();
    hipError_t e = hipMalloc(&p->d_plan, sizeof(SurPlanDev));
    if (e == hipSuccess) e = hipMalloc(&p->d_tiles, sizeof(SurTile) * std::max<size_t>(1, tiles.size()));
    if (e == hipSuccess) e = hipMemcpy(p->d_plan, &h, sizeof(h), hipMemcpyHostToDevice);
    if (e == hipSuccess && !tiles.empty())
        e = hipMemcpy(p->d_tiles, tiles.data(), sizeof(SurTile) * tiles.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        hipFree(p->d_plan);
        hipFree(p->d_tiles);
        delete p;
        return fail(SFH_ERR_HIP, "surrogate plan: %s", hipGetErrorString(e));
    }
    *out = p;
    return SFH_OK;
}

extern "C" int sfh_surrogate_plan_destroy(sfh_sur_plan p)
{
    if (!p) return fail(SFH_ERR_INVALID, "null plan");
    hipDeviceSynchronize();
    hipFree(p->d_plan);
    hipFree(p->d_tiles);
    delete p;
    return SFH_OK;
}

extern "C" int sfh_surrogate_run(sfh_sur_plan p, sfh_ocean o, const double* d_states, uint8_t* d_ood, int8_t* d_pos,
                                 double* d_ft, void* stream)
{
    if (!p || !d_states || !d_ft) return fail(SFH_ERR_INVALID, "null argument");
    int rc = check_ocean(o);
    if (rc != SFH_OK) return rc;
    if (p->ntiles == 0) return SFH_OK;
    hipStream_t s = (hipStream_t)stream;
    const OceanK oc = make_oceank(o);
    if (p->TR == 64)
        hipLaunchKernelGGL(k_sur_step<64>, dim3(p->ntiles), dim3(256), 0, s, p->d_plan, p->d_tiles, p->batch->d_units,
                           d_states, oc, d_ood, d_pos, d_ft);
    else
        hipLaunchKernelGGL(k_sur_step<32>, dim3(p->ntiles), dim3(256), 0, s, p->d_plan, p->d_tiles, p->batch->d_units,
                           d_states, oc, d_ood, d_pos, d_ft);
    HIPCHK(hipGetLastError());
    return SFH_OK;
}

extern "C" int sfh_surrogate_merge(sfh_sur_plan p, const uint8_t* d_ood, const double* d_exact, double* d_ft,
                                   float* d_err, void* stream)
{
    if (!p || !d_exact || !d_ft) return fail(SFH_ERR_INVALID, "null argument");
    const uint32_t n = p->batch->nunits;
    if (n == 0) return SFH_OK;
    hipLaunchKernelGGL(k_sur_merge, dim3((n + 127) / 128), dim3(128), 0, (hipStream_t)stream, n, d_ood, d_exact, d_ft,
                       d_err);
    HIPCHK(hipGetLastError());
    return SFH_OK;
}

extern "C" int sfh_surrogate_set_table(sfh_batch b, uint32_t first, uint32_t count, const float* d_table,
                                       uint32_t ns, uint32_t ntheta, uint32_t nphi)
{
    if (!b) return fail(SFH_ERR_INVALID, "null batch");
    if (b->aero) return fail(SFH_ERR_INVALID, "the surrogate models hydrodynamic batches, not aerodynamic ones");
    if ((uint64_t)first + count > b->nunits) return fail(SFH_ERR_SIZE, "unit range out of the batch");
    if (d_table && (ns < 2 || ntheta < 2 || nphi < 2)) return fail(SFH_ERR_SIZE, "table dimensions must be >= 2");
    if (count == 0) return SFH_OK;
    std::vector<UnitStatic> u(count);
    HIPCHK(hipMemcpy(u.data(), b->d_units + first, sizeof(UnitStatic) * count, hipMemcpyDeviceToHost));
    for (UnitStatic& x : u) {
        x.hst = d_table;
        x.hs = d_table ? ns : 0;
        x.hth = d_table ? ntheta : 0;
        x.hph = d_table ? nphi : 0;
    }
    HIPCHK(hipMemcpy(b->d_units + first, u.data(), sizeof(UnitStatic) * count, hipMemcpyHostToDevice));
    return SFH_OK;
}
