// Per-face hydrodynamics integrand, FP32, evaluated in the body's C (mesh)
// frame.  Shared by the gfx950 kernels (device) and the host-side unit harness.
//
// Reference: SolidEntity::ComputeHydrodynamicForcesSurface (SolidEntity.cpp:1289-1721)
// and ComputeHydrodynamicForcesSubmerged (SolidEntity.cpp:1723-1797).  The
// reference transforms every vertex to world coordinates in float; here the
// same quantities are formed in the mesh frame, where a rigid transform leaves
// cross products, dot products, lengths, triple products and affine
// combinations (clip points, centroids) unchanged.  World quantities are
// recovered once per body in FP64: forces and moments are rotated by R, the
// moment about the centre of gravity is M + dt x F.  This keeps every FP32
// operand at the scale of the body instead of its world position.
#pragma once

#ifdef __HIPCC__
#define SFH_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define SFH_HD inline
#endif

namespace sfh {

// Square root and reciprocal: the hardware's single instructions on the
// device (about 1 ulp), the libm functions on the host.
#ifdef __HIP_DEVICE_COMPILE__
SFH_HD float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
SFH_HD float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
SFH_HD float frsq(float x) { return __builtin_amdgcn_rsqf(x); }
#else
SFH_HD float fsqrt(float x) { return sqrtf(x); }
SFH_HD float frcp(float x) { return 1.0f / x; }
SFH_HD float frsq(float x) { return 1.0f / sqrtf(x); }
#endif

struct f3 { float x, y, z; };

SFH_HD f3 mk3(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
SFH_HD f3 operator+(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
SFH_HD f3 operator-(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
SFH_HD f3 operator*(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
SFH_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
SFH_HD f3 cross(f3 a, f3 b) { return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }

// Thresholds spelled as in the reference (float literals).
#define SFH_DEGENERATE_N2 1e-12f   // |fn|^2 below this: face skipped (1386 etc.)
#define SFH_APPROACH_VN  -1e-12f   // quadratic drag when vc.n below this (1776)
#define SFH_SKIN_VT2      1e-9f    // skin friction when |vt|^2 above this (1784)

// Per-body, per-step constants of the face loop, all in the mesh frame.
// The velocity of the fluid relative to a face centroid c (mesh frame) is
// vc = u - (v + w x (c + dt)) = -(vb + w x c) with vb = v + w x dt - u.
struct FaceBody {
    f3 vb;    // R^T * (v + w x (t_C - t_CG) - u)
    f3 w;     // R^T * angular velocity
    f3 o;     // R^T * (p0 - t_C), p0 = (t_CG.x, t_CG.y, 0): volume apex
};

// Accumulator element type: double on the device (FP64 reduction), float or
// double in host harnesses.
template <typename T>
struct Acc {
    T Fb[3], Mb[3];      // pressure buoyancy (waves): force, moment about the mesh origin
    T Fdq[3], Mdq[3];    // quadratic drag
    T Fdf[3], Mdf[3];    // skin friction
    T CB[3];             // sum of tetra centroid (relative to o) * 6V
    T V6;                // sum of 6V
    T Swet;              // wetted area
};

template <typename T>
SFH_HD void acc_zero(Acc<T>& a)
{
    for (int k = 0; k < 3; ++k) a.Fb[k] = a.Mb[k] = a.Fdq[k] = a.Mdq[k] = a.Fdf[k] = a.Mdf[k] = a.CB[k] = T(0);
    a.V6 = T(0);
    a.Swet = T(0);
}

// Accumulators of the submerged loop (drag only).
template <typename T>
struct AccDrag {
    T Fdq[3], Mdq[3], Fdf[3], Mdf[3];
};

template <typename T>
SFH_HD void acc_zero(AccDrag<T>& a)
{
    for (int k = 0; k < 3; ++k) a.Fdq[k] = a.Mdq[k] = a.Fdf[k] = a.Mdf[k] = T(0);
}

template <typename T>
SFH_HD void add3(T* acc, f3 x)
{
    acc[0] += (T)x.x; acc[1] += (T)x.y; acc[2] += (T)x.z;
}

// Fluid velocity that varies over the body (jets and pipes, Ocean.cpp:183-201)
// at a mesh-frame point given in FP64, in the mesh frame: a functor with `on`
// false when there is none (the uniform part is folded into FaceBody::vb).
struct NoFlow {
    static constexpr bool on = false;
    SFH_HD f3 operator()(double, double, double) const { return mk3(0.f, 0.f, 0.f); }
};

// Drag integrand of one (possibly clipped) face: centroid c (mesh frame), unit
// normal n1, area A (SolidEntity.cpp:1660-1683, 1767-1790).  With vm = -vc:
// quadratic q = vc |vc| (-vc.n1) A = vm (-|vm| (vm.n1) A), applied when
// vc.n1 < -1e-12; tangential vt = vc - n1 (vc.n1) = n1 (vm.n1) - vm, skin
// s = vt A when |vt|^2 > 1e-9.  A varying fluid velocity uf at the centroid
// (when Flow::on) enters as vm -= uf.
template <bool ON, class AccT>
SFH_HD void face_drag(const FaceBody& B, f3 c, f3 n1, float A, f3 uf, AccT& acc)
{
    f3 vm = B.vb + cross(B.w, c);
    if (ON) vm = vm - uf;
    float vn = dot(vm, n1);
    if (vn > -SFH_APPROACH_VN) {
        float m = fsqrt(dot(vm, vm));
        f3 q = vm * (-(m * vn * A));
        add3(acc.Fdq, q);
        add3(acc.Mdq, cross(c, q));
    }
    f3 vt = n1 * vn - vm;
    if (dot(vt, vt) > SFH_SKIN_VT2) {
        f3 sk = vt * A;
        add3(acc.Fdf, sk);
        add3(acc.Mdf, cross(c, sk));
    }
}

template <class AccT>
SFH_HD void face_drag(const FaceBody& B, f3 c, f3 n1, float A, AccT& acc)
{
    face_drag<false>(B, c, n1, A, mk3(0.f, 0.f, 0.f), acc);
}

// Submerged face (no clipping, no buoyancy): SolidEntity.cpp:1749-1791.  A
// varying current is sampled at the centroid formed in FP64 (its gradients
// can be steep: a jet's Gaussian, a pipe's wall profile).
template <class Flow, class AccT>
SFH_HD void face_submerged(const FaceBody& B, f3 p1, f3 p2, f3 p3, const Flow& flow, AccT& acc)
{
    f3 e1 = p2 - p1, e2 = p3 - p1;
    f3 n = cross(e1, e2);
    float n2 = dot(n, n);
    if (n2 < SFH_DEGENERATE_N2) return;
    float inv = frsq(n2);
    f3 c = p1 + (e1 + e2) * (1.0f / 3.0f);
    f3 uf = mk3(0.f, 0.f, 0.f);
    if (Flow::on)
        uf = flow(((double)p1.x + (double)p2.x + (double)p3.x) / 3.0, ((double)p1.y + (double)p2.y + (double)p3.y) / 3.0,
                  ((double)p1.z + (double)p2.z + (double)p3.z) / 3.0);
    face_drag<Flow::on>(B, c, n * inv, 0.5f * (n2 * inv), uf, acc);
}

template <class AccT>
SFH_HD void face_submerged(const FaceBody& B, f3 p1, f3 p2, f3 p3, AccT& acc)
{
    face_submerged(B, p1, p2, p3, NoFlow(), acc);
}

// Aerodynamic face (SolidEntity.cpp:1884-1912): drag of the normal velocity
// only, vn = (vc.n1) n1, applied when n1.vn < -1e-12, with the reference's
// integrand `vn * vn.length() * A` (SolidEntity.cpp:1913).  vn is a glm::vec3
// and glm's member length() is the GLSL one, the number of components (3),
// not the Euclidean norm (glm::length(vn)), so the term is q = 3 A vn: linear
// in the normal velocity.  With vm = -vc and s = vm.n1: vn = -s n1, the
// condition is s > 1e-12 and q = n1 (-3 s A).  Accumulated in the
// quadratic-drag slot; there is no skin friction.  The wind (the atmosphere's
// velocity fields) is sampled at the centroid as for the currents.
#define SFH_GLM_VEC3_LENGTH 3.0f
template <class Flow, class AccT>
SFH_HD void face_aero(const FaceBody& B, f3 p1, f3 p2, f3 p3, const Flow& flow, AccT& acc)
{
    f3 e1 = p2 - p1, e2 = p3 - p1;
    f3 n = cross(e1, e2);
    float n2 = dot(n, n);
    if (n2 < SFH_DEGENERATE_N2) return;
    float inv = frsq(n2);
    f3 n1 = n * inv;
    float A = 0.5f * (n2 * inv);
    f3 c = p1 + (e1 + e2) * (1.0f / 3.0f);
    f3 vm = B.vb + cross(B.w, c);
    if (Flow::on)
        vm = vm - flow(((double)p1.x + (double)p2.x + (double)p3.x) / 3.0,
                       ((double)p1.y + (double)p2.y + (double)p3.y) / 3.0,
                       ((double)p1.z + (double)p2.z + (double)p3.z) / 3.0);
    float s = dot(vm, n1);
    if (s > -SFH_APPROACH_VN) {
        f3 q = n1 * (-(s * SFH_GLM_VEC3_LENGTH * A));
        add3(acc.Fdq, q);
        add3(acc.Mdq, cross(c, q));
    }
}

// Result of clipping one face against the waterline: the wet polygon as a
// base vertex plus an offset to its centroid, its unit normal and area, and
// whether the face contributes.
struct Wet {
    f3 base, off;   // centroid = base + off
    f3 n1;
    float A;
    bool ok;
};

// Tetra term of the submerged-volume integral for a triangle with normal
// vector n (|n| = twice its area) through the point base + q:
// 6V = (base - o + q) . n, and the centroid term (3(base - o) + s)/4 * 6V where
// s is the sum of the triangle's offsets from base.  CB false: the centroid
// term is not wanted (a surface with waves takes the buoyancy from the
// pressure integral, the volume only for Vsub).
template <typename T, bool CB = true>
SFH_HD void tetra_term(f3 bo, f3 q, f3 n, f3 s, Acc<T>& acc)
{
    float v6 = dot(bo + q, n);
    if (CB) {
        f3 g = (bo * 3.0f + s) * 0.25f;
        add3(acc.CB, g * v6);
    }
    acc.V6 += (T)v6;
}

// tetra_term with q = 0 (the polygon's first point is its base vertex):
// the same values without the three additions of zero.
template <typename T, bool CB = true>
SFH_HD void tetra_term0(f3 bo, f3 n, f3 s, Acc<T>& acc)
{
    float v6 = dot(bo, n);
    if (CB) {
        f3 g = (bo * 3.0f + s) * 0.25f;
        add3(acc.CB, g * v6);
    }
    acc.V6 += (T)v6;
}

// Surface face: waterline clip (7 cases, SolidEntity.cpp:1362-1647), tetra
// volume, face properties.  d0..d2 are the vertex depths (>= 0 is wet).  The
// tetra volume is accumulated before the degeneracy test, as in the reference.
//
// Every polygon is carried as a wet base vertex plus offsets built from mesh
// edges: a clip point is base + (edge * t), so edges, normals and areas of thin
// wet slivers keep FP32 precision relative to the sliver, not to the body.
template <typename T, bool CB = true>
SFH_HD Wet face_clip(const FaceBody& B, f3 p1, f3 p2, f3 p3, float d0, float d1, float d2, Acc<T>& acc)
{
    Wet r;
    r.ok = false;
    f3 base, fv1, fv2, fv3, fv4, csum;
    bool quad = false;
    if (d0 < 0.f) {
        if (d1 < 0.f) {                 // A: only p3 wet (1364-1398)
            base = p3;
            f3 q1 = (p1 - p3) * (d2 / (fabsf(d0) + d2));
            f3 q2 = (p2 - p3) * (d2 / (fabsf(d1) + d2));
            fv1 = q2 - q1; fv2 = q1 * -1.0f;
            f3 n = cross(fv1, fv2);
            tetra_term<T, CB>(base - B.o, q1, n, q1 + q2, acc);
            csum = q1 + q2;
        } else if (d2 < 0.f) {          // B: only p2 wet (1399-1433)
            base = p2;
            f3 q1 = (p1 - p2) * (d1 / (fabsf(d0) + d1));
            f3 q3 = (p3 - p2) * (d1 / (fabsf(d2) + d1));
            fv1 = q1 * -1.0f; fv2 = q3 - q1;
            f3 n = cross(fv1, fv2);
            tetra_term<T, CB>(base - B.o, q1, n, q1 + q3, acc);
            csum = q1 + q3;
        } else {                        // C: only p1 dry, quad p1' p2 p3 p4 (1434-1482)
            base = p2;
            f3 e23 = p3 - p2;
            f3 q1 = (p1 - p2) * (d1 / (fabsf(d0) + d1));
            f3 q4 = e23 + (p1 - p3) * (d2 / (fabsf(d0) + d2));
            f3 bo = base - B.o;
            tetra_term<T, CB>(bo, q1, cross(q1 * -1.0f, e23 - q1), q1 + e23, acc);      // (p1', p2, p3)
            tetra_term<T, CB>(bo, q1, cross(e23 - q1, q4 - q1), q1 + e23 + q4, acc);   // (p1', p3, p4)
            fv1 = q1 * -1.0f; fv2 = q4 - q1; fv3 = e23 * -1.0f; fv4 = q4 - e23;
            csum = q1 + e23 + q4;
            quad = true;
        }
    } else if (d1 < 0.f) {
        if (d2 < 0.f) {                 // D: only p1 wet (1486-1520)
            base = p1;
            f3 q2 = (p2 - p1) * (d0 / (fabsf(d1) + d0));
            f3 q3 = (p3 - p1) * (d0 / (fabsf(d2) + d0));
            fv1 = q2; fv2 = q3;
            f3 n = cross(fv1, fv2);
            tetra_term0<T, CB>(base - B.o, n, q2 + q3, acc);
            csum = q2 + q3;
        } else {                        // E: only p2 dry, quad p1 p2' p4 p3 (1521-1568)
            base = p1;
            f3 q3 = p3 - p1;
            f3 q2 = (p2 - p1) * (d0 / (fabsf(d1) + d0));
            f3 q4 = q3 + (p2 - p3) * (d2 / (fabsf(d1) + d2));
            f3 bo = base - B.o;
            tetra_term0<T, CB>(bo, cross(q2, q3), q2 + q3, acc);     // (p1, p2', p3)
            tetra_term<T, CB>(bo, q2, cross(q4 - q2, q3 - q2), q2 + q4 + q3, acc);     // (p2', p4, p3)
            fv1 = q2; fv2 = q3; fv3 = q2 - q3; fv4 = q4 - q3;
            csum = q2 + q3 + q4;
            quad = true;
        }
    } else if (d2 < 0.f) {              // F: only p3 dry, quad p1 p2 p3' p4 (1570-1617)
        base = p1;
        f3 q2 = p2 - p1;
        f3 q3 = q2 + (p3 - p2) * (d1 / (fabsf(d2) + d1));
        f3 q4 = (p3 - p1) * (d0 / (fabsf(d2) + d0));
        f3 bo = base - B.o;
        tetra_term0<T, CB>(bo, cross(q2, q3), q2 + q3, acc);         // (p1, p2, p3')
        tetra_term0<T, CB>(bo, cross(q3, q4), q3 + q4, acc);         // (p1, p3', p4)
        fv1 = q2; fv2 = q4; fv3 = q2 - q3; fv4 = q4 - q3;
        csum = q2 + q3 + q4;
        quad = true;
    } else {                            // G: all wet (1618-1647)
        base = p1;
        fv1 = p2 - p1; fv2 = p3 - p1;
        f3 n = cross(fv1, fv2);
        tetra_term0<T, CB>(base - B.o, n, fv1 + fv2, acc);
        csum = fv1 + fv2;
    }
    f3 n = cross(fv1, fv2);
    float n2 = dot(n, n);
    if (n2 < SFH_DEGENERATE_N2) return r;
    float inv = frsq(n2);
    float len = n2 * inv;
    r.n1 = n * inv;
    r.base = base;
    if (quad) {
        f3 n34 = cross(fv3, fv4);
        r.A = 0.5f * (len + fsqrt(dot(n34, n34)));
        r.off = csum * 0.25f;
    } else {
        r.A = 0.5f * len;
        r.off = csum * (1.0f / 3.0f);
    }
    r.ok = true;
    return r;
}

// All-wet face (case G of face_clip, every vertex depth >= 0): the same
// operations in the same order, without the case analysis.
template <typename T, bool CB = true>
SFH_HD Wet face_allwet(const FaceBody& B, f3 p1, f3 p2, f3 p3, Acc<T>& acc)
{
    Wet r;
    r.ok = false;
    const f3 base = p1;
    const f3 fv1 = p2 - p1, fv2 = p3 - p1;
    f3 n = cross(fv1, fv2);
    tetra_term0<T, CB>(base - B.o, n, fv1 + fv2, acc);
    const f3 csum = fv1 + fv2;
    n = cross(fv1, fv2);
    float n2 = dot(n, n);
    if (n2 < SFH_DEGENERATE_N2) return r;
    float inv = frsq(n2);
    float len = n2 * inv;
    r.n1 = n * inv;
    r.base = base;
    r.A = 0.5f * len;
    r.off = csum * (1.0f / 3.0f);
    r.ok = true;
    return r;
}

}  // namespace sfh
