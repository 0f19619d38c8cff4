// Host-side per-body work: the hull preprocessing (once per hull) — the mesh's
// physical properties (ComputePhysicalProperties, GeometryFileUtil.cpp:
// 335-563) and the fluid-dynamics proxy with its drag coefficients
// (SolidEntity::ComputeFluidDynamicsApprox, SolidEntity.cpp:698-1039) — and
// the proxy's aerodynamic correction (SolidEntity::CorrectAerodynamicForces,
// SolidEntity.cpp:1920-1948).
//
// Everything is FP64 with the reference's operation order, so the results
// follow the reference's Scalar = double arithmetic: Bullet's conventions are
// kept where they change the last bits (a vector divided by a scalar is
// multiplied by the reciprocal; a matrix inverse is by cofactors).  The
// minimum-volume ellipsoid iteration evaluates the sums u_j, v_j once per
// iteration instead of once per point: the same sums in the same order.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/sfhydro.h"

int sfh_internal_fail(int code, const char* msg);   // sfh_hydro.hip: sets sfh_last_error

namespace {

struct V3 {
    double x, y, z;
    double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    double& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};

V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
V3 operator/(V3 a, double s) { return a * (1.0 / s); }   // btVector3: times the reciprocal
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double len2(V3 a) { return dot(a, a); }
V3 normalized(V3 a) { return a / std::sqrt(len2(a)); }
V3 vabs(V3 a) { return {std::fabs(a.x), std::fabs(a.y), std::fabs(a.z)}; }

struct M3 {
    double m[3][3];
    V3 row(int r) const { return {m[r][0], m[r][1], m[r][2]}; }
    V3 operator*(V3 v) const { return {dot(row(0), v), dot(row(1), v), dot(row(2), v)}; }
    double cofac(int r1, int c1, int r2, int c2) const { return m[r1][c1] * m[r2][c2] - m[r1][c2] * m[r2][c1]; }
    double det() const { return dot(row(0), cross(row(1), row(2))); }   // btTriple
    M3 inverse() const                                                  // btMatrix3x3::inverse
    {
        const V3 co = {cofac(1, 1, 2, 2), cofac(1, 2, 2, 0), cofac(1, 0, 2, 1)};
        const double s = 1.0 / dot(row(0), co);
        return M3{{{co.x * s, cofac(0, 2, 2, 1) * s, cofac(0, 1, 1, 2) * s},
                   {co.y * s, cofac(0, 0, 2, 2) * s, cofac(0, 2, 1, 0) * s},
                   {co.z * s, cofac(0, 1, 2, 0) * s, cofac(0, 0, 1, 1) * s}}};
    }
};

M3 basis_gl(const double* g)   // OpenGL column-major 4x4 -> rows
{
    M3 b;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) b.m[r][c] = g[4 * c + r];
    return b;
}

// btMatrix3x3::setRotation (btMatrix3x3.h:215-282, scalar path)
M3 from_quat(double qx, double qy, double qz, double qw)
{
    const double d = qx * qx + qy * qy + qz * qz + qw * qw;
    const double s = 2.0 / d;
    const double xs = qx * s, ys = qy * s, zs = qz * s;
    const double wx = qw * xs, wy = qw * ys, wz = qw * zs;
    const double xx = qx * xs, xy = qx * ys, xz = qx * zs;
    const double yy = qy * ys, yz = qy * zs, zz = qz * zs;
    return M3{{{1.0 - (yy + zz), xy - wz, xz + wy}, {xy + wz, 1.0 - (xx + zz), yz - wx}, {xz - wy, yz + wx, 1.0 - (xx + yy)}}};
}

// Quaternion(yaw, pitch, roll) with BT_EULER_DEFAULT_ZYX (the reference's
// Bullet flags): setEulerZYX (btQuaternion.h:140-155).
M3 euler_zyx(double yawZ, double pitchY, double rollX)
{
    const double hy = yawZ * 0.5, hp = pitchY * 0.5, hr = rollX * 0.5;
    const double cy = std::cos(hy), sy = std::sin(hy), cp = std::cos(hp), sp = std::sin(hp);
    const double cr = std::cos(hr), sr = std::sin(hr);
    return from_quat(sr * cp * cy - cr * sp * sy, cr * sp * cy + sr * cp * sy, cr * cp * sy - sr * sp * cy,
                     cr * cp * cy + sr * sp * sy);
}

void to_gl(const M3& b, V3 t, double* g)
{
    for (int k = 0; k < 16; ++k) g[k] = 0.0;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) g[4 * c + r] = b.m[r][c];
    g[12] = t.x;
    g[13] = t.y;
    g[14] = t.z;
    g[15] = 1.0;
}

V3 vertex(const float* xyz, uint32_t i) { return {(double)xyz[3 * i], (double)xyz[3 * i + 1], (double)xyz[3 * i + 2]}; }

// Pjk sums of a tetra fan over the faces, vertices offset by `off` along the
// face normal (shells) and taken about the CG (GeometryFileUtil.cpp:411-446).
M3 solid_inertia(const float* xyz, const uint32_t* idx, uint32_t nf, double density, V3 CG, double off)
{
    double Pxx = 0, Pyy = 0, Pzz = 0, Pxy = 0, Pxz = 0, Pyz = 0;
    for (uint32_t f = 0; f < nf; ++f) {
        V3 v1 = vertex(xyz, idx[3 * f]), v2 = vertex(xyz, idx[3 * f + 1]), v3 = vertex(xyz, idx[3 * f + 2]);
        if (off != 0.0) {
            const V3 n = normalized(cross(v2 - v1, v3 - v1));
            v1 = v1 + n * off / 2.0 - CG;
            v2 = v2 + n * off / 2.0 - CG;
            v3 = v3 + n * off / 2.0 - CG;
        } else {
            v1 = v1 - CG;
            v2 = v2 - CG;
            v3 = v3 - CG;
        }
        const double V6 = dot(v1, cross(v2, v3));
        Pxx += V6 * 2 * (v1.x * v1.x + v2.x * v2.x + v3.x * v3.x + v1.x * v2.x + v1.x * v3.x + v2.x * v3.x);
        Pyy += V6 * 2 * (v1.y * v1.y + v2.y * v2.y + v3.y * v3.y + v1.y * v2.y + v1.y * v3.y + v2.y * v3.y);
        Pzz += V6 * 2 * (v1.z * v1.z + v2.z * v2.z + v3.z * v3.z + v1.z * v2.z + v1.z * v3.z + v2.z * v3.z);
        Pxy += V6 * (2 * (v1.x * v1.y + v2.x * v2.y + v3.x * v3.y) + v1.x * v2.y + v1.y * v2.x + v1.x * v3.y +
                     v1.y * v3.x + v2.x * v3.y + v2.y * v3.x);
        Pxz += V6 * (2 * (v1.x * v1.z + v2.x * v2.z + v3.x * v3.z) + v1.x * v2.z + v1.z * v2.x + v1.x * v3.z +
                     v1.z * v3.x + v2.x * v3.z + v2.z * v3.x);
        Pyz += V6 * (2 * (v1.y * v1.z + v2.y * v2.z + v3.y * v3.z) + v1.y * v2.z + v1.z * v2.y + v1.y * v3.z +
                     v1.z * v3.y + v2.y * v3.z + v2.z * v3.y);
    }
    const double k = density / 120.0;   // 1/20 from the formula, 1/6 from the volume
    Pxx *= k; Pyy *= k; Pzz *= k; Pxy *= k; Pxz *= k; Pyz *= k;
    return M3{{{Pyy + Pzz, -Pxy, -Pxz}, {-Pxy, Pxx + Pzz, -Pyz}, {-Pxz, -Pyz, Pxx + Pyy}}};
}

// FindInertialAxis (GeometryFileUtil.cpp:572-603)
V3 inertial_axis(const M3& I, double value)
{
    M3 L = I;
    for (int k = 0; k < 3; ++k) L.m[k][k] = I.m[k][k] - value;
    const V3 c0 = cross(L.row(0), L.row(1)), c1 = cross(L.row(0), L.row(2)), c2 = cross(L.row(1), L.row(2));
    if (len2(c0) >= len2(c1)) return len2(c0) >= len2(c2) ? normalized(c0) : normalized(c2);
    return len2(c1) >= len2(c2) ? normalized(c1) : normalized(c2);
}

// IsDiagonal (GeometryFileUtil.cpp:605-619)
bool is_diagonal(const M3& A)
{
    const double mn = std::min(std::min(std::fabs(A.m[0][0]), std::fabs(A.m[1][1])), std::fabs(A.m[2][2]));
    const double t = mn / 1e6;
    return std::fabs(A.m[0][1]) < t && std::fabs(A.m[0][2]) < t && std::fabs(A.m[1][0]) < t &&
           std::fabs(A.m[1][2]) < t && std::fabs(A.m[2][0]) < t && std::fabs(A.m[2][1]) < t;
}

// SolidEntity::LambKFactor (SolidEntity.cpp:1028-1039), as written there.
double lamb_k(double r1, double r2)
{
    const double e = 1.0 - r2 * r2 / r1;
    const double elog = (1.0 + e) / (1.0 - e);
    if (elog > 0.0) {
        const double alpha0 = 2.0 * (1.0 - e * e) / (e * e) * (0.5 * std::log((1.0 + e) / (1.0 - e)) - e);
        return alpha0 / (2.0 - alpha0);
    }
    return 1.0;
}

void set_coefficients(sfh_fluid_approx* out, V3 Cd)
{
    // SetHydrodynamicCoefficients (SolidEntity.cpp:192-198): Cf = 0.1 Cd
    const V3 Cf = Cd * 0.1;
    if (Cd.x >= 0 && Cd.y >= 0 && Cd.z >= 0)
        for (int k = 0; k < 3; ++k) out->Cd[k] = Cd[k];
    if (Cf.x >= 0 && Cf.y >= 0 && Cf.z >= 0)
        for (int k = 0; k < 3; ++k) out->Cf[k] = Cf[k];
}

// ComputeSphericalApprox (SolidEntity.cpp:720-758)
void sphere_approx(const std::vector<V3>& x, V3 P_CB, double rho, sfh_fluid_approx* out)
{
    double r = 0.0;
    for (const V3& p : x) r = std::max(r, len2(p));   // squared radius
    out->type = SFH_APPROX_SPHERE;
    out->params[0] = std::sqrt(r);
    const double m = 2.0 * M_PI * rho * r * r * r / 3.0;   // the reference's: r is still the squared radius
    for (int k = 0; k < 3; ++k) { out->added_mass[k] = m; out->added_inertia[k] = 0.0; }
    M3 I = {{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}};
    to_gl(I, P_CB, out->T_CG2H);
    set_coefficients(out, V3{1, 1, 1});
}

// ComputeCylindricalApprox (SolidEntity.cpp:760-847)
void cylinder_approx(const std::vector<V3>& x, V3 P_CB, const M3& B_CG2O, double rho, sfh_fluid_approx* out)
{
    double r[3] = {0, 0, 0};
    for (const V3& p : x) {
        double d = std::sqrt(p.y * p.y + p.z * p.z);
        r[0] = d > r[0] ? d : r[0];
        d = std::sqrt(p.x * p.x + p.z * p.z);
        r[1] = d > r[1] ? d : r[1];
        d = std::sqrt(p.x * p.x + p.y * p.y);
        r[2] = d > r[2] ? d : r[2];
    }
    int axis = 2;
    if (r[0] <= r[1] && r[0] <= r[2]) axis = 0;
    else if (r[1] <= r[0] && r[1] <= r[2]) axis = 1;
    double l_2 = 0.0;
    for (const V3& p : x) {
        const double d = std::fabs(p[axis]);
        l_2 = d > l_2 ? d : l_2;
    }
    out->type = SFH_APPROX_CYLINDER;
    out->params[0] = r[axis];
    out->params[1] = l_2 * 2.0;
    const M3 BH = axis == 0 ? euler_zyx(0, M_PI_2, 0) : (axis == 1 ? euler_zyx(0, 0, M_PI_2)
                                                                     : M3{{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}});
    to_gl(BH, P_CB, out->T_CG2H);
    const double p0 = out->params[0], p1 = out->params[1];
    const double m1 = rho * M_PI * p0 * p0;        // parallel to the axis
    const double m2 = rho * M_PI * p0 * p0 * p1;   // perpendicular
    const double I1 = 0.0;
    const double I2 = 1.0 / 12.0 * M_PI * rho * p1 * p1 * std::pow(p0, 3.0);
    const V3 am = vabs(BH * V3{m2, m2, m1}), ai = vabs(BH * V3{I2, I2, I1});
    for (int k = 0; k < 3; ++k) { out->added_mass[k] = am[k]; out->added_inertia[k] = ai[k]; }
    set_coefficients(out, vabs(B_CG2O.inverse() * (BH * V3{0.5, 0.5, 1.0})));
}

// ComputeEllipsoidalApprox (SolidEntity.cpp:849-1026): P. Kumar, E. A.
// Yildirim, Computing minimum-volume enclosing axis-aligned ellipsoids,
// J. Optim. Theory Appl. 136 (2008) 211-228, as the reference runs it.
void ellipsoid_approx(const std::vector<V3>& x, V3 P_CB, const M3& B_CG2O, double rho, sfh_fluid_approx* out)
{
    const size_t n = x.size();
    std::vector<V3> x0;
    for (int k = 0; k < 3; ++k) {
        // std::minmax_element: the first smallest and the last largest
        size_t imin = 0, imax = 0;
        for (size_t i = 1; i < n; ++i) {
            if (x[i][k] < x[imin][k]) imin = i;
            if (!(x[i][k] < x[imax][k])) imax = i;
        }
        x0.push_back(x[imin]);
        x0.push_back(x[imax]);
    }
    auto same = [](V3 a, V3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; };
    std::vector<double> sigma(n, 0.0);
    for (size_t i = 0; i < n; ++i)
        for (const V3& q : x0)
            if (same(x[i], q)) { sigma[i] = 1.0 / 6.0; break; }
    double u[3], v[3];
    auto sums = [&]() {
        for (int j = 0; j < 3; ++j) {
            double su = 0.0, sv = 0.0;
            for (size_t i = 0; i < n; ++i) {
                su += sigma[i] * x[i][j] * x[i][j];
                sv += sigma[i] * x[i][j];
            }
            u[j] = su;
            v[j] = sv;
        }
    };
    auto lambda = [&](size_t i) {
        double s = 0.0;
        for (int j = 0; j < 3; ++j) s += (x[i][j] - v[j]) * (x[i][j] - v[j]) / (3 * (u[j] - v[j] * v[j]));
        return s;
    };
    auto argmax_lambda = [&](double& best) {   // std::max_element: the first maximum
        size_t is = 0;
        best = lambda(0);
        for (size_t i = 1; i < n; ++i) {
            const double l = lambda(i);
            if (l > best) { best = l; is = i; }
        }
        return is;
    };
    sums();
    double lmax;
    size_t iStar = argmax_lambda(lmax);
    double epsilon = lmax - 1.0;
    const double errorTol = 0.2;
    const double epsilonTol = std::pow(1.0 + errorTol, 2.0 / 3.0) - 1.0;
    const size_t maxIter = 10;
    size_t k = 0;
    while (epsilon > epsilonTol && k < maxIter) {
        x0.push_back(x[iStar]);
        const double beta = epsilon / (double(3 + 1) * (1.0 + epsilon));
        for (double& s : sigma) s = (1.0 - beta) * s;
        sigma[iStar] += beta;
        sums();
        iStar = argmax_lambda(lmax);
        epsilon = lmax - 1.0;
        ++k;
    }
    V3 c, d;
    if (k == 0) {
        c = {(x0[0].x + x0[1].x) / 2.0, (x0[2].y + x0[3].y) / 2.0, (x0[4].z + x0[5].z) / 2.0};
        d = {std::fabs(x0[0].x - c.x), std::fabs(x0[2].y - c.y), std::fabs(x0[4].z - c.z)};
    } else {
        c = {v[0], v[1], v[2]};
        for (int j = 0; j < 3; ++j) {
            d[j] = 1.0 / (3.0 * (u[j] - v[j] * v[j]));
            d[j] = 1.0 / std::sqrt(d[j]);
        }
    }
    (void)c;
    out->type = SFH_APPROX_ELLIPSOID;
    out->iterations = (int32_t)k;
    for (int j = 0; j < 3; ++j) out->params[j] = d[j];
    const double* a = out->params;
    const double r12 = (a[1] + a[2]) / 2.0;
    out->added_mass[0] = lamb_k(a[0], r12) * 4.0 / 3.0 * M_PI * rho * a[0] * r12 * r12;
    out->added_mass[1] = 4.0 / 3.0 * M_PI * rho * a[2] * a[2] * a[0];
    out->added_mass[2] = 4.0 / 3.0 * M_PI * rho * a[1] * a[1] * a[0];
    out->added_inertia[0] = 0.0;
    out->added_inertia[1] = 1.0 / 12.0 * M_PI * rho * a[1] * a[1] * std::pow(a[0], 3.0);
    out->added_inertia[2] = 1.0 / 12.0 * M_PI * rho * a[2] * a[2] * std::pow(a[0], 3.0);
    M3 I = {{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}};
    to_gl(I, P_CB, out->T_CG2H);
    V3 Cd = {1.0 / a[0], 1.0 / a[1], 1.0 / a[2]};
    const double maxCd = std::max(std::max(Cd.x, Cd.y), Cd.z);
    Cd = Cd / maxCd;
    set_coefficients(out, vabs(B_CG2O.inverse() * Cd));
}

}  // namespace

extern "C" int sfh_correct_aero(int type, const double params[3], const double T_CG2H[16], double Fda[3],
                                double Tda[3])
{
    if (!T_CG2H || !Fda || !Tda || (type >= SFH_APPROX_CYLINDER && !params))
        return sfh_internal_fail(SFH_ERR_INVALID, "null argument");
    V3 F = {Fda[0], Fda[1], Fda[2]};
    V3 d = basis_gl(T_CG2H).inverse() * F;
    // btVector3::safeNormalize (btVector3.h:286-299)
    const double l2 = len2(d);
    d = l2 >= 2.220446049250313e-16 * 2.220446049250313e-16 ? d / std::sqrt(l2) : V3{1.0, 0.0, 0.0};
    double cor = 1.0;
    if (type == SFH_APPROX_CYLINDER) {
        cor = dot(V3{0.5, 0.5, 1.0}, d);
    } else if (type == SFH_APPROX_ELLIPSOID) {
        V3 Cd = {1.0 / params[0], 1.0 / params[1], 1.0 / params[2]};
        Cd = Cd / std::max(std::max(Cd.x, Cd.y), Cd.z);
        cor = dot(Cd, d);
    }
    cor = std::fabs(cor);
    for (int k = 0; k < 3; ++k) {
        Fda[k] *= cor;
        Tda[k] *= cor;
    }
    return SFH_OK;
}

extern "C" int sfh_compute_physical_properties(const float* xyz, uint32_t nv, const uint32_t* idx, uint32_t nf,
                                               double thickness, double density, sfh_mesh_properties* out)
{
    if (!xyz || !idx || !out) return sfh_internal_fail(SFH_ERR_INVALID, "null argument");
    for (uint64_t i = 0; i < 3ull * nf; ++i)
        if (idx[i] >= nv) return sfh_internal_fail(SFH_ERR_INVALID, "vertex index out of range");
    V3 CG = {0, 0, 0};
    double volume = 0.0, surface = 0.0;
    for (uint32_t f = 0; f < nf; ++f) {
        const V3 v1 = vertex(xyz, idx[3 * f]), v2 = vertex(xyz, idx[3 * f + 1]), v3 = vertex(xyz, idx[3 * f + 2]);
        if (thickness > 0.0) {   // shell
            const double A = std::sqrt(len2(cross(v2 - v1, v3 - v1))) / 2.0;
            surface += A;
            const V3 triCG = (v1 + v2 + v3) / 3.0;
            const double triVolume = A * thickness;
            CG = CG + triCG * triVolume;
            volume += triVolume;
        } else {                 // solid: signed tetrahedra from the origin
            surface += std::sqrt(len2(cross(v2 - v1, v3 - v1))) / 2.0;
            const V3 tetraCG = (v1 + v2 + v3) / 4.0;
            const double tetraVolume6 = dot(v1, cross(v2, v3));
            CG = CG + tetraCG * tetraVolume6;
            volume += tetraVolume6;
        }
    }
    CG = volume > 0.0 ? CG / volume : V3{0, 0, 0};
    if (!(thickness > 0.0)) volume /= 6.0;
    M3 I;
    if (thickness > 0.0) {
        // outer solid minus inner solid (the reference doubts this itself)
        const M3 Io = solid_inertia(xyz, idx, nf, density, CG, thickness);
        const M3 Ii = solid_inertia(xyz, idx, nf, density, CG, -thickness);
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) I.m[r][c] = Io.m[r][c] - Ii.m[r][c];
    } else {
        I = solid_inertia(xyz, idx, nf, density, CG, 0.0);
    }
    out->mass = volume * density;
    out->volume = volume;
    out->surface = surface;
    for (int k = 0; k < 3; ++k) out->CG[k] = CG[k];
    V3 Ipri = {I.m[0][0], I.m[1][1], I.m[2][2]};
    M3 Irot = {{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}};
    if (!is_diagonal(I)) {
        const double T = I.m[0][0] + I.m[1][1] + I.m[2][2];
        const double II = I.m[0][0] * I.m[1][1] + I.m[0][0] * I.m[2][2] + I.m[1][1] * I.m[2][2] -
                          I.m[0][1] * I.m[0][1] - I.m[0][2] * I.m[0][2] - I.m[1][2] * I.m[1][2];
        const double U = std::sqrt(T * T - 3.0 * II) / 3.0;
        const double theta = std::acos((-2.0 * T * T * T + 9.0 * T * II - 27.0 * I.det()) / (54.0 * U * U * U));
        const double A = T / 3.0 - 2.0 * U * std::cos(theta / 3.0);
        const double B = T / 3.0 - 2.0 * U * std::cos(theta / 3.0 - 2.0 * M_PI / 3.0);
        const double C = T / 3.0 - 2.0 * U * std::cos(theta / 3.0 + 2.0 * M_PI / 3.0);
        Ipri = {A, B, C};
        const V3 axis1 = inertial_axis(I, A);
        V3 axis2 = inertial_axis(I, B);
        const V3 axis3 = cross(axis1, axis2);
        axis2 = cross(axis3, axis1);
        Irot = M3{{{axis1.x, axis2.x, axis3.x}, {axis1.y, axis2.y, axis3.y}, {axis1.z, axis2.z, axis3.z}}};
    }
    for (int k = 0; k < 3; ++k) out->Ipri[k] = Ipri[k];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) out->Irot[3 * r + c] = Irot.m[r][c];
    return SFH_OK;
}

extern "C" int sfh_compute_fluid_dynamics_approx(const float* xyz, uint32_t nv, int type, const double T_CG2C[16],
                                                 const double P_CB[3], const double T_CG2O[16], double rho,
                                                 sfh_fluid_approx* out)
{
    if (!xyz || !T_CG2C || !P_CB || !T_CG2O || !out) return sfh_internal_fail(SFH_ERR_INVALID, "null argument");
    if (type < SFH_APPROX_AUTO || type > SFH_APPROX_ELLIPSOID)
        return sfh_internal_fail(SFH_ERR_INVALID, "approximation type must be 0..3 (AUTO, SPHERE, CYLINDER, ELLIPSOID)");
    if (nv < 2) return sfh_internal_fail(SFH_ERR_SIZE, "at least two vertices");   // the reference returns early
    memset(out, 0, sizeof(*out));
    const M3 B = basis_gl(T_CG2C);
    const V3 t = {T_CG2C[12], T_CG2C[13], T_CG2C[14]};
    const V3 pcb = {P_CB[0], P_CB[1], P_CB[2]};
    std::vector<V3> x(nv);
    for (uint32_t i = 0; i < nv; ++i) x[i] = (B * vertex(xyz, i) + t) - pcb;   // CG frame, about the CoB
    const M3 BO = basis_gl(T_CG2O);
    switch (type) {
        case SFH_APPROX_SPHERE: sphere_approx(x, pcb, rho, out); break;
        case SFH_APPROX_CYLINDER: cylinder_approx(x, pcb, BO, rho, out); break;
        default: ellipsoid_approx(x, pcb, BO, rho, out); break;
    }
    return SFH_OK;
}
