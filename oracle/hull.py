"""CPU restatement of the reference's hull preprocessing, in FP64 Python.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's
sfh_compute_physical_properties / sfh_compute_fluid_dynamics_approx.  Never
used by the product (stonefish_amd/).

Follows, operation by operation:
  ComputePhysicalProperties   Library/src/utils/GeometryFileUtil.cpp:335-563
  FindInertialAxis, IsDiagonal GeometryFileUtil.cpp:572-619
  ComputeFluidDynamicsApprox  Library/src/entities/SolidEntity.cpp:698-718
  ComputeSphericalApprox      SolidEntity.cpp:720-758
  ComputeCylindricalApprox    SolidEntity.cpp:760-847
  ComputeEllipsoidalApprox    SolidEntity.cpp:849-1026
  LambKFactor                 SolidEntity.cpp:1028-1039
  SetHydrodynamicCoefficients SolidEntity.cpp:192-198
  CorrectAerodynamicForces    SolidEntity.cpp:1920-1948
Bullet conventions (BT_USE_DOUBLE_PRECISION, BT_EULER_DEFAULT_ZYX):
btVector3 division multiplies by the reciprocal (btVector3.h), cofactor
inverse (btMatrix3x3.h:1093-1103), setRotation (btMatrix3x3.h:215-282),
setEulerZYX (btQuaternion.h:140-155).  Scalar loops in plain Python keep the
reference's summation order.
"""
from __future__ import annotations

import math

AUTO, SPHERE, CYLINDER, ELLIPSOID = 0, 1, 2, 3


def _add(a, b):
    return [a[0] + b[0], a[1] + b[1], a[2] + b[2]]


def _sub(a, b):
    return [a[0] - b[0], a[1] - b[1], a[2] - b[2]]


def _mul(a, s):
    return [a[0] * s, a[1] * s, a[2] * s]


def _div(a, s):
    return _mul(a, 1.0 / s)


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def _cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def _normalized(a):
    return _div(a, math.sqrt(_dot(a, a)))


def _mv(M, v):
    return [_dot(M[0], v), _dot(M[1], v), _dot(M[2], v)]


def _inverse(M):
    def cof(r1, c1, r2, c2):
        return M[r1][c1] * M[r2][c2] - M[r1][c2] * M[r2][c1]
    co = [cof(1, 1, 2, 2), cof(1, 2, 2, 0), cof(1, 0, 2, 1)]
    s = 1.0 / _dot(M[0], co)
    return [[co[0] * s, cof(0, 2, 2, 1) * s, cof(0, 1, 1, 2) * s],
            [co[1] * s, cof(0, 0, 2, 2) * s, cof(0, 2, 1, 0) * s],
            [co[2] * s, cof(0, 1, 2, 0) * s, cof(0, 0, 1, 1) * s]]


def _basis(gl):
    return [[float(gl[4 * c + r]) for c in range(3)] for r in range(3)]


def _quat_matrix(x, y, z, w):
    d = x * x + y * y + z * z + w * w
    s = 2.0 / d
    xs, ys, zs = x * s, y * s, z * s
    wx, wy, wz = w * xs, w * ys, w * zs
    xx, xy, xz = x * xs, x * ys, x * zs
    yy, yz, zz = y * ys, y * zs, z * zs
    return [[1.0 - (yy + zz), xy - wz, xz + wy], [xy + wz, 1.0 - (xx + zz), yz - wx], [xz - wy, yz + wx, 1.0 - (xx + yy)]]


def euler_zyx_matrix(yaw_z, pitch_y, roll_x):
    """Matrix of Quaternion(yaw, pitch, roll) under BT_EULER_DEFAULT_ZYX."""
    hy, hp, hr = yaw_z * 0.5, pitch_y * 0.5, roll_x * 0.5
    cy, sy, cp, sp, cr, sr = math.cos(hy), math.sin(hy), math.cos(hp), math.sin(hp), math.cos(hr), math.sin(hr)
    return _quat_matrix(sr * cp * cy - cr * sp * sy, cr * sp * cy + sr * cp * sy, cr * cp * sy - sr * sp * cy,
                        cr * cp * cy + sr * sp * sy)


def _inertia(V, F, density, CG, off):
    Pxx = Pyy = Pzz = Pxy = Pxz = Pyz = 0.0
    for f in F:
        v1, v2, v3 = (list(map(float, V[i])) for i in f)
        if off != 0.0:
            n = _normalized(_cross(_sub(v2, v1), _sub(v3, v1)))
            h = _mul(n, off)
            h = _div(h, 2.0)
            v1, v2, v3 = (_sub(_add(v, h), CG) for v in (v1, v2, v3))
        else:
            v1, v2, v3 = (_sub(v, CG) for v in (v1, v2, v3))
        V6 = _dot(v1, _cross(v2, v3))
        X = [v1[0], v2[0], v3[0]]
        Y = [v1[1], v2[1], v3[1]]
        Z = [v1[2], v2[2], v3[2]]

        def sq(a):
            return a[0] * a[0] + a[1] * a[1] + a[2] * a[2] + a[0] * a[1] + a[0] * a[2] + a[1] * a[2]

        def pr(a, b):
            return (2 * (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) + a[0] * b[1] + a[1] * b[0] + a[0] * b[2]
                    + a[2] * b[0] + a[1] * b[2] + a[2] * b[1])
        Pxx += V6 * 2 * sq(X)
        Pyy += V6 * 2 * sq(Y)
        Pzz += V6 * 2 * sq(Z)
        Pxy += V6 * pr(X, Y)
        Pxz += V6 * pr(X, Z)
        Pyz += V6 * pr(Y, Z)
    k = density / 120.0
    Pxx, Pyy, Pzz, Pxy, Pxz, Pyz = (p * k for p in (Pxx, Pyy, Pzz, Pxy, Pxz, Pyz))
    return [[Pyy + Pzz, -Pxy, -Pxz], [-Pxy, Pxx + Pzz, -Pyz], [-Pxz, -Pyz, Pxx + Pyy]]


def _axis(I, value):
    L = [row[:] for row in I]
    for k in range(3):
        L[k][k] = I[k][k] - value
    c = [_cross(L[0], L[1]), _cross(L[0], L[2]), _cross(L[1], L[2])]
    n2 = [_dot(x, x) for x in c]
    if n2[0] >= n2[1]:
        return _normalized(c[0] if n2[0] >= n2[2] else c[2])
    return _normalized(c[1] if n2[1] >= n2[2] else c[2])


def _is_diagonal(A):
    mn = min(min(abs(A[0][0]), abs(A[1][1])), abs(A[2][2]))
    t = mn / 1e6
    return all(abs(A[r][c]) < t for r in range(3) for c in range(3) if r != c)


def physical_properties(V, F, thickness=0.0, density=1000.0):
    """ComputePhysicalProperties: dict with mass, volume, surface, CG, Ipri, Irot."""
    CG = [0.0, 0.0, 0.0]
    volume = surface = 0.0
    for f in F:
        v1, v2, v3 = (list(map(float, V[i])) for i in f)
        if thickness > 0.0:
            A = math.sqrt(_dot(*(2 * [_cross(_sub(v2, v1), _sub(v3, v1))]))) / 2.0
            surface += A
            tri = _div(_add(_add(v1, v2), v3), 3.0)
            tv = A * thickness
            CG = _add(CG, _mul(tri, tv))
            volume += tv
        else:
            c = _cross(_sub(v2, v1), _sub(v3, v1))
            surface += math.sqrt(_dot(c, c)) / 2.0
            tet = _div(_add(_add(v1, v2), v3), 4.0)
            v6 = _dot(v1, _cross(v2, v3))
            CG = _add(CG, _mul(tet, v6))
            volume += v6
    CG = _div(CG, volume) if volume > 0.0 else [0.0, 0.0, 0.0]
    if not thickness > 0.0:
        volume /= 6.0
    if thickness > 0.0:
        Io = _inertia(V, F, density, CG, thickness)
        Ii = _inertia(V, F, density, CG, -thickness)
        I = [[Io[r][c] - Ii[r][c] for c in range(3)] for r in range(3)]
    else:
        I = _inertia(V, F, density, CG, 0.0)
    Ipri = [I[0][0], I[1][1], I[2][2]]
    Irot = [[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]]
    if not _is_diagonal(I):
        T = I[0][0] + I[1][1] + I[2][2]
        II = (I[0][0] * I[1][1] + I[0][0] * I[2][2] + I[1][1] * I[2][2] - I[0][1] * I[0][1] - I[0][2] * I[0][2]
              - I[1][2] * I[1][2])
        U = math.sqrt(T * T - 3.0 * II) / 3.0
        det = _dot(I[0], _cross(I[1], I[2]))
        theta = math.acos((-2.0 * T * T * T + 9.0 * T * II - 27.0 * det) / (54.0 * U * U * U))
        A = T / 3.0 - 2.0 * U * math.cos(theta / 3.0)
        B = T / 3.0 - 2.0 * U * math.cos(theta / 3.0 - 2.0 * math.pi / 3.0)
        C = T / 3.0 - 2.0 * U * math.cos(theta / 3.0 + 2.0 * math.pi / 3.0)
        Ipri = [A, B, C]
        a1 = _axis(I, A)
        a2 = _axis(I, B)
        a3 = _cross(a1, a2)
        a2 = _cross(a3, a1)
        Irot = [[a1[0], a2[0], a3[0]], [a1[1], a2[1], a3[1]], [a1[2], a2[2], a3[2]]]
    return {"mass": volume * density, "volume": volume, "surface": surface, "CG": CG, "Ipri": Ipri, "Irot": Irot}


def lamb_k(r1, r2):
    e = 1.0 - r2 * r2 / r1
    elog = (1.0 + e) / (1.0 - e)
    if elog > 0.0:
        alpha0 = 2.0 * (1.0 - e * e) / (e * e) * (0.5 * math.log((1.0 + e) / (1.0 - e)) - e)
        return alpha0 / (2.0 - alpha0)
    return 1.0


def _coefficients(Cd):
    out = {}
    Cf = _mul(Cd, 0.1)
    if all(c >= 0 for c in Cd):
        out["Cd"] = Cd
    if all(c >= 0 for c in Cf):
        out["Cf"] = Cf
    return out


def fluid_dynamics_approx(V, kind=AUTO, T_CG2C=None, P_CB=(0.0, 0.0, 0.0), T_CG2O=None, rho=1000.0):
    """ComputeFluidDynamicsApprox of a rigid body with physics-mesh vertices V."""
    eye = [1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0]
    a = T_CG2C if T_CG2C is not None else eye
    B, t = _basis(a), [float(a[12]), float(a[13]), float(a[14])]
    BO = _basis(T_CG2O if T_CG2O is not None else eye)
    pcb = [float(p) for p in P_CB]
    x = [_sub(_add(_mv(B, list(map(float, v))), t), pcb) for v in V]
    n = len(x)
    ident = [[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]]
    if kind == SPHERE:
        r = 0.0
        for p in x:
            rc = _dot(p, p)
            if rc > r:
                r = rc
        m = 2 * math.pi * rho * r * r * r / 3.0
        res = {"type": SPHERE, "iterations": 0, "params": [math.sqrt(r), 0.0, 0.0], "H": ident,
               "added_mass": [m, m, m], "added_inertia": [0.0, 0.0, 0.0]}
        res.update(_coefficients([1.0, 1.0, 1.0]))
    elif kind == CYLINDER:
        r = [0.0, 0.0, 0.0]
        for p in x:
            for ax, (i, j) in enumerate(((1, 2), (0, 2), (0, 1))):
                d = math.sqrt(p[i] * p[i] + p[j] * p[j])
                r[ax] = d if d > r[ax] else r[ax]
        if r[0] <= r[1] and r[0] <= r[2]:
            axis = 0
        elif r[1] <= r[0] and r[1] <= r[2]:
            axis = 1
        else:
            axis = 2
        l_2 = 0.0
        for p in x:
            d = abs(p[axis])
            l_2 = d if d > l_2 else l_2
        H = [euler_zyx_matrix(0, math.pi / 2, 0), euler_zyx_matrix(0, 0, math.pi / 2), ident][axis]
        p0, p1 = r[axis], l_2 * 2.0
        m1 = rho * math.pi * p0 * p0
        m2 = rho * math.pi * p0 * p0 * p1
        I2 = 1.0 / 12.0 * math.pi * rho * p1 * p1 * p0 ** 3
        res = {"type": CYLINDER, "iterations": 0, "params": [p0, p1, 0.0], "H": H,
               "added_mass": [abs(c) for c in _mv(H, [m2, m2, m1])],
               "added_inertia": [abs(c) for c in _mv(H, [I2, I2, 0.0])]}
        res.update(_coefficients([abs(c) for c in _mv(_inverse(BO), _mv(H, [0.5, 0.5, 1.0]))]))
    else:
        x0 = []
        for k in range(3):
            col = [p[k] for p in x]
            lo = min(range(n), key=lambda i: (col[i], i))               # the first smallest
            hi = max(range(n), key=lambda i: (col[i], i))               # the last largest
            x0 += [x[lo], x[hi]]
        sigma = [1.0 / 6.0 if any(p == q for q in x0) else 0.0 for p in x]

        def uv():
            u, v = [], []
            for j in range(3):
                su = sv = 0.0
                for i in range(n):
                    su += sigma[i] * x[i][j] * x[i][j]
                    sv += sigma[i] * x[i][j]
                u.append(su)
                v.append(sv)
            return u, v

        def lam(i, u, v):
            s = 0.0
            for j in range(3):
                s += (x[i][j] - v[j]) * (x[i][j] - v[j]) / (3 * (u[j] - v[j] * v[j]))
            return s

        def best(u, v):
            ls = [lam(i, u, v) for i in range(n)]
            i = max(range(n), key=lambda k: (ls[k], -k))                 # the first maximum
            return i, ls[i]

        u, v = uv()
        istar, l = best(u, v)
        eps = l - 1.0
        tol = (1.0 + 0.2) ** (2.0 / 3.0) - 1.0
        k = 0
        while eps > tol and k < 10:
            x0.append(x[istar])
            beta = eps / (4.0 * (1.0 + eps))
            sigma = [(1.0 - beta) * s for s in sigma]
            sigma[istar] += beta
            u, v = uv()
            istar, l = best(u, v)
            eps = l - 1.0
            k += 1
        if k == 0:
            c = [(x0[0][0] + x0[1][0]) / 2.0, (x0[2][1] + x0[3][1]) / 2.0, (x0[4][2] + x0[5][2]) / 2.0]
            d = [abs(x0[0][0] - c[0]), abs(x0[2][1] - c[1]), abs(x0[4][2] - c[2])]
        else:
            d = []
            for j in range(3):
                dj = 1.0 / (3.0 * (u[j] - v[j] * v[j]))
                d.append(1.0 / math.sqrt(dj))
        r12 = (d[1] + d[2]) / 2.0
        res = {"type": ELLIPSOID, "iterations": k, "params": d, "H": ident,
               "added_mass": [lamb_k(d[0], r12) * 4.0 / 3.0 * math.pi * rho * d[0] * r12 * r12,
                              4.0 / 3.0 * math.pi * rho * d[2] * d[2] * d[0],
                              4.0 / 3.0 * math.pi * rho * d[1] * d[1] * d[0]],
               "added_inertia": [0.0, 1.0 / 12.0 * math.pi * rho * d[1] * d[1] * d[0] ** 3,
                                 1.0 / 12.0 * math.pi * rho * d[2] * d[2] * d[0] ** 3]}
        Cd = [1.0 / d[0], 1.0 / d[1], 1.0 / d[2]]
        Cd = _div(Cd, max(max(Cd[0], Cd[1]), Cd[2]))
        res.update(_coefficients([abs(c) for c in _mv(_inverse(BO), Cd)]))
    H = res.pop("H")
    res["T_CG2H"] = [H[0][0], H[1][0], H[2][0], 0.0, H[0][1], H[1][1], H[2][1], 0.0,
                     H[0][2], H[1][2], H[2][2], 0.0, pcb[0], pcb[1], pcb[2], 1.0]
    return res


def correct_aero(kind, params, T_CG2H, Fda, Tda):
    """CorrectAerodynamicForces: (Fda, Tda) scaled by |Cd . safeNormalize(inv(H) Fda)|."""
    d = _mv(_inverse(_basis(T_CG2H)), [float(x) for x in Fda])
    l2 = _dot(d, d)
    eps = 2.220446049250313e-16
    d = _div(d, math.sqrt(l2)) if l2 >= eps * eps else [1.0, 0.0, 0.0]
    cor = 1.0
    if kind == CYLINDER:
        cor = _dot([0.5, 0.5, 1.0], d)
    elif kind == ELLIPSOID:
        Cd = [1.0 / params[0], 1.0 / params[1], 1.0 / params[2]]
        Cd = _div(Cd, max(max(Cd[0], Cd[1]), Cd[2]))
        cor = _dot(Cd, d)
    cor = abs(cor)
    return [f * cor for f in Fda], [t * cor for t in Tda]
