// Driver for oracle/_ref (TEST INFRASTRUCTURE ONLY): exposes the reference's own
// Bullet LinearMath routines, compiled from the reference checkout's headers
// with BT_USE_DOUBLE_PRECISION, through a C ABI so the tests can pin the
// oracle's LinearMath restatement against them.  The driver code is this
// repository's; the routines it calls are the reference's.
#include "LinearMath/btTransform.h"
#include "LinearMath/btAabbUtil2.h"
#include "LinearMath/btQuaternion.h"

extern "C" {

void btref_quat_to_gl(const double q[4], const double origin[3], double m[16])
{
    btTransform t(btQuaternion(q[0], q[1], q[2], q[3]), btVector3(origin[0], origin[1], origin[2]));
    t.getOpenGLMatrix(m);
}

void btref_gl_mul(const double a[16], const double b[16], double out[16])
{
    btTransform ta, tb;
    ta.setFromOpenGLMatrix(a);
    tb.setFromOpenGLMatrix(b);
    (ta * tb).getOpenGLMatrix(out);
}

void btref_gl_inverse(const double a[16], double out[16])
{
    btTransform ta;
    ta.setFromOpenGLMatrix(a);
    ta.inverse().getOpenGLMatrix(out);
}

void btref_basis_inverse(const double m[16], double inv[9])
{
    btTransform t;
    t.setFromOpenGLMatrix(m);
    btMatrix3x3 r = t.getBasis().inverse();
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) inv[3 * i + j] = r[i][j];
}

void btref_safe_normalize(double v[3])
{
    btVector3 x(v[0], v[1], v[2]);
    x.safeNormalize();
    v[0] = x.x(); v[1] = x.y(); v[2] = x.z();
}

void btref_transform_aabb(const double lmin[3], const double lmax[3], double margin,
                          const double m[16], double omin[3], double omax[3])
{
    btTransform t;
    t.setFromOpenGLMatrix(m);
    btVector3 a, b;
    btTransformAabb(btVector3(lmin[0], lmin[1], lmin[2]), btVector3(lmax[0], lmax[1], lmax[2]), margin, t, a, b);
    for (int k = 0; k < 3; ++k) { omin[k] = a[k]; omax[k] = b[k]; }
}

// btMatrix3x3 * btVector3 (row dot products), as CorrectHydrodynamicForces uses
void btref_basis_mul_vec(const double inv9[9], const double x[3], double out[3])
{
    btMatrix3x3 M(inv9[0], inv9[1], inv9[2], inv9[3], inv9[4], inv9[5], inv9[6], inv9[7], inv9[8]);
    btVector3 r = M * btVector3(x[0], x[1], x[2]);
    out[0] = r.x(); out[1] = r.y(); out[2] = r.z();
}

}
