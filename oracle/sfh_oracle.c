/*
 * CPU oracle for the Stonefish hydrodynamics path (TEST INFRASTRUCTURE ONLY:
 * used by tests/, the smoke check and bench.py's CPU-baseline leg, never by the
 * product).  See sfh_oracle.h for the reference files it restates.
 */
#define _GNU_SOURCE
#include "sfh_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <float.h>

/* ------------------------------------------------------------------ */
/* LinearMath restatement (double precision, BT_USE_DOUBLE_PRECISION)   */
/* ------------------------------------------------------------------ */

/* basis element (row r, col c) of an OpenGL column-major matrix */
#define GLB(m, r, c) ((m)[(c) * 4 + (r)])

/* btMatrix3x3::setRotation, scalar path (btMatrix3x3.h:215-282), then
 * btTransform::getOpenGLMatrix (btTransform.h:136-143). */
void orc_quat_to_gl(const double q[4], const double origin[3], double m[16])
{
    double d = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    double s = 2.0 / d;
    double xs = q[0] * s, ys = q[1] * s, zs = q[2] * s;
    double wx = q[3] * xs, wy = q[3] * ys, wz = q[3] * zs;
    double xx = q[0] * xs, xy = q[0] * ys, xz = q[0] * zs;
    double yy = q[1] * ys, yz = q[1] * zs, zz = q[2] * zs;
    double b[3][3] = {{1.0 - (yy + zz), xy - wz, xz + wy},
                      {xy + wz, 1.0 - (xx + zz), yz - wx},
                      {xz - wy, yz + wx, 1.0 - (xx + yy)}};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) GLB(m, r, c) = b[r][c];
    m[3] = m[7] = m[11] = 0.0;
    m[12] = origin[0]; m[13] = origin[1]; m[14] = origin[2]; m[15] = 1.0;
}

/* btTransform operator* (btTransform.h:231-235): basis = a.basis * b.basis,
 * origin = a(b.origin). */
void orc_gl_mul(const double a[16], const double b[16], double out[16])
{
    double r[16];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            GLB(r, i, j) = GLB(a, i, 0) * GLB(b, 0, j) + GLB(a, i, 1) * GLB(b, 1, j) + GLB(a, i, 2) * GLB(b, 2, j);
    for (int i = 0; i < 3; ++i)
        r[12 + i] = (GLB(a, i, 0) * b[12] + GLB(a, i, 1) * b[13] + GLB(a, i, 2) * b[14]) + a[12 + i];
    r[3] = r[7] = r[11] = 0.0;
    r[15] = 1.0;
    memcpy(out, r, sizeof(r));
}

/* btTransform::inverse: basis transposed, origin = inv * -origin. */
void orc_gl_inverse(const double a[16], double out[16])
{
    double r[16];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) GLB(r, i, j) = GLB(a, j, i);
    for (int i = 0; i < 3; ++i)
        r[12 + i] = GLB(r, i, 0) * -a[12] + GLB(r, i, 1) * -a[13] + GLB(r, i, 2) * -a[14];
    r[3] = r[7] = r[11] = 0.0;
    r[15] = 1.0;
    memcpy(out, r, sizeof(r));
}

static inline double cofac(const double* m, int r1, int c1, int r2, int c2)
{
    return GLB(m, r1, c1) * GLB(m, r2, c2) - GLB(m, r1, c2) * GLB(m, r2, c1);
}

/* btMatrix3x3::inverse (btMatrix3x3.h:1093-1103), cofactor form; result
 * row-major. */
void orc_basis_inverse(const double m[16], double inv[9])
{
    double co[3] = {cofac(m, 1, 1, 2, 2), cofac(m, 1, 2, 2, 0), cofac(m, 1, 0, 2, 1)};
    double det = GLB(m, 0, 0) * co[0] + GLB(m, 0, 1) * co[1] + GLB(m, 0, 2) * co[2];
    double s = 1.0 / det;
    inv[0] = co[0] * s; inv[1] = cofac(m, 0, 2, 2, 1) * s; inv[2] = cofac(m, 0, 1, 1, 2) * s;
    inv[3] = co[1] * s; inv[4] = cofac(m, 0, 0, 2, 2) * s; inv[5] = cofac(m, 0, 2, 1, 0) * s;
    inv[6] = co[2] * s; inv[7] = cofac(m, 0, 1, 2, 0) * s; inv[8] = cofac(m, 0, 0, 1, 1) * s;
}

/* btVector3::safeNormalize (btVector3.h:286-299), SIMD_EPSILON = DBL_EPSILON. */
void orc_safe_normalize(double v[3])
{
    double l2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    if (l2 >= DBL_EPSILON * DBL_EPSILON) {
        /* btVector3::operator/= multiplies by the reciprocal (btVector3.h) */
        double inv = 1.0 / sqrt(l2);
        v[0] *= inv; v[1] *= inv; v[2] *= inv;
    } else {
        v[0] = 1.0; v[1] = 0.0; v[2] = 0.0;
    }
}

/* btTransformAabb (btAabbUtil2.h:182-196); btCompoundShape::getAabb
 * (btCompoundShape.cpp:161-181) is the same computation with the compound's
 * margin. */
void orc_transform_aabb(const double lmin[3], const double lmax[3], double margin,
                        const double m[16], double omin[3], double omax[3])
{
    double h[3], c[3], wc[3], e[3];
    for (int k = 0; k < 3; ++k) {
        h[k] = 0.5 * (lmax[k] - lmin[k]) + margin;
        c[k] = 0.5 * (lmax[k] + lmin[k]);
    }
    for (int i = 0; i < 3; ++i) {
        wc[i] = (GLB(m, i, 0) * c[0] + GLB(m, i, 1) * c[1] + GLB(m, i, 2) * c[2]) + m[12 + i];
        e[i] = fabs(GLB(m, i, 0)) * h[0] + fabs(GLB(m, i, 1)) * h[1] + fabs(GLB(m, i, 2)) * h[2];
    }
    for (int k = 0; k < 3; ++k) {
        omin[k] = wc[k] - e[k];
        omax[k] = wc[k] + e[k];
    }
}

/* ------------------------------------------------------------------ */
/* velocity fields (Uniform.cpp:51-54, Jet.cpp:54-75, Pipe.cpp:57-78)   */
/* ------------------------------------------------------------------ */

static void current_at(const orc_current* c, const double p[3], double out[3])
{
    out[0] = out[1] = out[2] = 0.0;
    if (c->type == ORC_CUR_UNIFORM) {
        out[0] = c->p[0]; out[1] = c->p[1]; out[2] = c->p[2];
    } else if (c->type == ORC_CUR_JET) {
        const double* n = c->p + 3;
        double cp[3] = {p[0] - c->p[0], p[1] - c->p[1], p[2] - c->p[2]};
        double x[3] = {cp[1] * n[2] - cp[2] * n[1], cp[2] * n[0] - cp[0] * n[2], cp[0] * n[1] - cp[1] * n[0]};
        double d = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
        double t = cp[0] * n[0] + cp[1] * n[1] + cp[2] * n[2];
        if (t < 0.0) return;
        double r = c->p[6], vout = c->p[7];
        double rr = 1.0 / 5.0 * (t + 5.0 * r);
        if (d >= rr) return;
        double vm = 10.0 * r / (t + 5.0 * r) * vout;
        double f = exp(-50.0 * d * d / (t * t));
        for (int k = 0; k < 3; ++k) out[k] = f * (vm * n[k]);
    } else if (c->type == ORC_CUR_PIPE) {
        const double* n = c->p + 3;
        double pp[3] = {p[0] - c->p[0], p[1] - c->p[1], p[2] - c->p[2]};
        double x[3] = {pp[1] * n[2] - pp[2] * n[1], pp[2] * n[0] - pp[0] * n[2], pp[0] * n[1] - pp[1] * n[0]};
        double d = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
        double t = pp[0] * n[0] + pp[1] * n[1] + pp[2] * n[2];
        double l = c->p[6], r1 = c->p[7], r2 = c->p[8], vin = c->p[9], gam = c->p[10];
        if (t < 0.0 || t > l) return;
        double r = r1 + (r2 - r1) * t / l;
        if (d >= r) return;
        double vm = r1 / r * vin;
        double f = pow(1.0 - d / r, gam);
        for (int k = 0; k < 3; ++k) out[k] = f * (vm * n[k]);
    }
}

/* Ocean::GetFluidVelocity(Vector3) (Ocean.cpp:183-196). */
void orc_fluid_velocity(const orc_ocean* o, const double p[3], double out[3])
{
    out[0] = out[1] = out[2] = 0.0;
    if (!o->currents_enabled) return;
    for (uint32_t i = 0; i < o->ncurrents; ++i) {
        if (!o->currents[i].enabled) continue;
        double u[3];
        current_at(&o->currents[i], p, u);
        out[0] += u[0]; out[1] += u[1]; out[2] += u[2];
    }
}

/* SolidEntity::CorrectHydrodynamicForces (SolidEntity.cpp:1263-1287), double. */
void orc_correct(const orc_ocean* o, double Fdq[3], double Tdq[3], double Fdf[3], double Tdf[3],
                 const double Cd[3], const double Cf[3], const double T_O[16])
{
    orc_correct_k(o, Fdq, Tdq, Fdf, Tdf, Cd, Cf, T_O, NULL);
}

void orc_correct_k(const orc_ocean* o, double Fdq[3], double Tdq[3], double Fdf[3], double Tdf[3],
                   const double Cd[3], const double Cf[3], const double T_O[16], double kout[4])
{
    double M[9];
    orc_basis_inverse(T_O, M);
    double* vecs[4] = {Fdq, Tdq, Fdf, Tdf};
    for (int q = 0; q < 4; ++q) {
        double* x = vecs[q];
        double d[3];
        for (int i = 0; i < 3; ++i) d[i] = M[3 * i] * x[0] + M[3 * i + 1] * x[1] + M[3 * i + 2] * x[2];
        orc_safe_normalize(d);
        const double* C = (q < 2) ? Cd : Cf;
        double c = fabs(d[0]) * C[0] + fabs(d[1]) * C[1] + fabs(d[2]) * C[2];
        double s = (q < 2) ? 0.5 * o->density * c : o->density * c;
        for (int i = 0; i < 3; ++i) x[i] = s * x[i];
        if (kout) kout[q] = s;
    }
}

/* ------------------------------------------------------------------ */
/* the per-face loops, instantiated in double and in float              */
/* ------------------------------------------------------------------ */

#define RT double
#define SUF(n) n##_64
#define RSQRT sqrt
#define RABS fabs
#define RMODF modf
#define RTRUNC trunc
#include "hydro_core.inc"
#undef RT
#undef SUF
#undef RSQRT
#undef RABS
#undef RMODF
#undef RTRUNC

#define RT float
#define SUF(n) n##_32
#define RSQRT sqrtf
#define RABS fabsf
#define RMODF modff
#define RTRUNC truncf
#include "hydro_core.inc"
#undef RT
#undef SUF
#undef RSQRT
#undef RABS
#undef RMODF
#undef RTRUNC

double orc_wave_height_64(const orc_ocean* o, double x, double y) { return wave_height_64(o, x, y); }
float orc_wave_height_32(const orc_ocean* o, float x, float y) { return wave_height_32(o, x, y); }
double orc_depth_64(const orc_ocean* o, const double p[3]) { return depth_64(o, mk_64(p[0], p[1], p[2])); }
float orc_depth_32(const orc_ocean* o, const float p[3]) { return depth_32(o, mk_32(p[0], p[1], p[2])); }

/* ------------------------------------------------------------------ */
/* Compound::ComputeHydrodynamicForces (Compound.cpp:291-437), double   */
/* ------------------------------------------------------------------ */

void orc_compound_forces_64(const orc_body* comp, const orc_part* parts, uint32_t nparts,
                            const orc_ocean* o, const double T_CG[16],
                            const double v[3], const double w[3], orc_result* res)
{
    if (!comp->hydro) {   /* mode gate (Compound.cpp:293): not classified */
        res->position = -1;
        return;
    }
    int pos = orc_classify_64(comp, o, T_CG);
    res->position = pos;
    if (pos == 0) {
        memset(res->Fb, 0, sizeof(double) * 18);
        res->Swet = res->Vsub = 0.0;
        res->written |= 0xFF;
        return;
    }
    double T_O[16];
    orc_gl_mul(T_CG, comp->T_CG2O, T_O);
    if (pos == 1) {
        if (comp->buoyant) {
            double cb[3], r[3];
            for (int k = 0; k < 3; ++k) res->Fb[k] = -comp->volume * o->density * o->gravity[k];
            for (int k = 0; k < 3; ++k)
                cb[k] = T_CG[k] * comp->P_CB[0] + T_CG[4 + k] * comp->P_CB[1] + T_CG[8 + k] * comp->P_CB[2] + T_CG[12 + k];
            for (int k = 0; k < 3; ++k) r[k] = cb[k] - T_CG[12 + k];
            res->Tb[0] = r[1] * res->Fb[2] - r[2] * res->Fb[1];
            res->Tb[1] = r[2] * res->Fb[0] - r[0] * res->Fb[2];
            res->Tb[2] = r[0] * res->Fb[1] - r[1] * res->Fb[0];
            res->written |= ORC_W_FB | ORC_W_TB;
        }
        double F[12] = {0};
        for (uint32_t i = 0; i < nparts; ++i) {
            const orc_part* pt = &parts[i];
            if (!pt->external || !pt->solid->hydro) continue;
            double t1[16], T_C[16], T_Op[16];
            orc_gl_mul(T_O, pt->origin, t1);
            orc_gl_mul(t1, pt->O2C, T_C);
            orc_gl_mul(T_O, pt->origin, T_Op);
            orc_result pr;
            memset(&pr, 0, sizeof(pr));
            orc_submerged_64(pt->solid->mesh, o, T_CG, T_C, v, w, &pr);
            orc_correct(o, pr.Fdq, pr.Tdq, pr.Fdf, pr.Tdf, pt->solid->Cd, pt->solid->Cf, T_Op);
            for (int k = 0; k < 3; ++k) {
                F[k] += pr.Fdq[k]; F[3 + k] += pr.Tdq[k]; F[6 + k] += pr.Fdf[k]; F[9 + k] += pr.Tdf[k];
            }
        }
        for (int k = 0; k < 3; ++k) {
            res->Fdq[k] = F[k]; res->Tdq[k] = F[3 + k]; res->Fdf[k] = F[6 + k]; res->Tdf[k] = F[9 + k];
        }
        res->Swet = comp->surface;
        res->Vsub = comp->volume;
        res->written |= ORC_W_FDQ | ORC_W_TDQ | ORC_W_FDF | ORC_W_TDF | ORC_W_SWET | ORC_W_VSUB;
        return;
    }
    /* crossing: settings {true, true}; clear and sum over parts */
    memset(res->Fb, 0, sizeof(double) * 18);
    res->Swet = res->Vsub = 0.0;
    res->written |= 0xFF;
    /* the part temporaries are declared once before the loop in the reference
     * (Compound.cpp:392-399), so outputs a part leaves untouched keep the
     * previous part's values */
    orc_result pr;
    memset(&pr, 0, sizeof(pr));
    for (uint32_t i = 0; i < nparts; ++i) {
        const orc_part* pt = &parts[i];
        if (!pt->solid->hydro) continue;
        double t1[16], T_C[16], T_Op[16];
        orc_gl_mul(T_O, pt->origin, t1);
        orc_gl_mul(t1, pt->O2C, T_C);
        orc_gl_mul(T_O, pt->origin, T_Op);
        int pbuoy = pt->solid->buoyant;
        if (pt->external) {
            orc_surface_64(1, pbuoy, pt->solid->mesh, o, T_CG, T_C, v, w, &pr);
            orc_correct(o, pr.Fdq, pr.Tdq, pr.Fdf, pr.Tdf, pt->solid->Cd, pt->solid->Cf, T_Op);
            for (int k = 0; k < 3; ++k) {
                res->Fb[k] += pr.Fb[k]; res->Tb[k] += pr.Tb[k];
                res->Fdq[k] += pr.Fdq[k]; res->Tdq[k] += pr.Tdq[k];
                res->Fdf[k] += pr.Fdf[k]; res->Tdf[k] += pr.Tdf[k];
            }
            res->Swet += pr.Swet;
            res->Vsub += pr.Vsub;
        } else if (pbuoy) {
            orc_surface_64(0, 1, pt->solid->mesh, o, T_CG, T_C, v, w, &pr);
            for (int k = 0; k < 3; ++k) { res->Fb[k] += pr.Fb[k]; res->Tb[k] += pr.Tb[k]; }
            res->Vsub += pr.Vsub;
        }
    }
}

/* ------------------------------------------------------------------ */
/* SolidEntity::ComputeAerodynamicForces static (SolidEntity.cpp:1866-1925) */
/* The face term is the reference's `vn * vn.length() * A`                */
/* (SolidEntity.cpp:1913) on a glm::vec3: glm's member length() is the    */
/* GLSL one and returns the component count, 3 (the Euclidean norm is the  */
/* free function glm::length), so q = 3 A vn.                              */
/* ------------------------------------------------------------------ */
#define ORC_GLM_VEC3_LENGTH 3.0

void orc_aero_64(const orc_mesh* mesh, double gas_density, const orc_current* wind, uint32_t nwind,
                 const double T_CG[16], const double T_C[16],
                 const double v_[3], const double w_[3], double Fda[3], double Tda[3], double S[2])
{
    if (S) S[0] = S[1] = 0.0;
    if (mesh == NULL) {
        Fda[0] = Fda[1] = Fda[2] = Tda[0] = Tda[1] = Tda[2] = 0.0;
        return;
    }
    v3_64 F = {0, 0, 0}, T = {0, 0, 0};
    mat4_64 TC = glmat_64(T_C);
    v3_64 v = mk_64(v_[0], v_[1], v_[2]), omega = mk_64(w_[0], w_[1], w_[2]);
    v3_64 p = mk_64(T_CG[12], T_CG[13], T_CG[14]);
    for (uint32_t i = 0; i < mesh->nf; ++i) {
        const uint32_t* f = mesh->faces + 3 * (size_t)i;
        v3_64 p1 = xform_64(&TC, mesh->verts + 3 * (size_t)f[0]);
        v3_64 p2 = xform_64(&TC, mesh->verts + 3 * (size_t)f[1]);
        v3_64 p3 = xform_64(&TC, mesh->verts + 3 * (size_t)f[2]);
        v3_64 fn = cross_64(sub_64(p2, p1), sub_64(p3, p1));
        double len = len2_64(fn);
        if (len < (double)1e-12f) continue;
        len = sqrt(len);
        v3_64 fn1 = divs_64(fn, len);
        double Ar = len / 2.0;
        v3_64 fc = divs_64(add_64(add_64(p1, p2), p3), 3.0);
        double u[3] = {0, 0, 0};
        double fcd[3] = {fc.x, fc.y, fc.z};
        for (uint32_t k = 0; k < nwind; ++k) {
            double uk[3];
            current_at(&wind[k], fcd, uk);
            u[0] += uk[0]; u[1] += uk[1]; u[2] += uk[2];
        }
        v3_64 r = sub_64(fc, p);
        v3_64 vc = sub_64(mk_64(u[0], u[1], u[2]), add_64(v, cross_64(omega, r)));
        v3_64 vn = mul_64(fn1, dot_64(vc, fn1));
        if (dot_64(fn1, vn) < (double)-1e-12f) {
            v3_64 q = mul_64(mul_64(vn, ORC_GLM_VEC3_LENGTH), Ar);
            v3_64 t = cross_64(r, q);
            F = add_64(F, q);
            T = add_64(T, t);
            if (S) {
                S[0] += 0.5 * gas_density * sqrt(len2_64(q));
                S[1] += 0.5 * gas_density * sqrt(len2_64(t));
            }
        }
    }
    Fda[0] = 0.5 * gas_density * F.x; Fda[1] = 0.5 * gas_density * F.y; Fda[2] = 0.5 * gas_density * F.z;
    Tda[0] = 0.5 * gas_density * T.x; Tda[1] = 0.5 * gas_density * T.y; Tda[2] = 0.5 * gas_density * T.z;
}

/* ------------------------------------------------------------------ */
/* batched CPU baseline: a pool of threads, one task per body per step  */
/* (SimulationManager.cpp:1619-1655, ThreadPool.hpp:42-149)             */
/* ------------------------------------------------------------------ */

typedef struct {
    const orc_body* bodies;
    uint32_t nbodies;
    const orc_ocean* o;
    const double* states;
    orc_result* out;
    uint32_t steps;
    int prec64;
    int nthreads;
    atomic_uint next;
    pthread_barrier_t bar;
} pool_t;

static void* pool_worker(void* arg)
{
    pool_t* P = (pool_t*)arg;
    for (uint32_t s = 0; s < P->steps; ++s) {
        pthread_barrier_wait(&P->bar);   /* step start */
        for (;;) {
            uint32_t i = atomic_fetch_add(&P->next, 1u);
            if (i >= P->nbodies) break;
            const double* st = P->states + 22 * (size_t)i;
            if (P->prec64) orc_body_forces_64(&P->bodies[i], P->o, st, st + 16, st + 19, &P->out[i]);
            else           orc_body_forces_32(&P->bodies[i], P->o, st, st + 16, st + 19, &P->out[i]);
        }
        pthread_barrier_wait(&P->bar);   /* waitAll */
    }
    return NULL;
}

static double batch_run(const orc_body* bodies, uint32_t nbodies, const orc_ocean* o,
                        const double* states, uint32_t steps, int nthreads, orc_result* out, int prec64)
{
    if (nthreads < 1) nthreads = 1;
    pool_t P;
    P.bodies = bodies; P.nbodies = nbodies; P.o = o; P.states = states; P.out = out;
    P.steps = steps; P.prec64 = prec64; P.nthreads = nthreads;
    atomic_init(&P.next, 0u);
    pthread_barrier_init(&P.bar, NULL, (unsigned)nthreads + 1);
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, pool_worker, &P);
    struct timespec a, b;
    clock_gettime(CLOCK_MONOTONIC, &a);
    for (uint32_t s = 0; s < steps; ++s) {
        atomic_store(&P.next, 0u);
        pthread_barrier_wait(&P.bar);
        pthread_barrier_wait(&P.bar);
    }
    clock_gettime(CLOCK_MONOTONIC, &b);
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    pthread_barrier_destroy(&P.bar);
    free(th);
    return (double)(b.tv_sec - a.tv_sec) + 1e-9 * (double)(b.tv_nsec - a.tv_nsec);
}

double orc_batch_forces_32(const orc_body* bodies, uint32_t nbodies, const orc_ocean* o,
                           const double* states, uint32_t steps, int nthreads, orc_result* out)
{
    return batch_run(bodies, nbodies, o, states, steps, nthreads, out, 0);
}

double orc_batch_forces_64(const orc_body* bodies, uint32_t nbodies, const orc_ocean* o,
                           const double* states, uint32_t steps, int nthreads, orc_result* out)
{
    return batch_run(bodies, nbodies, o, states, steps, nthreads, out, 1);
}
