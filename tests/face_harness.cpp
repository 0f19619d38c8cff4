// Host build of the shared per-face integrand (stonefish_amd/csrc/sfh_face.h)
// for the CPU tests: evaluates one body over a flat sea in the mesh frame with
// FP64 accumulation and rotates the sums to world, the same reduction the GPU
// epilogue performs.  Test infrastructure only.
#include <cstdint>
#include <cstring>

#include "../stonefish_amd/csrc/sfh_face.h"

using namespace sfh;

static void rows(const double* m, double* R) {
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) R[3 * r + c] = m[4 * c + r];
}
static void mtv(const double* R, const double* x, double* y) {
    for (int i = 0; i < 3; ++i) y[i] = R[i] * x[0] + R[3 + i] * x[1] + R[6 + i] * x[2];
}
static void mv(const double* R, const double* x, double* y) {
    for (int i = 0; i < 3; ++i) y[i] = R[3 * i] * x[0] + R[3 * i + 1] * x[1] + R[3 * i + 2] * x[2];
}
static void cr(const double* a, const double* b, double* c) {
    c[0] = a[1] * b[2] - a[2] * b[1]; c[1] = a[2] * b[0] - a[0] * b[2]; c[2] = a[0] * b[1] - a[1] * b[0];
}

// mode 1: submerged loop, mode 2: surface loop (flat sea, damping and buoyancy on).
// out: 24 doubles in the sfhydro output layout (raw, before Correct).
extern "C" void harness_body(int mode, const float* verts, const uint32_t* faces, uint32_t nf,
                             const double* T_CG, const double* T_C, const double* v, const double* w,
                             double rho, double g, double* out)
{
    double R[9], tc[3] = {T_C[12], T_C[13], T_C[14]}, tcg[3] = {T_CG[12], T_CG[13], T_CG[14]}, y[3], x[3];
    rows(T_C, R);
    FaceBody B;
    double vl[3], wl[3], wd[3];
    mtv(R, v, vl);
    mtv(R, w, wl); B.w = mk3((float)wl[0], (float)wl[1], (float)wl[2]);
    double dt[3];
    for (int i = 0; i < 3; ++i) x[i] = tc[i] - tcg[i];
    mtv(R, x, dt);
    cr(wl, dt, wd);
    B.vb = mk3((float)(vl[0] + wd[0]), (float)(vl[1] + wd[1]), (float)(vl[2] + wd[2]));
    x[0] = tcg[0] - tc[0]; x[1] = tcg[1] - tc[1]; x[2] = -tc[2];
    mtv(R, x, y); B.o = mk3((float)y[0], (float)y[1], (float)y[2]);
    double rz[3] = {R[6], R[7], R[8]}, tz = tc[2];
    auto depth = [&](f3 p) { return (float)(tz + rz[0] * p.x + rz[1] * p.y + rz[2] * p.z); };
    Acc<double> A;
    acc_zero(A);
    for (uint32_t f = 0; f < nf; ++f) {
        const float* a = verts + 3 * faces[3 * f];
        const float* b = verts + 3 * faces[3 * f + 1];
        const float* c = verts + 3 * faces[3 * f + 2];
        f3 p1 = mk3(a[0], a[1], a[2]), p2 = mk3(b[0], b[1], b[2]), p3 = mk3(c[0], c[1], c[2]);
        if (mode == 1) { face_submerged(B, p1, p2, p3, A); continue; }
        float d0 = depth(p1), d1 = depth(p2), d2 = depth(p3);
        if (d0 < 0.f && d1 < 0.f && d2 < 0.f) continue;
        Wet wt = face_clip(B, p1, p2, p3, d0, d1, d2, A);
        if (!wt.ok) continue;
        face_drag(B, wt.base + wt.off, wt.n1, wt.A, A);
        A.Swet += wt.A;
    }
    memset(out, 0, 24 * sizeof(double));
    const double* F[2] = {A.Fdq, A.Fdf};
    const double* M[2] = {A.Mdq, A.Mdf};
    for (int q = 0; q < 2; ++q) {
        double t[3], m[3];
        cr(dt, F[q], t);
        for (int k = 0; k < 3; ++k) m[k] = M[q][k] + t[k];
        mv(R, F[q], out + 6 + 6 * q);
        mv(R, m, out + 9 + 6 * q);
    }
    if (mode == 2) {
        out[18] = A.Swet;
        if (A.V6 > (double)1e-9f) {
            double Vs = A.V6 / 6.0, cbl[3], cb[3], r[3], Fb[3] = {0, 0, -Vs * rho * g};
            for (int k = 0; k < 3; ++k) cbl[k] = A.CB[k] / A.V6;
            mv(R, cbl, cb);
            r[0] = cb[0]; r[1] = cb[1]; r[2] = cb[2] - tcg[2];
            Fb[0] = -Vs * rho * 0.0; Fb[1] = -Vs * rho * 0.0;
            out[0] = Fb[0]; out[1] = Fb[1]; out[2] = Fb[2];
            cr(r, Fb, out + 3);
            out[19] = Vs;
        }
    }
}

// Host build of the surrogate's GELU (stonefish_amd/csrc/sfm_gelu.h).
#include "../stonefish_amd/csrc/sfm_gelu.h"

extern "C" void harness_gelu(const float* z, uint32_t n, float* g, float* dg)
{
    for (uint32_t i = 0; i < n; ++i) {
        g[i] = sfm::gelu(z[i]);
        dg[i] = sfm::gelu_grad(z[i]);
    }
}
