// sfhydro.hpp: C++ host interface with the reference's call signatures, over
// the C ABI of sfhydro.h (header-only; link libsfhydro.so).
//
// Mirrors, with the same names, argument order and meaning:
//   sf::HydrodynamicsSettings                        Library/include/entities/forcefields/Ocean.h:36-40
//   SolidEntity::ComputeHydrodynamicForcesSurface    Library/include/entities/SolidEntity.h:135-137
//   SolidEntity::ComputeHydrodynamicForcesSubmerged  Library/include/entities/SolidEntity.h:152-153
//   SolidEntity::CorrectHydrodynamicForces           Library/include/entities/SolidEntity.h:113-114
//   SolidEntity::ComputeAerodynamicForces (static)   Library/include/entities/SolidEntity.h:180-181
//   SolidEntity::CorrectAerodynamicForces            Library/include/entities/SolidEntity.h:167
// The value types are minimal stand-ins for the reference's (Scalar = double,
// Vector3, Transform with getOpenGLMatrix, a Mesh of Vertex/Face); a Stonefish
// build passes its own types through the small adaptors at the bottom instead.
// Error behaviour: the reference's statics return void; these return the
// sfhydro status and also throw sfh::Error when constructed with throwing on.
#pragma once

#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "sfhydro.h"

namespace sfh {

using Scalar = double;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline int check(int rc)
{
    if (rc != SFH_OK) throw Error(rc, sfh_last_error());
    return rc;
}

struct Vector3 {
    Scalar v[3] = {0, 0, 0};
    Vector3() = default;
    Vector3(Scalar x, Scalar y, Scalar z) { v[0] = x; v[1] = y; v[2] = z; }
    Scalar x() const { return v[0]; }
    Scalar y() const { return v[1]; }
    Scalar z() const { return v[2]; }
    Scalar& operator[](int i) { return v[i]; }
    Scalar operator[](int i) const { return v[i]; }
    void setZero() { v[0] = v[1] = v[2] = 0; }
};

// Rigid transform: row-major basis and origin, as btTransform.
struct Transform {
    Scalar basis[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    Vector3 origin;
    static Transform identity() { return Transform(); }
    // btTransform::getOpenGLMatrix (btTransform.h:136-143)
    void getOpenGLMatrix(Scalar* m) const
    {
        for (int c = 0; c < 3; ++c) {
            for (int r = 0; r < 3; ++r) m[4 * c + r] = basis[r][c];
            m[4 * c + 3] = 0;
        }
        m[12] = origin[0]; m[13] = origin[1]; m[14] = origin[2]; m[15] = 1;
    }
};

struct HydrodynamicsSettings {
    bool dampingForces = true;
    bool reallisticBuoyancy = true;
};

// Device-resident copy of a triangle mesh (the reference's const Mesh*).
class Mesh {
public:
    struct Vertex { float pos[3]; float normal[3]; };   // PlainMesh vertex (24 B)
    struct Face { uint32_t vertexID[3]; };

    Mesh(const std::vector<Vertex>& vertices, const std::vector<Face>& faces)
    {
        static_assert(sizeof(Vertex) == 24 && sizeof(Face) == 12, "layout");
        check(sfh_mesh_create(vertices.empty() ? nullptr : vertices[0].pos, sizeof(Vertex),
                              (uint32_t)vertices.size(), faces.empty() ? nullptr : faces[0].vertexID,
                              (uint32_t)faces.size(), &h_));
    }
    // packed xyz / index arrays
    Mesh(const float* xyz, uint32_t nv, const uint32_t* ids, uint32_t nf)
    {
        check(sfh_mesh_create(xyz, 12, nv, ids, nf, &h_));
    }
    ~Mesh() { if (h_) sfh_mesh_destroy(h_); }
    Mesh(const Mesh&) = delete;
    Mesh& operator=(const Mesh&) = delete;
    sfh_mesh handle() const { return h_; }

private:
    sfh_mesh h_ = nullptr;
};

// Liquid, free surface and currents (the reference's Ocean*).
class Ocean {
public:
    explicit Ocean(Scalar density = 1000.0, const Vector3& gravity = Vector3(0, 0, 9.81))
    {
        check(sfh_ocean_create(density, gravity.v, &h_));
    }
    ~Ocean() { if (h_) sfh_ocean_destroy(h_); }
    Ocean(const Ocean&) = delete;
    Ocean& operator=(const Ocean&) = delete;
    // fftData-style texels: n x n x channels floats, layer l = channel l of size lengths[l]
    void setWaves(const float* data, uint32_t n, uint32_t channels, uint32_t layers, const float* lengths)
    {
        check(sfh_ocean_set_waves(h_, data, 1, n, channels, layers, lengths));
    }
    void addUniformCurrent(const Vector3& v, bool enabled = true)
    {
        double p[12] = {v[0], v[1], v[2]};
        check(sfh_ocean_add_current(h_, 0, enabled ? 1 : 0, p));
    }
    void EnableCurrents() { check(sfh_ocean_enable_currents(h_, 1)); }
    void DisableCurrents() { check(sfh_ocean_enable_currents(h_, 0)); }
    sfh_ocean handle() const { return h_; }

protected:
    sfh_ocean h_ = nullptr;
};

// Gas and wind fields (the reference's Atmosphere*): every wind field is
// summed, as Atmosphere::GetFluidVelocity does (Atmosphere.cpp:146-152).
class Atmosphere : public Ocean {
public:
    explicit Atmosphere(Scalar density = 1.225, const Vector3& gravity = Vector3(0, 0, 9.81))
        : Ocean(density, gravity)
    {
        EnableCurrents();
    }
    void addUniformWind(const Vector3& v) { addUniformCurrent(v, true); }
};

struct SolidEntity {
    // SolidEntity.cpp:1289-1721.  The reference's `Renderable& debug` output
    // (DEBUG_HYDRO outlines) has no counterpart.
    static int ComputeHydrodynamicForcesSurface(const HydrodynamicsSettings& settings, const Mesh* mesh, Ocean* ocn,
                                                const Transform& T_CG, const Transform& T_C, const Vector3& v,
                                                const Vector3& omega, Vector3& Fb, Vector3& Tb, Vector3& Fdq,
                                                Vector3& Tdq, Vector3& Fdf, Vector3& Tdf, Scalar& Swet, Scalar& Vsub)
    {
        double tcg[16], tc[16];
        T_CG.getOpenGLMatrix(tcg);
        T_C.getOpenGLMatrix(tc);
        sfh_settings s = {settings.dampingForces ? 1 : 0, settings.reallisticBuoyancy ? 1 : 0};
        return sfh_forces_surface(&s, mesh ? mesh->handle() : nullptr, ocn->handle(), tcg, tc, v.v, omega.v, Fb.v,
                                  Tb.v, Fdq.v, Tdq.v, Fdf.v, Tdf.v, &Swet, &Vsub);
    }

    // SolidEntity.cpp:1723-1797.
    static int ComputeHydrodynamicForcesSubmerged(const Mesh* mesh, Ocean* ocn, const Transform& T_CG,
                                                  const Transform& T_C, const Vector3& v, const Vector3& omega,
                                                  Vector3& Fdq, Vector3& Tdq, Vector3& Fdf, Vector3& Tdf)
    {
        double tcg[16], tc[16];
        T_CG.getOpenGLMatrix(tcg);
        T_C.getOpenGLMatrix(tc);
        return sfh_forces_submerged(mesh ? mesh->handle() : nullptr, ocn->handle(), tcg, tc, v.v, omega.v, Fdq.v,
                                    Tdq.v, Fdf.v, Tdf.v);
    }

    // SolidEntity.cpp:1865-1923.
    static int ComputeAerodynamicForces(const Mesh* mesh, Atmosphere* atm, const Transform& T_CG,
                                        const Transform& T_C, const Vector3& v, const Vector3& omega,
                                        Vector3& Fda, Vector3& Tda)
    {
        double tcg[16], tc[16];
        T_CG.getOpenGLMatrix(tcg);
        T_C.getOpenGLMatrix(tc);
        return sfh_forces_aero(mesh ? mesh->handle() : nullptr, atm->handle(), tcg, tc, v.v, omega.v, Fda.v, Tda.v);
    }

    // SolidEntity.cpp:1925-1948; the entity's fdApproxType, fdApproxParams and T_CG2H.
    static int CorrectAerodynamicForces(int fdApproxType, const Scalar fdApproxParams[3], const Transform& T_CG2H,
                                        Vector3& Fda, Vector3& Tda)
    {
        double h[16];
        T_CG2H.getOpenGLMatrix(h);
        return sfh_correct_aero(fdApproxType, fdApproxParams, h, Fda.v, Tda.v);
    }

    // SolidEntity.cpp:1263-1287.
    static int CorrectHydrodynamicForces(Ocean* ocn, Vector3& Fdq, Vector3& Tdq, Vector3& Fdf, Vector3& Tdf,
                                         const Vector3& fdCd, const Vector3& fdCf, const Transform& T_O)
    {
        double to[16];
        T_O.getOpenGLMatrix(to);
        return sfh_correct(ocn->handle(), Fdq.v, Tdq.v, Fdf.v, Tdf.v, fdCd.v, fdCf.v, to);
    }
};

// Many bodies per call: the tick-level replacement of the per-body pool
// fan-out (SimulationManager.cpp:1619-1655).
class BodyBatch {
public:
    explicit BodyBatch(const std::vector<sfh_body_desc>& bodies, uint32_t faces_per_lane = 0)
    {
        sfh_batch_opts o = {faces_per_lane, 0, 0};
        check(sfh_batch_create(bodies.data(), (uint32_t)bodies.size(), &o, &h_));
        n_ = (uint32_t)bodies.size();
    }
    ~BodyBatch() { if (h_) sfh_batch_destroy(h_); }
    BodyBatch(const BodyBatch&) = delete;
    BodyBatch& operator=(const BodyBatch&) = delete;
    // states: n x SFH_STATE_DOUBLES; out: n x SFH_OUT_DOUBLES (host memory)
    void compute(Ocean& ocn, const std::vector<double>& states, std::vector<double>& out)
    {
        out.resize((size_t)n_ * SFH_OUT_DOUBLES);
        check(sfh_batch_compute_host(h_, ocn.handle(), states.data(), out.data()));
    }
    sfh_batch handle() const { return h_; }

private:
    sfh_batch h_ = nullptr;
    uint32_t n_ = 0;
};

// Adaptors for a Stonefish build: any transform type with
// getOpenGLMatrix(double*) and any vector type with getX/getY/getZ.
template <class TransformT>
inline void to_gl(const TransformT& t, double m[16]) { t.getOpenGLMatrix(m); }

template <class VectorT>
inline void to_array(const VectorT& v, double out[3])
{
    out[0] = v.getX(); out[1] = v.getY(); out[2] = v.getZ();
}

}  // namespace sfh
