// Drives the C++ host interface (include/sfhydro.hpp) for the GPU tests:
// reads a mesh and a list of states, calls SolidEntity::ComputeHydrodynamicForces
// Surface / Submerged / CorrectHydrodynamicForces through it, writes the outputs.
//
// input (binary):  u32 nv, u32 nf, f32 xyz[3 nv], u32 ids[3 nf], u32 ns,
//                  then per state f64 T_CG[16], T_C[16], v[3], w[3]
// output (binary): per state f64 surface[20] (Fb Tb Fdq Tdq Fdf Tdf Swet Vsub),
//                  submerged[12] (Fdq Tdq Fdf Tdf), corrected[12]
#include <cstdio>
#include <vector>

#include "../../include/sfhydro.hpp"

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s input.bin output.bin\n", argv[0]);
        return 2;
    }
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    uint32_t nv, nf, ns;
    if (std::fread(&nv, 4, 1, f) != 1 || std::fread(&nf, 4, 1, f) != 1) return 2;
    std::vector<float> xyz(3 * (size_t)nv);
    std::vector<uint32_t> ids(3 * (size_t)nf);
    if (std::fread(xyz.data(), 4, xyz.size(), f) != xyz.size()) return 2;
    if (std::fread(ids.data(), 4, ids.size(), f) != ids.size()) return 2;
    if (std::fread(&ns, 4, 1, f) != 1) return 2;
    std::vector<double> st(38 * (size_t)ns);
    if (std::fread(st.data(), 8, st.size(), f) != st.size()) return 2;
    std::fclose(f);

    try {
        sfh::check(sfh_init(0));
        // exercise the AoS Vertex path (24-byte stride, as PlainMesh)
        std::vector<sfh::Mesh::Vertex> verts(nv);
        for (uint32_t i = 0; i < nv; ++i)
            for (int k = 0; k < 3; ++k) { verts[i].pos[k] = xyz[3 * i + k]; verts[i].normal[k] = 0.f; }
        std::vector<sfh::Mesh::Face> faces(nf);
        for (uint32_t i = 0; i < nf; ++i)
            for (int k = 0; k < 3; ++k) faces[i].vertexID[k] = ids[3 * i + k];
        sfh::Mesh mesh(verts, faces);
        sfh::Ocean ocn(1000.0, sfh::Vector3(0, 0, 9.81));
        sfh::Atmosphere atm(1.2);
        atm.addUniformWind(sfh::Vector3(2.0, -1.0, 0.5));
        std::vector<double> out;
        for (uint32_t s = 0; s < ns; ++s) {
            const double* p = st.data() + 38 * (size_t)s;
            sfh::Transform Tcg, Tc;
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) { Tcg.basis[r][c] = p[4 * c + r]; Tc.basis[r][c] = p[16 + 4 * c + r]; }
            Tcg.origin = sfh::Vector3(p[12], p[13], p[14]);
            Tc.origin = sfh::Vector3(p[28], p[29], p[30]);
            sfh::Vector3 v(p[32], p[33], p[34]), w(p[35], p[36], p[37]);
            sfh::Vector3 Fb, Tb, Fdq, Tdq, Fdf, Tdf;
            sfh::Scalar Swet = 0, Vsub = 0;
            sfh::HydrodynamicsSettings settings;
            sfh::check(sfh::SolidEntity::ComputeHydrodynamicForcesSurface(settings, &mesh, &ocn, Tcg, Tc, v, w, Fb,
                                                                           Tb, Fdq, Tdq, Fdf, Tdf, Swet, Vsub));
            for (const sfh::Vector3* x : {&Fb, &Tb, &Fdq, &Tdq, &Fdf, &Tdf})
                for (int k = 0; k < 3; ++k) out.push_back((*x)[k]);
            out.push_back(Swet);
            out.push_back(Vsub);
            sfh::Vector3 a, b, c, d;
            sfh::check(sfh::SolidEntity::ComputeHydrodynamicForcesSubmerged(&mesh, &ocn, Tcg, Tc, v, w, a, b, c, d));
            for (const sfh::Vector3* x : {&a, &b, &c, &d})
                for (int k = 0; k < 3; ++k) out.push_back((*x)[k]);
            sfh::Vector3 Cd(1.0, 0.5, 0.25), Cf(0.1, 0.05, 0.025);
            sfh::check(sfh::SolidEntity::CorrectHydrodynamicForces(&ocn, a, b, c, d, Cd, Cf, Tcg));
            for (const sfh::Vector3* x : {&a, &b, &c, &d})
                for (int k = 0; k < 3; ++k) out.push_back((*x)[k]);
            sfh::Vector3 Fda, Tda;
            sfh::check(sfh::SolidEntity::ComputeAerodynamicForces(&mesh, &atm, Tcg, Tc, v, w, Fda, Tda));
            for (const sfh::Vector3* x : {&Fda, &Tda})
                for (int k = 0; k < 3; ++k) out.push_back((*x)[k]);
            const sfh::Scalar params[3] = {0.9, 0.35, 0.3};
            sfh::check(sfh::SolidEntity::CorrectAerodynamicForces(SFH_APPROX_ELLIPSOID, params, Tc, Fda, Tda));
            for (const sfh::Vector3* x : {&Fda, &Tda})
                for (int k = 0; k < 3; ++k) out.push_back((*x)[k]);
        }
        FILE* g = std::fopen(argv[2], "wb");
        if (!g) return 2;
        std::fwrite(out.data(), 8, out.size(), g);
        std::fclose(g);
    } catch (const sfh::Error& e) {
        std::fprintf(stderr, "sfhydro error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
