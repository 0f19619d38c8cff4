/*
 * CPU oracle for the Stonefish geometry-based hydrodynamics path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the
 * Python package) links, imports or calls this code.  Only the test suite,
 * the smoke check and the CPU-baseline leg of bench.py use it, and only as the
 * checker / the timed CPU baseline.
 *
 * Two flavours of one restatement of the reference algorithm:
 *   *_64 : everything in double precision.  This is the parity target that the
 *          GPU results are compared against.
 *   *_32 : the reference's own precision: float transforms (glMatrixFromTransform),
 *          float per-face arithmetic and float accumulation, double epilogue
 *          and double CorrectHydrodynamicForces.  This is the CPU timing
 *          baseline and a cross-check of how far the reference's own float path
 *          is from *_64.
 *
 * Reference files followed (paths relative to the reference checkout):
 *   Library/src/entities/SolidEntity.cpp:1239-1261  CheckBodyFluidPosition
 *   Library/src/entities/SolidEntity.cpp:1263-1287  CorrectHydrodynamicForces
 *   Library/src/entities/SolidEntity.cpp:1289-1721  ComputeHydrodynamicForcesSurface
 *   Library/src/entities/SolidEntity.cpp:1723-1797  ComputeHydrodynamicForcesSubmerged
 *   Library/src/entities/SolidEntity.cpp:1799-1850  ComputeHydrodynamicForces
 *   Library/src/entities/SolidEntity.cpp:1852-1956  aerodynamic per-face loop
 *   Library/src/entities/solids/Compound.cpp:291-437 Compound::ComputeHydrodynamicForces
 *   Library/src/entities/forcefields/Ocean.cpp:149-201 GetDepth / GetFluidVelocity
 *   Library/src/entities/forcefields/{Uniform,Jet,Pipe}.cpp GetVelocityAtPoint
 *   Library/src/graphics/OpenGLRealOcean.cpp:280-331 wave-height bilinear query
 *   Library/src/graphics/OpenGLDataStructs.cpp:148-171 double -> float boundary
 *   3rdparty/LinearMath/btMatrix3x3.h:215-282 (setRotation, scalar path),
 *     1093-1103 (inverse); btVector3.h:286-299 (safeNormalize);
 *     btTransform.h:91-94, 136-143, 231-235; btAabbUtil2.h:182-196;
 *   3rdparty/BulletCollision/CollisionShapes/btCompoundShape.cpp:161-181
 *
 * Transforms are 16 doubles in OpenGL column-major order, exactly what
 * btTransform::getOpenGLMatrix writes: m[0..2] basis column 0, m[4..6] column 1,
 * m[8..10] column 2, m[12..14] origin, m[3]=m[7]=m[11]=0, m[15]=1.
 */
#ifndef SFH_ORACLE_H
#define SFH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_CUR_UNIFORM = 0, ORC_CUR_JET = 1, ORC_CUR_PIPE = 2 };

typedef struct {
    int32_t type;     /* ORC_CUR_* */
    int32_t enabled;  /* VelocityField::isEnabled */
    /* uniform: p[0..2] = v
     * jet:     p[0..2] = c, p[3..5] = n (unit), p[6] = r, p[7] = vout
     * pipe:    p[0..2] = p1, p[3..5] = n (unit), p[6] = l, p[7] = r1, p[8] = r2,
     *          p[9] = vin, p[10] = gamma                                   */
    double p[12];
} orc_current;

typedef struct {
    double density;          /* Ocean::getLiquid().density */
    double gravity[3];       /* SimulationManager::getGravity(), NED (0,0,+g) */
    int32_t has_waves;       /* Ocean::hasWaves() */
    int32_t currents_enabled;
    /* wave height data: N x N texels, `channels` interleaved floats per texel,
     * layer l reads channel l and has physical size wave_len[l] metres.
     * height(x,y) = - sum_l bilinear(layer l, x/len_l, y/len_l). */
    const float* wave_data;
    uint32_t wave_n;
    uint32_t wave_channels;
    uint32_t wave_layers;
    uint32_t ncurrents;
    float wave_len[4];
    const orc_current* currents;
} orc_ocean;

typedef struct {
    const float* verts;     /* 3 floats per vertex: Vertex::pos */
    const uint32_t* faces;  /* 3 vertex ids per face: Face::vertexID */
    uint32_t nv;
    uint32_t nf;
} orc_mesh;

/* output write mask bits: which outputs an entry point assigned (the reference
 * leaves some outputs untouched, e.g. buoyancy when Vsub <= 1e-9) */
enum {
    ORC_W_FB = 1, ORC_W_TB = 2, ORC_W_FDQ = 4, ORC_W_TDQ = 8,
    ORC_W_FDF = 16, ORC_W_TDF = 32, ORC_W_SWET = 64, ORC_W_VSUB = 128
};

typedef struct {
    double Fb[3], Tb[3], Fdq[3], Tdq[3], Fdf[3], Tdf[3];
    double Swet, Vsub;
    uint32_t written;
    int32_t position;   /* body-level entry points: 0 outside, 1 inside, 2 crossing, -1 not hydrodynamic */
    /* Error-analysis extras (not outputs of the reference):
     * S[6]  condition scale of Fb, Tb, Fdq, Tdq, Fdf, Tdf: the sum over faces of
     *       the norms of the per-face contributions (for the flat-sea volume
     *       buoyancy: rho |g| sum|6V|/6 and that times the largest lever arm;
     *       for the pressure buoyancy with waves A (|depth| + the polygon's
     *       extent from its centroid), times |r| for its moment; for the
     *       quadratic drag the size of the term's operands, |vc|^2 A and
     *       |r| |vc|^2 A, over the faces the flow meets or passes
     *       tangentially);
     * raw[12] Fdq, Tdq, Fdf, Tdf before CorrectHydrodynamicForces;
     * k[4]  the scalars CorrectHydrodynamicForces multiplied them by (1 if not
     *       applied). */
    double S[6];
    double raw[12];
    double k[4];
} orc_result;

/* ---- LinearMath restatement (double) -------------------------------- */
void orc_quat_to_gl(const double q[4] /* x,y,z,w */, const double origin[3], double m[16]);
void orc_gl_mul(const double a[16], const double b[16], double out[16]);
void orc_gl_inverse(const double a[16], double out[16]);
void orc_basis_inverse(const double m[16], double inv9[9]);  /* row-major, cofactor */
void orc_safe_normalize(double v[3]);
void orc_transform_aabb(const double lmin[3], const double lmax[3], double margin,
                        const double m[16], double omin[3], double omax[3]);

/* ---- ocean queries --------------------------------------------------- */
double orc_wave_height_64(const orc_ocean* o, double x, double y);
float  orc_wave_height_32(const orc_ocean* o, float x, float y);
double orc_depth_64(const orc_ocean* o, const double p[3]);
float  orc_depth_32(const orc_ocean* o, const float p[3]);
void   orc_fluid_velocity(const orc_ocean* o, const double p[3], double out[3]);

/* ---- the two static per-face loops ---------------------------------- */
/* settings: damping = HydrodynamicsSettings::dampingForces,
 *           buoyancy = HydrodynamicsSettings::reallisticBuoyancy.
 * `res` is in/out: fields not assigned by the reference keep their values. */
void orc_surface_64(int damping, int buoyancy, const orc_mesh* mesh, const orc_ocean* o,
                    const double T_CG[16], const double T_C[16],
                    const double v[3], const double w[3], orc_result* res);
void orc_surface_32(int damping, int buoyancy, const orc_mesh* mesh, const orc_ocean* o,
                    const double T_CG[16], const double T_C[16],
                    const double v[3], const double w[3], orc_result* res);
void orc_submerged_64(const orc_mesh* mesh, const orc_ocean* o,
                      const double T_CG[16], const double T_C[16],
                      const double v[3], const double w[3], orc_result* res);
void orc_submerged_32(const orc_mesh* mesh, const orc_ocean* o,
                      const double T_CG[16], const double T_C[16],
                      const double v[3], const double w[3], orc_result* res);

void orc_correct(const orc_ocean* o, double Fdq[3], double Tdq[3], double Fdf[3], double Tdf[3],
                 const double Cd[3], const double Cf[3], const double T_O[16]);
/* Same, also returning the four scalars applied (may be NULL). */
void orc_correct_k(const orc_ocean* o, double Fdq[3], double Tdq[3], double Fdf[3], double Tdf[3],
                   const double Cd[3], const double Cf[3], const double T_O[16], double k[4]);

/* ---- body level (SolidEntity::ComputeHydrodynamicForces) ------------- */
typedef struct {
    const orc_mesh* mesh;
    double T_CG2C[16];      /* getCTransform() = getCGTransform() * T_CG2C */
    double T_CG2O[16];      /* getOTransform() = getCGTransform() * T_CG2O */
    /* AABB of the rigid body's collision shape in the CG frame, i.e. the
     * btCompoundShape local AABB: btTransformAabb(child AABB, child margin,
     * T_CG2C) for a single solid, the union over external parts for a compound */
    double aabb_min[3], aabb_max[3];
    double Cd[3], Cf[3];
    double volume, surface;
    double P_CB[3];         /* centre of buoyancy in the CG frame */
    int32_t buoyant;        /* isBuoyant() */
    int32_t hydro;          /* phy.mode is FLOATING or SUBMERGED */
} orc_body;

int  orc_classify_64(const orc_body* b, const orc_ocean* o, const double T_CG[16]);
int  orc_classify_32(const orc_body* b, const orc_ocean* o, const double T_CG[16]);
void orc_body_forces_64(const orc_body* b, const orc_ocean* o, const double T_CG[16],
                        const double v[3], const double w[3], orc_result* res);
void orc_body_forces_32(const orc_body* b, const orc_ocean* o, const double T_CG[16],
                        const double v[3], const double w[3], orc_result* res);

/* ---- compound (Compound::ComputeHydrodynamicForces) ------------------ */
typedef struct {
    const orc_body* solid;  /* part solid: mesh, Cd/Cf, buoyant, hydro */
    double origin[16];      /* CompoundPart::origin */
    double O2C[16];         /* part solid getO2CTransform() */
    int32_t external;
} orc_part;

void orc_compound_forces_64(const orc_body* comp, const orc_part* parts, uint32_t nparts,
                            const orc_ocean* o, const double T_CG[16],
                            const double v[3], const double w[3], orc_result* res);

/* ---- aerodynamics (SolidEntity::ComputeAerodynamicForces static) ------ */
/* wind: Atmosphere::GetFluidVelocity sums every wind field (no enabled flag);
 * Fda/Tda are the raw integrals times 0.5*rho_gas, before CorrectAerodynamicForces.
 * S (may be NULL): sums of the norms of the per-face force and torque terms. */
void orc_aero_64(const orc_mesh* mesh, double gas_density, const orc_current* wind, uint32_t nwind,
                 const double T_CG[16], const double T_C[16],
                 const double v[3], const double w[3], double Fda[3], double Tda[3], double S[2]);

/* ---- batched CPU baseline: one task per body on a thread pool -------- */
/* states: per body 22 doubles: T_CG[16], v[3], w[3]; the C transform is
 * T_CG * T_CG2C as the reference's getCTransform().
 * Returns elapsed wall seconds for `steps` passes over all bodies. */
double orc_batch_forces_32(const orc_body* bodies, uint32_t nbodies, const orc_ocean* o,
                           const double* states, uint32_t steps, int nthreads, orc_result* out);
double orc_batch_forces_64(const orc_body* bodies, uint32_t nbodies, const orc_ocean* o,
                           const double* states, uint32_t steps, int nthreads, orc_result* out);

#ifdef __cplusplus
}
#endif
#endif
