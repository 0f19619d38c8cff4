/*
 * sfhydro: MI355X (gfx950) implementation of Stonefish's geometry-based
 * hydrodynamics behind a C ABI.
 *
 * The entry points replace the per-face loops of SolidEntity and the body-level
 * driver around them.  Each declaration names the reference interface it
 * stands in for (paths relative to the reference checkout).  All entry points
 * return an int status: SFH_OK (0) or a negative SFH_ERR_* code; the message of
 * the last error on the calling thread is returned by sfh_last_error().
 *
 * Conventions (as in the reference):
 *   - world frame NED, z down, gravity (0, 0, +g);
 *   - transforms are 16 doubles in OpenGL column-major order, exactly what
 *     btTransform::getOpenGLMatrix writes (basis columns in [0..2], [4..6],
 *     [8..10], origin in [12..14]);
 *   - velocities are world-frame doubles;
 *   - drag outputs of the two per-face entry points are raw integrals (before
 *     CorrectHydrodynamicForces), as in the reference.
 */
#ifndef SFHYDRO_H
#define SFHYDRO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SFH_OK 0
#define SFH_ERR_INVALID (-1)     /* null handle, bad argument */
#define SFH_ERR_SIZE (-2)        /* size out of range, buffer too small */
#define SFH_ERR_HIP (-3)         /* HIP runtime error */
#define SFH_ERR_NODEVICE (-4)    /* no gfx950 device */
#define SFH_ERR_UNSUPPORTED (-5) /* configuration the kernels do not implement */

typedef struct sfh_mesh_t* sfh_mesh;
typedef struct sfh_ocean_t* sfh_ocean;
typedef struct sfh_batch_t* sfh_batch;

/* Message of the last failed call on this thread ("" if none). */
const char* sfh_last_error(void);

/* Selects the HIP device used by this thread's later calls. */
int sfh_init(int device);

/* ---- meshes ------------------------------------------------------------ */

/* Uploads a triangle mesh to device memory.
 * Replaces: the `const Mesh*` argument of the per-face loops
 * (OpenGLDataStructs.h:190-310; Mesh::getVertexPos at 207-210 / 259-262).
 * vertex_pos points at Vertex::pos of vertex 0; vertex_stride is the byte
 * distance between vertices: sizeof(Vertex) = 24, sizeof(TexturableVertex) =
 * 44, or 12 for packed xyz.  face_ids holds Face::vertexID, 3 per face.  The
 * device copy is packed: 12 B per vertex, 12 B per face. */
int sfh_mesh_create(const void* vertex_pos, size_t vertex_stride, uint32_t nv,
                    const uint32_t* face_ids, uint32_t nf, sfh_mesh* out);
int sfh_mesh_destroy(sfh_mesh mesh);
/* Vertex/face counts and the float AABB of the vertices (mesh frame). */
int sfh_mesh_info(sfh_mesh mesh, uint32_t* nv, uint32_t* nf, float aabb_min[3], float aabb_max[3]);

/* ---- ocean ------------------------------------------------------------- */

/* Liquid density and gravity.
 * Replaces: Ocean::getLiquid().density (Ocean.h) and
 * SimulationManager::getGravity() (SimulationManager.cpp:662-665). */
int sfh_ocean_create(double density, const double gravity[3], sfh_ocean* out);
int sfh_ocean_destroy(sfh_ocean ocean);

/* Wave height data for Ocean::GetDepth (Ocean.cpp:149-168) through
 * OpenGLRealOcean::ComputeWaveHeight (OpenGLRealOcean.cpp:280-331): n x n
 * texels, `channels` interleaved floats per texel (4 for the reference's RGBA
 * fftData), layer l reads channel l and spans lengths[l] metres (the
 * reference: 2 layers, gridSizes.x and .y).  height(x,y) = - sum over layers
 * of the periodic bilinear interpolation.  data == NULL or layers == 0 makes
 * the surface flat (hasWaves() false).  host_data: 1 if `data` is host memory,
 * 0 if device memory. */
int sfh_ocean_set_waves(sfh_ocean ocean, const float* data, int host_data, uint32_t n,
                        uint32_t channels, uint32_t layers, const float* lengths);

/* Velocity fields summed by Ocean::GetFluidVelocity (Ocean.cpp:183-201),
 * in the order added:
 *   type 0 Uniform (Uniform.cpp:51-54): params[0..2] = v;
 *   type 1 Jet (Jet.cpp:54-75): params[0..2] = outlet centre c,
 *          params[3..5] = unit axis n, params[6] = r, params[7] = vout;
 *   type 2 Pipe (Pipe.cpp:57-78): params[0..2] = p1, params[3..5] = unit
 *          axis (p2 - p1) / l, params[6] = l, params[7] = r1, params[8] = r2,
 *          params[9] = vin, params[10] = gamma.
 * Jets and pipes (at most 4) are evaluated at every face centroid; uniform
 * currents fold into the body's constant velocity.  Errors: SFH_ERR_UNSUPPORTED
 * for another type, SFH_ERR_SIZE for a fifth jet or pipe, SFH_ERR_INVALID for
 * an axis that is not a unit vector or a pipe length <= 0.  With jets or pipes
 * enabled the surrogate flags every unit out of distribution (its features
 * carry the uniform current only). */
int sfh_ocean_add_current(sfh_ocean ocean, int type, int enabled, const double params[12]);
/* Ocean::EnableCurrents / DisableCurrents (Ocean.cpp:203-211). */
int sfh_ocean_enable_currents(sfh_ocean ocean, int enabled);

/* ---- reference-shaped per-body calls (host memory, synchronous) ----------- */

typedef struct {
    int32_t damping_forces;      /* HydrodynamicsSettings::dampingForces */
    int32_t realistic_buoyancy;  /* HydrodynamicsSettings::reallisticBuoyancy */
} sfh_settings;

/* Replaces: SolidEntity::ComputeHydrodynamicForcesSurface (SolidEntity.h:135-137,
 * SolidEntity.cpp:1289-1721).  Outputs the reference leaves untouched (buoyancy
 * when the submerged volume is <= 1e-9, drag when damping is off) keep the
 * caller's values.  mesh == NULL zeroes the outputs as the reference does. */
int sfh_forces_surface(const sfh_settings* settings, sfh_mesh mesh, sfh_ocean ocean,
                       const double T_CG[16], const double T_C[16],
                       const double v[3], const double omega[3],
                       double Fb[3], double Tb[3], double Fdq[3], double Tdq[3],
                       double Fdf[3], double Tdf[3], double* Swet, double* Vsub);

/* Replaces: SolidEntity::ComputeHydrodynamicForcesSubmerged (SolidEntity.h:152-153,
 * SolidEntity.cpp:1723-1797). */
int sfh_forces_submerged(sfh_mesh mesh, sfh_ocean ocean,
                         const double T_CG[16], const double T_C[16],
                         const double v[3], const double omega[3],
                         double Fdq[3], double Tdq[3], double Fdf[3], double Tdf[3]);

/* Replaces: SolidEntity::ComputeAerodynamicForces(mesh, atm, ...) (SolidEntity.h:180,
 * SolidEntity.cpp:1865-1918): every face whose normal relative velocity vn
 * points into it adds the reference's `vn * vn.length() * A`
 * (SolidEntity.cpp:1913), where glm's vec3::length() is the component count 3,
 * so 3 A vn (no skin friction); the sums times 0.5 rho_gas.  The atmosphere is an
 * sfh_ocean made with the gas density, flat, with every wind field added
 * enabled and sfh_ocean_enable_currents(h, 1): Atmosphere::GetFluidVelocity
 * (Atmosphere.cpp:146-152) sums all of them.  mesh == NULL zeroes the outputs
 * as the reference does. */
int sfh_forces_aero(sfh_mesh mesh, sfh_ocean atmosphere, const double T_CG[16], const double T_C[16],
                    const double v[3], const double omega[3], double Fda[3], double Tda[3]);

/* Replaces: SolidEntity::CorrectAerodynamicForces (SolidEntity.h:167,
 * SolidEntity.cpp:1920-1948), host FP64: scales Fda and Tda by
 * |Cd . safeNormalize(inverse(basis of T_CG2H) Fda)| with Cd of the body's
 * proxy (sfh_compute_fluid_dynamics_approx: type, params, T_CG2H); 1 for a
 * sphere. */
int sfh_correct_aero(int type, const double params[3], const double T_CG2H[16], double Fda[3], double Tda[3]);

/* Replaces: SolidEntity::CorrectHydrodynamicForces (SolidEntity.h:113-114,
 * SolidEntity.cpp:1263-1287).  In/out drag vectors; T_O = getOTransform(). */
int sfh_correct(sfh_ocean ocean, double Fdq[3], double Tdq[3], double Fdf[3], double Tdf[3],
                const double Cd[3], const double Cf[3], const double T_O[16]);

/* ---- per-body evaluation that allocates nothing --------------------------- */

#define SFH_NOALLOC_SURFACE 0    /* as sfh_forces_surface */
#define SFH_NOALLOC_SUBMERGED 1  /* as sfh_forces_submerged */

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  Computes what sfh_forces_surface (mode
 * SFH_NOALLOC_SURFACE, with `settings`) or sfh_forces_submerged (mode
 * SFH_NOALLOC_SUBMERGED, settings ignored) computes, with the same kernels,
 * asynchronously on `stream` (a hipStream_t, NULL for the default stream),
 * from device memory:
 *   d_state: 38 doubles, T_CG[16] T_C[16] v[3] omega[3] (OpenGL order);
 *   d_out:   SFH_OUT_DOUBLES doubles laid out as a batch output row; slots
 *            the reference leaves untouched keep their contents, the
 *            written mask and position are stored in slots 20 and 21.
 * Its work buffers live in one file-scope device array of under 256 KB per
 * GPU (held until the process ends), sized for meshes of at most 524,288
 * faces; a larger mesh returns SFH_ERR_SIZE.  It calls
 * no hipMalloc, hipFree or allocating routine.  Every call on one GPU shares
 * that array: calls that overlap on one GPU (several streams or threads)
 * race, so the caller orders them.  mesh must not be NULL. */
int sfh_forces_noalloc(const sfh_settings* settings, int mode, sfh_mesh mesh, sfh_ocean ocean,
                       const double* d_state, double* d_out, void* stream);

/* ---- batched body-level evaluation ------------------------------------- */

/* Static description of one body: what SolidEntity::ComputeHydrodynamicForces
 * (SolidEntity.cpp:1799-1850) reads from the entity besides its state. */
typedef struct {
    sfh_mesh mesh;            /* getPhysicsMesh() */
    double T_CG2C[16];        /* getCTransform() = getCGTransform() * T_CG2C */
    double T_CG2O[16];        /* getOTransform() = getCGTransform() * T_CG2O */
    double aabb_min[3];       /* collision-shape AABB in the CG frame: the  */
    double aabb_max[3];       /* btCompoundShape local AABB (getAABB, 215-226) */
    double Cd[3], Cf[3];      /* fdCd, fdCf */
    double volume, surface;   /* volume, surface */
    double P_CB[3];           /* P_CB (CG frame) */
    int32_t buoyant;          /* isBuoyant() */
    int32_t hydro;            /* phy.mode is FLOATING or SUBMERGED */
} sfh_body_desc;

/* Per-body state: 22 doubles, T_CG[16] (getCGTransform), v[3], omega[3]. */
#define SFH_STATE_DOUBLES 22
/* Per-body output: 24 doubles: Fb[3] Tb[3] Fdq[3] Tdq[3] Fdf[3] Tdf[3] Swet
 * Vsub, then the written mask (bit 0 Fb, 1 Tb, 2 Fdq, 3 Tdq, 4 Fdf, 5 Tdf,
 * 6 Swet, 7 Vsub) and the body/fluid position (0 outside, 1 inside,
 * 2 crossing, -1 for a body that is not hydrodynamic: the mode gate returns
 * before classifying it), both stored as doubles, and 2 spare. */
#define SFH_OUT_DOUBLES 24

typedef struct {
    uint32_t faces_per_lane;  /* chunk = 64 lanes x faces_per_lane faces; 0 = default */
    uint32_t waves;           /* wavefronts of the face launch; 0 = default */
    uint32_t split_passes;    /* 0: one face launch runs the submerged and the crossing
                                 chunks; 1: a launch for each, each at its own occupancy */
} sfh_batch_opts;

/* Replaces the per-body pool fan-out of SimulationManager::SimulationTickCallback
 * (SimulationManager.cpp:1619-1655) + Ocean::ApplyFluidForces (Ocean.cpp:219-268)
 * + SolidEntity::ComputeHydrodynamicForces: one batch holds many bodies and
 * evaluates all of them in one pass on one stream.  Cuts every body's faces
 * into chunks (runs of consecutive faces, each owned by one body). */
int sfh_batch_create(const sfh_body_desc* bodies, uint32_t n, const sfh_batch_opts* opts, sfh_batch* out);
int sfh_batch_destroy(sfh_batch batch);
/* A new faces_per_lane rebuilds the batch's chunks: graphs captured from
 * the batch before (sfh_batch_capture, or by the caller around its compute)
 * must be captured again. */
int sfh_batch_set_opts(sfh_batch batch, const sfh_batch_opts* opts);
int sfh_batch_get_opts(sfh_batch batch, sfh_batch_opts* opts);

typedef struct {
    uint32_t faces_per_lane;   /* chunk of 64 x faces_per_lane faces */
    uint32_t waves_per_cu;     /* wavefronts of the face launch per compute unit */
    uint32_t split_passes;     /* as in sfh_batch_opts */
    uint32_t pad_;
    double step_ms;            /* average time of the whole step (every launch) */
    double face_ms;            /* of the face launches, last timed step */
} sfh_tune_point;

/* No counterpart in the reference (its only knob is the thread pool size,
 * SimulationApp.cpp:47).  Autotuner of the batched step: times the step on
 * the caller's own inputs (d_states; d_active as in sfh_batch_compute_masked,
 * or NULL) for faces_per_lane in {4, 8, 16, 32} x waves_per_cu in {8, 16,
 * 24, 32, 48, 64, 96, 128, 192, 256} (a face launch has at most one
 * wavefront per chunk) x split_passes in {0, 1} (0 only for an aerodynamic
 * batch), reps timed launches each (0 = 5), and leaves the batch set to the
 * fastest.  d_out is written.  The measured points (up to
 * max_points) go to `points`, their number to *npoints.  Synchronous; graphs
 * captured from the batch before must be captured again. */
int sfh_batch_autotune(sfh_batch batch, sfh_ocean ocean, const double* d_states, const uint8_t* d_active,
                       double* d_out, void* stream, uint32_t reps, sfh_tune_point* points,
                       uint32_t max_points, uint32_t* npoints);
/* nbodies: rows of the state and output arrays (bodies, or compounds);
 * nchunks, nfaces: the face pass's work (over the parts of compounds). */
int sfh_batch_info(sfh_batch batch, uint32_t* nbodies, uint32_t* nchunks, uint64_t* nfaces);

/* ---- compounds ----------------------------------------------------------- */

/* One part of a compound (Compound.h CompoundPart + the part solid). */
typedef struct {
    sfh_mesh mesh;            /* part solid getPhysicsMesh() */
    double origin[16];        /* CompoundPart::origin (compound O frame -> part O frame) */
    double O2C[16];           /* part solid getO2CTransform() */
    double Cd[3], Cf[3];      /* part solid getHydrodynamicCoefficients() */
    int32_t external;         /* CompoundPart::isExternal */
    int32_t buoyant;          /* part solid isBuoyant() */
    int32_t hydro;            /* part solid mode is FLOATING or SUBMERGED */
} sfh_part_desc;

/* A compound body: `body` carries the compound's T_CG2O, CG-frame AABB,
 * volume, surface, P_CB, buoyant and hydro (its mesh, T_CG2C, Cd, Cf are not
 * used); parts[0 .. nparts-1] in the compound's order. */
typedef struct {
    sfh_body_desc body;
    const sfh_part_desc* parts;
    uint32_t nparts;
} sfh_compound_desc;

/* Replaces Compound::ComputeHydrodynamicForces (Compound.cpp:291-437) for a
 * batch of compounds: every part is a body of the face pass (classified
 * with its compound's AABB, T_C = T_CG * T_CG2O * origin * O2C, corrected
 * with its own Cd, Cf and T_O * origin), then one lane per compound composes
 * the part results in part order.  The batch's states and outputs are per
 * compound (n rows); every batch entry point applies. */
int sfh_compound_batch_create(const sfh_compound_desc* comps, uint32_t n, const sfh_batch_opts* opts,
                              sfh_batch* out);

/* ---- aerodynamic batches ------------------------------------------------ */

/* One body of Atmosphere::ApplyFluidForces (Atmosphere.cpp:160-195): what
 * SolidEntity::ComputeAerodynamicForces(atm) (SolidEntity.cpp:1852-1863) and
 * CorrectAerodynamicForces (1925-1948) read from the entity. */
typedef struct {
    sfh_mesh mesh;            /* getPhysicsMesh() (NULL: zero drag) */
    double T_CG2C[16];        /* getCTransform() = getCGTransform() * T_CG2C */
    double T_CG2H[16];        /* the drag proxy's frame (its basis only is used) */
    double approx_params[3];  /* fdApproxParams */
    int32_t approx_type;      /* fdApproxType: SFH_APPROX_* */
    int32_t aero;             /* phy.mode == AERODYNAMIC (else the row is untouched) */
} sfh_aero_desc;

/* One part of Compound::ComputeAerodynamicForces (Compound.cpp:439-464). */
typedef struct {
    sfh_mesh mesh;            /* part solid getPhysicsMesh() */
    double origin[16];        /* CompoundPart::origin */
    double O2C[16];           /* part solid getO2CTransform() */
    double T_CG2H[16];        /* part solid's drag proxy frame */
    double approx_params[3];  /* part solid fdApproxParams */
    int32_t approx_type;      /* part solid fdApproxType */
    int32_t external;         /* CompoundPart::isExternal (internal parts add nothing) */
} sfh_aero_part_desc;

typedef struct {
    double T_CG2O[16];        /* the compound's getOTransform() = getCGTransform() * T_CG2O */
    int32_t aero;             /* the compound's phy.mode == AERODYNAMIC */
    int32_t pad_;
    const sfh_aero_part_desc* parts;
    uint32_t nparts;
} sfh_aero_compound_desc;

/* Replaces the per-body loop of Atmosphere::ApplyFluidForces (recompute
 * step) + SolidEntity::ComputeAerodynamicForces(atm): a batch whose compute
 * runs the aerodynamic face loop for every body with aero set, then
 * CorrectAerodynamicForces.  Evaluate with sfh_batch_compute* / capture /
 * replay on an atmosphere (an sfh_ocean as for sfh_forces_aero): each output
 * row holds Fda in [6..8] and Tda in [9..11], written mask 12 and position 1;
 * rows of bodies without aero keep their values (mask 0, position -1).  The
 * surrogate entry points reject these batches. */
int sfh_aero_batch_create(const sfh_aero_desc* bodies, uint32_t n, const sfh_batch_opts* opts, sfh_batch* out);

/* Compound::ComputeAerodynamicForces for a batch of compounds: every external
 * part runs the face loop with T_C = T_CG * T_CG2O * origin * O2C and its own
 * CorrectAerodynamicForces; the compound's row is their sum (rows as above). */
int sfh_aero_compound_batch_create(const sfh_aero_compound_desc* comps, uint32_t n, const sfh_batch_opts* opts,
                                   sfh_batch* out);

/* Evaluates every body (classification, per-face loop, FP64 reduction,
 * buoyancy, CorrectHydrodynamicForces).  d_states: n x SFH_STATE_DOUBLES,
 * d_out: n x SFH_OUT_DOUBLES, both device memory; `stream` is a hipStream_t
 * (NULL = default stream).  Asynchronous.  Outputs the reference leaves
 * untouched keep the values d_out held. */
int sfh_batch_compute(sfh_batch batch, sfh_ocean ocean, const double* d_states, double* d_out, void* stream);

/* Same with host buffers; synchronous.  Its device copies of the buffers are
 * allocated on the first call and kept with the batch. */
int sfh_batch_compute_host(sfh_batch batch, sfh_ocean ocean, const double* states, double* out);

/* sfh_batch_compute restricted to the bodies with d_active[i] != 0 (device,
 * n bytes; NULL = all).  The face pass evaluates only the active bodies'
 * chunks; the output rows of inactive bodies are not written.  This is
 * the exact fallback of the surrogate (the bodies outside its training range)
 * and its audit. */
int sfh_batch_compute_masked(sfh_batch batch, sfh_ocean ocean, const double* d_states,
                             const uint8_t* d_active, double* d_out, void* stream);

/* Captures sfh_batch_compute(batch, ocean, d_states, d_out) on `stream` into a
 * HIP graph held by the batch, then sfh_batch_replay launches it. */
int sfh_batch_capture(sfh_batch batch, sfh_ocean ocean, const double* d_states, double* d_out, void* stream);
int sfh_batch_replay(sfh_batch batch, void* stream);

/* Average durations (ms) over the last `count` sfh_batch_compute calls made
 * with timing on (HIP events on the launch stream): face_ms of the face
 * launches, total_ms of the whole step (every launch). */
int sfh_batch_timing(sfh_batch batch, int enable);
int sfh_batch_kernel_ms(sfh_batch batch, double* face_ms, double* total_ms, uint32_t* count);

/* ---- wave field builder ------------------------------------------------ */

/* One sinusoid of the wave elevation field:
 *   eta(x, y) = sum_k a_k cos(2 pi (mx_k x + my_k y) / L + phase_k)
 * with integer wavenumbers so that the field is periodic on the L x L tile. */
typedef struct {
    float amplitude;
    float phase;
    int32_t mx, my;
} sfh_wave_component;

/* Builds the n x n elevation grid (1 channel, texel (i, j) at x = i L / n,
 * y = j L / n) into device memory d_grid.  This is the input generator of the
 * wave-field workload; it feeds sfh_ocean_set_waves like the reference's
 * fftData. */
int sfh_wave_grid_build(const sfh_wave_component* comps, uint32_t ncomp, uint32_t n,
                        float* d_grid, void* stream);

/* ---- tick driver ----------------------------------------------------------- */

/* Replaces the geometry-based force step of SimulationManager::
 * SimulationTickCallback (SimulationManager.cpp:1583-1655) with
 * Ocean::ApplyFluidForces (Ocean.cpp:219-268) or Atmosphere::ApplyFluidForces
 * (Atmosphere.cpp:160-195) for every body of a batch: on ticks where
 * fdCounter % fdPrescaler == 0 (the counter starts at 0 and advances every
 * step) the bodies with d_overlap[i] != 0 (the fluid ghost's overlapping
 * pairs; NULL = all) are recomputed into d_out, which holds the entity
 * members between ticks (zero it once, as the entities start); every tick
 * d_applied (n x 6 doubles, device) receives the force and torque the
 * entities apply: Fb + Fdq + Fdf and Tb + Tdq + Tdf (ApplyHydrodynamicForces,
 * SolidEntity.cpp:1958-1962), or Fda and Tda for an aerodynamic batch, and
 * zeros for rows the fluid does not overlap.  Asynchronous on `stream`: the
 * host can run other work until it synchronises; *recomputed (may be NULL)
 * tells whether this tick recomputed.  The batch must outlive the driver. */
typedef struct sfh_tick_t* sfh_tick;
int sfh_tick_create(sfh_batch batch, uint32_t prescaler, sfh_tick* out);
int sfh_tick_step(sfh_tick tick, sfh_ocean fluid, const double* d_states, const uint8_t* d_overlap,
                  double* d_out, double* d_applied, int* recomputed, void* stream);
int sfh_tick_destroy(sfh_tick tick);

/* ---- headless FFT wave producer ----------------------------------------- */

/* Replaces OpenGLOcean::GenerateWavesSpectrum (OpenGLOcean.cpp:1025-1046) with
 * the sea state of OpenGLRealOcean (OpenGLRealOcean.cpp:43-45: wind = 5 state
 * + 2, omega = 5 exp(-state) + 0.2, A = 1) and OpenGLOcean's constants
 * (km = 370, cm = 0.23, propagating waves): the h0(k) spectra of the four
 * nested grids on the n x n texels, n = 2^passes, as the reference's
 * texSpectrum12 / texSpectrum34 (RGBA per texel: grid 1 re, im, grid 2 re, im
 * and grids 3, 4).  grid_sizes NULL = the reference's (893, 101, 21, 11) m.
 * Random phases from the reference's LCG (SystemUtil.hpp:207-217) started at
 * seed 1234, as for the first ocean of a process.  Host code. */
int sfh_wave_spectrum(float state, uint32_t passes, const float grid_sizes[4], float* spectrum12,
                      float* spectrum34);

/* Replaces OpenGLOcean::Simulate's wave passes (OpenGLOcean.cpp:455-530:
 * oceanInit.frag, oceanFFTX.frag, oceanFFTY.frag, the butterfly table of
 * ComputeButterflyLookupTable, 879-925) and the readback of the height layer
 * into fftData (OpenGLRealOcean.cpp:333-356), without OpenGL.  create uploads
 * the spectra (rounded to binary16 like the reference's RGBA16F textures; n =
 * 2^passes, 1 <= passes <= 11) and the butterfly table; step writes the wave
 * heights at time t (seconds) into d_heights, device memory of n x n x 4
 * floats (texel (x, y) at (y n + x) 4: the heights of grids 1-4), which is
 * the reference's layer-0 layout: sfh_ocean_set_waves(ocean, d_heights, 0,
 * n, 4, 2, grid_sizes) gives Ocean::GetDepth's two-grid query.  Three
 * launches on `stream`, no allocation. */
typedef struct sfh_wave_fft_t* sfh_wave_fft;
int sfh_wave_fft_create(uint32_t passes, const float grid_sizes[4], const float* spectrum12,
                        const float* spectrum34, sfh_wave_fft* out);
int sfh_wave_fft_step(sfh_wave_fft w, float t, float* d_heights, void* stream);
int sfh_wave_fft_info(sfh_wave_fft w, uint32_t* n);
int sfh_wave_fft_destroy(sfh_wave_fft w);

/* ---- surrogate glue (no counterpart in the reference) -------------------- */

/* The learned surrogate (sfmlp.h) sees 13 features of a body's state, all in
 * the CG frame:
 *   0-2   linear velocity relative to the current at the CG (R^T (v - u))
 *   3-5   angular velocity (R^T omega)
 *   6-8   direction of gravity (R^T g / |g|)
 *   9     depth of the centre of buoyancy below the local surface
 *   10-12 the local surface's unit normal, towards increasing depth
 * where the local surface is the plane fitted to the wave field (the heights
 * of Ocean::GetDepth) over the body's horizontal footprint: a 3 x 3 stencil
 * at offsets of the world AABB's half extents (at least one texel of the
 * first layer), weights (1/4, 1/2, 1/4) per axis; a flat sea gives the
 * CoB's depth and the normal of gravity's direction.
 * and predicts the total force and torque about the CG, CG frame,
 *   F = R^T (Fb + Fdq + Fdf), T = R^T (Tb + Tdq + Tdf) after CorrectHydrodynamicForces,
 * relative to a hydrostatic estimate (Fb0, Tb0: the analytic buoyancy of the
 * immersed fraction of the body, P_CB x Fb0) and in units of the body's
 * force and torque scales sF = rho g V + rho A/4 (u^3 + u), sT = sF Lc
 * (u = |v_rel| + |omega| Lc, Lc the largest half extent of the CG-frame AABB):
 *   network output = ((F - Fb0) / sF, (T - Tb0) / sT), then standardised. */
#define SFH_SUR_FEATURES 13
#define SFH_SUR_ROW 32        /* network input row: 13 features, zero-padded */

typedef struct {
    float in_shift[SFH_SUR_FEATURES];   /* network input = (feature - shift) * scale */
    float in_scale[SFH_SUR_FEATURES];
    float in_lo[SFH_SUR_FEATURES];      /* training range: a feature outside it makes */
    float in_hi[SFH_SUR_FEATURES];      /* the body out-of-distribution               */
    float out_shift[6];                 /* CG-frame F, T = network output * scale + shift */
    float out_scale[6];
    float gate_band;                    /* error-aware gate (0: off): a body crossing the   */
                                        /* surface whose depth feature (feature 9, the CB's */
                                        /* depth below the local surface) is below          */
                                        /* in_lo[9] + gate_band is flagged like an          */
                                        /* out-of-distribution one, so the exact kernel     */
                                        /* answers for the barely wet bodies whose small    */
                                        /* forces the network fits worst                    */
} sfh_surrogate_norm;

/* Features of bodies first .. first+count-1 of the batch (states: the whole
 * batch's state array, device).  Every output is optional (NULL) and indexed
 * from 0:
 *   d_raw      count x 13 FP32 features
 *   d_x        rows x 32 BF16 network inputs (rows >= count, a multiple of 32;
 *              rows past count are zero); needs norm
 *   d_ood      count bytes, 1 where a feature is outside [in_lo, in_hi] or not
 *              finite, or the gate flags the body (gate_band), and the body is
 *              in the water; needs norm
 *   d_pos      count int8: the body's position (CheckBodyFluidPosition on its
 *              AABB: 0 outside, 1 inside, 2 crossing; -1 not hydrodynamic)
 * and, from the exact outputs d_exact (the batch's n x SFH_OUT_DOUBLES array;
 * fields whose written bit is clear count as zero):
 *   d_raw_target  count x 6 FP32 ((F - Fb0) / sF, (T - Tb0) / sT)
 *   d_target      count x 6 FP32 the same standardised (network targets); needs norm */
int sfh_surrogate_features(sfh_batch batch, sfh_ocean ocean, const double* d_states, uint32_t first,
                           uint32_t count, uint32_t rows, const sfh_surrogate_norm* norm,
                           float* d_raw, uint16_t* d_x, uint8_t* d_ood, int8_t* d_pos,
                           const double* d_exact, float* d_raw_target, float* d_target, void* stream);

/* World-frame total force and torque (count x 6 doubles, d_ft) of bodies
 * first .. first+count-1 from the network outputs d_y (count x 6 FP32): bodies
 * flagged in d_ood (optional) take the exact totals from d_exact instead, and
 * bodies out of the water (position 0) or not hydrodynamic get zero force and
 * torque; d_pos (optional) receives the positions.
 * With d_exact and d_err (count x 2 FP32, optional), also the audit errors of
 * the network against the exact result: |F - F_exact| / |F_exact| and
 * |T - T_exact| / |T_exact| (0 when both are 0). */
int sfh_surrogate_apply(sfh_batch batch, sfh_ocean ocean, const double* d_states, uint32_t first, uint32_t count,
                        const sfh_surrogate_norm* norm, const float* d_y, const uint8_t* d_ood,
                        int8_t* d_pos, const double* d_exact, double* d_ft, float* d_err, void* stream);

/* Hydrostatic table of a hull, the surrogate's baseline for bodies crossing
 * the surface (without one, the analytic buoyancy scaled by the immersed
 * fraction of the AABB's height is used): d_table (device, FP32) holds
 * ns x ntheta x nphi entries of 6 floats, F / (rho g V) and T / (rho g V Lc)
 * in the CG frame, of the hull at rest on a flat surface, against
 *   s      = (CoB depth + g^.(aabb_c - P_CB)) / (g^.|aabb_h|) on ns nodes over [-1, 1],
 *   theta  = acos(g^_z) on ntheta nodes over [0, pi],
 *   phi    = atan2(g^_y, g^_x) on nphi periodic nodes over [-pi, pi),
 * g^ the unit gravity direction in the CG frame (trilinear lookup).  The
 * table is built with the exact kernel (surrogate.py build_hydrostatic_table)
 * and must stay alive with the batch; NULL removes it. */
int sfh_surrogate_set_table(sfh_batch batch, uint32_t first, uint32_t count, const float* d_table,
                            uint32_t ns, uint32_t ntheta, uint32_t nphi);

/* One-launch surrogate step.  A plan binds a batch whose units are grouped
 * by hull (counts[g] consecutive units per group, g < 16) to one network and
 * one normalisation per group; the network's device weights come from
 * sfm_weights (sfmlp.h).  The plan must not outlive the batch or the networks. */
typedef struct sfh_sur_plan_t* sfh_sur_plan;
typedef struct {
    const uint16_t* weights[5];   /* BF16 [out x in], padded widths {32,256,256,256,256,32},
                                     in the fragment order of sfm_weights */
    const float* bias[5];
} sfh_net_ptrs;

int sfh_surrogate_plan_create(sfh_batch batch, uint32_t ngroups, const uint32_t* counts,
                              const sfh_surrogate_norm* norms, const sfh_net_ptrs* nets, sfh_sur_plan* out);
int sfh_surrogate_plan_destroy(sfh_sur_plan plan);

/* Every unit of the batch in one launch (tiles of 32 or 64 units per
 * workgroup): the features, the out-of-distribution flag (d_ood, optional)
 * and position (d_pos, optional) as sfh_surrogate_features gives them, the
 * network (BF16 MFMA, activations in LDS) and the decode to world-frame F, T
 * (d_ft, n x 6) as sfh_surrogate_apply gives it. */
int sfh_surrogate_run(sfh_sur_plan plan, sfh_ocean ocean, const double* d_states, uint8_t* d_ood, int8_t* d_pos,
                      double* d_ft, void* stream);

/* After an exact evaluation d_exact (n x SFH_OUT_DOUBLES, e.g. by
 * sfh_batch_compute_masked on the OOD units): with d_err (n x 2, optional)
 * the surrogate's F and T errors against it for every unit, then the exact
 * totals replace d_ft for the units flagged in d_ood (optional). */
int sfh_surrogate_merge(sfh_sur_plan plan, const uint8_t* d_ood, const double* d_exact, double* d_ft, float* d_err,
                        void* stream);

/* ---- hull preprocessing (host, once per hull) -------------------------------- */

/* ComputePhysicalProperties (GeometryFileUtil.cpp:335-563): mass, volume,
 * surface, CG, principal moments Ipri and principal axes Irot (row-major,
 * axes as columns) of a closed triangle mesh (thickness <= 0: solid) or of a
 * shell of the given thickness.  xyz: nv x 3 floats (mesh frame), idx:
 * nf x 3 vertex indices. */
typedef struct {
    double mass, volume, surface;
    double CG[3];
    double Ipri[3];
    double Irot[9];
} sfh_mesh_properties;

int sfh_compute_physical_properties(const float* xyz, uint32_t nv, const uint32_t* idx, uint32_t nf,
                                    double thickness, double density, sfh_mesh_properties* out);

/* SolidEntity::ComputeFluidDynamicsApprox (SolidEntity.cpp:698-1039): the
 * body's fluid-dynamics proxy from its physics-mesh vertices, as the
 * reference computes it for a rigid body.
 *   type: SFH_APPROX_AUTO / _ELLIPSOID (the Kumar-Yildirim minimum-volume
 *         axis-aligned ellipsoid, at most 10 iterations, 20 % volume
 *         tolerance), _SPHERE, _CYLINDER (GeometryApproxType order);
 *   T_CG2C, T_CG2O: OpenGL column-major 4x4 transforms (CG -> collision
 *         geometry, CG -> origin); P_CB: centre of buoyancy, CG frame;
 *   rho: the liquid's density (the reference uses 1000 without an ocean).
 * Outputs the proxy type and parameters (sphere r; cylinder r, length;
 * ellipsoid semi-axes), its frame T_CG2H, added mass and inertia, and the
 * drag and skin-friction coefficients Cd, Cf (origin frame, Cf = 0.1 Cd) of
 * SetHydrodynamicCoefficients.  The reference's quirks are kept: the sphere's
 * added mass uses the squared radius in place of the radius, and Lamb's k
 * factor takes e = 1 - r2^2 / r1. */
#define SFH_APPROX_AUTO 0
#define SFH_APPROX_SPHERE 1
#define SFH_APPROX_CYLINDER 2
#define SFH_APPROX_ELLIPSOID 3
typedef struct {
    int32_t type;          /* SFH_APPROX_SPHERE / _CYLINDER / _ELLIPSOID */
    int32_t iterations;    /* ellipsoid: iterations run */
    double params[3];
    double T_CG2H[16];
    double added_mass[3], added_inertia[3];
    double Cd[3], Cf[3];
} sfh_fluid_approx;

int sfh_compute_fluid_dynamics_approx(const float* xyz, uint32_t nv, int type, const double T_CG2C[16],
                                      const double P_CB[3], const double T_CG2O[16], double rho,
                                      sfh_fluid_approx* out);

/* ---- rollout harness (no counterpart in the hydrodynamics path) ------------ */

/* Rigid-body data of identical bodies: mass and principal moments of inertia
 * about the CG frame's axes. */
typedef struct {
    double mass;
    double inertia[3];
} sfh_rigid;

/* Advances count body states (d_states, count x SFH_STATE_DOUBLES, device) by
 * dt under their world-frame total force and torque d_ft (count x 6) plus
 * gravity (the ocean's): semi-implicit Euler of the Newton-Euler equations,
 * orientation by the exact rotation of omega dt.  With d_ood (optional,
 * needs d_exact: n x SFH_OUT_DOUBLES exact outputs) the exact totals first
 * replace d_ft for the flagged bodies, as sfh_surrogate_merge does, in the
 * same launch.  The reference integrates bodies with Bullet (btRigidBody),
 * outside this path; this step drives the mixed exact/surrogate rollout. */
int sfh_rollout_integrate(sfh_ocean ocean, double* d_states, uint32_t count, double* d_ft, const uint8_t* d_ood,
                          const double* d_exact, const sfh_rigid* rigid, double dt, void* stream);

#ifdef __cplusplus
}
#endif
#endif
