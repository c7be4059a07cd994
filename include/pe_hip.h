/*
 * pe_hip.h -- C-ABI of the MI355X-native HCSITS rigid-body time step.
 *
 * One context = one pe world (bodies, materials, collision-system settings) whose
 * state lives in HBM; pehip_step advances it with hand-written gfx950 kernels:
 * hash-grid broadphase, sphere/box/plane narrowphase, contact-graph colouring,
 * coloured non-linear block Gauss-Seidel (NBGS) relaxation with the
 * InelasticCoulombContactByDecoupling model, and semi-implicit integration.
 * Arithmetic is FP64 throughout.
 *
 * Conventions: every call returns pehip_status; no exception crosses the ABI.
 * The caller owns host arrays (copied during the call); the library owns all
 * device memory.  A context is used by one host thread at a time and owns one
 * HIP stream.  Arrays are structure-of-arrays in the caller's body order, which
 * is pe's BodyStorage order (creation order, pe/core/rigidbody/BodyStorage.h:487-492).
 * Reference citations are relative to the pe source tree.
 */
#ifndef PE_HIP_H
#define PE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pehip_ctx pehip_ctx;

typedef enum {
  PEHIP_OK = 0,
  PEHIP_INVALID_ARG = 1,
  PEHIP_OUT_OF_MEMORY = 2,
  PEHIP_CAPACITY_RETRY = 3,
  PEHIP_HIP_ERROR = 4,
  PEHIP_NONFINITE = 5,
  PEHIP_NOT_SUPPORTED = 6
} pehip_status;

typedef enum { PEHIP_SPHERE = 1, PEHIP_BOX = 2, PEHIP_PLANE = 3, PEHIP_CAPSULE = 4, PEHIP_UNION = 5 } pehip_shape;

/* CollisionSystem<HCSITS>::RelaxationModel, same numbering
 * (pe/core/collisionsystem/HardContactSemiImplicitTimesteppingSolvers.h:177-184). */
typedef enum {
  PEHIP_INELASTIC_FRICTIONLESS = 0,
  PEHIP_APPROX_INELASTIC_COULOMB_DECOUPLING = 1,
  PEHIP_APPROX_INELASTIC_COULOMB_ORTHO = 2,
  PEHIP_INELASTIC_COULOMB_DECOUPLING = 3,
  PEHIP_INELASTIC_COULOMB_ORTHO = 4,
  PEHIP_INELASTIC_GMD = 5
} pehip_relaxation_model;

typedef enum {
  /* Contacts are coloured so that no two contacts of one colour share a non-fixed
   * body; each colour is relaxed in parallel.  The production mode. */
  PEHIP_SWEEP_COLOURED = 0,
  /* One wavefront relaxes the contacts strictly sequentially, in the order given by
   * pehip_set_contact_order (verification mode: pe's sweep order, HCSITS.h:2341). */
  PEHIP_SWEEP_REFERENCE_ORDER = 1
} pehip_sweep_mode;

typedef struct {
  uint32_t max_iterations;     /* setMaxIterations, HCSITS.h:793 (default 100, HCSITS.h:421) */
  int32_t relaxation_model;    /* setRelaxationModel, HCSITS.h:814 */
  double erp;                  /* setErrorReductionParameter, HCSITS.h:840 (default 0.7, :420) */
  double abort_threshold;      /* <= 0: always max_iterations sweeps, as pe (HCSITS.h:1935-1946
                                  is compiled out); > 0: stop when the L-inf impulse change of
                                  a sweep drops below it (the reference's 1e-7, HCSITS.h:424) */
  double gravity[3];           /* World::setGravity, World.h:754, 767 */
  double damping;              /* World::setDamping, World.h:736 (1 = off) */
  double relaxation_param;     /* setRelaxationParameter, HCSITS.h:770 (MPI-only, stored) */
  uint32_t max_sub_iterations; /* golden-section steps, HCSITS.h:423 (default 20) */
  int32_t sweep_mode;          /* pehip_sweep_mode */
  int32_t exact_trig;          /* 1: correctly rounded sin/cos/atan2/acos in the golden-section
                                  search and the orientation update, reproducing the host libm
                                  (glibc) wherever it is correctly rounded; 0: device libm.  Always
                                  on in PEHIP_SWEEP_REFERENCE_ORDER. */
  int32_t sweep_schedule;      /* coloured mode only -- pehip_sweep_schedule */
} pehip_params;

typedef enum {
  /* abort_threshold <= 0: dataflow when the step's contacts fit it (65,536), else colour launches;
   * abort_threshold > 0: one workgroup for up to 512 contacts of relaxation models 0 and 1,
   * else colour launches */
  PEHIP_SCHEDULE_AUTO = 0,
  /* one kernel launch per colour and sweep */
  PEHIP_SCHEDULE_COLOUR_LAUNCHES = 1,
  /* all sweeps in one cooperative launch; each contact waits only for the earlier updates of its
   * own bodies (same results as the colour launches, bit for bit); abort_threshold must be <= 0 */
  PEHIP_SCHEDULE_DATAFLOW = 2,
  /* all sweeps in one launch of one workgroup, a barrier between colours and the abort test after
   * every sweep on the device (same results as the colour launches, bit for bit); for small steps */
  PEHIP_SCHEDULE_SINGLE_BLOCK = 3,
  /* all sweeps of each island of the contact graph (the non-fixed bodies joined by contacts) in
   * one wavefront (up to 64 contacts, several islands per wavefront) or one workgroup (up to 256),
   * bodies in LDS, one lane per contact; larger islands by dataflow (same results as the colour
   * launches, bit for bit); abort_threshold must be <= 0 */
  PEHIP_SCHEDULE_ISLANDS = 4
} pehip_sweep_schedule;

/* Fills p with pe's defaults (HCSITS.h:408-431, Settings.cpp:46-48). */
void pehip_default_params(pehip_params* p);

/* Replaces the CollisionSystem<HCSITS> singleton constructor (HCSITS.h:408-431) and
 * World (src/core/World.cpp:65).  device = HIP device ordinal. */
pehip_status pehip_create(const pehip_params* params, int device, pehip_ctx** out);
/* Ends the world (World::~World, src/core/World.cpp:88) and frees the context's device memory. */
pehip_status pehip_destroy(pehip_ctx* ctx);
/* All settings of pehip_params at once: the CollisionSystem and World setters each field cites
 * (HCSITS.h:770-840, World.h:736-767), validated as they validate; the next step uses them. */
pehip_status pehip_set_params(pehip_ctx* ctx, const pehip_params* params);
pehip_status pehip_get_params(const pehip_ctx* ctx, pehip_params* params);

/* Material pair tables already combined on the host: n x n row-major static and
 * dynamic friction (Materials.cpp:237-243 combine csf_i + csf_j), replacing
 * Material::get{Static,Dynamic}Friction (pe/core/Materials.h:500-537).  The tables may grow
 * between steps (pe only ever adds materials); with bodies uploaded, n below the current count is
 * refused (the bodies' material IDs index the tables).  A failure keeps the previous tables. */
pehip_status pehip_set_materials(pehip_ctx* ctx, uint32_t n, const double* csf_pair,
                                 const double* cdf_pair);

/* Uploads all bodies, replacing World::add -> CollisionSystem::add (World.cpp:262-277,
 * HCSITS.h:1305-1325).  Per body i:
 *   shape[i]           pehip_shape
 *   sid[i]             system id (UniqueID.h:143); sphere-sphere contacts order their bodies by it
 *   material[i]        row of the pair tables
 *   fixed[i]           1 = fixed (planes are always fixed)
 *   param[3i..]        sphere: radius,0,0; box: side lengths; capsule: radius,length,0;
 *                      union: bounding radius about its centre of mass (largest |rpos| + member
 *                      radius), smallest member radius, 0 (used to size the hash grids);
 *                      plane: unused
 *   pos[3i], q[4i], v[3i], w[3i]   state (R is derived from q as R = q.toRotationMatrix())
 *   inv_mass[i], inertia_body[9i], inv_inertia_body[9i]   mass properties (body frame)
 *   plane_normal[3i], plane_d[i]   plane parameters, host-computed (meaningful for planes)
 * Any previous bodies are replaced.  A failure after validation (out of device memory) leaves
 * the context without bodies. */
pehip_status pehip_upload_bodies(pehip_ctx* ctx, uint32_t n, const uint8_t* shape,
                                 const uint64_t* sid, const uint32_t* material,
                                 const uint8_t* fixed, const double* param, const double* pos,
                                 const double* q, const double* v, const double* w,
                                 const double* inv_mass, const double* inertia_body,
                                 const double* inv_inertia_body, const double* plane_normal,
                                 const double* plane_d);

/* The spheres contained in the union bodies (pe's Union of Spheres, pe/core/rigidbody/Union.h;
 * createGranularParticle, Union.cpp:2376-2402), replacing Union::add (Union.cpp:1482-1545).  Called
 * after pehip_upload_bodies whenever that call holds a union; until then pehip_step and
 * pehip_detect fail.  Per member k, grouped by union in body order and, within a union, in the
 * union's order:
 *   owner[k]     body index of its union (non-decreasing; every union owns at least one member)
 *   sid[k]       the member's own system id (sphere-sphere contacts order by it)
 *   material[k]  row of the pair tables (a union's contacts take their members' materials)
 *   radius[k]    sphere radius
 *   rpos[3k]     position in the union's body frame (RigidBody::calcRelPosition, RigidBody.h:1370)
 *   gpos[3k]     global position
 * The union bodies themselves carry the union's centre of mass, orientation, velocities and
 * mass properties (UnionBase::calcCenterOfMass / calcInertia, UnionBase.cpp:240-360). */
pehip_status pehip_upload_members(pehip_ctx* ctx, uint32_t nm, const uint32_t* owner,
                                  const uint64_t* sid, const uint32_t* material, const double* radius,
                                  const double* rpos, const double* gpos);
/* Global positions of the union members (they follow their unions each step, Sphere.cpp:1052). */
pehip_status pehip_download_members(pehip_ctx* ctx, uint32_t nm, double* gpos);

/* Overwrites the dynamic state of all n bodies (checkpoint restore; RigidBody::setPosition,
 * setOrientation, setLinearVel, setAngularVel).  The members of a union whose pose changes follow
 * it as pe's setters move them: a new orientation places them at pos + R rpos (Union.cpp:306-322,
 * Sphere.cpp:1059), a new position alone translates them (Union.cpp:251-272, Sphere.cpp:1031);
 * pehip_upload_members afterwards places them explicitly. */
pehip_status pehip_upload_state(pehip_ctx* ctx, uint32_t n, const double* pos, const double* q,
                                const double* v, const double* w, const double* plane_normal,
                                const double* plane_d);

/* External force and torque for the next step (RigidBody::addForce/addTorque);
 * reset after the step as Settings::forceReset (Settings.cpp:46). */
pehip_status pehip_set_forces(pehip_ctx* ctx, uint32_t n, const double* force,
                              const double* torque);

/* No counterpart in pe (a test hook: pe relaxes its contacts in HashGrids' emission order).
 * Reference-order mode: the order in which the next step relaxes its contacts, as
 * body-index pairs (body1, body2) in pe's contact emission order.  Contacts of one
 * pair keep their narrowphase order.  Used only with PEHIP_SWEEP_REFERENCE_ORDER;
 * consumed by the next pehip_step. */
pehip_status pehip_set_contact_order(pehip_ctx* ctx, uint32_t n, const uint32_t* body1,
                                     const uint32_t* body2);

/* World::run(steps, dt) (src/core/World.cpp:374-378): nsteps x World::simulationStep
 * (World.cpp:330-357) -> CollisionSystem<HCSITS>::simulationStep (HCSITS.h:1472-1531).
 * dt <= 0 -> PEHIP_INVALID_ARG (World.cpp:333-334). */
pehip_status pehip_step(pehip_ctx* ctx, double dt, uint32_t nsteps);

/* pehip_step that allocates nothing (no counterpart in pe): for callers that cannot allocate
 * device memory -- code under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation.  The step's per-contact arrays (raw and solver contacts, colouring, sweep order,
 * dataflow ranks) live in one static device array of the library (12 MiB per device, held until
 * the process ends) instead of the context's own buffers, so a step may hold at most 16,384
 * contacts: a step that detects more returns PEHIP_INVALID_ARG before changing the state.
 * Coloured sweep mode only, profiling off (PEHIP_NOT_SUPPORTED otherwise).  Every context on a
 * device shares that device's array: the contact, sweep-order and impulse readbacks of a context
 * show the last pehip_step_noalloc on the device until the context's next pehip_step, and calls
 * that overlap on one device (from several host threads or streams) race.  Bodies, materials and
 * all other per-body buffers are allocated beforehand by the upload calls, as for pehip_step. */
pehip_status pehip_step_noalloc(pehip_ctx* ctx, double dt, uint32_t nsteps);

/* Detection only (HashGrids::findContacts, HashGrids.h:1121-1314) on the current state,
 * without resolving or integrating; the contacts are then read with pehip_get_contacts. */
pehip_status pehip_detect(pehip_ctx* ctx);

/* Every body's dynamic state as the RigidBody getters read it (getPosition, getLinearVel,
 * getAngularVel, getQuaternion, pe/core/rigidbody/RigidBody.h:146-151) and the planes' normal and
 * displacement (Plane::getNormal, getDisplacement), in the layout of pehip_upload_state.  Any
 * output may be NULL. */
pehip_status pehip_download_state(pehip_ctx* ctx, uint32_t n, double* pos, double* q,
                                  double* v, double* w, double* plane_normal, double* plane_d);

/* The collision system's active contacts (contacts_, HCSITS.h:291) of the last step or
 * detection, in the library's order (deterministic; pe's is HashGrids' emission order).
 * cap = capacity of the arrays; *n receives the number of contacts (if *n > cap only
 * cap entries are written).  Any output pointer may be NULL. */
pehip_status pehip_get_contacts(pehip_ctx* ctx, size_t cap, size_t* n, uint32_t* body1,
                                uint32_t* body2, double* pos, double* normal, double* dist,
                                double* mu);
/* No counterpart in pe (its sweeps follow the contacts' order).
 * Sweep order of the last step: slot k of the sweep relaxes contact contact_index[k]
 * (pehip_get_contacts numbering).  Coloured mode: colour c occupies slots
 * [colour_offsets[c], colour_offsets[c+1]); *ncolours receives the colour count and
 * colour_offsets needs room for it + 1 entries (at most 129).  Any pointer may be NULL. */
pehip_status pehip_get_sweep_order(pehip_ctx* ctx, size_t cap, size_t* n, uint32_t* contact_index,
                                   uint32_t* ncolours, uint32_t* colour_offsets);

/* Solved impulses of the last step (p_, HCSITS.h:327, world frame as pe keeps them,
 * HCSITS.h:2083), in pehip_get_contacts order. */
pehip_status pehip_get_impulses(pehip_ctx* ctx, size_t cap, size_t* n, double* p);

typedef struct {
  uint64_t num_contacts;      /* CollisionSystem::getNumberOfContacts, HCSITS.h:737 */
  double max_penetration;     /* getMaximumPenetration, HCSITS.h:715 */
  uint32_t iterations;        /* sweeps run in the last step */
  uint32_t num_colours;       /* colours of the last step's contact graph */
  double last_delta;          /* L-inf impulse change of the last sweep */
  double kinetic_energy;      /* 0.5 m v^2 + 0.5 w.(I_w w) over non-fixed bodies, after the step */
  uint64_t fixed_contacts;    /* contacts of the last step that involve a fixed body */
  double sweep_bytes;         /* algorithmic HBM bytes of the last step's sweeps:
                                 sweeps * (296 * contacts + 200 * non-fixed contact ends) */
  /* profiling (pehip_set_profiling): phase times of the last step, and totals since
   * profiling was enabled of the sweep-kernel launches, their summed durations
   * (one HIP event pair per launch on the context stream) and algorithmic bytes */
  double t_detection_ms, t_colouring_ms, t_setup_ms, t_solve_ms, t_integrate_ms;
  uint64_t prof_steps;
  uint64_t prof_sweep_launches;
  double prof_sweep_kernel_ms;
  double prof_sweep_bytes;
  int32_t schedule_used;      /* schedule of the last step's sweeps: PEHIP_SCHEDULE_COLOUR_LAUNCHES,
                                 _DATAFLOW, _SINGLE_BLOCK or _ISLANDS; -1 for reference order or no
                                 sweeps */
  int32_t reserved;
} pehip_stats;
pehip_status pehip_get_stats(pehip_ctx* ctx, pehip_stats* out);
/* The same record without the kinetic-energy reduction (no kernel launch, no synchronisation):
 * kinetic_energy keeps the value of the last pehip_get_stats.  For per-step readout in loops
 * (CollisionSystem::getNumberOfContacts / getMaximumPenetration, HCSITS.h:715-737). */
pehip_status pehip_get_step_stats(pehip_ctx* ctx, pehip_stats* out);

/* Phase timing and per-sweep-launch timing with HIP events on the context's stream
 * (off by default); enabling resets the profiling totals.  pe's counterpart is its per-phase
 * wall-clock timers (HCSITS.h:351), run inside pe_PROFILING_SECTION blocks (HCSITS.h:1479-1531). */
pehip_status pehip_set_profiling(pehip_ctx* ctx, int enable);

/* No counterpart in pe (a test hook).  Evaluates the device trig of the solver on n arguments
 * (checks against a host libm):
 * fn 0 = sin(x), 1 = cos(x), 2 = atan2(y, x), 3 = acos(x), 4 / 5 = sin(x) / cos(x) as the
 * golden-section search evaluates them (one sincos); exact selects the correctly
 * rounded implementation (see pehip_params.exact_trig).  y may be NULL unless fn = 2. */
pehip_status pehip_eval_trig(pehip_ctx* ctx, int fn, int exact, size_t n, const double* x,
                             const double* y, double* out);

/* No counterpart in pe.  The context's HIP stream (hipStream_t), for callers that time or
 * order work on it. */
void* pehip_stream(pehip_ctx* ctx);

/* Human-readable message of the last failing call on this context (or of the last
 * failing pehip_create when ctx is NULL). */
const char* pehip_last_error(const pehip_ctx* ctx);

/* Library version string (no counterpart in pe). */
const char* pehip_version(void);

/* ------------------------------------------------------------------------------------------
 * Differentiable time step (no counterpart in pe).  A forward run keeps checkpoints of the
 * state; the reverse pass re-runs each segment between checkpoints and differentiates every
 * step through its converged contact impulses at fixed active set (each contact keeps its mode:
 * separating, static or sliding), solving the transposed linearised fixed-point system with
 * Gauss-Seidel sweeps in the forward's colouring.  Supported: spheres and planes, relaxation
 * models 0, 1 and 3, coloured mode, no damping.  Derivatives are of the converged solve, so the
 * forward should run to a tight abort_threshold with enough max_iterations.  Model 3 is
 * differentiated at the exact minimiser on the friction circle, which its golden-section search
 * approximates: raise max_sub_iterations (about 80) for a forward run that meets it to rounding.
 * ------------------------------------------------------------------------------------------ */
typedef struct {
  uint64_t steps;          /* steps differentiated */
  uint64_t total_sweeps;   /* adjoint sweeps summed over the steps */
  uint64_t max_contacts;   /* largest contact count of a step */
  double max_last_delta;   /* largest final L-inf change of an adjoint sweep */
} pehip_adjoint_stats;

/* World::run(nsteps, dt) as pehip_step, keeping a checkpoint of the state every
 * checkpoint_every steps (the first at the start).  A run that fails part-way is dropped:
 * pehip_adjoint_backward then reports that there is no differentiable run. */
pehip_status pehip_adjoint_forward(pehip_ctx* ctx, double dt, uint32_t nsteps, uint32_t checkpoint_every);

/* Reverse pass of the last pehip_adjoint_forward.  Inputs: dL/d(final positions, linear and
 * angular velocities), 3n doubles each (NULL = 0).  Outputs (any may be NULL): dL/d(initial
 * positions, velocities, angular velocities), and dL/dmu = the sum over all contacts of all steps
 * of dL/d(contact friction) -- the derivative with respect to the friction coefficient when every
 * contact's friction is that one coefficient.  adjoint_sweeps bounds the sweeps per step;
 * adjoint_tol > 0 stops a step's sweeps when the L-inf change of the adjoint impulses drops below
 * it.  The context is left at the forward run's final state, and with the parameters it had
 * before the call, also when the pass fails. */
pehip_status pehip_adjoint_backward(pehip_ctx* ctx, const double* dL_dpos, const double* dL_dv,
                                    const double* dL_dw, double* dL_dpos0, double* dL_dv0,
                                    double* dL_dw0, double* dL_dmu, uint32_t adjoint_sweeps,
                                    double adjoint_tol, pehip_adjoint_stats* stats);

/* ------------------------------------------------------------------------------------------
 * Batched scenes.  Many independent pe worlds (scenes) of uneven body counts advance together:
 * one workgroup runs one scene's whole time step (detection on a per-scene hash grid in LDS,
 * colouring, contact caching, the NBGS sweeps with workgroup barriers between colours,
 * integration), and a persistent grid of workgroups pulls scenes from a work queue, largest
 * scene first.  Each scene follows the same semantics as a pehip_ctx in coloured mode.  pe
 * has no batched mode: a scene is what one pe process simulates (World + CollisionSystem
 * singletons, World.cpp:65, HCSITS.h:118).
 * ------------------------------------------------------------------------------------------ */
typedef struct pehip_batch pehip_batch;

/* Launch configuration of the batched step (chosen by pehip_batch_autotune or set by hand). */
typedef struct {
  uint32_t threads;          /* workgroup size per scene: 64, 128, 256 or 512 */
  uint32_t blocks_per_cu;    /* persistent workgroups per compute unit (1..16) */
  uint32_t steps_per_launch; /* time steps a workgroup advances a scene before taking the next */
  uint32_t use_graph;        /* 1: replay the launches of pehip_batch_step from a captured hipGraph */
} pehip_batch_tuning;

typedef struct {
  uint32_t nscenes;
  uint64_t num_bodies;
  uint64_t contact_capacity;   /* total contact slots over all scenes */
  uint64_t num_contacts;       /* contacts of the last step, summed over scenes */
  uint32_t max_colours;        /* largest colour count of a scene in the last step */
  uint64_t steps_done_min;     /* fewest steps any scene has completed since upload */
  uint32_t capacity_retries;   /* contact-capacity growths since upload */
  uint32_t launches;           /* kernel launches of the last pehip_batch_step */
  double sweep_bytes;          /* algorithmic sweep bytes of the last step over all scenes */
} pehip_batch_stats;

pehip_status pehip_batch_create(const pehip_params* params, int device, pehip_batch** out);
pehip_status pehip_batch_destroy(pehip_batch* b);
pehip_status pehip_batch_set_params(pehip_batch* b, const pehip_params* params);
/* Shared material pair tables (as pehip_set_materials). */
pehip_status pehip_batch_set_materials(pehip_batch* b, uint32_t n, const double* csf_pair,
                                       const double* cdf_pair);
/* Uploads all scenes: scene s owns bodies [scene_offset[s], scene_offset[s+1]) of the body
 * arrays, which have the layout of pehip_upload_bodies.  A scene holds at most 4096 bodies and
 * 16 planes; its grid bodies must be of similar size (one hash-grid level per scene, cell edge
 * = the largest grid-body bounding box).  contacts_per_body sizes each scene's contact slots
 * (capacity = contacts_per_body * bodies + 64; grown on demand, see pehip_batch_step).  A failure
 * after validation (out of device memory) leaves the batch without scenes. */
pehip_status pehip_batch_upload(pehip_batch* b, uint32_t nscenes, const uint32_t* scene_offset,
                                const uint8_t* shape, const uint64_t* sid, const uint32_t* material,
                                const uint8_t* fixed, const double* param, const double* pos,
                                const double* q, const double* v, const double* w,
                                const double* inv_mass, const double* inertia_body,
                                const double* inv_inertia_body, const double* plane_normal,
                                const double* plane_d, double contacts_per_body);
/* Per-scene friction: every contact of scene s gets static_mu[s] (or dynamic_mu[s] when its
 * tangential speed exceeds the friction threshold) instead of the material tables.  NULL
 * pointers restore the material tables. */
pehip_status pehip_batch_set_scene_friction(pehip_batch* b, uint32_t nscenes, const double* static_mu,
                                            const double* dynamic_mu);
/* Overwrites the dynamic state of all bodies (layout of pehip_upload_state). */
pehip_status pehip_batch_upload_state(pehip_batch* b, uint64_t n, const double* pos, const double* q,
                                      const double* v, const double* w, const double* plane_normal,
                                      const double* plane_d);
/* Advances every scene by nsteps time steps of dt.  A scene whose contacts outgrow its slots
 * stops where it is; the library then grows that scene's capacity and finishes its steps, so
 * the call returns with every scene advanced (PEHIP_CAPACITY_RETRY if the grown capacity still
 * does not suffice; PEHIP_OUT_OF_MEMORY if the device cannot hold it, which leaves the batch
 * without scenes). */
pehip_status pehip_batch_step(pehip_batch* b, double dt, uint32_t nsteps);
pehip_status pehip_batch_download_state(pehip_batch* b, uint64_t n, double* pos, double* q, double* v,
                                        double* w, double* plane_normal, double* plane_d);
/* Per-scene results of the last step: contacts, colours, sweeps run, max penetration, steps
 * completed since upload.  Any pointer may be NULL. */
pehip_status pehip_batch_scene_stats(pehip_batch* b, uint32_t nscenes, uint32_t* contacts, uint32_t* colours,
                                     uint32_t* iterations, double* max_penetration, uint32_t* steps_done);
/* Contacts of scene s from its last step (body indices are global, detection order). */
pehip_status pehip_batch_get_contacts(pehip_batch* b, uint32_t scene, size_t cap, size_t* n, uint32_t* body1,
                                      uint32_t* body2, double* pos, double* normal, double* dist, double* mu);
/* Sweep order of scene s in its last step (as pehip_get_sweep_order). */
pehip_status pehip_batch_get_sweep_order(pehip_batch* b, uint32_t scene, size_t cap, size_t* n,
                                         uint32_t* contact_index, uint32_t* ncolours, uint32_t* colour_offsets);
pehip_status pehip_batch_get_stats(pehip_batch* b, pehip_batch_stats* out);
pehip_status pehip_batch_get_tuning(pehip_batch* b, pehip_batch_tuning* out);
pehip_status pehip_batch_set_tuning(pehip_batch* b, const pehip_batch_tuning* t);
/* Autotuner: times trial_steps steps of dt for each launch configuration from the current state
 * (restored after every trial, so the scenes do not advance), keeps the fastest and returns it.
 * Results do not depend on the configuration. */
pehip_status pehip_batch_autotune(pehip_batch* b, double dt, uint32_t trial_steps, pehip_batch_tuning* best);
/* The batch's HIP stream (hipStream_t). */
void* pehip_batch_stream(pehip_batch* b);
const char* pehip_batch_last_error(const pehip_batch* b);

#ifdef __cplusplus
}
#endif
#endif
