/*
 * CPU oracle for the HCSITS rigid-body time step -- TEST INFRASTRUCTURE ONLY.
 *
 * This is a single-threaded C++ restatement of the pe reference's time step
 * (HashGrids broadphase, MaxContacts sphere/box/plane narrowphase, HCSITS
 * InelasticCoulombContactByDecoupling NBGS relaxation, semi-implicit
 * integration) written to check the HIP product path.  Only tests/, the
 * smoke() entry point and bench.py's cpu_baseline leg may load it.  The
 * product library never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the pe source tree).
 */
#ifndef PE_ORACLE_API_H
#define PE_ORACLE_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_world orc_world;

enum { ORC_SPHERE = 1, ORC_BOX = 2, ORC_PLANE = 3, ORC_CAPSULE = 4, ORC_UNION = 5 };

/* Relaxation models, numbered as the reference's RelaxationModel enum
 * (pe/core/collisionsystem/HardContactSemiImplicitTimesteppingSolvers.h:177-184). */
enum {
  ORC_INELASTIC_FRICTIONLESS = 0,
  ORC_APPROX_INELASTIC_COULOMB_DECOUPLING = 1,
  ORC_APPROX_INELASTIC_COULOMB_ORTHO = 2,
  ORC_INELASTIC_COULOMB_DECOUPLING = 3,
  ORC_INELASTIC_COULOMB_ORTHO = 4,
  ORC_INELASTIC_GMD = 5
};

orc_world* orc_create(void);
void orc_destroy(orc_world* w);

/* Settings: World::setGravity/setDamping (pe/core/World.h:715-767),
 * CollisionSystem setters (HCSITS.h:793-840).  abort_threshold <= 0 keeps the
 * reference behaviour (always max_iterations sweeps; the check is compiled out
 * at HCSITS.h:1935-1946); > 0 enables that check, for converged comparisons. */
void orc_set_gravity(orc_world* w, double gx, double gy, double gz);
void orc_set_damping(orc_world* w, double damping);
void orc_set_max_iterations(orc_world* w, size_t n);
void orc_set_relaxation_model(orc_world* w, int model);
void orc_set_erp(orc_world* w, double erp);
void orc_set_abort_threshold(orc_world* w, double thr);
/* Trig of the golden-section search and of the orientation update: 0 = the C library's
 * (pe's behaviour, default), 1 = correctly rounded (113-bit quad precision rounded once). */
void orc_set_trig_mode(orc_world* w, int correctly_rounded);
/* fn 0 sin(x), 1 cos(x), 2 atan2(y, x), 3 acos(x) in the given trig mode. */
double orc_trig(int fn, double x, double y, int correctly_rounded);

/* Materials.cpp:175-247: returns the material id. */
int orc_create_material(orc_world* w, double density, double cor, double csf, double cdf);
/* Overrides one pair entry of the static and dynamic friction tables (Materials.h:746). */
void orc_set_pair_friction(orc_world* w, int m1, int m2, double csf, double cdf);

/* Body factories; return the body index (= BodyStorage order = creation order). */
int orc_create_sphere(orc_world* w, double x, double y, double z, double r, int mat);
int orc_create_box(orc_world* w, double x, double y, double z, double lx, double ly, double lz, int mat);
int orc_create_capsule(orc_world* w, double x, double y, double z, double r, double length, int mat);
int orc_create_plane(orc_world* w, double nx, double ny, double nz, double d, int mat);
/* A union of n spheres (offsets disp[3n] before placement, radii, materials) whose centre of mass
 * is placed at (x, y, z); returns the union's body index, -1 if n < 1. */
int orc_create_union(orc_world* w, double x, double y, double z, int n, const double* disp,
                     const double* radius, const int* mat);
/* Union members (spheres), grouped by union in creation order; any output may be NULL. */
size_t orc_num_members(const orc_world* w);
void orc_get_members(const orc_world* w, uint32_t* owner, uint64_t* sid, int* mat, double* radius,
                     double* rpos, double* gpos);
void orc_set_members(orc_world* w, const double* gpos);

size_t orc_num_bodies(const orc_world* w);

/* Full dynamic state, SoA, n = orc_num_bodies.  normal/d are meaningful for planes. */
void orc_get_state(const orc_world* w, double* pos, double* q, double* v, double* om,
                   double* normal, double* d);
/* Overwrites the dynamic state (R is recomputed from q, AABBs are recomputed).  For
 * planes normal/d are taken as given (they carry the libm history of the reference). */
void orc_set_state(orc_world* w, const double* pos, const double* q, const double* v,
                   const double* om, const double* normal, const double* d);
/* Force and torque on every body (world frame, 3 per body) for the next step only
 * (RigidBody::addForce/addTorque; reset after the step as Settings::forceReset says). */
void orc_set_forces(orc_world* w, const double* force, const double* torque);
void orc_get_body_info(const orc_world* w, int* shape, uint64_t* sid, int* mat, int* fixed,
                       double* param /* 3 per body: radius | lengths | 0 */,
                       double* inv_mass, double* inv_inertia_body /* 9 per body */);
void orc_get_aabb(const orc_world* w, double* aabb /* 6 per body */);

/* World::simulationStep (src/core/World.cpp:330-357) repeated nsteps times.
 * Returns 0 on success, -1 on invalid dt. */
int orc_step(orc_world* w, double dt, int nsteps);

/* For the next step only: relax the detected contacts in the given (body1, body2) order
 * instead of the emission order (contacts of one pair keep their narrowphase order).
 * Lets the oracle replay a sweep order chosen elsewhere, e.g. colour by colour. */
void orc_set_sweep_order(orc_world* w, size_t n, const uint32_t* b1, const uint32_t* b2);
/* As orc_set_sweep_order, but entry k names the rank[k]-th contact (in narrowphase order) of
 * the pair (b1[k], b2[k]) -- for orders that interleave the contacts of one pair. */
void orc_set_sweep_order_ranked(orc_world* w, size_t n, const uint32_t* b1, const uint32_t* b2,
                                const uint32_t* rank);

/* Contacts of the last step, in the reference's emission order. */
size_t orc_num_contacts(const orc_world* w);
void orc_get_contacts(const orc_world* w, uint32_t* b1, uint32_t* b2, double* pos,
                      double* normal, double* dist, double* mu);
/* Solved impulses p (world frame, 3 per contact) of the last step; zero for the contacts of an
 * orc_detect call, which has solved nothing (one entry per contact orc_num_contacts counts). */
void orc_get_impulses(const orc_world* w, double* p);

/* Runs only the detection phase on the current state (no state change) with
 * either the HashGrids replica (exhaustive = 0) or ExhaustiveSearch
 * (exhaustive = 1, ExhaustiveSearch.h:212-232).  Returns the contact count;
 * the contacts are then read with orc_get_contacts.  The HashGrids replica's
 * bookkeeping does advance, as in the reference. */
size_t orc_detect(orc_world* w, int exhaustive);
/* AABB-overlapping candidate pairs that survive the both-fixed filter
 * (HashGrids.h:1351-1358), brute force.  Writes up to cap pairs (i<j). */
size_t orc_candidate_pairs(const orc_world* w, uint32_t* a, uint32_t* b, size_t cap);

/* Stats of the last step. */
typedef struct {
  size_t num_contacts;
  double max_penetration;
  size_t iterations;
  double last_delta;
  double t_detection, t_contact_caching, t_body_caching, t_solving, t_integration;
} orc_stats;
void orc_get_stats(const orc_world* w, orc_stats* s);

/* Exposed single-contact primitives for closed-form unit tests. */
void orc_quat_from_axis_angle(const double axis[3], double angle, double q[4]);
void orc_quat_to_rotation(const double q[4], double R[9]);
void orc_quat_mul(const double a[4], const double b[4], double out[4]);
void orc_quat_rotate(const double q[4], const double v[3], double out[3]);
void orc_mat3_inverse(const double m[9], double out[9]);

#ifdef __cplusplus
}
#endif
#endif
