/*
 * ne_oracle.h -- CPU restatement of NativeEngine's per-step hot path.
 *
 * ORACLE / TEST INFRASTRUCTURE ONLY.  This is the checker the HIP path is
 * compared against; only tests/, the smoke check and bench.py's cpu_baseline
 * leg may load it.  The product never links it.
 *
 * Single-threaded C99, IEEE double, no fused multiply-add: every arithmetic
 * step follows the reference's operation order (cited per function in
 * ne_oracle.c) so results are bit-identical to the reference on the inputs it
 * can express (cubic box, 8 solver iterations).  Extensions (per-axis box
 * extents, solver_iterations != 8) follow the same formulas.
 *
 * Body and parameter layouts are the ones of include/ne_gpu.h.
 */
#ifndef NE_ORACLE_H
#define NE_ORACLE_H

#include <stdint.h>
#include "../include/ne_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ne_oracle ne_oracle;

int ne_oracle_create(ne_oracle **out);
int ne_oracle_destroy(ne_oracle *o);
int ne_oracle_set_params(ne_oracle *o, const ne_world_params *p);
int ne_oracle_set_bodies(ne_oracle *o, const ne_bodies *b);
int ne_oracle_set_state(ne_oracle *o, const double *x, const double *v, const double *q, const double *w);
int ne_oracle_step(ne_oracle *o, int32_t nsteps);
int ne_oracle_reset_warm_cache(ne_oracle *o);
int ne_oracle_get_state(ne_oracle *o, double *x, double *v, double *q, double *w);
int ne_oracle_get_body_aux(ne_oracle *o, uint32_t *flags, double *sleep_timer,
                           double *inv_mass, double *inv_inertia);
int ne_oracle_get_contacts(ne_oracle *o, uint32_t cap, uint32_t *count, uint32_t *i, uint32_t *j,
                           double *normal, double *point, double *depth, double *normal_impulse);
int ne_oracle_get_pairs(ne_oracle *o, uint32_t cap, uint32_t *count, uint32_t *i, uint32_t *j);
int ne_oracle_pair_contacts(ne_oracle *o, uint32_t a, uint32_t b, uint32_t cap, uint32_t *count,
                            double *normal, double *point, double *depth);
int ne_oracle_candidate_count(ne_oracle *o, uint32_t *count);
int ne_oracle_diagnostics(ne_oracle *o, double *p, double *L, double *ke);
int ne_oracle_min_image(ne_oracle *o, const double *d, double *out);
int ne_oracle_wrap(ne_oracle *o, const double *p, double *out);
/* Largest EPA face list / horizon seen in this process (workspace sizing). */
void ne_oracle_epa_stats(int32_t *max_faces, int32_t *max_horizon);
int ne_oracle_tape(ne_oracle *o, int32_t mode);
int ne_oracle_set_joints(ne_oracle *o, uint32_t n, const ne_distance_joint *joints);
int ne_oracle_set_meshes(ne_oracle *o, uint32_t n, const ne_mesh *meshes);
int ne_oracle_mesh_bvh(ne_oracle *o, uint32_t m, int32_t *order, uint32_t node_cap, uint32_t *nnodes,
                       double *bounds, int32_t *links);
void ne_oracle_epa_stats_reset(int32_t faces, int32_t horizon);
/* Facade contact events after every step (PhysicsWorld::processContacts,
 * src/physics_world.cpp:253-302); kinds NE_EVENT_*, pairs in (i, j) order. */
int ne_oracle_set_events(ne_oracle *o, int32_t enable);
int ne_oracle_get_events(ne_oracle *o, int32_t kind, uint32_t cap, uint32_t *count, uint32_t *i, uint32_t *j,
                         double *point, double *normal, double *depth);
void ne_oracle_epa_counts(int64_t *calls, int64_t *big);
/* General joints (World::joints / solveJoints, src/world.hpp:130,
 * src/world.cpp:206-315), as ne_gpu_set_general_joints / _get_general_joints. */
int ne_oracle_set_general_joints(ne_oracle *o, uint32_t n, const ne_joint *joints);
int ne_oracle_get_general_joints(ne_oracle *o, uint32_t n, int32_t *broken, double *applied, double *motor);
/* Scene queries (src/physics_world.cpp:146-251), the signatures of
 * ne_gpu_raycast / ne_gpu_raycast_all / ne_gpu_overlap. */
int ne_oracle_raycast(ne_oracle *o, uint32_t n, const double *origin, const double *dir, const double *max_dist,
                      const uint32_t *mask, const double *cast_radius, int32_t *body, double *distance,
                      double *point, double *normal);
int ne_oracle_raycast_all(ne_oracle *o, uint32_t n, const double *origin, const double *dir, const double *max_dist,
                          const uint32_t *mask, uint32_t cap, uint32_t *offsets, int32_t *body, double *distance,
                          double *point, double *normal);
int ne_oracle_overlap(ne_oracle *o, uint32_t n, const ne_query_shape *shapes, const uint32_t *mask, int32_t contacts,
                      uint32_t cap, uint32_t *offsets, int32_t *body, double *normal, double *depth);

#ifdef __cplusplus
}
#endif

#endif /* NE_ORACLE_H */
