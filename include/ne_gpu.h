/*
 * ne_gpu.h -- C ABI of the MI355X (gfx950) rigid-body step.
 *
 * This is the drop-in seam for NativeEngine's per-step hot path.  The reference
 * exposes that path as the C++ value type `ne::World` (src/world.hpp:73-161)
 * whose `step()` (src/world.cpp:133-146) runs integrate -> collide ->
 * wrap/walls on `std::vector<Body>`.  This header replaces it with a context
 * that keeps the bodies resident in HBM as structure-of-arrays FP64 and steps
 * them with hand-written HIP kernels.  `include/ne_world.hpp` rebuilds the
 * reference's `ne::World` class on top of these entry points.
 *
 * Conventions
 *   - Every function returns an int status: NE_OK (0) or a negative NE_ERR_*;
 *     ne_gpu_step and ne_gpu_rollout can also return the positive warning
 *     NE_WARN_EPA_CAPPED.
 *     Nothing throws across this boundary.
 *   - A context is single-threaded: one host thread, one HIP stream.
 *   - All pointers are host pointers unless the name says `device`.
 *   - Vectors are packed xyz triples (n*3 doubles); quaternions are packed
 *     (w, x, y, z) quadruples (n*4 doubles), the reference's `Q` order
 *     (src/vmath.hpp:47-48).
 *   - Body index == identity, as in the reference (warm-start keys, pair order).
 */
#ifndef NE_GPU_H
#define NE_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define NE_OK                 0
#define NE_ERR_INVALID      (-1)   /* bad argument / unsupported configuration */
#define NE_ERR_HIP          (-2)   /* a HIP runtime call failed               */
#define NE_ERR_CAPACITY     (-3)   /* a work buffer still overflowed after the
                                      step's own reruns; state is undefined,
                                      re-upload and retry */
#define NE_ERR_NONFINITE    (-4)   /* a non-finite position/velocity was found */
#define NE_ERR_UNSUPPORTED  (-5)   /* shape / feature not on the GPU path yet  */
#define NE_ERR_NO_DEVICE    (-6)   /* no HIP device visible                    */
/* Positive: the call completed, with a warning. */
#define NE_WARN_EPA_CAPPED    1    /* ne_gpu_step: at least one EPA run stopped
                                      at the workspace's face cap
                                      (ne_gpu_config.epa_face_cap), so its
                                      contact can differ from the reference's;
                                      ne_gpu_epa_stats counts them */

/* ---- shapes: the reference's `enum class Shape` order (src/body.hpp:24) -- */
#define NE_SHAPE_SPHERE    0
#define NE_SHAPE_CYLINDER  1
#define NE_SHAPE_BOX       2
#define NE_SHAPE_CAPSULE   3
#define NE_SHAPE_CONVEX    4
#define NE_SHAPE_PLANE     5
#define NE_SHAPE_MESH      6
#define NE_SHAPE_COMPOUND  7

/* ---- body flag bits (the reference's bool members, src/body.hpp:106-127) - */
#define NE_FLAG_KINEMATIC  1u
#define NE_FLAG_SENSOR     2u
#define NE_FLAG_DYNAMIC    4u
#define NE_FLAG_SLEEPING   8u
#define NE_FLAG_CCD       16u

/* ---- facade contact events (PhysicsWorld::processContacts,
 *      src/physics_world.cpp:253-302): the callback each list feeds ------- */
#define NE_EVENT_CONTACT        0   /* onContact: every solid contact of the step */
#define NE_EVENT_BEGIN          1   /* onBegin: solid pairs new this step          */
#define NE_EVENT_STAY           2   /* onStay: solid pairs also in the last step   */
#define NE_EVENT_END            3   /* onEnd: last step's solid pairs gone now (last step's contact) */
#define NE_EVENT_TRIGGER_ENTER  4   /* onTrigEnter: pairs with a sensor, new       */
#define NE_EVENT_TRIGGER_STAY   5   /* onTrigStay                                  */
#define NE_EVENT_TRIGGER_EXIT   6   /* onTrigExit (last step's contact)            */

/* ---- material combine modes (src/world.hpp:85) --------------------------- */
#define NE_COMBINE_AVERAGE        0
#define NE_COMBINE_MIN            1
#define NE_COMBINE_MAX            2
#define NE_COMBINE_MULTIPLY       3
#define NE_COMBINE_GEOMETRIC_MEAN 4

/*
 * World parameters: the public scalar fields of `ne::World`
 * (src/world.hpp:75-126).  Extensions beyond the reference:
 *   half[3]            per-axis half extents (the reference's Box is a cube,
 *                      src/box.hpp:16-18; pass three equal values for parity)
 *   solver_iterations  the reference hard-codes 8 (src/world.cpp:572)
 */
typedef struct ne_world_params {
    double  half[3];
    int32_t periodic;
    int32_t open_boundary;
    double  dt;
    double  gravity[3];
    double  restitution;
    double  friction;
    double  contact_beta;
    int32_t restitution_combine;
    int32_t friction_combine;
    int32_t warm_start;
    int32_t force_brute_force;   /* accepted for parity; results never depend on it */
    int32_t solver_iterations;
    int32_t sleep_enabled;
    double  sleep_lin_vel, sleep_ang_vel, sleep_time, wake_vel;
} ne_world_params;

/*
 * Bodies as flat host arrays (n entries each).  Mirrors `ne::Body`
 * (src/body.hpp:41-129).  Nullable arrays take the reference's defaults:
 * force/torque 0, friction/restitution -1 (inherit), damping 0, gravity_scale 1,
 * layer 0, mask 0xFFFFFFFF, flags NE_FLAG_DYNAMIC, no hull, no exclusions.
 */
typedef struct ne_bodies {
    uint32_t        n;
    const double   *x, *v, *q, *w;            /* n*3, n*3, n*4, n*3            */
    const double   *force, *torque;           /* n*3 (nullable)                */
    const double   *inv_mass;                 /* n                             */
    const double   *inv_inertia;              /* n*3 body-frame diagonal       */
    const int32_t  *shape;                    /* n, NE_SHAPE_*                 */
    const double   *radius, *half_height;     /* n                             */
    const double   *half_extents;             /* n*3 (plane: world normal)     */
    const double   *friction, *restitution;   /* n (nullable)                  */
    const double   *linear_damping, *angular_damping, *gravity_scale; /* n (nullable) */
    const uint32_t *layer, *mask;             /* n (nullable)                  */
    const uint32_t *flags;                    /* n (nullable)                  */
    const double   *sleep_timer;              /* n (nullable)                  */
    const double   *inv_mass_store;           /* n (nullable)                  */
    const double   *inv_inertia_store;        /* n*3 (nullable)                */
    const uint32_t *hull_offsets;             /* n+1 (nullable): CSR into hull_verts */
    const double   *hull_verts;               /* 3 doubles per vertex          */
    const uint32_t *nocollide_offsets;        /* n+1 (nullable): CSR           */
    const uint32_t *nocollide;                /* partner body indices          */
} ne_bodies;

/* A distance joint (src/world.hpp:20-26): bodies a and b joined at body-frame
 * anchors local_a / local_b.  stiffness == 0: a rigid link (bilateral
 * velocity constraint with Baumgarte length correction, solved after the
 * contacts); stiffness > 0: a spring-damper applied as a force before
 * integration. */
typedef struct ne_distance_joint {
    uint32_t a, b;
    double   local_a[3], local_b[3];
    double   rest, stiffness, damping;
} ne_distance_joint;

/* A general joint (`ne::Joint`, src/world.hpp:31-54): ball, hinge, fixed or
 * slider between bodies a and b, with anchors and axes in each body's frame
 * and ref_rel = qA * conj(qB) at creation (PhysicsWorld::addGeneralJoint,
 * src/physics_world.hpp:408-423, computes them from world anchors).  Hinges
 * may drive a motor (clamped impulse per step) and hold an angle limit;
 * breakable joints break when a step's accumulated impulse exceeds
 * break_impulse; broken is the persistent state (a broken joint is skipped). */
#define NE_JOINT_BALL    0
#define NE_JOINT_HINGE   1
#define NE_JOINT_FIXED   2
#define NE_JOINT_SLIDER  3

typedef struct ne_joint {
    int32_t  type;                 /* NE_JOINT_*                            */
    uint32_t a, b;
    int32_t  use_motor, use_limit, breakable, broken;
    double   local_a[3], local_b[3];
    double   axis_a[3], axis_b[3];
    double   ref_rel[4];           /* w, x, y, z                            */
    double   motor_speed, max_motor_impulse;
    double   lower, upper;         /* hinge angle limits (rad)              */
    double   break_impulse;
} ne_joint;

/* A static triangle mesh attached to a body of shape NE_SHAPE_MESH
 * (`Body::mesh`, src/body.hpp:62; `MeshData`, src/mesh.hpp:14-93): vertices
 * in the body frame, triangles as vertex-index triples.  A heightfield
 * (`PhysicsWorld::addHeightfield`, src/physics_world.hpp:96-109) is such a mesh
 * with the grid's two triangles per cell. */
typedef struct ne_mesh {
    uint32_t       body;      /* index of the NE_SHAPE_MESH body            */
    uint32_t       nverts;
    const double  *verts;     /* nverts*3                                   */
    uint32_t       ntris;
    const int32_t *tris;      /* ntris*3, each in [0, nverts)               */
} ne_mesh;

/* Context creation options. Zero fields take defaults. */
typedef struct ne_gpu_config {
    int32_t  device;            /* HIP device ordinal                         */
    uint32_t pair_capacity;     /* candidate pairs per step (default 8*n+4096) */
    uint32_t contact_capacity;  /* constraints per step (default 4*n+4096)    */
    double   cell_scale;        /* grid cell >= cell_scale * 2*maxR (>= 1; 0 = 1) */
    int32_t  broadphase;        /* NE_BROADPHASE_*: the neighbourhood walk variant */
    uint32_t cells_per_body;    /* cell-table capacity per body (0 = 8)       */
    uint32_t epa_face_cap;      /* faces per EPA run in the large workspace
                                   (0 = 131072; at least 64, at most 2^26).
                                   The reference's EPA grows its face list
                                   without bound; a run that needs more than
                                   the cap stops at its closest face and the
                                   step returns NE_WARN_EPA_CAPPED.  The
                                   workspace is allocated at min(cap, 131072)
                                   faces a run for 64 runs (about 0.4 GiB at
                                   the default) and grows past that on use */
} ne_gpu_config;

/* Broadphase walk variants (same candidate rows; a tuning knob). */
#define NE_BROADPHASE_AUTO   0
#define NE_BROADPHASE_WALK   1   /* one thread per body walks 27 cells in global memory */
#define NE_BROADPHASE_BRICK  2   /* one workgroup per 8x8x4-cell brick, halo staged in LDS */
#define NE_BROADPHASE_CELLS  3   /* one thread per cell (coalesced neighbour rows) */
/* NE_BROADPHASE_AUTO is the per-body walk; ne_gpu_autotune measures the others. */

typedef struct ne_gpu_ctx ne_gpu_ctx;

/* Replaces constructing an `ne::World` (src/world.hpp:73). */
int ne_gpu_create(const ne_gpu_config *cfg, ne_gpu_ctx **out);
int ne_gpu_destroy(ne_gpu_ctx *ctx);

/* Sets the World's public scalars (src/world.hpp:75-126). */
int ne_gpu_set_params(ne_gpu_ctx *ctx, const ne_world_params *p);

/* Replaces assigning `World::bodies` (src/world.hpp:128): uploads all bodies
 * (shape, mass, material and state).  Keeps the warm-start cache (the
 * reference's cache survives body edits); call ne_gpu_reset_warm_cache to drop
 * it, as `World::resetWarmCache` (src/world.hpp:140). */
int ne_gpu_set_bodies(ne_gpu_ctx *ctx, const ne_bodies *b);

/* Replaces assigning `World::distanceJoints` (src/world.hpp:129): the joints
 * act in the order given, as the reference's vector does.  Bodies must be
 * uploaded first; ne_gpu_set_bodies clears the joints. */
int ne_gpu_set_joints(ne_gpu_ctx *ctx, uint32_t n, const ne_distance_joint *joints);

/* Replaces assigning `World::joints` (src/world.hpp:130): the general joints,
 * solved after the rigid distance joints by `World::solveJoints`
 * (src/world.cpp:206-315), 8 sequential-impulse iterations over the joints
 * in the order given.  Bodies must be uploaded first; ne_gpu_set_bodies
 * clears them.  The hinge limit's angle is 2 atan2(.) of the relative
 * rotation (world.cpp:300); atan2 here is rounded to nearest, which the host
 * C library's is not always (glibc 2.35: off by one unit in the last place
 * for about 4 in 10,000 arguments), so a step whose active hinge limit meets
 * such an argument differs from the reference's in the last bits; every
 * other joint term is the reference's arithmetic. */
int ne_gpu_set_general_joints(ne_gpu_ctx *ctx, uint32_t n, const ne_joint *joints);

/* The general joints' runtime state after the last step, n entries each (any
 * may be NULL): broken, the impulse accumulated in that step's solve (the
 * break test's appliedImpulse) and the motor accumulator. */
int ne_gpu_get_general_joints(ne_gpu_ctx *ctx, uint32_t n, int32_t *broken, double *applied_impulse,
                              double *motor_acc);

/* Replaces building `MeshData` and assigning `Body::mesh` (src/mesh.hpp:38-43,
 * src/physics_world.hpp:83-93): each mesh's bounding-volume tree is built
 * here, once, by the reference's median split (leaves of at most 4
 * triangles), and the meshes stay resident.  Mesh bodies must be static
 * (zero inverse mass and inertia).  A mesh body with no mesh has no contacts,
 * as in the reference.  Bodies must be uploaded first; ne_gpu_set_bodies
 * clears the meshes. */
int ne_gpu_set_meshes(ne_gpu_ctx *ctx, uint32_t n, const ne_mesh *meshes);

/* The tree built for mesh `m` (the index in the last ne_gpu_set_meshes call):
 * `order` (ntris triangle indices, tree order), then per node, in the
 * reference's creation order, `bounds` (lo xyz, hi xyz) and `links` (left,
 * right, start, count; -1 children for a leaf).  Call with NULL arrays to
 * get the node count.  For checking against `MeshData::order` / `nodes`. */
int ne_gpu_mesh_bvh(ne_gpu_ctx *ctx, uint32_t m, int32_t *order, uint32_t node_cap,
                    uint32_t *nnodes, double *bounds, int32_t *links);

/* Uploads only the dynamic state (x, v, q, w; any may be NULL) of the current
 * bodies. */
int ne_gpu_set_state(ne_gpu_ctx *ctx, const double *x, const double *v,
                     const double *q, const double *w);

/* Replaces `World::step()` (src/world.cpp:133-146), nsteps times, resident:
 * spring joints, integrate, the CCD pass of bodies flagged NE_FLAG_CCD
 * (World::ccdPass, :658-696), collide, rigid distance joints, general joints
 * (World::solveJoints, :206-315), wrap or walls, sleeping.
 * Work buffers (candidate pairs, constraints, the large EPA workspace) are
 * sized from the bodies; when one overflows, the call enlarges it, restores
 * the state it started from (a device-side copy taken at the start of every
 * call) and runs again, up to 4 times, so results never depend on capacity.
 * Returns NE_WARN_EPA_CAPPED (> 0) when the steps completed but an EPA run
 * reached the face cap of ne_gpu_config.epa_face_cap. */
int ne_gpu_step(ne_gpu_ctx *ctx, int32_t nsteps);

/* Downloads x (n*3), v (n*3), q (n*4), w (n*3); any may be NULL. */
int ne_gpu_get_state(ne_gpu_ctx *ctx, double *x, double *v, double *q, double *w);

/* Downloads per-body sleep state: flags (n), sleep_timer (n), inv_mass (n),
 * inv_inertia (n*3); any may be NULL. */
int ne_gpu_get_body_aux(ne_gpu_ctx *ctx, uint32_t *flags, double *sleep_timer,
                        double *inv_mass, double *inv_inertia);

/* Downloads `Body::invMassStore` (n) and `Body::invInertiaStore` (n*3)
 * (src/body.hpp:107-110): the masses a sleeping body gets back when it wakes
 * (World::wake, src/world.cpp:317-323).  Either may be NULL. */
int ne_gpu_get_sleep_store(ne_gpu_ctx *ctx, double *inv_mass_store, double *inv_inertia_store);

/* `World::resetWarmCache()` (src/world.hpp:140). */
int ne_gpu_reset_warm_cache(ne_gpu_ctx *ctx);

/* `PhysicsWorld::saveState()` (src/physics_world.cpp:25-39): the bodies in
 * the reference's "NES1" bytes -- magic 0x4E455331, body count (uint32), then
 * per body x, q (w,x,y,z), v, w, invMass, invInertiaBody, invMassStore,
 * invInertiaStore, sleepTimer (doubles) and sleeping (one byte), packed, then
 * the facade's joint count (uint32; 0, since the general joints are not on
 * this path).  Packed on the device.  *size is always set to the snapshot's
 * length; the bytes are written when out is non-NULL, and a cap below *size
 * is NE_ERR_INVALID. */
int ne_gpu_save_state(ne_gpu_ctx *ctx, void *out, uint64_t cap, uint64_t *size);

/* ne_gpu_save_state without any device allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  The same
 * bytes as ne_gpu_save_state, packed into a static 8 MiB device buffer (one
 * per GPU, held until the process ends) instead of a context allocation, so
 * at most 47,393 bodies: more is NE_ERR_INVALID, with *size set.  Calls that
 * overlap on one GPU share that buffer and race. */
int ne_gpu_save_state_noalloc(ne_gpu_ctx *ctx, void *out, uint64_t cap, uint64_t *size);

/* `PhysicsWorld::loadState()` (src/physics_world.cpp:41-64): restores every
 * body field of an NES1 snapshot and resets the warm cache.  NE_ERR_INVALID
 * (nothing changed) for a wrong magic, another body count, facade joints, or
 * too few bytes -- where the reference returns false, possibly after loading
 * some bodies.  NE_ERR_UNSUPPORTED (nothing changed) when a static body that
 * the grid leaves out would become movable or asleep: upload such a world
 * with ne_gpu_set_bodies. */
int ne_gpu_load_state(ne_gpu_ctx *ctx, const void *data, uint64_t size);

/* The warm-start cache (`World::warmCache_`, src/world.hpp:157): the last
 * warm-started step's normal impulses, one entry per contact, ordered by
 * pair key (i << 32 | j, src/world.cpp:34-36) then contact slot.  Export with
 * NULL arrays to get the count; import replaces the cache (keys must be in
 * that order).  Lets a copied World keep its cache, as the reference's value
 * semantics do. */
int ne_gpu_export_warm_cache(ne_gpu_ctx *ctx, uint32_t cap, uint32_t *count,
                             uint64_t *keys, double *normal_impulse);
int ne_gpu_import_warm_cache(ne_gpu_ctx *ctx, uint32_t count, const uint64_t *keys,
                             const double *normal_impulse);

/* Parity taps for the last completed step.  The contact list is the
 * reference's constraint list (src/world.cpp:535-553) in its order: body pair
 * (i<j), normal b->a, contact point, depth, and the converged normal impulse.
 * Pass NULL arrays to query the count only; `cap` bounds the copy. */
int ne_gpu_get_contacts(ne_gpu_ctx *ctx, uint32_t cap, uint32_t *count,
                        uint32_t *i, uint32_t *j, double *normal, double *point,
                        double *depth, double *normal_impulse);

/* Distinct hit pairs (pairs with >=1 contact) of the last step, in (i,j) order. */
int ne_gpu_get_pairs(ne_gpu_ctx *ctx, uint32_t cap, uint32_t *count,
                     uint32_t *i, uint32_t *j);

/* Facade contact events: `PhysicsWorld::processContacts`
 * (src/physics_world.cpp:253-302), run after every step while enabled, on the
 * device.  The pairs i < j of the post-step state that pass the facade's
 * filters (a pair with a sensor always, else not two immovable bodies;
 * layers; no-collide lists) and overlap under detectContact (one contact,
 * no box-box manifold) are found with the step's own grid and narrow-phase
 * kernels and diffed against the last step's.  Enabling (or uploading
 * bodies) empties the last-step lists, as the facade's maps start empty.
 * While enabled, ne_gpu_step takes one step per call (NE_ERR_INVALID
 * otherwise) and HIP-graph replay is off.  Worlds with mesh bodies are
 * NE_ERR_UNSUPPORTED (the reference's processContacts runs detectContact on
 * a mesh body as a cylinder of its radius and half height), as are batched
 * worlds. */
int ne_gpu_set_events(ne_gpu_ctx *ctx, int32_t enable);

/* The last step's events of one kind (NE_EVENT_*), in (i, j) order -- the
 * facade's handle order when handles follow body indices; the facade maps
 * indices to handles.  Contacts: normal b -> a, point, depth; END and
 * TRIGGER_EXIT carry the previous step's contact.  NULL arrays query the
 * count; `cap` bounds the copy. */
int ne_gpu_get_events(ne_gpu_ctx *ctx, int32_t kind, uint32_t cap, uint32_t *count, uint32_t *i, uint32_t *j,
                      double *point, double *normal, double *depth);

/* ---- scene queries: PhysicsWorld (src/physics_world.cpp:146-251) ---------
 * Batched over many rays or query shapes, against the current state on the
 * device (after the last ne_gpu_step).  Bodies are found by index; the facade
 * maps index -> handle and userData.  A mask passes body b when
 * mask & (1 << (b.layer & 31)) (layerAllowed, physics_world.cpp:67); a NULL
 * mask array passes every layer.  Each periodic body is tested at its
 * minimum image nearest the ray's origin (or the query shape).  One world
 * only: batched worlds are NE_ERR_UNSUPPORTED. */

/* Nearest hit of each of n rays (PhysicsWorld::raycast, :146-164), or of n
 * swept spheres when cast_radius is not NULL (PhysicsWorld::sphereCast,
 * :229-251: every body inflated by the ray's radius, the point moved back
 * onto the real surface).  origin and dir are n*3 (dir is normalised here),
 * max_dist n.  A body wins when its distance is strictly below the best so
 * far, in body order, exactly as the reference's loop.  Only spheres, boxes,
 * cylinders and capsules are hit (rayVsBody, src/queries.hpp:61-122).
 * Outputs per ray: body (-1: no hit, and distance, point and normal 0),
 * distance, point n*3, outward normal n*3. */
int ne_gpu_raycast(ne_gpu_ctx *ctx, uint32_t n, const double *origin, const double *dir,
                   const double *max_dist, const uint32_t *mask, const double *cast_radius,
                   int32_t *body, double *distance, double *point, double *normal);

/* Every hit of each of n rays within max_dist (PhysicsWorld::raycastAll,
 * :166-188), ray k's hits at [offsets[k], offsets[k+1]) sorted by distance
 * (stable: equal distances keep body order -- the reference's std::sort order
 * whenever a ray's distances differ or it has at most 16 hits).  offsets has
 * n+1 entries and is always written; when offsets[n] > cap nothing else is
 * and the call returns NE_ERR_CAPACITY (call again with cap >= offsets[n]). */
int ne_gpu_raycast_all(ne_gpu_ctx *ctx, uint32_t n, const double *origin, const double *dir,
                       const double *max_dist, const uint32_t *mask, uint32_t cap, uint32_t *offsets,
                       int32_t *body, double *distance, double *point, double *normal);

/* A query body for ne_gpu_overlap: a shape at a pose, as a `Body` the
 * reference would build (makeSphere / makeBox for overlapSphere /
 * overlapBox, any Body for overlapContacts).  Fields as in ne_bodies;
 * verts: the hull's nverts body-frame vertices for NE_SHAPE_CONVEX. */
typedef struct ne_query_shape {
    int32_t       shape;              /* NE_SHAPE_SPHERE .. NE_SHAPE_PLANE     */
    uint32_t      nverts;
    double        x[3];
    double        q[4];               /* w, x, y, z                            */
    double        half_extents[3];    /* box; plane: normal                    */
    double        radius, half_height;
    const double *verts;              /* nverts*3 (convex only)                */
} ne_query_shape;

/* Bodies overlapping each of n query shapes under detectContact(query, body)
 * (PhysicsWorld::overlapSphere / overlapBox, :190-212, for sphere and box
 * shapes; contacts != 0: overlapContacts, :214-227, which also skips sensors
 * and returns each overlap's normal (body -> query) and depth).  Query k's
 * bodies, in body order, at [offsets[k], offsets[k+1]); normal (3 per hit)
 * and depth may be NULL.  Capacity as ne_gpu_raycast_all.  Worlds with mesh
 * bodies are NE_ERR_UNSUPPORTED (detectContact would take a mesh body for a
 * cylinder).  May return NE_WARN_EPA_CAPPED (contacts only). */
int ne_gpu_overlap(ne_gpu_ctx *ctx, uint32_t n, const ne_query_shape *shapes, const uint32_t *mask,
                   int32_t contacts, uint32_t cap, uint32_t *offsets, int32_t *body, double *normal,
                   double *depth);

/* Diagnostics of `World` (src/world.cpp:704-729): sums over invMass>0 bodies
 * in index order.  p[3], L[3] and ke may be NULL. */
int ne_gpu_diagnostics(ne_gpu_ctx *ctx, double *p, double *L, double *ke);

/* Counters of the last step: [candidate pairs, mesh triangle candidates,
 * constraints, islands, constraints in the largest island, gridded bodies,
 * grid cells, fault bits].  While islands of more than 16 constraints hold
 * at least 3/4 of the constraints, eager steps solve every constraint as one
 * dataflow without building islands (every 32nd step builds them to check):
 * the island counts then read 0. */
int ne_gpu_get_counters(ne_gpu_ctx *ctx, uint32_t out[8]);

/* The HIP stream the context launches on (as a void*), for event timing. */
int ne_gpu_get_stream(ne_gpu_ctx *ctx, void **stream);

/* Kernel ids for ne_gpu_time_kernel (one per step stage). */
#define NE_KERNEL_NONE        0
#define NE_KERNEL_INTEGRATE   1
#define NE_KERNEL_CELL_COUNT  2
#define NE_KERNEL_CELL_FILL   3
#define NE_KERNEL_CAND_COUNT  4
#define NE_KERNEL_CAND_WRITE  5
#define NE_KERNEL_NARROW      6
#define NE_KERNEL_CONS_BUILD  7
#define NE_KERNEL_SOLVE       8
#define NE_KERNEL_BOUNDARY    9
#define NE_KERNEL_STEP       10   /* the whole step, first to last launch */
#define NE_KERNEL_RAY        11   /* the ray kernel of ne_gpu_raycast */

/* Brackets every launch of one kernel with HIP events on the context's
 * stream during the following ne_gpu_step (or, for NE_KERNEL_RAY,
 * ne_gpu_raycast) calls (NE_KERNEL_NONE disables),
 * and resets the accumulated time.  For live roofline measurement. */
int ne_gpu_time_kernel(ne_gpu_ctx *ctx, int32_t kernel_id);

/* Total milliseconds and launch count accumulated for the timed kernel. */
int ne_gpu_kernel_time(ne_gpu_ctx *ctx, double *total_ms, uint32_t *launches);

/* Device arithmetic probe: out[t] = sqrt(a) (op 0), a/b (1), floor(a) (2),
 * a*b+a (3), atan2(a, b) rounded to nearest as the hinge limit takes it (4),
 * the device library's atan2(a, b) (5), computed by a kernel -- checks the
 * IEEE ground truth and the two atan2s against the host's. */
int ne_gpu_fp_probe(ne_gpu_ctx *ctx, int32_t op, uint32_t n, const double *a,
                    const double *b, double *out);

/* Batched worlds.  No counterpart in the reference, which steps one World at
 * a time; this steps many independent Worlds in one set of launches.  Splits
 * the uploaded bodies into `nworlds` periodic worlds: world w holds bodies
 * [offsets[w], offsets[w+1]) (offsets[0] = 0, offsets[nworlds] = n), has box
 * half extents half[3w..3w+2], and default restitution / friction
 * restitution[w] / friction[w] (the World::restitution and World::friction of
 * src/world.hpp:91-92; NULL keeps ne_world_params').  dt, gravity, combine
 * modes, contact beta, iterations and warm start are shared.  Each world's
 * results equal stepping that world alone; pair, contact and warm-cache
 * indices are global body indices.  Requires periodic params.
 * ne_gpu_set_bodies returns to one world, as does nworlds = 0. */
int ne_gpu_set_worlds(ne_gpu_ctx *ctx, uint32_t nworlds, const uint32_t *offsets,
                      const double *half, const double *restitution,
                      const double *friction);

/* EPA workspace use during the last ne_gpu_step call.  Most EPA runs fit a
 * small per-thread workspace; the rest (near-degenerate contacts, where the
 * reference's vectors grow to thousands of faces) are rerun with a
 * global-memory workspace that grows up to ne_gpu_config.epa_face_cap
 * faces -- `reruns` counts them.  A run that would need more stops at the
 * closest face found so far, as at the reference's 64-iteration limit, where
 * the reference would keep growing its vectors -- `truncated` counts those
 * (the only results that can differ from the reference's; ne_gpu_step then
 * returns NE_WARN_EPA_CAPPED). */
int ne_gpu_epa_stats(ne_gpu_ctx *ctx, uint32_t *reruns, uint32_t *truncated);

/* The EPA work of the last step, as the lengths of the tiers' queues (no
 * counterpart in the reference: measurement, and the tests' proof that every
 * tier ran).  out[0]: convex pairs GJK found overlapping (each starts in
 * tier A, one lane per run); out[1]: runs past tier A (into tier B, 4 lanes
 * a run); out[2]: past B (into C, 8 lanes); out[3]: past C (into the
 * one-wave tier); out[4]: runs rerun by the one-block tier (pairs and mesh
 * candidates); out[5]: mesh triangle candidates GJK found overlapping (each
 * starts in the mesh tier A); out[6]: mesh runs past tier A (into the group
 * tier); out[7]: mesh runs past the group tier (into the one-wave tier). */
int ne_gpu_epa_queues(ne_gpu_ctx *ctx, uint32_t out[8]);

/* Replays each step from a captured HIP graph (enable = 1) instead of
 * launching its kernels one by one.  Results are identical.  Steps that apply
 * uploaded forces, and steps with kernel timing on, run eagerly. */
int ne_gpu_use_graph(ne_gpu_ctx *ctx, int32_t enable);

/* Autotuner: from the current state, times `steps` steps for every
 * combination of cell scale (1, 1.25, 1.5), neighbourhood walk
 * (NE_BROADPHASE_WALK, NE_BROADPHASE_CELLS) and graph replay (off, on),
 * restoring bodies and warm cache before each trial and at the end, then
 * keeps the fastest.  Results do not depend on the choice. */
typedef struct ne_tune_result {
    double   cell_scale;
    int32_t  broadphase;
    int32_t  graph;
    double   ms_per_step;
    uint32_t trials;
} ne_tune_result;
int ne_gpu_autotune(ne_gpu_ctx *ctx, int32_t steps, ne_tune_result *out);

/* Differentiable rollouts.  No counterpart in the reference.  A rollout runs
 * `nsteps` steps from the current state (after replacing the velocities by
 * v0 and the worlds' default friction / restitution by mu / e, when given),
 * starting with an empty warm cache, keeps a checkpoint every
 * `checkpoint_every` steps and writes the final positions to x_out.
 * ne_gpu_rollout_backward then takes d loss / d x_final and returns
 * d loss / d v0 and d loss / d (mu, e) of each world, by a hand-written
 * adjoint of integrate, the constraint build, the warm start and the
 * fixed-iteration impulse solve, with each step's contact set (pairs,
 * normals, points, depths) held fixed; it recomputes each checkpointed
 * segment forward, keeping each recomputed step's constraint set and solve
 * tape in device memory for the segment's reverse sweep (the tape up to
 * 2 GiB a step), and leaves the world at the rollout's start.
 * v0, x_out, grad_x and grad_v0 are n x 3 doubles in host or device memory
 * (copied with hipMemcpyDefault); mu, e, grad_mu and grad_e are host arrays
 * with one entry per world (1 without ne_gpu_set_worlds).  Periodic worlds
 * without applied forces only.  ne_gpu_rollout returns NE_WARN_EPA_CAPPED
 * (the rollout and its checkpoints complete) when a step reached the EPA
 * face cap. */
int ne_gpu_rollout(ne_gpu_ctx *ctx, const double *v0, const double *mu, const double *e,
                   int32_t nsteps, int32_t checkpoint_every, double *x_out);
int ne_gpu_rollout_backward(ne_gpu_ctx *ctx, const double *grad_x, double *grad_v0,
                            double *grad_mu, double *grad_e);

/* Last error message of this context (static storage, never NULL). */
const char *ne_gpu_last_error(ne_gpu_ctx *ctx);

#ifdef __cplusplus
}
#endif

#endif /* NE_GPU_H */
