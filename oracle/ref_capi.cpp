// ref_capi.cpp -- test-infrastructure harness around the UNMODIFIED reference.
//
// ORACLE / TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile into
// oracle/_ref/libne_ref.so from the reference's own sources (src/world.cpp plus
// its headers) where they lie; nothing from the reference is copied.  Only the
// tests, the smoke check and bench.py's cpu_baseline leg load it.
//
// It drives the reference through its public API only:
//   - `ne::World` public fields and `step()`            (src/world.hpp:73-161)
//   - `World::distanceJoints`                           (src/world.hpp:20-26, 129)
//   - `MeshData` built by its own `build()` and attached as `Body::mesh`
//                                                       (src/mesh.hpp:14-93,
//                                                        src/body.hpp:62)
//   - `captureContacts` / `debugContacts` for the per-step constraint list
//                                                       (src/world.hpp:135-136,
//                                                        src/world.cpp:552)
//   - the factories' Body fields                        (src/body.hpp:41-129)
//   - the factories themselves, for their mass and inertia bits
//                                                       (src/body.hpp:206-345)
//   - `detectContact` / `boxBoxManifold`                (src/detect.hpp:33,
//                                                        src/manifold.hpp:90)
//   - `totalLinearMomentum/AngularMomentum/Kinetic`     (src/world.cpp:704-729)
//   - `PhysicsWorld::saveState` / `loadState` over a copy of the world, for
//     the "NES1" snapshot bytes                         (src/physics_world.cpp:22-64,
//                                                        physics_world.hpp:379)
//   - `PhysicsWorld::step` with its contact callbacks set, for the facade's
//     contact / trigger events                          (physics_world.hpp:218-246,
//                                                        physics_world.cpp:253-302)
// and exposes them through the same flat-array layout as include/ne_gpu.h so a
// test can feed one scene to the GPU, the C restatement and the reference.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

#include "world.hpp"
#include "detect.hpp"
#include "manifold.hpp"
#include "physics_world.hpp"

#include "../include/ne_gpu.h"

namespace {

struct RefEvent {
    int kind;   // NE_EVENT_*
    ne::ContactInfo c;
};

struct RefWorld {
    ne::World w;
    std::vector<std::shared_ptr<const ne::MeshData>> meshes;   // in ne_ref_set_meshes order
    std::unique_ptr<ne::PhysicsWorld> pw;                       // the facade, once events are on
    std::vector<RefEvent> events;                                // the last facade step's callbacks
};

ne::V3 v3(const double* p, uint32_t i) { return ne::V3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }

int build_bodies(const ne_bodies* b, std::vector<ne::Body>& out) {
    out.clear();
    out.resize(b->n);
    for (uint32_t i = 0; i < b->n; ++i) {
        ne::Body& B = out[i];
        B.x = v3(b->x, i);
        B.v = v3(b->v, i);
        B.q = ne::Q{b->q[4 * i], b->q[4 * i + 1], b->q[4 * i + 2], b->q[4 * i + 3]};
        B.w = v3(b->w, i);
        if (b->force) B.force = v3(b->force, i);
        if (b->torque) B.torque = v3(b->torque, i);
        B.invMass = b->inv_mass[i];
        B.invInertiaBody = v3(b->inv_inertia, i);
        int s = b->shape[i];
        if (s < 0 || s > NE_SHAPE_MESH) return NE_ERR_UNSUPPORTED;
        B.shape = static_cast<ne::Shape>(s);
        B.radius = b->radius[i];
        B.halfHeight = b->half_height[i];
        B.halfExtents = v3(b->half_extents, i);
        if (b->hull_offsets) {
            for (uint32_t k = b->hull_offsets[i]; k < b->hull_offsets[i + 1]; ++k)
                B.vertices.push_back(v3(b->hull_verts, k));
        }
        if (b->friction) B.friction = b->friction[i];
        if (b->restitution) B.restitution = b->restitution[i];
        if (b->linear_damping) B.linearDamping = b->linear_damping[i];
        if (b->angular_damping) B.angularDamping = b->angular_damping[i];
        if (b->gravity_scale) B.gravityScale = b->gravity_scale[i];
        if (b->layer) B.layer = b->layer[i];
        if (b->mask) B.mask = b->mask[i];
        uint32_t fl = b->flags ? b->flags[i] : NE_FLAG_DYNAMIC;
        B.kinematic = (fl & NE_FLAG_KINEMATIC) != 0;
        B.sensor = (fl & NE_FLAG_SENSOR) != 0;
        B.dynamic = (fl & NE_FLAG_DYNAMIC) != 0;
        B.sleeping = (fl & NE_FLAG_SLEEPING) != 0;
        B.ccd = (fl & NE_FLAG_CCD) != 0;
        if (b->sleep_timer) B.sleepTimer = b->sleep_timer[i];
        if (b->inv_mass_store) B.invMassStore = b->inv_mass_store[i];
        if (b->inv_inertia_store) B.invInertiaStore = v3(b->inv_inertia_store, i);
        if (b->nocollide_offsets) {
            for (uint32_t k = b->nocollide_offsets[i]; k < b->nocollide_offsets[i + 1]; ++k)
                B.noCollide.push_back(b->nocollide[k]);
        }
        B.id = (int)i;
    }
    return NE_OK;
}

ne::World::Combine combine_of(int32_t m) { return static_cast<ne::World::Combine>(m); }

}  // namespace

extern "C" {

int ne_ref_create(void** out) {
    if (!out) return NE_ERR_INVALID;
    *out = new RefWorld();
    return NE_OK;
}

int ne_ref_destroy(void* h) {
    delete static_cast<RefWorld*>(h);
    return NE_OK;
}

// The reference can only express a cubic box and 8 solver iterations
// (src/box.hpp:16-18, src/world.cpp:572); anything else is refused.
int ne_ref_set_params(void* h, const ne_world_params* p) {
    RefWorld* r = static_cast<RefWorld*>(h);
    if (!r || !p) return NE_ERR_INVALID;
    if (p->half[0] != p->half[1] || p->half[0] != p->half[2]) return NE_ERR_UNSUPPORTED;
    if (p->solver_iterations != 8) return NE_ERR_UNSUPPORTED;
    ne::World& w = r->w;
    w.box.half = p->half[0];
    w.box.periodic = p->periodic != 0;
    w.openBoundary = p->open_boundary != 0;
    w.dt = p->dt;
    w.gravity = ne::V3{p->gravity[0], p->gravity[1], p->gravity[2]};
    w.restitution = p->restitution;
    w.friction = p->friction;
    w.contactBeta = p->contact_beta;
    w.restitutionCombine = combine_of(p->restitution_combine);
    w.frictionCombine = combine_of(p->friction_combine);
    w.warmStart = p->warm_start != 0;
    w.forceBruteForce = p->force_brute_force != 0;
    w.sleepEnabled = p->sleep_enabled != 0;
    w.sleepLinVel = p->sleep_lin_vel;
    w.sleepAngVel = p->sleep_ang_vel;
    w.sleepTime = p->sleep_time;
    w.wakeVel = p->wake_vel;
    return NE_OK;
}

int ne_ref_set_bodies(void* h, const ne_bodies* b) {
    RefWorld* r = static_cast<RefWorld*>(h);
    if (!r || !b) return NE_ERR_INVALID;
    return build_bodies(b, r->w.bodies);
}

int ne_ref_set_joints(void* h, uint32_t n, const ne_distance_joint* j) {
    RefWorld* r = static_cast<RefWorld*>(h);
    if (!r || (n && !j)) return NE_ERR_INVALID;
    r->w.distanceJoints.clear();
    for (uint32_t k = 0; k < n; ++k) {
        ne::DistanceJoint d;
        d.a = j[k].a;
        d.b = j[k].b;
        d.localA = v3(j[k].local_a, 0);
        d.localB = v3(j[k].local_b, 0);
        d.rest = j[k].rest;
        d.stiffness = j[k].stiffness;
        d.damping = j[k].damping;
        r->w.distanceJoints.push_back(d);
    }
    return NE_OK;
}

int ne_ref_set_meshes(void* h, uint32_t n, const ne_mesh* ms) {
    RefWorld* r = static_cast<RefWorld*>(h);
    if (!r || (n && !ms)) return NE_ERR_INVALID;
    r->meshes.clear();
    for (uint32_t m = 0; m < n; ++m) {
        if (ms[m].body >= r->w.bodies.size() || r->w.bodies[ms[m].body].shape != ne::Shape::Mesh) return NE_ERR_INVALID;
        auto md = std::make_shared<ne::MeshData>();
        for (uint32_t k = 0; k < ms[m].nverts; ++k) md->verts.push_back(v3(ms[m].verts, k));
        for (uint32_t t = 0; t < ms[m].ntris; ++t)
            md->tris.push_back({ms[m].tris[3 * t], ms[m].tris[3 * t + 1], ms[m].tris[3 * t + 2]});
        md->build();
        r->w.bodies[ms[m].body].mesh = md;
        r->meshes.push_back(md);
    }
    return NE_OK;
}

int ne_ref_mesh_bvh(void* h, uint32_t m, int32_t* order, uint32_t node_cap, uint32_t* nnodes, double* bounds,
                    int32_t* links) {
    RefWorld* r = static_cast<RefWorld*>(h);
    if (!r || m >= r->meshes.size()) return NE_ERR_INVALID;
    const ne::MeshData& md = *r->meshes[m];
    if (nnodes) *nnodes = (uint32_t)md.nodes.size();
    if (order) for (size_t k = 0; k < md.order.size(); ++k) order[k] = md.order[k];
    if (bounds && links && node_cap >= md.nodes.size())
        for (size_t k = 0; k < md.nodes.size(); ++k) {
            const auto& d = md.nodes[k];
            bounds[6 * k] = d.lo.x; bounds[6 * k + 1] = d.lo.y; bounds[6 * k + 2] = d.lo.z;
            bounds[6 * k + 3] = d.hi.x; bounds[6 * k + 4] = d.hi.y; bounds[6 * k + 5] = d.hi.z;
            links[4 * k] = d.left; links[4 * k + 1] = d.right; links[4 * k + 2] = d.start; links[4 * k + 3] = d.count;
        }
    return NE_OK;
}

int ne_ref_set_capture(void* h, int on) {
    RefWorld* r = static_cast<RefWorld*>(h);
    if (!r) return NE_ERR_INVALID;
    r->w.captureContacts = on != 0;
    return NE_OK;
}

int ne_ref_step(void* h, int32_t nsteps) {
    RefWorld* r = static_cast<RefWorld*>(h);
    if (!r) return NE_ERR_INVALID;
    for (int32_t s = 0; s < nsteps; ++s) r->w.step();
    return NE_OK;
}

int ne_ref_reset_warm_cache(void* h) {
    RefWorld* r = static_cast<RefWorld*>(h);
    if (!r) return NE_ERR_INVALID;
    r->w.resetWarmCache();
    return NE_OK;
}

int ne_ref_get_state(void* h, double* x, double* v, double* q, double* w) {
    RefWorld* r = static_cast<RefWorld*>(h);
    if (!r) return NE_ERR_INVALID;
    const auto& B = r->w.bodies;
    for (size_t i = 0; i < B.size(); ++i) {
        if (x) { x[3 * i] = B[i].x.x; x[3 * i + 1] = B[i].x.y; x[3 * i + 2] = B[i].x.z; }
        if (v) { v[3 * i] = B[i].v.x; v[3 * i + 1] = B[i].v.y; v[3 * i + 2] = B[i].v.z; }
        if (q) { q[4 * i] = B[i].q.w; q[4 * i + 1] = B[i].q.x; q[4 * i + 2] = B[i].q.y; q[4 * i + 3] = B[i].q.z; }
        if (w) { w[3 * i] = B[i].w.x; w[3 * i + 1] = B[i].w.y; w[3 * i + 2] = B[i].w.z; }
    }
    return NE_OK;
}

int ne_ref_get_body_aux(void* h, uint32_t* flags, double* sleep_timer, double* inv_mass,
                        double* inv_inertia) {
    RefWorld* r = static_cast<RefWorld*>(h);
    if (!r) return NE_ERR_INVALID;
    const auto& B = r->w.bodies;
    for (size_t i = 0; i < B.size(); ++i) {
        if (flags)
            flags[i] = (B[i].kinematic ? NE_FLAG_KINEMATIC : 0u) | (B[i].sensor ? NE_FLAG_SENSOR : 0u) |
                       (B[i].dynamic ? NE_FLAG_DYNAMIC : 0u) | (B[i].sleeping ? NE_FLAG_SLEEPING : 0u) |
                       (B[i].ccd ? NE_FLAG_CCD : 0u);
        if (sleep_timer) sleep_timer[i] = B[i].sleepTimer;
        if (inv_mass) inv_mass[i] = B[i].invMass;
        if (inv_inertia) {
            inv_inertia[3 * i] = B[i].invInertiaBody.x;
            inv_inertia[3 * i + 1] = B[i].invInertiaBody.y;
            inv_inertia[3 * i + 2] = B[i].invInertiaBody.z;
        }
    }
    return NE_OK;
}

// The last step's constraint list as captured by the reference itself
// (src/world.cpp:552): i, j, normal (b->a), point, depth, in constraint order.
int ne_ref_get_contacts(void* h, uint32_t cap, uint32_t* count, uint32_t* ci, uint32_t* cj,
                        double* normal, double* point, double* depth) {
    RefWorld* r = static_cast<RefWorld*>(h);
    if (!r || !count) return NE_ERR_INVALID;
    const auto& D = r->w.debugContacts;
    *count = (uint32_t)D.size();
    uint32_t m = (uint32_t)D.size() < cap ? (uint32_t)D.size() : cap;
    for (uint32_t k = 0; k < m; ++k) {
        if (ci) ci[k] = (uint32_t)D[k].i;
        if (cj) cj[k] = (uint32_t)D[k].j;
        if (normal) { normal[3 * k] = D[k].normal.x; normal[3 * k + 1] = D[k].normal.y; normal[3 * k + 2] = D[k].normal.z; }
        if (point) { point[3 * k] = D[k].point.x; point[3 * k + 1] = D[k].point.y; point[3 * k + 2] = D[k].point.z; }
        if (depth) depth[k] = D[k].depth;
    }
    return NE_OK;
}

// Narrow phase of one pair (a, b) of the current bodies, the way the
// reference's collide() gathers it for non-mesh shapes (src/world.cpp:122-129):
// detectContact, then the box-box clipping manifold, falling back to the single
// detected point when the manifold is empty.  Returns the contact count.
int ne_ref_pair_contacts(void* h, uint32_t a, uint32_t b, uint32_t cap, uint32_t* count,
                         double* normal, double* point, double* depth) {
    RefWorld* r = static_cast<RefWorld*>(h);
    if (!r || !count) return NE_ERR_INVALID;
    const auto& B = r->w.bodies;
    if (a >= B.size() || b >= B.size()) return NE_ERR_INVALID;
    std::vector<ne::Contact> out;
    ne::Contact c = ne::detectContact(B[a], B[b], r->w.box);
    if (c.hit) {
        bool done = false;
        if (B[a].shape == ne::Shape::Box && B[b].shape == ne::Shape::Box) {
            ne::Manifold mf = ne::boxBoxManifold(B[a], B[b], c.normal, r->w.box);
            for (int p = 0; p < mf.count; ++p) out.push_back({true, c.normal, mf.point[p], mf.depth[p]});
            done = mf.count > 0;
        }
        if (!done) out.push_back(c);
    }
    *count = (uint32_t)out.size();
    uint32_t m = (uint32_t)out.size() < cap ? (uint32_t)out.size() : cap;
    for (uint32_t k = 0; k < m; ++k) {
        if (normal) { normal[3 * k] = out[k].normal.x; normal[3 * k + 1] = out[k].normal.y; normal[3 * k + 2] = out[k].normal.z; }
        if (point) { point[3 * k] = out[k].point.x; point[3 * k + 1] = out[k].point.y; point[3 * k + 2] = out[k].point.z; }
        if (depth) depth[k] = out[k].overlap;
    }
    return NE_OK;
}

int ne_ref_diagnostics(void* h, double* p, double* L, double* ke) {
    RefWorld* r = static_cast<RefWorld*>(h);
    if (!r) return NE_ERR_INVALID;
    if (p) { ne::V3 t = r->w.totalLinearMomentum(); p[0] = t.x; p[1] = t.y; p[2] = t.z; }
    if (L) { ne::V3 t = r->w.totalAngularMomentum(); L[0] = t.x; L[1] = t.y; L[2] = t.z; }
    if (ke) *ke = r->w.totalKinetic();
    return NE_OK;
}

// Box::minImage / Box::wrap of the current world's box (src/box.hpp:24-41).
int ne_ref_min_image(void* h, const double* d, double* out) {
    RefWorld* r = static_cast<RefWorld*>(h);
    if (!r) return NE_ERR_INVALID;
    ne::V3 m = r->w.box.minImage(ne::V3{d[0], d[1], d[2]});
    out[0] = m.x; out[1] = m.y; out[2] = m.z;
    return NE_OK;
}

int ne_ref_wrap(void* h, const double* p, double* out) {
    RefWorld* r = static_cast<RefWorld*>(h);
    if (!r) return NE_ERR_INVALID;
    ne::V3 m = r->w.box.wrap(ne::V3{p[0], p[1], p[2]});
    out[0] = m.x; out[1] = m.y; out[2] = m.z;
    return NE_OK;
}

// The reference's snapshot of the world's bodies: a PhysicsWorld holding a
// copy of the world (no facade joints) writes it.  *size is always set; the
// bytes are copied when cap is enough.
int ne_ref_save_state(void* h, void* out, uint64_t cap, uint64_t* size) {
    RefWorld* r = static_cast<RefWorld*>(h);
    if (!r || !size) return NE_ERR_INVALID;
    ne::PhysicsWorld pw;
    pw.world() = r->w;
    const std::vector<char> b = pw.saveState();
    *size = b.size();
    if (out && cap >= b.size()) std::memcpy(out, b.data(), b.size());
    return NE_OK;
}

// The reference's loadState on a copy of the world, copied back on success
// (bodies from the snapshot, warm cache reset).
int ne_ref_load_state(void* h, const void* data, uint64_t size) {
    RefWorld* r = static_cast<RefWorld*>(h);
    if (!r || (size && !data)) return NE_ERR_INVALID;
    ne::PhysicsWorld pw;
    pw.world() = r->w;
    const char* p = static_cast<const char*>(data);
    if (!pw.loadState(std::vector<char>(p, p + size))) return NE_ERR_INVALID;
    r->w = pw.world();
    return NE_OK;
}

// The reference's factories (src/body.hpp:206-345) on the given sizes:
// sphere dims = (r), cylinder (r, height), box (hx, hy, hz), capsule
// (r, halfHeight), convex the vertices (dims unused), plane dims = the normal.
// Writes invMass, invInertiaBody (3), halfExtents (3), radius, halfHeight.
int ne_ref_make_body(int32_t shape, const double* dims, uint32_t nverts, const double* verts, double density,
                     double* inv_mass, double* inv_inertia, double* half_extents, double* radius,
                     double* half_height) {
    const ne::V3 x{0.0, 0.0, 0.0};
    const ne::Q q{1.0, 0.0, 0.0, 0.0};
    ne::Body b;
    switch (shape) {
        case NE_SHAPE_SPHERE: b = ne::makeSphere(0, x, dims[0], density); break;
        case NE_SHAPE_CYLINDER: b = ne::makeCylinder(0, x, q, dims[0], dims[1], density); break;
        case NE_SHAPE_BOX: b = ne::makeBox(0, x, q, ne::V3{dims[0], dims[1], dims[2]}, density); break;
        case NE_SHAPE_CAPSULE: b = ne::makeCapsule(0, x, q, dims[0], dims[1], density); break;
        case NE_SHAPE_CONVEX: {
            std::vector<ne::V3> vs;
            for (uint32_t k = 0; k < nverts; ++k) vs.push_back(v3(verts, k));
            b = ne::makeConvex(0, x, q, vs, density);
            break;
        }
        case NE_SHAPE_PLANE: b = ne::makePlane(0, x, ne::V3{dims[0], dims[1], dims[2]}); break;
        default: return NE_ERR_INVALID;
    }
    *inv_mass = b.invMass;
    inv_inertia[0] = b.invInertiaBody.x; inv_inertia[1] = b.invInertiaBody.y; inv_inertia[2] = b.invInertiaBody.z;
    half_extents[0] = b.halfExtents.x; half_extents[1] = b.halfExtents.y; half_extents[2] = b.halfExtents.z;
    *radius = b.radius;
    *half_height = b.halfHeight;
    return NE_OK;
}

// Facade events: moves the world into a PhysicsWorld (one handle per body,
// handle = index + 1, as add() hands them out in order) with every contact
// and trigger callback recording into the event list.  The world must have
// no distance joints (PhysicsWorld::step rebuilds them from its own).
int ne_ref_facade_begin(void* h) {
    RefWorld* r = static_cast<RefWorld*>(h);
    if (!r || !r->w.distanceJoints.empty()) return NE_ERR_INVALID;
    r->pw = std::make_unique<ne::PhysicsWorld>();
    for (size_t k = 0; k < r->w.bodies.size(); ++k) r->pw->addSphere(ne::V3{0.0, 0.0, 0.0}, 0.5, 1.0);
    r->pw->world() = r->w;
    auto rec = [r](int kind) { return [r, kind](const ne::ContactInfo& c) { r->events.push_back({kind, c}); }; };
    r->pw->setContactCallback(rec(NE_EVENT_CONTACT));
    r->pw->setContactBeginCallback(rec(NE_EVENT_BEGIN));
    r->pw->setContactStayCallback(rec(NE_EVENT_STAY));
    r->pw->setContactEndCallback(rec(NE_EVENT_END));
    r->pw->setTriggerEnterCallback(rec(NE_EVENT_TRIGGER_ENTER));
    r->pw->setTriggerStayCallback(rec(NE_EVENT_TRIGGER_STAY));
    r->pw->setTriggerExitCallback(rec(NE_EVENT_TRIGGER_EXIT));
    return NE_OK;
}

// PhysicsWorld::step, nsteps times; the event list keeps the last step's.
int ne_ref_facade_step(void* h, int32_t nsteps) {
    RefWorld* r = static_cast<RefWorld*>(h);
    if (!r || !r->pw) return NE_ERR_INVALID;
    for (int32_t t = 0; t < nsteps; ++t) {
        r->events.clear();
        r->pw->step();
    }
    r->w = r->pw->world();
    return NE_OK;
}

// The last facade step's events of one kind, in callback order, as body
// indices (handle - 1).
int ne_ref_facade_events(void* h, int32_t kind, uint32_t cap, uint32_t* count, uint32_t* ci, uint32_t* cj,
                         double* point, double* normal, double* depth) {
    RefWorld* r = static_cast<RefWorld*>(h);
    if (!r || !count) return NE_ERR_INVALID;
    uint32_t m = 0;
    for (const RefEvent& e : r->events) {
        if (e.kind != kind) continue;
        if (m < cap) {
            if (ci) ci[m] = e.c.a - 1;
            if (cj) cj[m] = e.c.b - 1;
            if (point) { point[3 * m] = e.c.point.x; point[3 * m + 1] = e.c.point.y; point[3 * m + 2] = e.c.point.z; }
            if (normal) { normal[3 * m] = e.c.normal.x; normal[3 * m + 1] = e.c.normal.y; normal[3 * m + 2] = e.c.normal.z; }
            if (depth) depth[m] = e.c.depth;
        }
        ++m;
    }
    *count = m;
    return NE_OK;
}

}  // extern "C"

// ---- scene queries (PhysicsWorld, src/physics_world.cpp:146-251) --------
// The queries run on a PhysicsWorld over the current world: the stepping
// facade when events are on, else a fresh facade holding a copy (handles
// 1..n in body order, so index = handle - 1).
namespace {
ne::PhysicsWorld& query_facade(RefWorld* r, std::unique_ptr<ne::PhysicsWorld>& tmp) {
    if (r->pw) return *r->pw;
    tmp = std::make_unique<ne::PhysicsWorld>();
    for (size_t k = 0; k < r->w.bodies.size(); ++k) tmp->addSphere(ne::V3{0.0, 0.0, 0.0}, 0.5, 1.0);
    tmp->world() = r->w;
    return *tmp;
}

void put_hit(const ne::RayHit& hit, uint32_t k, int32_t* body, double* distance, double* point, double* normal) {
    body[k] = hit.hit ? (int32_t)hit.body - 1 : -1;
    distance[k] = hit.distance;
    point[3 * k] = hit.point.x; point[3 * k + 1] = hit.point.y; point[3 * k + 2] = hit.point.z;
    normal[3 * k] = hit.normal.x; normal[3 * k + 1] = hit.normal.y; normal[3 * k + 2] = hit.normal.z;
}
}  // namespace

extern "C" {
int ne_ref_raycast(void* h, uint32_t n, const double* origin, const double* dir, const double* max_dist,
                   const uint32_t* mask, const double* cast_radius, int32_t* body, double* distance, double* point,
                   double* normal) {
    RefWorld* r = static_cast<RefWorld*>(h);
    if (!r) return NE_ERR_INVALID;
    std::unique_ptr<ne::PhysicsWorld> tmp;
    ne::PhysicsWorld& pw = query_facade(r, tmp);
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t m = mask ? mask[k] : 0xFFFFFFFFu;
        ne::RayHit hit = cast_radius ? pw.sphereCast(v3(origin, k), cast_radius[k], v3(dir, k), max_dist[k], m)
                                     : pw.raycast(v3(origin, k), v3(dir, k), max_dist[k], m);
        put_hit(hit, k, body, distance, point, normal);
    }
    return NE_OK;
}

int ne_ref_raycast_all(void* h, uint32_t n, const double* origin, const double* dir, const double* max_dist,
                       const uint32_t* mask, uint32_t cap, uint32_t* offsets, int32_t* body, double* distance,
                       double* point, double* normal) {
    RefWorld* r = static_cast<RefWorld*>(h);
    if (!r) return NE_ERR_INVALID;
    std::unique_ptr<ne::PhysicsWorld> tmp;
    ne::PhysicsWorld& pw = query_facade(r, tmp);
    std::vector<std::vector<ne::RayHit>> all(n);
    offsets[0] = 0;
    for (uint32_t k = 0; k < n; ++k) {
        all[k] = pw.raycastAll(v3(origin, k), v3(dir, k), max_dist[k], mask ? mask[k] : 0xFFFFFFFFu);
        offsets[k + 1] = offsets[k] + (uint32_t)all[k].size();
    }
    if (offsets[n] > cap) return NE_ERR_CAPACITY;
    for (uint32_t k = 0; k < n; ++k)
        for (size_t e = 0; e < all[k].size(); ++e) put_hit(all[k][e], offsets[k] + (uint32_t)e, body, distance, point, normal);
    return NE_OK;
}

// contacts == 0: overlapSphere (sphere shapes) / overlapBox (box shapes);
// contacts != 0: overlapContacts with the shape as the query Body.
int ne_ref_overlap(void* h, uint32_t n, const ne_query_shape* s, const uint32_t* mask, int32_t contacts, uint32_t cap,
                   uint32_t* offsets, int32_t* body, double* normal, double* depth) {
    RefWorld* r = static_cast<RefWorld*>(h);
    if (!r) return NE_ERR_INVALID;
    std::unique_ptr<ne::PhysicsWorld> tmp;
    ne::PhysicsWorld& pw = query_facade(r, tmp);
    std::vector<std::vector<ne::PhysicsWorld::OverlapHit>> all(n);
    offsets[0] = 0;
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t m = mask ? mask[k] : 0xFFFFFFFFu;
        const ne::V3 x{s[k].x[0], s[k].x[1], s[k].x[2]};
        const ne::Q q{s[k].q[0], s[k].q[1], s[k].q[2], s[k].q[3]};
        const ne::V3 he{s[k].half_extents[0], s[k].half_extents[1], s[k].half_extents[2]};
        if (!contacts) {
            std::vector<ne::BodyId> ids;
            if (s[k].shape == NE_SHAPE_SPHERE) ids = pw.overlapSphere(x, s[k].radius, m);
            else if (s[k].shape == NE_SHAPE_BOX) ids = pw.overlapBox(x, q, he, m);
            else return NE_ERR_INVALID;
            for (ne::BodyId id : ids) {
                ne::PhysicsWorld::OverlapHit oh;
                oh.body = id;
                all[k].push_back(oh);
            }
        } else {
            ne::Body b;
            b.shape = static_cast<ne::Shape>(s[k].shape);
            b.x = x;
            b.q = q;
            b.halfExtents = he;
            b.radius = s[k].radius;
            b.halfHeight = s[k].half_height;
            for (uint32_t v = 0; v < s[k].nverts; ++v) b.vertices.push_back(v3(s[k].verts, v));
            all[k] = pw.overlapContacts(b, m);
        }
        offsets[k + 1] = offsets[k] + (uint32_t)all[k].size();
    }
    if (offsets[n] > cap) return NE_ERR_CAPACITY;
    for (uint32_t k = 0; k < n; ++k)
        for (size_t e = 0; e < all[k].size(); ++e) {
            const uint32_t at = offsets[k] + (uint32_t)e;
            body[at] = (int32_t)all[k][e].body - 1;
            if (normal) { normal[3 * at] = all[k][e].normal.x; normal[3 * at + 1] = all[k][e].normal.y; normal[3 * at + 2] = all[k][e].normal.z; }
            if (depth) depth[at] = all[k][e].depth;
        }
    return NE_OK;
}
}  // extern "C"

// ---- general joints (World::joints, src/world.hpp:31-54, 130) -----------
extern "C" {
int ne_ref_set_general_joints(void* h, uint32_t n, const ne_joint* js) {
    RefWorld* r = static_cast<RefWorld*>(h);
    if (!r || (n && !js)) return NE_ERR_INVALID;
    r->w.joints.clear();
    for (uint32_t k = 0; k < n; ++k) {
        const ne_joint& s = js[k];
        if (s.a >= r->w.bodies.size() || s.b >= r->w.bodies.size() || s.type < 0 || s.type > NE_JOINT_SLIDER)
            return NE_ERR_INVALID;
        ne::Joint j;
        j.type = static_cast<ne::JointType>(s.type);
        j.a = s.a;
        j.b = s.b;
        j.localA = v3(s.local_a, 0);
        j.localB = v3(s.local_b, 0);
        j.axisA = v3(s.axis_a, 0);
        j.axisB = v3(s.axis_b, 0);
        j.refRel = ne::Q{s.ref_rel[0], s.ref_rel[1], s.ref_rel[2], s.ref_rel[3]};
        j.useMotor = s.use_motor != 0;
        j.motorSpeed = s.motor_speed;
        j.maxMotorImpulse = s.max_motor_impulse;
        j.useLimit = s.use_limit != 0;
        j.lower = s.lower;
        j.upper = s.upper;
        j.breakable = s.breakable != 0;
        j.breakImpulse = s.break_impulse;
        j.broken = s.broken != 0;
        r->w.joints.push_back(j);
    }
    return NE_OK;
}

int ne_ref_get_general_joints(void* h, uint32_t n, int32_t* broken, double* applied, double* motor) {
    RefWorld* r = static_cast<RefWorld*>(h);
    if (!r || n > r->w.joints.size()) return NE_ERR_INVALID;
    for (uint32_t k = 0; k < n; ++k) {
        if (broken) broken[k] = r->w.joints[k].broken ? 1 : 0;
        if (applied) applied[k] = r->w.joints[k].appliedImpulse;
        if (motor) motor[k] = r->w.joints[k].motorAcc;
    }
    return NE_OK;
}

// The frames PhysicsWorld::addGeneralJoint computes (src/physics_world.hpp:
// 408-423) for bodies at poses (xa, qa), (xb, qb), in the reference's own
// V3 / Q arithmetic; anchor NULL: addFixedJoint's midpoint of the centres.
int ne_ref_joint_frames(const double* xa, const double* qa, const double* xb, const double* qb, const double* anchor,
                        const double* axis, ne_joint* out) {
    const ne::V3 XA = v3(xa, 0), XB = v3(xb, 0);
    const ne::Q QA{qa[0], qa[1], qa[2], qa[3]}, QB{qb[0], qb[1], qb[2], qb[3]};
    const ne::V3 A = anchor ? v3(anchor, 0) : (XA + XB) * 0.5;
    const ne::V3 la = QA.inverseRotate(A - XA), lb = QB.inverseRotate(A - XB);
    const ne::V3 ax = v3(axis, 0).normalized();
    const ne::V3 aa = QA.inverseRotate(ax), ab = QB.inverseRotate(ax);
    const ne::Q rr = QA * QB.conjugate();
    const double* src[4] = {&la.x, &lb.x, &aa.x, &ab.x};
    double* dst[4] = {out->local_a, out->local_b, out->axis_a, out->axis_b};
    for (int k = 0; k < 4; ++k)
        for (int c = 0; c < 3; ++c) dst[k][c] = src[k][c];
    out->ref_rel[0] = rr.w; out->ref_rel[1] = rr.x; out->ref_rel[2] = rr.y; out->ref_rel[3] = rr.z;
    return NE_OK;
}
}  // extern "C"
