// ne_gpu_session.hpp -- steps any World-like object on the MI355X.
//
// A header-only C++17 adapter over include/ne_gpu.h for code that already has
// the reference's `ne::World` (src/world.hpp:73-161) or a type with the same
// public fields: it flattens `bodies` into the C ABI's arrays, runs the step
// on the GPU and writes the state back.  It names no reference type, so it can
// be included next to the reference's own headers:
//
//     #include "world.hpp"            // the reference
//     #include "ne_gpu_session.hpp"   // this adapter
//     void World::step() { static thread_local ne_gpu::Session s; s.step(*this); }
//
// Errors from the GPU path become std::runtime_error (the reference's step()
// has no error channel); a pair/contact buffer overflow is retried with the
// enlarged buffers the library allocates on overflow.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "ne_gpu.h"

namespace ne_gpu {

// Adapter usable with any World-like type that has the reference's public
// fields (the reference's own `ne::World` included): flattens its bodies into
// the C ABI's arrays, steps on the GPU, and writes the state back.
class Session {
public:
    explicit Session(int device = 0) {
        ne_gpu_config cfg{};
        cfg.device = device;
        int rc = ne_gpu_create(&cfg, &ctx_);
        if (rc != NE_OK) throw std::runtime_error("ne_gpu_create failed: " + std::to_string(rc));
    }
    ~Session() { if (ctx_) ne_gpu_destroy(ctx_); }
    Session(const Session&) = delete;
    Session& operator=(const Session&) = delete;
    ne_gpu_ctx* context() const { return ctx_; }

    template <class W>
    void step(W& w, int nsteps = 1, int iterations = 8) {
        for (int attempt = 0; attempt < 4; ++attempt) {
            upload(w, iterations);
            int rc = ne_gpu_step(ctx_, nsteps);
            if (rc == NE_OK || rc == NE_WARN_EPA_CAPPED) {   // the warning: the steps ran, see epaCapped()
                epaCapped_ = rc == NE_WARN_EPA_CAPPED;
                download(w);
                return;
            }
            if (rc != NE_ERR_CAPACITY) fail(rc);
        }
        fail(NE_ERR_CAPACITY);
    }
    // Whether the last step() had an EPA run stop at the workspace's face cap
    // (its contact can then differ from the reference's; see NE_WARN_EPA_CAPPED).
    bool epaCapped() const { return epaCapped_; }
    void resetWarmCache() { check(ne_gpu_reset_warm_cache(ctx_)); }
    void exportWarmCache(std::vector<std::uint64_t>& keys, std::vector<double>& imp) {
        std::uint32_t n = 0;
        check(ne_gpu_export_warm_cache(ctx_, 0, &n, nullptr, nullptr));
        keys.resize(n);
        imp.resize(n);
        if (n) check(ne_gpu_export_warm_cache(ctx_, n, &n, keys.data(), imp.data()));
    }
    void importWarmCache(const std::vector<std::uint64_t>& keys, const std::vector<double>& imp) {
        check(ne_gpu_import_warm_cache(ctx_, static_cast<std::uint32_t>(keys.size()), keys.data(), imp.data()));
    }

private:
    ne_gpu_ctx* ctx_ = nullptr;
    bool epaCapped_ = false;
    std::vector<double> x_, v_, q_, w_, f_, t_, im_, ii_, r_, hh_, he_, fr_, re_, ld_, ad_, gs_, hull_;
    std::vector<double> st_, ims_, iis_;                  // sleep timer, stored inverse mass / inertia
    std::vector<ne_mesh> meshes_;
    std::vector<std::vector<double>> meshVerts_;
    std::vector<std::vector<std::int32_t>> meshTris_;
    std::vector<std::uint32_t> fl_;
    std::vector<std::int32_t> shape_;
    std::vector<std::uint32_t> layer_, mask_, flags_, hoff_, noff_, nc_;
    std::vector<ne_distance_joint> joints_;
    std::vector<ne_joint> gjoints_;
    std::vector<std::int32_t> gjBroken_;
    std::vector<double> gjApplied_, gjMotor_;

    [[noreturn]] void fail(int rc) {
        throw std::runtime_error("ne_gpu step failed (" + std::to_string(rc) + "): " + ne_gpu_last_error(ctx_));
    }
    void check(int rc) { if (rc != NE_OK) fail(rc); }

    template <class Vec>
    static void put3(std::vector<double>& a, size_t i, const Vec& v) { a[3 * i] = v.x; a[3 * i + 1] = v.y; a[3 * i + 2] = v.z; }

    template <class W>
    void upload(W& w, int iterations) {
        ne_world_params p{};
        const auto& bx = w.box;
        H3 h = halvesOf(bx);
        p.half[0] = h.x; p.half[1] = h.y; p.half[2] = h.z;
        p.periodic = bx.periodic ? 1 : 0;
        p.open_boundary = w.openBoundary ? 1 : 0;
        p.dt = w.dt;
        p.gravity[0] = w.gravity.x; p.gravity[1] = w.gravity.y; p.gravity[2] = w.gravity.z;
        p.restitution = w.restitution;
        p.friction = w.friction;
        p.contact_beta = w.contactBeta;
        p.restitution_combine = static_cast<int32_t>(w.restitutionCombine);
        p.friction_combine = static_cast<int32_t>(w.frictionCombine);
        p.warm_start = w.warmStart ? 1 : 0;
        p.force_brute_force = w.forceBruteForce ? 1 : 0;
        p.solver_iterations = iterations;
        p.sleep_enabled = w.sleepEnabled ? 1 : 0;
        p.sleep_lin_vel = w.sleepLinVel; p.sleep_ang_vel = w.sleepAngVel;
        p.sleep_time = w.sleepTime; p.wake_vel = w.wakeVel;
        check(ne_gpu_set_params(ctx_, &p));
        const size_t n = w.bodies.size();
        for (auto* a : {&x_, &v_, &w_, &f_, &t_, &ii_, &he_}) a->assign(3 * n, 0.0);
        q_.assign(4 * n, 0.0);
        for (auto* a : {&im_, &r_, &hh_, &fr_, &re_, &ld_, &ad_, &gs_, &st_, &ims_}) a->assign(n, 0.0);
        iis_.assign(3 * n, 0.0);
        shape_.assign(n, 0);
        layer_.assign(n, 0); mask_.assign(n, 0); flags_.assign(n, 0);
        hoff_.assign(n + 1, 0); noff_.assign(n + 1, 0);
        hull_.clear(); nc_.clear();
        for (size_t i = 0; i < n; ++i) {
            const auto& b = w.bodies[i];
            put3(x_, i, b.x); put3(v_, i, b.v); put3(w_, i, b.w); put3(f_, i, b.force); put3(t_, i, b.torque);
            put3(ii_, i, b.invInertiaBody); put3(he_, i, b.halfExtents);
            q_[4 * i] = b.q.w; q_[4 * i + 1] = b.q.x; q_[4 * i + 2] = b.q.y; q_[4 * i + 3] = b.q.z;
            im_[i] = b.invMass; r_[i] = b.radius; hh_[i] = b.halfHeight;
            fr_[i] = b.friction; re_[i] = b.restitution; ld_[i] = b.linearDamping; ad_[i] = b.angularDamping;
            gs_[i] = b.gravityScale;
            shape_[i] = static_cast<std::int32_t>(b.shape);
            layer_[i] = b.layer; mask_[i] = b.mask;
            flags_[i] = (b.kinematic ? NE_FLAG_KINEMATIC : 0u) | (b.sensor ? NE_FLAG_SENSOR : 0u) |
                        (b.dynamic ? NE_FLAG_DYNAMIC : 0u) | (b.sleeping ? NE_FLAG_SLEEPING : 0u) |
                        (b.ccd ? NE_FLAG_CCD : 0u);
            for (const auto& vv : b.vertices) { hull_.push_back(vv.x); hull_.push_back(vv.y); hull_.push_back(vv.z); }
            hoff_[i + 1] = static_cast<std::uint32_t>(hull_.size() / 3);
            for (auto k : b.noCollide) nc_.push_back(k);
            noff_[i + 1] = static_cast<std::uint32_t>(nc_.size());
            putSleep(b, i, 0);
        }
        if (hull_.empty()) hull_.assign(3, 0.0);
        if (nc_.empty()) nc_.assign(1, 0u);
        ne_bodies B{};
        B.n = static_cast<std::uint32_t>(n);
        B.x = x_.data(); B.v = v_.data(); B.q = q_.data(); B.w = w_.data();
        B.force = f_.data(); B.torque = t_.data();
        B.inv_mass = im_.data(); B.inv_inertia = ii_.data(); B.shape = shape_.data();
        B.radius = r_.data(); B.half_height = hh_.data(); B.half_extents = he_.data();
        B.friction = fr_.data(); B.restitution = re_.data();
        B.linear_damping = ld_.data(); B.angular_damping = ad_.data(); B.gravity_scale = gs_.data();
        B.layer = layer_.data(); B.mask = mask_.data(); B.flags = flags_.data();
        B.hull_offsets = hoff_.data(); B.hull_verts = hull_.data();
        B.nocollide_offsets = noff_.data(); B.nocollide = nc_.data();
        B.sleep_timer = st_.data(); B.inv_mass_store = ims_.data(); B.inv_inertia_store = iis_.data();
        check(ne_gpu_set_bodies(ctx_, &B));
        uploadJoints(w, 0);
        uploadGeneralJoints(w, 0);
        uploadMeshes(w, 0);
    }

    // Body::sleepTimer / invMassStore / invInertiaStore (src/body.hpp:107-110),
    // when the body type has them
    template <class B>
    auto putSleep(const B& b, size_t i, int) -> decltype(b.sleepTimer, b.invMassStore, b.invInertiaStore.x, void()) {
        st_[i] = b.sleepTimer; ims_[i] = b.invMassStore; put3(iis_, i, b.invInertiaStore);
    }
    template <class B>
    void putSleep(const B&, size_t, long) {}
    template <class B>
    auto getSleep(B& b, size_t i, int) -> decltype(b.sleepTimer, b.invMassStore, b.invInertiaStore.x, void()) {
        b.sleeping = (fl_[i] & NE_FLAG_SLEEPING) != 0;
        b.sleepTimer = st_[i];
        b.invMass = im_[i];
        b.invInertiaBody.x = ii_[3 * i]; b.invInertiaBody.y = ii_[3 * i + 1]; b.invInertiaBody.z = ii_[3 * i + 2];
        b.invMassStore = ims_[i];
        b.invInertiaStore.x = iis_[3 * i]; b.invInertiaStore.y = iis_[3 * i + 1]; b.invInertiaStore.z = iis_[3 * i + 2];
    }
    template <class B>
    void getSleep(B&, size_t, long) {}

    // Body::mesh (src/body.hpp:62) of the Shape::Mesh bodies, when the body
    // type has it.  The bodies upload drops the meshes, so they go up again.
    template <class W>
    auto uploadMeshes(W& w, int) -> decltype(w.bodies[0].mesh->verts[0].x, w.bodies[0].mesh->tris[0].a, void()) {
        meshes_.clear(); meshVerts_.clear(); meshTris_.clear();
        for (size_t i = 0; i < w.bodies.size(); ++i) {
            const auto& b = w.bodies[i];
            if (static_cast<int>(b.shape) != NE_SHAPE_MESH || !b.mesh) continue;
            std::vector<double> vv;
            std::vector<std::int32_t> tt;
            for (const auto& p : b.mesh->verts) { vv.push_back(p.x); vv.push_back(p.y); vv.push_back(p.z); }
            for (const auto& t : b.mesh->tris) { tt.push_back(t.a); tt.push_back(t.b); tt.push_back(t.c); }
            meshVerts_.push_back(std::move(vv));
            meshTris_.push_back(std::move(tt));
            ne_mesh m{};
            m.body = static_cast<std::uint32_t>(i);
            m.nverts = static_cast<std::uint32_t>(b.mesh->verts.size());
            m.ntris = static_cast<std::uint32_t>(b.mesh->tris.size());
            meshes_.push_back(m);
        }
        for (size_t k = 0; k < meshes_.size(); ++k) {
            meshes_[k].verts = meshVerts_[k].data();
            meshes_[k].tris = meshTris_[k].data();
        }
        if (!meshes_.empty()) check(ne_gpu_set_meshes(ctx_, static_cast<std::uint32_t>(meshes_.size()), meshes_.data()));
    }
    template <class W>
    void uploadMeshes(W&, long) {}

    // World::distanceJoints (src/world.hpp:20-26, 129), when the type has them
    template <class W>
    auto uploadJoints(W& w, int) -> decltype(w.distanceJoints.size(), void()) {
        joints_.resize(w.distanceJoints.size());
        for (size_t k = 0; k < joints_.size(); ++k) {
            const auto& d = w.distanceJoints[k];
            ne_distance_joint& j = joints_[k];
            j.a = static_cast<std::uint32_t>(d.a);
            j.b = static_cast<std::uint32_t>(d.b);
            j.local_a[0] = d.localA.x; j.local_a[1] = d.localA.y; j.local_a[2] = d.localA.z;
            j.local_b[0] = d.localB.x; j.local_b[1] = d.localB.y; j.local_b[2] = d.localB.z;
            j.rest = d.rest; j.stiffness = d.stiffness; j.damping = d.damping;
        }
        check(ne_gpu_set_joints(ctx_, static_cast<std::uint32_t>(joints_.size()), joints_.data()));
    }
    template <class W>
    void uploadJoints(W&, long) {}

    // World::joints (src/world.hpp:28-53, 130), when the type has them: the
    // general joints with their broken flags; the step's accumulators start
    // from zero on the device as in World::solveJoints.
    template <class W>
    auto uploadGeneralJoints(W& w, int) -> decltype(w.joints.size(), void()) {
        gjoints_.assign(w.joints.size(), ne_joint{});
        for (size_t k = 0; k < gjoints_.size(); ++k) {
            const auto& s = w.joints[k];
            ne_joint& j = gjoints_[k];
            j.type = static_cast<std::int32_t>(s.type);
            j.a = static_cast<std::uint32_t>(s.a);
            j.b = static_cast<std::uint32_t>(s.b);
            j.use_motor = s.useMotor ? 1 : 0;
            j.use_limit = s.useLimit ? 1 : 0;
            j.breakable = s.breakable ? 1 : 0;
            j.broken = s.broken ? 1 : 0;
            j.local_a[0] = s.localA.x; j.local_a[1] = s.localA.y; j.local_a[2] = s.localA.z;
            j.local_b[0] = s.localB.x; j.local_b[1] = s.localB.y; j.local_b[2] = s.localB.z;
            j.axis_a[0] = s.axisA.x; j.axis_a[1] = s.axisA.y; j.axis_a[2] = s.axisA.z;
            j.axis_b[0] = s.axisB.x; j.axis_b[1] = s.axisB.y; j.axis_b[2] = s.axisB.z;
            j.ref_rel[0] = s.refRel.w; j.ref_rel[1] = s.refRel.x; j.ref_rel[2] = s.refRel.y; j.ref_rel[3] = s.refRel.z;
            j.motor_speed = s.motorSpeed;
            j.max_motor_impulse = s.maxMotorImpulse;
            j.lower = s.lower;
            j.upper = s.upper;
            j.break_impulse = s.breakImpulse;
        }
        check(ne_gpu_set_general_joints(ctx_, static_cast<std::uint32_t>(gjoints_.size()), gjoints_.data()));
    }
    template <class W>
    void uploadGeneralJoints(W&, long) {}

    template <class W>
    auto downloadGeneralJoints(W& w, int) -> decltype(w.joints.size(), void()) {
        const size_t n = w.joints.size();
        if (n == 0) return;
        gjBroken_.assign(n, 0);
        gjApplied_.assign(n, 0.0);
        gjMotor_.assign(n, 0.0);
        check(ne_gpu_get_general_joints(ctx_, static_cast<std::uint32_t>(n), gjBroken_.data(), gjApplied_.data(),
                                        gjMotor_.data()));
        for (size_t k = 0; k < n; ++k) {
            w.joints[k].broken = gjBroken_[k] != 0;
            w.joints[k].appliedImpulse = gjApplied_[k];
            w.joints[k].motorAcc = gjMotor_[k];
        }
    }
    template <class W>
    void downloadGeneralJoints(W&, long) {}

    template <class W>
    void download(W& w) {
        const size_t n = w.bodies.size();
        check(ne_gpu_get_state(ctx_, x_.data(), v_.data(), q_.data(), w_.data()));
        fl_.assign(n, 0u);
        check(ne_gpu_get_body_aux(ctx_, fl_.data(), st_.data(), im_.data(), ii_.data()));
        check(ne_gpu_get_sleep_store(ctx_, ims_.data(), iis_.data()));
        for (size_t i = 0; i < n; ++i) {
            auto& b = w.bodies[i];
            b.prevX = b.x;            // world.cpp:135 snapshot (pre-step pose)
            b.prevQ = b.q;
            b.x.x = x_[3 * i]; b.x.y = x_[3 * i + 1]; b.x.z = x_[3 * i + 2];
            b.v.x = v_[3 * i]; b.v.y = v_[3 * i + 1]; b.v.z = v_[3 * i + 2];
            b.w.x = w_[3 * i]; b.w.y = w_[3 * i + 1]; b.w.z = w_[3 * i + 2];
            b.q.w = q_[4 * i]; b.q.x = q_[4 * i + 1]; b.q.y = q_[4 * i + 2]; b.q.z = q_[4 * i + 3];
            b.force.x = b.force.y = b.force.z = 0.0;   // integrate clears the accumulators
            b.torque.x = b.torque.y = b.torque.z = 0.0;
            getSleep(b, i, 0);
        }
        downloadGeneralJoints(w, 0);
    }

    struct H3 { double x, y, z; };
    template <class B>
    static auto halvesOf(const B& b) -> decltype(b.halves(), H3{}) { auto h = b.halves(); return H3{h.x, h.y, h.z}; }
    template <class B, class... Ignored>
    static H3 halvesOf(const B& b, Ignored...) { return H3{b.half, b.half, b.half}; }
};

}  // namespace ne_gpu
