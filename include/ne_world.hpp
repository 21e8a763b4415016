// ne_world.hpp -- C++17 `ne::World` whose step() runs on the MI355X.
//
// Mirrors the public interface of the reference's `ne::World`
// (src/world.hpp:73-161): the same public fields with the same meaning and
// defaults, `step()`, `wake()`, `resetWarmCache()` and the three `total*`
// diagnostics, over value types with the reference's member names (`V3`,
// `Q`, `Box`, `Body`, factories from src/body.hpp:206-345).  Code written
// against the reference's World compiles against this header unchanged for
// the hot path it uses.
//
// `step()` is the compat mode of the drop-in: it uploads `bodies` (the
// reference lets callers edit them freely between steps), runs one step on
// the GPU through include/ne_gpu.h and downloads the state back, so results
// are the reference's bit for bit.  `stepMany(k)` does the same around k
// resident steps (what throughput runs use).  The warm-start cache lives on
// the device and survives body edits exactly as the reference's does.
//
// Differences, by design:
//   - step() throws std::runtime_error if the GPU path cannot run the scene
//     (no device, a shape or feature not on the GPU path); the reference's
//     step() cannot fail.  A pair/contact buffer overflow is retried
//     transparently with larger buffers.
//   - Box carries per-axis half extents (`halfXYZ`), and `solverIterations`
//     is a field (the reference fixes 8, src/world.cpp:572).
#pragma once

#include <cmath>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "ne_gpu.h"
#include "ne_gpu_session.hpp"

namespace ne {

constexpr double PI = 3.14159265358979323846;

struct V3 {
    double x = 0.0, y = 0.0, z = 0.0;
    V3() = default;
    V3(double x_, double y_, double z_) : x(x_), y(y_), z(z_) {}
    V3 operator+(const V3& o) const { return {x + o.x, y + o.y, z + o.z}; }
    V3 operator-(const V3& o) const { return {x - o.x, y - o.y, z - o.z}; }
    V3 operator-() const { return {-x, -y, -z}; }
    V3 operator*(double s) const { return {x * s, y * s, z * s}; }
    V3 operator/(double s) const { return {x / s, y / s, z / s}; }
    V3& operator+=(const V3& o) { x += o.x; y += o.y; z += o.z; return *this; }
    V3& operator-=(const V3& o) { x -= o.x; y -= o.y; z -= o.z; return *this; }
    double dot(const V3& o) const { return x * o.x + y * o.y + z * o.z; }
    V3 cross(const V3& o) const { return {y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x}; }
    double norm2() const { return x * x + y * y + z * z; }
    double norm() const { return std::sqrt(norm2()); }
    V3 normalized() const {
        double n = norm();
        return (n > 1e-300) ? V3{x / n, y / n, z / n} : V3{0, 0, 0};
    }
};
inline V3 operator*(double s, const V3& v) { return v * s; }

struct Q {
    double w = 1.0, x = 0.0, y = 0.0, z = 0.0;
    Q() = default;
    Q(double w_, double x_, double y_, double z_) : w(w_), x(x_), y(y_), z(z_) {}
    Q operator*(const Q& o) const {
        return {w * o.w - x * o.x - y * o.y - z * o.z, w * o.x + x * o.w + y * o.z - z * o.y,
                w * o.y - x * o.z + y * o.w + z * o.x, w * o.z + x * o.y - y * o.x + z * o.w};
    }
    Q operator+(const Q& o) const { return {w + o.w, x + o.x, y + o.y, z + o.z}; }
    Q operator*(double s) const { return {w * s, x * s, y * s, z * s}; }
    double norm() const { return std::sqrt(w * w + x * x + y * y + z * z); }
    Q normalized() const {
        double n = norm();
        return (n > 1e-300) ? Q{w / n, x / n, y / n, z / n} : Q{1, 0, 0, 0};
    }
    Q conjugate() const { return {w, -x, -y, -z}; }
    V3 rotate(const V3& v) const {
        V3 qv{x, y, z};
        V3 t = 2.0 * qv.cross(v);
        return v + w * t + qv.cross(t);
    }
    V3 inverseRotate(const V3& v) const { return conjugate().rotate(v); }
};

// Cubic domain of half-edge `half` like the reference (src/box.hpp:16-42);
// `halfXYZ` (if any component is > 0) gives per-axis extents instead.
struct Box {
    double half = 12.0;
    bool periodic = false;
    V3 halfXYZ{0.0, 0.0, 0.0};
    double edge() const { return 2.0 * half; }
    V3 halves() const { return (halfXYZ.x > 0 || halfXYZ.y > 0 || halfXYZ.z > 0) ? halfXYZ : V3{half, half, half}; }
    V3 minImage(V3 d) const {
        if (!periodic) return d;
        const V3 h = halves();
        const double L[3] = {2.0 * h.x, 2.0 * h.y, 2.0 * h.z};
        d.x -= L[0] * std::floor(d.x / L[0] + 0.5);
        d.y -= L[1] * std::floor(d.y / L[1] + 0.5);
        d.z -= L[2] * std::floor(d.z / L[2] + 0.5);
        return d;
    }
    V3 wrap(V3 p) const {
        if (!periodic) return p;
        const V3 h = halves();
        p.x -= 2.0 * h.x * std::floor((p.x + h.x) / (2.0 * h.x));
        p.y -= 2.0 * h.y * std::floor((p.y + h.y) / (2.0 * h.y));
        p.z -= 2.0 * h.z * std::floor((p.z + h.z) / (2.0 * h.z));
        return p;
    }
};

enum class Shape { Sphere, Cylinder, Box, Capsule, Convex, Plane, Mesh, Compound };

// A static triangle mesh (src/mesh.hpp:14-25): vertices and index triples in
// the mesh body's frame.  The GPU library builds the bounding-volume tree
// itself when the mesh is uploaded (by the reference's median split), so
// build() has nothing to do here and the tree is not kept on the host.
struct MeshData {
    struct Tri { int a, b, c; };
    std::vector<V3> verts;
    std::vector<Tri> tris;
    void build() {}
};

struct Body {
    V3 x, v;
    Q q{1, 0, 0, 0};
    V3 w, force, torque;
    double invMass = 0.0;
    V3 invInertiaBody;
    Shape shape = Shape::Sphere;
    double radius = 1.0, halfHeight = 0.0;
    V3 halfExtents;
    std::vector<V3> vertices;
    int id = -1;
    double friction = -1.0, restitution = -1.0;
    double linearDamping = 0.0, angularDamping = 0.0, gravityScale = 1.0;
    std::uint32_t layer = 0, mask = 0xFFFFFFFFu;
    std::vector<std::uint32_t> noCollide;
    std::uint64_t userData = 0;
    bool dynamic = true, sleeping = false;
    double sleepTimer = 0.0, invMassStore = 0.0;
    V3 invInertiaStore;
    std::shared_ptr<const MeshData> mesh;   // Shape::Mesh only
    bool kinematic = false, sensor = false, ccd = false;
    V3 prevX;
    Q prevQ{1, 0, 0, 0};

    bool isSphere() const { return shape == Shape::Sphere; }
    V3 applyInvInertiaWorld(const V3& u) const {
        V3 ub = q.inverseRotate(u);
        return q.rotate(V3{ub.x * invInertiaBody.x, ub.y * invInertiaBody.y, ub.z * invInertiaBody.z});
    }
    V3 inertiaBody() const {
        return {invInertiaBody.x > 0 ? 1.0 / invInertiaBody.x : 0.0, invInertiaBody.y > 0 ? 1.0 / invInertiaBody.y : 0.0,
                invInertiaBody.z > 0 ? 1.0 / invInertiaBody.z : 0.0};
    }
    V3 angularMomentum() const {
        V3 I = inertiaBody(), wb = q.inverseRotate(w);
        return q.rotate(V3{wb.x * I.x, wb.y * I.y, wb.z * I.z});
    }
    double kinetic() const {
        double lin = 0.5 * (invMass > 0 ? 1.0 / invMass : 0.0) * v.norm2();
        V3 I = inertiaBody(), wb = q.inverseRotate(w);
        V3 Lb{wb.x * I.x, wb.y * I.y, wb.z * I.z};
        return lin + 0.5 * wb.dot(Lb);
    }
};

// Factories with the reference's mass/inertia formulas (src/body.hpp:206-345).
inline Body makeSphere(int id, const V3& x, double r, double density) {
    Body b; b.shape = Shape::Sphere; b.id = id; b.x = x; b.radius = r; b.halfHeight = 0.0;
    double m = density * (4.0 / 3.0) * PI * r * r * r;
    b.invMass = (m > 0) ? 1.0 / m : 0.0;
    double I = (2.0 / 5.0) * m * r * r, iI = (I > 0) ? 1.0 / I : 0.0;
    b.invInertiaBody = {iI, iI, iI};
    return b;
}
inline Body makeCylinder(int id, const V3& x, const Q& q, double r, double height, double density) {
    Body b; b.shape = Shape::Cylinder; b.id = id; b.x = x; b.q = q; b.radius = r; b.halfHeight = 0.5 * height;
    double m = density * PI * r * r * height;
    b.invMass = (m > 0) ? 1.0 / m : 0.0;
    double Iyy = 0.5 * m * r * r, Ixx = (1.0 / 12.0) * m * (3.0 * r * r + height * height);
    b.invInertiaBody = {Ixx > 0 ? 1.0 / Ixx : 0.0, Iyy > 0 ? 1.0 / Iyy : 0.0, Ixx > 0 ? 1.0 / Ixx : 0.0};
    return b;
}
inline Body makeBox(int id, const V3& x, const Q& q, const V3& he, double density) {
    Body b; b.shape = Shape::Box; b.id = id; b.x = x; b.q = q; b.halfExtents = he;
    const double hx = he.x, hy = he.y, hz = he.z;
    double m = density * 8.0 * hx * hy * hz;
    b.invMass = (m > 0) ? 1.0 / m : 0.0;
    double Ix = (1.0 / 3.0) * m * (hy * hy + hz * hz), Iy = (1.0 / 3.0) * m * (hx * hx + hz * hz),
           Iz = (1.0 / 3.0) * m * (hx * hx + hy * hy);
    b.invInertiaBody = {Ix > 0 ? 1.0 / Ix : 0.0, Iy > 0 ? 1.0 / Iy : 0.0, Iz > 0 ? 1.0 / Iz : 0.0};
    return b;
}
inline Body makeCapsule(int id, const V3& x, const Q& q, double r, double halfHeight, double density) {
    Body b; b.shape = Shape::Capsule; b.id = id; b.x = x; b.q = q; b.radius = r; b.halfHeight = halfHeight;
    const double L = 2.0 * halfHeight, mc = density * PI * r * r * L, mh = density * (4.0 / 3.0) * PI * r * r * r;
    const double m = mc + mh;
    b.invMass = (m > 0) ? 1.0 / m : 0.0;
    double Iyy = 0.5 * mc * r * r + 0.4 * mh * r * r;
    double Ixx = mc * (0.25 * r * r + L * L / 12.0) + mh * (0.4 * r * r + 0.25 * L * L);
    b.invInertiaBody = {Ixx > 0 ? 1.0 / Ixx : 0.0, Iyy > 0 ? 1.0 / Iyy : 0.0, Ixx > 0 ? 1.0 / Ixx : 0.0};
    return b;
}
inline Body makeConvex(int id, const V3& x, const Q& q, std::vector<V3> verts, double density) {
    Body b; b.shape = Shape::Convex; b.id = id; b.x = x; b.q = q; b.vertices = std::move(verts);
    V3 lo{1e30, 1e30, 1e30}, hi{-1e30, -1e30, -1e30};
    for (const V3& v : b.vertices) {
        lo = V3{std::fmin(lo.x, v.x), std::fmin(lo.y, v.y), std::fmin(lo.z, v.z)};
        hi = V3{std::fmax(hi.x, v.x), std::fmax(hi.y, v.y), std::fmax(hi.z, v.z)};
    }
    V3 he = (hi - lo) * 0.5;
    const double hx = std::fmax(he.x, 1e-4), hy = std::fmax(he.y, 1e-4), hz = std::fmax(he.z, 1e-4);
    double m = density * 8.0 * hx * hy * hz;
    b.invMass = (m > 0) ? 1.0 / m : 0.0;
    double Ix = (1.0 / 3.0) * m * (hy * hy + hz * hz), Iy = (1.0 / 3.0) * m * (hx * hx + hz * hz),
           Iz = (1.0 / 3.0) * m * (hx * hx + hy * hy);
    b.invInertiaBody = {Ix > 0 ? 1.0 / Ix : 0.0, Iy > 0 ? 1.0 / Iy : 0.0, Iz > 0 ? 1.0 / Iz : 0.0};
    return b;
}
inline Body makePlane(int id, const V3& point, const V3& normal) {
    Body b; b.shape = Shape::Plane; b.id = id; b.x = point; b.halfExtents = normal.normalized();
    b.invMass = 0.0; b.invInertiaBody = {0, 0, 0}; b.dynamic = false;
    return b;
}
// A static mesh placed by (point, q) (src/body.hpp:319-330).
inline Body makeMesh(int id, const V3& point, const Q& q, std::shared_ptr<const MeshData> data) {
    Body b; b.shape = Shape::Mesh; b.id = id; b.x = point; b.q = q; b.mesh = std::move(data);
    b.invMass = 0.0; b.invInertiaBody = {0, 0, 0}; b.dynamic = false;
    return b;
}



// A distance joint (src/world.hpp:20-26): stiffness 0 = a rigid link,
// stiffness > 0 = a spring-damper.
struct DistanceJoint {
    std::size_t a = 0, b = 0;
    V3 localA, localB;
    double rest = 0.0;
    double stiffness = 0.0;
    double damping = 0.0;
};

// A general joint (src/world.hpp:28-53), solved by sequential impulse after
// the rigid distance joints: Ball pins the anchors together, Hinge and Fixed
// add angular rows, Slider frees one translation axis.  The runtime fields
// (broken, appliedImpulse, motorAcc) come back after every step.
enum class JointType { Ball, Hinge, Fixed, Slider };

struct Joint {
    JointType type = JointType::Ball;
    std::size_t a = 0, b = 0;
    V3 localA, localB;
    V3 axisA{0, 1, 0}, axisB{0, 1, 0};
    Q refRel{1, 0, 0, 0};
    bool useMotor = false;
    double motorSpeed = 0.0;
    double maxMotorImpulse = 0.0;
    bool useLimit = false;
    double lower = 0.0, upper = 0.0;
    bool breakable = false;
    double breakImpulse = 0.0;
    bool broken = false;
    double appliedImpulse = 0.0;
    double motorAcc = 0.0;
};

// The reference's World (src/world.hpp:73-161), stepping on the GPU.
class World {
public:
    Box box;
    double dt = 1.0 / 120.0;
    double restitution = 1.0;
    double friction = 0.0;
    double contactBeta = 0.2;
    V3 gravity;
    enum class Combine { Average, Min, Max, Multiply, GeometricMean };
    Combine restitutionCombine = Combine::Max;
    Combine frictionCombine = Combine::GeometricMean;
    double effFriction(const Body& b) const { return b.friction < 0 ? friction : b.friction; }
    double effRestitution(const Body& b) const { return b.restitution < 0 ? restitution : b.restitution; }
    static bool layersCollide(const Body& a, const Body& b) {
        return (a.mask & (1u << (b.layer & 31u))) && (b.mask & (1u << (a.layer & 31u)));
    }
    static bool pairExcluded(const Body& a, std::uint32_t j) {
        for (std::uint32_t k : a.noCollide) if (k == j) return true;
        return false;
    }
    bool sleepEnabled = false;
    double sleepLinVel = 0.05, sleepAngVel = 0.05, sleepTime = 0.5, wakeVel = 0.5;
    bool openBoundary = false;
    bool warmStart = true;
    bool forceBruteForce = false;
    int solverIterations = 8;          // extension: the reference fixes 8
    std::vector<Body> bodies;
    std::vector<DistanceJoint> distanceJoints;
    std::vector<Joint> joints;

    World() = default;
    // Value semantics like the reference's World: a copy carries the bodies
    // and the warm-start cache, and then evolves independently.
    World(const World& o) { *this = o; }
    World& operator=(const World& o) {
        if (this == &o) return *this;
        box = o.box; dt = o.dt; restitution = o.restitution; friction = o.friction; contactBeta = o.contactBeta;
        gravity = o.gravity; restitutionCombine = o.restitutionCombine; frictionCombine = o.frictionCombine;
        sleepEnabled = o.sleepEnabled; sleepLinVel = o.sleepLinVel; sleepAngVel = o.sleepAngVel;
        sleepTime = o.sleepTime; wakeVel = o.wakeVel; openBoundary = o.openBoundary; warmStart = o.warmStart;
        forceBruteForce = o.forceBruteForce; solverIterations = o.solverIterations; bodies = o.bodies;
        distanceJoints = o.distanceJoints;
        joints = o.joints;
        session_.reset();
        cacheKeys_ = o.cacheKeys_;
        cacheImp_ = o.cacheImp_;
        if (o.session_) o.session_->exportWarmCache(cacheKeys_, cacheImp_);
        havePending_ = o.session_ != nullptr || o.havePending_;
        return *this;
    }

    void step() { stepMany(1); }
    void stepMany(int k) {
        if (!session_) {
            session_ = std::make_unique<ne_gpu::Session>(0);
            if (havePending_) session_->importWarmCache(cacheKeys_, cacheImp_);
            havePending_ = false;
        }
        session_->step(*this, k, solverIterations);
    }
    static void wake(Body& b) {
        if (!b.sleeping) return;
        b.invMass = b.invMassStore;
        b.invInertiaBody = b.invInertiaStore;
        b.sleeping = false;
        b.sleepTimer = 0.0;
    }
    void resetWarmCache() {
        if (session_) session_->resetWarmCache();
        cacheKeys_.clear();
        cacheImp_.clear();
        havePending_ = false;
    }

    V3 totalLinearMomentum() const {
        V3 p;
        for (const Body& b : bodies) if (b.invMass > 0.0) p += b.v * (1.0 / b.invMass);
        return p;
    }
    V3 totalAngularMomentum() const {
        V3 L;
        for (const Body& b : bodies) {
            if (b.invMass <= 0.0) continue;
            double m = 1.0 / b.invMass;
            L += b.x.cross(b.v * m);
            L += b.angularMomentum();
        }
        return L;
    }
    double totalKinetic() const {
        double e = 0.0;
        for (const Body& b : bodies) if (b.invMass > 0.0) e += b.kinetic();
        return e;
    }

private:
    std::unique_ptr<ne_gpu::Session> session_;      // device context, created on first step
    std::vector<std::uint64_t> cacheKeys_;       // warm cache carried by a copy until its first step
    std::vector<double> cacheImp_;
    bool havePending_ = false;
};

}  // namespace ne
