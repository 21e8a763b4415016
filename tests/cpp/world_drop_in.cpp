// world_drop_in.cpp -- the C++ drop-in `ne::World` (include/ne_world.hpp) on the
// GPU, exercised the way the reference's own selftest drives its World
// (src/selftest.cpp).  Built and run by tests/test_cpp_drop_in.py.
//
//   world_drop_in checks           restated selftest scenarios, exit code = failures
//   world_drop_in scene IN OUT K   steps a flat-array scene K times, writes x v q w
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <vector>

#include "ne_world.hpp"

using namespace ne;

static int g_fail = 0;
static void check(bool ok, const char* name) {
    std::printf("  [%s] %s\n", ok ? "PASS" : "FAIL", name);
    if (!ok) ++g_fail;
}
static bool closeV(const V3& a, const V3& b, double tol) { return (a - b).norm() <= tol; }

// splitmix64: a portable generator for the scenes below
struct Rng {
    std::uint64_t s;
    explicit Rng(std::uint64_t seed) : s(seed) {}
    std::uint64_t next() {
        std::uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uni(double a, double b) { return a + (b - a) * ((next() >> 11) * (1.0 / 9007199254740992.0)); }
};

static void headon() {        // selftest.cpp:52-78
    World w;
    w.box.half = 100.0; w.box.periodic = false; w.restitution = 1.0; w.dt = 1e-3;
    Body a = makeSphere(0, V3{-2, 0, 0}, 1.0, 1.0), b = makeSphere(1, V3{2, 0, 0}, 1.0, 1.0);
    a.v = V3{1, 0, 0}; b.v = V3{-1, 0, 0};
    w.bodies = {a, b};
    V3 p0 = w.totalLinearMomentum(); double e0 = w.totalKinetic();
    w.stepMany(20000);
    check(closeV(p0, w.totalLinearMomentum(), 1e-9), "head-on: linear momentum conserved");
    check(std::fabs(e0 - w.totalKinetic()) <= 1e-6, "head-on: kinetic energy conserved (e=1)");
    check(w.bodies[0].v.x < 0 && w.bodies[1].v.x > 0, "head-on: equal masses exchange velocity sign");
}

static void transit() {       // selftest.cpp:84-107
    World w;
    w.box.half = 5.0; w.box.periodic = true; w.dt = 1e-2;
    Body a = makeSphere(0, V3{4.9, 0, 0}, 0.5, 1.0);
    a.v = V3{1, 0, 0};
    w.bodies = {a};
    V3 prev = w.bodies[0].x, unwrapped = prev;
    for (int s = 0; s < 300; ++s) {
        w.step();
        V3 cur = w.bodies[0].x;
        unwrapped += w.box.minImage(cur - prev);
        prev = cur;
    }
    check(closeV(w.bodies[0].v, V3{1, 0, 0}, 1e-12), "periodic: velocity unchanged crossing face");
    check(std::fabs(unwrapped.x - 7.9) < 1e-9, "periodic: unwrapped trajectory continuous");
    check(std::fabs(w.bodies[0].x.x) <= 5.0 + 1e-9, "periodic: stays in primary cell");
}

static World gas(bool periodic, bool brute) {   // selftest.cpp:166-181 shape, portable RNG
    World w;
    w.box.half = 8.0; w.box.periodic = periodic; w.restitution = 1.0; w.dt = 5e-3; w.forceBruteForce = brute;
    Rng r(999);
    for (int i = 0; i < 120; ++i) {
        Body b = makeSphere(i, V3{r.uni(-6, 6), r.uni(-6, 6), r.uni(-6, 6)}, 0.7, 1.0);
        b.v = V3{r.uni(-2, 2), r.uni(-2, 2), r.uni(-2, 2)};
        b.w = V3{r.uni(-2, 2), r.uni(-2, 2), r.uni(-2, 2)};
        w.bodies.push_back(b);
    }
    return w;
}

static bool same(const World& a, const World& b) {
    for (size_t i = 0; i < a.bodies.size(); ++i) {
        const Body &x = a.bodies[i], &y = b.bodies[i];
        if (x.x.x != y.x.x || x.x.y != y.x.y || x.x.z != y.x.z || x.v.x != y.v.x || x.v.y != y.v.y ||
            x.v.z != y.v.z || x.q.w != y.q.w || x.w.x != y.w.x)
            return false;
    }
    return true;
}

static void determinism_and_copies() {   // selftest.cpp:183-197, 744-774, 223-240
    World a = gas(true, false), b = gas(true, false);
    for (int s = 0; s < 200; ++s) { a.step(); b.step(); }
    check(same(a, b), "determinism: identical runs are bit-for-bit equal");
    World g = gas(true, false), br = gas(true, true);
    g.stepMany(300);
    br.stepMany(300);
    check(same(g, br), "broadphase: grid result is bit-identical to brute force");
    // a copy carries the warm cache and then evolves like the original
    World c = a;
    for (int s = 0; s < 100; ++s) { a.step(); c.step(); }
    check(same(a, c), "copy: a copied World evolves bit-identically to its source");
    World e = gas(true, false);
    V3 p0 = e.totalLinearMomentum(); double e0 = e.totalKinetic();
    e.stepMany(2000);
    check(closeV(e.totalLinearMomentum(), p0, 1e-6), "periodic gas: linear momentum conserved");
    check(std::fabs(e.totalKinetic() - e0) / e0 < 0.02, "periodic gas: kinetic energy conserved to 2%");
}

static void cross_boundary() {   // selftest.cpp:248-267
    auto run = [](bool periodic) {
        World w;
        w.dt = 1.0 / 240.0; w.restitution = 1.0; w.box.periodic = periodic; w.box.half = periodic ? 5.0 : 50.0;
        Body a = makeSphere(0, periodic ? V3{4.4, 0, 0} : V3{-0.6, 0, 0}, 0.5, 1.0);
        Body b = makeSphere(1, periodic ? V3{-4.4, 0, 0} : V3{0.6, 0, 0}, 0.5, 1.0);
        a.v = V3{1, 0, 0}; b.v = V3{-1, 0, 0};
        w.bodies = {a, b};
        w.stepMany(400);
        return w;
    };
    World fp = run(false), pr = run(true);
    check(std::fabs(pr.bodies[0].v.x - fp.bodies[0].v.x) < 1e-9 && std::fabs(pr.bodies[1].v.x - fp.bodies[1].v.x) < 1e-9,
          "periodic: cross-boundary collision equals free space");
    check(std::fabs(pr.bodies[0].v.x + 1.0) < 1e-6 && std::fabs(pr.bodies[1].v.x - 1.0) < 1e-6,
          "periodic: cross-boundary head-on exchanges velocity");
}

static void stacking() {   // selftest.cpp:663-697 on the World directly
    World w;
    w.box.half = 1e6; w.box.periodic = false; w.dt = 1.0 / 240.0; w.gravity = V3{0, -10, 0};
    w.restitution = 0.0; w.friction = 0.6;
    Body floor = makeBox(0, V3{0, -1, 0}, Q{1, 0, 0, 0}, V3{20, 1, 20}, 1.0);
    floor.invMass = 0.0; floor.invInertiaBody = V3{0, 0, 0}; floor.dynamic = false;
    w.bodies = {floor, makeBox(1, V3{0, 1.05, 0}, Q{1, 0, 0, 0}, V3{1, 1, 1}, 1.0),
                makeBox(2, V3{0, 3.10, 0}, Q{1, 0, 0, 0}, V3{1, 1, 1}, 1.0)};
    w.stepMany(4000);
    double yl = w.bodies[1].x.y, yh = w.bodies[2].x.y;
    check(yl > 0.85 && yl < 1.2, "stacking: lower box rests on floor (~y=1)");
    check(yh > 2.7 && yh < 3.3, "stacking: upper box rests on lower (~y=3)");
    check(w.bodies[2].v.norm() < 0.3, "stacking: stack is at rest");
}

static void joints() {   // selftest.cpp:779-823
    {   // rigid rod of length 3 from a static anchor; the bob starts level and swings
        World w;
        w.box.half = 1e6; w.box.periodic = false; w.dt = 1.0 / 240.0; w.gravity = V3{0, -10, 0};
        Body anchor = makeSphere(0, V3{0, 5, 0}, 0.1, 1.0);
        anchor.invMass = 0.0; anchor.invInertiaBody = V3{0, 0, 0}; anchor.dynamic = false;
        w.bodies = {anchor, makeSphere(1, V3{3, 5, 0}, 1.0, 1.0)};
        DistanceJoint j; j.a = 0; j.b = 1; j.rest = 3.0;
        w.distanceJoints = {j};
        double maxErr = 0.0, minY = 5.0;
        for (int s = 0; s < 2000; ++s) {
            w.step();
            maxErr = std::max(maxErr, std::fabs((w.bodies[1].x - w.bodies[0].x).norm() - 3.0));
            minY = std::min(minY, w.bodies[1].x.y);
        }
        check(maxErr < 0.15, "joint: rigid rod keeps pendulum length ~constant");
        check(minY < 3.0, "joint: pendulum actually swings down under gravity");
    }
    {   // spring k = 200, c = 20, rest 2: settles where k (dist - rest) = m g
        World w;
        w.box.half = 1e6; w.box.periodic = false; w.dt = 1.0 / 240.0; w.gravity = V3{0, -10, 0};
        Body anchor = makeSphere(0, V3{0, 5, 0}, 0.1, 1.0);
        anchor.invMass = 0.0; anchor.invInertiaBody = V3{0, 0, 0}; anchor.dynamic = false;
        w.bodies = {anchor, makeSphere(1, V3{0, 3, 0}, 1.0, 1.0)};
        DistanceJoint j; j.a = 0; j.b = 1; j.rest = 2.0; j.stiffness = 200.0; j.damping = 20.0;
        w.distanceJoints = {j};
        w.stepMany(6000);
        const double mass = 4.18879;   // density 1, r = 1
        const double len = (w.bodies[1].x - w.bodies[0].x).norm();
        check(std::fabs((len - 2.0) - mass * 10.0 / 200.0) < 0.05, "joint: spring settles at the gravity-balanced stretch");
        check(w.bodies[1].v.norm() < 0.1, "joint: damped spring comes to rest");
    }
}

// PhysicsWorld::addGeneralJoint (src/physics_world.hpp:408-423) on indices:
// anchors and axis in each body's frame, the current relative orientation.
static Joint makeJoint(const World& w, JointType type, std::size_t a, std::size_t b, const V3& anchor, const V3& axis) {
    const Body& ba = w.bodies[a];
    const Body& bb = w.bodies[b];
    Joint j;
    j.type = type;
    j.a = a; j.b = b;
    j.localA = ba.q.inverseRotate(anchor - ba.x);
    j.localB = bb.q.inverseRotate(anchor - bb.x);
    const V3 ax = axis.normalized();
    j.axisA = ba.q.inverseRotate(ax);
    j.axisB = bb.q.inverseRotate(ax);
    j.refRel = ba.q * bb.q.conjugate();
    return j;
}

static Body makeStaticBody(Body b) {   // PhysicsWorld::makeStatic (src/physics_world.hpp:112-120)
    b.invMass = 0.0; b.invInertiaBody = V3{0, 0, 0};
    b.v = V3{}; b.w = V3{};
    b.dynamic = false; b.kinematic = false;
    return b;
}

static World jointWorld(double gy) {
    World w;
    w.box.half = 1000.0; w.box.periodic = false; w.dt = 1.0 / 240.0; w.gravity = V3{0, gy, 0};
    return w;
}

static void general_joints() {   // selftest.cpp:1088-1175, World::joints on the GPU
    {   // (a) fixed joint beside a static anchor
        World w = jointWorld(-10.0);
        w.bodies = {makeStaticBody(makeBox(0, V3{0, 5, 0}, Q{1, 0, 0, 0}, V3{0.5, 0.5, 0.5}, 1.0)),
                    makeBox(1, V3{1, 5, 0}, Q{1, 0, 0, 0}, V3{0.5, 0.5, 0.5}, 1.0)};
        w.joints = {makeJoint(w, JointType::Fixed, 0, 1, (w.bodies[0].x + w.bodies[1].x) * 0.5, V3{0, 1, 0})};
        w.stepMany(600);
        const V3 p = w.bodies[1].x;
        check(std::fabs(p.x - 1.0) <= 0.08 && std::fabs(p.y - 5.0) <= 0.15 && std::fabs(p.z) <= 0.08,
              "fixed joint: holds the relative pose under gravity");
    }
    {   // (b) ball joint: a bob hangs below a static anchor
        World w = jointWorld(-10.0);
        w.bodies = {makeStaticBody(makeSphere(0, V3{0, 5, 0}, 0.2, 1.0)), makeSphere(1, V3{0, 3, 0}, 0.5, 1.0)};
        w.joints = {makeJoint(w, JointType::Ball, 0, 1, V3{0, 5, 0}, V3{0, 1, 0})};
        for (int s = 0; s < 800; ++s) w.step();   // compat mode: joints go up and come back every step
        const V3 p = w.bodies[1].x;
        check(std::fabs((p - V3{0, 5, 0}).norm() - 2.0) <= 0.05, "ball joint: anchor distance preserved");
        check(p.y < 5.0, "ball joint: hangs below the anchor");
    }
    {   // (c) hinge with a motor drives rotation about its axis only
        World w = jointWorld(0.0);
        w.bodies = {makeStaticBody(makeBox(0, V3{0, 0, 0}, Q{1, 0, 0, 0}, V3{0.2, 0.2, 0.2}, 1.0)),
                    makeBox(1, V3{1, 0, 0}, Q{1, 0, 0, 0}, V3{1, 0.1, 0.1}, 1.0)};
        Joint h = makeJoint(w, JointType::Hinge, 0, 1, V3{0, 0, 0}, V3{0, 1, 0});
        h.useMotor = true; h.motorSpeed = 2.0; h.maxMotorImpulse = 200.0;
        w.joints = {h};
        w.stepMany(240);
        check(std::fabs(w.bodies[1].w.y) > 1.0, "hinge motor: drives rotation about the axis");
        check(std::fabs(w.bodies[1].x.z) > 0.2, "hinge motor: arm swept around the axis");
        check(w.joints[0].motorAcc != 0.0, "hinge motor: the step's motor impulse comes back");
    }
    {   // (d) breakable: enough load severs the joint and the body falls
        World w = jointWorld(-30.0);
        w.bodies = {makeStaticBody(makeSphere(0, V3{0, 5, 0}, 0.2, 1.0)),
                    makeBox(1, V3{0, 3, 0}, Q{1, 0, 0, 0}, V3{1, 1, 1}, 50.0)};
        Joint j = makeJoint(w, JointType::Ball, 0, 1, V3{0, 5, 0}, V3{0, 1, 0});
        j.breakable = true; j.breakImpulse = 5.0;
        w.joints = {j};
        w.stepMany(300);
        check(w.joints[0].broken, "breakable joint: breaks under heavy load");
        check(w.bodies[1].x.y < 2.0, "breakable joint: body falls after breaking");
    }
    {   // (e) slider: free along its axis, locked across it
        World w = jointWorld(0.0);
        w.bodies = {makeStaticBody(makeBox(0, V3{0, 0, 0}, Q{1, 0, 0, 0}, V3{0.2, 0.2, 0.2}, 1.0)),
                    makeBox(1, V3{0, 0, 0}, Q{1, 0, 0, 0}, V3{0.5, 0.5, 0.5}, 1.0)};
        w.joints = {makeJoint(w, JointType::Slider, 0, 1, V3{0, 0, 0}, V3{1, 0, 0})};
        w.bodies[1].v = V3{2, 3, 1};
        w.stepMany(240);
        const V3 p = w.bodies[1].x;
        check(std::fabs(p.x) > 0.5, "slider: translates along the axis");
        check(std::fabs(p.y) < 0.05 && std::fabs(p.z) < 0.05, "slider: locked perpendicular to the axis");
    }
    {   // (f) a bridge of hinges stays planar
        World w = jointWorld(-10.0);
        w.bodies.push_back(makeStaticBody(makeBox(0, V3{-5, 6, 0}, Q{1, 0, 0, 0}, V3{1, 1, 2}, 1.0)));
        std::vector<Joint> js;
        for (int i = 0; i < 8; ++i) {
            const double x = -4.5 + i * 1.0;
            w.bodies.push_back(makeBox(i + 1, V3{x, 6, 0}, Q{1, 0, 0, 0}, V3{0.5, 0.12, 2}, 1.0));
            js.push_back(makeJoint(w, JointType::Hinge, i, i + 1, V3{x - 0.5, 6, 0}, V3{0, 0, 1}));
        }
        w.bodies.push_back(makeStaticBody(makeBox(9, V3{3, 6, 0}, Q{1, 0, 0, 0}, V3{1, 1, 2}, 1.0)));
        js.push_back(makeJoint(w, JointType::Hinge, 8, 9, V3{2.5, 6, 0}, V3{0, 0, 1}));
        w.joints = js;
        w.stepMany(1200);
        double maxTilt = 0.0;
        for (int i = 1; i <= 8; ++i) {
            const V3 lz = w.bodies[i].q.rotate(V3{0, 0, 1});
            maxTilt = std::max(maxTilt, std::acos(std::min(1.0, std::fabs(lz.z))) * 180.0 / 3.14159265358979323846);
        }
        check(maxTilt < 5.0, "hinge chain: a bridge of hinges stays planar (no twisting)");
    }
}

static World floorWorld() {
    World w;
    w.box.half = 1e6; w.box.periodic = false; w.dt = 1.0 / 240.0; w.gravity = V3{0, -10, 0};
    w.restitution = 0.0;
    return w;
}

static Body staticBox(int id, const V3& x, const V3& he) {
    Body b = makeBox(id, x, Q{1, 0, 0, 0}, he, 1.0);
    b.invMass = 0.0; b.invInertiaBody = V3{0, 0, 0}; b.dynamic = false;
    return b;
}

static void sleeping() {   // selftest.cpp:703-738, compat-mode step() each step
    World w = floorWorld();
    w.sleepEnabled = true;
    w.bodies = {staticBox(0, V3{0, -1, 0}, V3{20, 1, 20}), makeSphere(1, V3{0, 3.0, 0}, 1.0, 1.0)};
    for (int s = 0; s < 2000; ++s) w.step();
    check(w.bodies[1].sleeping, "sleeping: a settled body falls asleep");
    check(w.bodies[1].invMass == 0.0 && w.bodies[1].invMassStore > 0.0, "sleeping: its mass is stored while asleep");
    const V3 rest = w.bodies[1].x;
    for (int s = 0; s < 2000; ++s) w.step();
    check((w.bodies[1].x - rest).norm() < 1e-6, "sleeping: an asleep body does not move");
    World::wake(w.bodies[1]);                       // PhysicsWorld::applyImpulse
    w.bodies[1].v += V3{5, 0, 0} * w.bodies[1].invMass;
    check(!w.bodies[1].sleeping, "sleeping: the impulse wakes the body");
    w.step();
    check(w.bodies[1].v.x > 0.1, "sleeping: the woken body responds to the impulse");

    World w2 = floorWorld();
    w2.sleepEnabled = true;
    w2.bodies = {staticBox(0, V3{0, -1, 0}, V3{20, 1, 20}), makeBox(1, V3{0, 1.0, 0}, Q{1, 0, 0, 0}, V3{1, 1, 1}, 1.0)};
    w2.stepMany(2000);
    check(w2.bodies[1].sleeping, "sleeping: the lower box sleeps");
    w2.bodies.push_back(makeBox(2, V3{0, 6.0, 0}, Q{1, 0, 0, 0}, V3{1, 1, 1}, 1.0));
    bool woke = false;
    for (int s = 0; s < 2000 && !woke; ++s) { w2.step(); woke = !w2.bodies[1].sleeping; }
    check(woke, "sleeping: a falling box wakes the sleeper it lands on");
}

static void meshes() {   // selftest.cpp:1300-1326, compat-mode step() each step
    World w = floorWorld();
    w.box.half = 1000.0; w.friction = 0.5;
    auto quad = std::make_shared<MeshData>();
    quad->verts = {{-10, 0, -10}, {10, 0, -10}, {10, 0, 10}, {-10, 0, 10}};
    quad->tris = {{0, 1, 2}, {0, 2, 3}};
    quad->build();
    w.bodies = {makeMesh(0, V3{0, 0, 0}, Q{1, 0, 0, 0}, quad), makeSphere(1, V3{1, 5, 1}, 0.5, 1.0),
                makeBox(2, V3{-3, 5, -2}, Q{1, 0, 0, 0}, V3{0.5, 0.5, 0.5}, 1.0)};
    for (int s = 0; s < 1500; ++s) w.step();
    check(std::fabs(w.bodies[1].x.y - 0.5) <= 0.05, "mesh: sphere rests on the mesh floor");
    check(w.bodies[1].v.norm() < 0.15, "mesh: sphere comes to rest");
    check(std::fabs(w.bodies[2].x.y - 0.5) <= 0.08, "mesh: box rests on the mesh floor");

    // a flat 5 x 5 heightfield (PhysicsWorld::addHeightfield, physics_world.hpp:96-109)
    World h = floorWorld();
    h.box.half = 1000.0; h.friction = 0.5;
    auto hf = std::make_shared<MeshData>();
    for (int r = 0; r < 5; ++r)
        for (int c = 0; c < 5; ++c) hf->verts.push_back(V3{-4.0 + c * 2.0, 0.0, -4.0 + r * 2.0});
    for (int r = 0; r + 1 < 5; ++r)
        for (int c = 0; c + 1 < 5; ++c) {
            const int i00 = r * 5 + c, i01 = i00 + 1, i10 = i00 + 5, i11 = i10 + 1;
            hf->tris.push_back({i00, i10, i11});
            hf->tris.push_back({i00, i11, i01});
        }
    h.bodies = {makeMesh(0, V3{0, 0, 0}, Q{1, 0, 0, 0}, hf), makeSphere(1, V3{0, 5, 0}, 0.5, 1.0)};
    h.stepMany(1500);
    check(std::fabs(h.bodies[1].x.y - 0.5) <= 0.06, "heightfield: sphere rests on the terrain");
}

// scene IN OUT K: IN holds 24 doubles of world parameters (n first), then
// per body 26 doubles (x3 v3 q4 w3 invMass invI3 shape radius halfHeight he3
// friction restitution hull-vertex count) and its hull vertices, then the
// meshes; OUT receives x v q w, the sleeping flag, sleep timer and invMass.
static int run_scene(const char* in, const char* out, int k) {
    std::ifstream f(in, std::ios::binary);
    if (!f) return 2;
    double hdr[24];
    f.read(reinterpret_cast<char*>(hdr), sizeof hdr);
    const size_t n = (size_t)hdr[0];
    World w;
    w.box.half = hdr[1]; w.box.halfXYZ = V3{hdr[1], hdr[2], hdr[3]}; w.box.periodic = hdr[4] != 0;
    w.openBoundary = hdr[5] != 0; w.dt = hdr[6]; w.gravity = V3{hdr[7], hdr[8], hdr[9]};
    w.restitution = hdr[10]; w.friction = hdr[11]; w.contactBeta = hdr[12];
    w.restitutionCombine = static_cast<World::Combine>((int)hdr[13]);
    w.frictionCombine = static_cast<World::Combine>((int)hdr[14]);
    w.warmStart = hdr[15] != 0; w.solverIterations = (int)hdr[16];
    w.sleepEnabled = hdr[17] != 0;
    const int nmesh = (int)hdr[18];
    std::vector<double> row(26);
    for (size_t i = 0; i < n; ++i) {
        f.read(reinterpret_cast<char*>(row.data()), sizeof(double) * row.size());
        Body b;
        b.x = V3{row[0], row[1], row[2]}; b.v = V3{row[3], row[4], row[5]};
        b.q = Q{row[6], row[7], row[8], row[9]}; b.w = V3{row[10], row[11], row[12]};
        b.invMass = row[13]; b.invInertiaBody = V3{row[14], row[15], row[16]};
        b.shape = static_cast<Shape>((int)row[17]); b.radius = row[18]; b.halfHeight = row[19];
        b.halfExtents = V3{row[20], row[21], row[22]}; b.friction = row[23]; b.restitution = row[24];
        int nv = (int)row[25];
        for (int t = 0; t < nv; ++t) {
            double v[3];
            f.read(reinterpret_cast<char*>(v), sizeof v);
            b.vertices.push_back(V3{v[0], v[1], v[2]});
        }
        if (b.shape == Shape::Mesh || b.invMass == 0.0) b.dynamic = false;
        w.bodies.push_back(b);
    }
    for (int m = 0; m < nmesh; ++m) {   // body, nverts, ntris, then verts and tris as doubles
        double mh[3];
        f.read(reinterpret_cast<char*>(mh), sizeof mh);
        auto md = std::make_shared<MeshData>();
        for (int t = 0; t < (int)mh[1]; ++t) {
            double v[3];
            f.read(reinterpret_cast<char*>(v), sizeof v);
            md->verts.push_back(V3{v[0], v[1], v[2]});
        }
        for (int t = 0; t < (int)mh[2]; ++t) {
            double v[3];
            f.read(reinterpret_cast<char*>(v), sizeof v);
            md->tris.push_back({(int)v[0], (int)v[1], (int)v[2]});
        }
        w.bodies[(size_t)mh[0]].mesh = md;
    }
    for (int s = 0; s < k; ++s) w.step();   // compat mode: upload, step, download each time
    std::ofstream o(out, std::ios::binary);
    for (const Body& b : w.bodies) {
        double r[16] = {b.x.x, b.x.y, b.x.z, b.v.x, b.v.y, b.v.z, b.q.w, b.q.x, b.q.y, b.q.z, b.w.x, b.w.y, b.w.z,
                        b.sleeping ? 1.0 : 0.0, b.sleepTimer, b.invMass};
        o.write(reinterpret_cast<const char*>(r), sizeof r);
    }
    return 0;
}

int main(int argc, char** argv) {
    try {
        if (argc >= 5 && std::strcmp(argv[1], "scene") == 0) return run_scene(argv[2], argv[3], std::atoi(argv[4]));
        std::printf("=== ne::World drop-in on the GPU ===\n");
        headon();
        transit();
        determinism_and_copies();
        cross_boundary();
        stacking();
        joints();
        general_joints();
        sleeping();
        meshes();
        std::printf("failures: %d\n", g_fail);
        return g_fail == 0 ? 0 : 1;
    } catch (const std::exception& e) {
        std::printf("exception: %s\n", e.what());
        return 3;
    }
}
