"""Small test scenes shared by the parity tests and the fixture generator."""
import numpy as np

from nativeengine_amd import _abi
from nativeengine_amd import scene as S


def features_scene(seed=0):
    """Exercises every per-body switch on the step's path: layers/masks,
    no-collide pairs, sensors, a kinematic pusher, damping, gravity scale,
    per-body friction/restitution with non-default combine modes, forces and
    torques, cylinders and capsules, a plane."""
    rng = np.random.default_rng(seed)
    bodies = []
    k = 0
    for a in range(4):
        for b in range(4):
            for c in range(3):
                x = (-2.4 + 1.6 * a + rng.uniform(-0.1, 0.1), -1.5 + 1.2 * c, -2.4 + 1.6 * b + rng.uniform(-0.1, 0.1))
                q = tuple(S.random_unit_quaternions(rng, 1)[0])
                kind = k % 5
                if kind == 0:
                    bd = S.make_sphere(k, x, rng.uniform(0.4, 0.7), 1.0)
                elif kind == 1:
                    bd = S.make_box(k, x, q, tuple(rng.uniform(0.3, 0.6, 3)), 1.0)
                elif kind == 2:
                    bd = S.make_capsule(k, x, q, 0.25, 0.35, 1.0)
                elif kind == 3:
                    bd = S.make_cylinder(k, x, q, 0.4, 0.8, 1.0)
                else:
                    bd = S.make_convex(k, x, q, S.hull12(rng, 0.55), 1.0)
                bd.v = tuple(rng.normal(0, 1.5, 3))
                bd.w = tuple(rng.normal(0, 1.0, 3))
                if k % 7 == 0:
                    bd.friction = float(rng.uniform(0, 1))
                if k % 5 == 1:
                    bd.restitution = float(rng.uniform(0, 1))
                if k % 11 == 3:
                    bd.linear_damping = 0.3
                    bd.angular_damping = 0.2
                if k % 13 == 4:
                    bd.gravity_scale = -0.5
                if k % 9 == 2:
                    bd.layer = 3
                    bd.mask = 0xFFFFFFFF & ~(1 << 1)
                if k % 17 == 5:
                    bd.layer = 1
                if k % 19 == 6:
                    bd.flags |= _abi.FLAG_SENSOR
                if k % 10 == 7:
                    bd.force = tuple(rng.normal(0, 3.0, 3))
                    bd.torque = tuple(rng.normal(0, 1.0, 3))
                bodies.append(bd)
                k += 1
    # no-collide pairs (both sides list each other)
    for (i, j) in [(0, 1), (3, 4), (10, 22)]:
        bodies[i].no_collide = list(bodies[i].no_collide) + [j]
        bodies[j].no_collide = list(bodies[j].no_collide) + [i]
    # kinematic pusher box: infinite mass, real velocity
    kin = S.make_box(len(bodies), (0.0, -2.6, 0.0), (1.0, 0.0, 0.0, 0.0), (1.2, 0.3, 1.2), 1.0)
    kin.inv_mass = 0.0
    kin.inv_inertia = (0.0, 0.0, 0.0)
    kin.flags = _abi.FLAG_KINEMATIC
    kin.v = (0.0, 0.8, 0.0)
    kin.w = (0.0, 0.3, 0.0)
    bodies.append(kin)
    # static floor plane
    bodies.append(S.make_plane(len(bodies), (0.0, -3.4, 0.0), (0.05, 1.0, -0.02)))
    s = S.Scene.from_bodies(bodies)
    s.params = S.Params(half=(4.0, 4.0, 4.0), periodic=False, dt=1.0 / 90.0, gravity=(0.0, -6.0, 0.0),
                        restitution=0.4, friction=0.3, restitution_combine=_abi.COMBINE_AVERAGE,
                        friction_combine=_abi.COMBINE_MIN, warm_start=True)
    return s


def fixture_scenes():
    """name -> (scene, steps, diagnostics every k steps or 0) for the golden
    fixtures."""
    out = {}
    out["gas256"] = (S.sphere_gas(256, seed=3, sigma=2.0), 30, 0)
    g = S.sphere_gas(300, L=11.0, seed=8, sigma=2.5)
    g.w[:] = np.random.default_rng(1).normal(0, 1, g.w.shape)
    out["densegas300"] = (g, 25, 0)
    out["pile120"] = (S.mixed_pile(120, half=(5.0, 5.0, 5.0), seed=1), 40, 0)
    p = S.mixed_pile(80, half=(4.0, 4.0, 4.0), seed=4, ground=False, periodic=True)
    p.params.gravity = (0.0, 0.0, 0.0)
    rng = np.random.default_rng(3)
    p.v[:] = rng.normal(0, 2.0, p.v.shape)
    p.w[:] = rng.normal(0, 1.0, p.w.shape)
    out["ppile80"] = (p, 30, 0)
    out["features"] = (features_scene(0), 40, 0)
    f2 = features_scene(1)
    f2.params.open_boundary = True
    f2.params.warm_start = False
    out["features_open_cold"] = (f2, 30, 0)
    out["gas4096"] = (elastic_gas_4096(), 1000, 100)
    out["ccd"] = (ccd_scene(0, periodic=False), 60, 0)
    out["ccd_periodic"] = (ccd_scene(1, periodic=True), 60, 0)
    for name, (js, steps) in joint_selftest_scenes().items():
        out[f"joints_{name}"] = (js, steps, 0)
    out["joints_mix"] = (joint_mix_scene(0), 300, 0)
    out["joints_mix_periodic"] = (joint_mix_scene(1, periodic=True), 300, 0)
    out["joints_limits"] = (joint_mix_scene(2, limits=True), 120, 0)
    return out


# Fixture scenes whose hinge limits go through atan2: the device library's
# atan2 can differ from the host's in the last bit, so the GPU matches these
# to a tolerance instead of bit for bit.
ATAN2_SCENES = {"joints_limits"}


def ccd_selftest_scene(ccd: bool):
    """selftest.cpp:1241-1257: a sphere of radius 0.2 fired at 600 units/s
    (10 units a step) at a static wall 0.1 thick."""
    wall = S.make_box(0, (0.0, 0.0, 0.0), (1.0, 0.0, 0.0, 0.0), (0.05, 5.0, 5.0), 1.0).make_static()
    ball = S.make_sphere(1, (-5.0, 0.0, 0.0), 0.2, 1.0)
    ball.v = (600.0, 0.0, 0.0)
    if ccd:
        ball.flags |= _abi.FLAG_CCD
    s = S.Scene.from_bodies([wall, ball])
    s.params = S.Params(half=(5000.0, 5000.0, 5000.0), periodic=False, dt=1.0 / 60.0, gravity=(0.0, -9.81, 0.0),
                        restitution=0.0)
    return s


def ccd_scene(seed=0, periodic=False):
    """Fast CCD bodies (spheres, boxes, capsules, cylinders, a hull) fired at
    thin static plates, a static cylinder and capsule, and each other (so
    later CCD bodies meet earlier ones' stopped positions), among slow CCD
    bodies that do not sweep, a sleeping CCD body, a sensor plate and a
    plate on a masked layer (both passed through), ordinary dynamic bodies,
    a ground box; periodic: obstacles across the boundary (minimum images)."""
    rng = np.random.default_rng(seed)
    b = []
    h = 8.0
    plates = [((3.0, 0.0, 0.0), (0.04, 2.5, 2.5)), ((-3.0, 0.0, 0.0), (0.04, 2.5, 2.5)),
              ((0.0, 0.0, 3.0), (2.5, 2.5, 0.04)), ((0.0, 3.5, 0.0), (2.5, 0.04, 2.5))]
    if periodic:
        plates += [((h - 0.05, 0.0, 0.0), (0.04, 2.0, 2.0))]
    for x, he in plates:
        b.append(S.make_box(len(b), x, (1.0, 0.0, 0.0, 0.0), he, 1.0).make_static())
    b.append(S.make_cylinder(len(b), (0.0, 0.0, -3.0), (1.0, 0.0, 0.0, 0.0), 0.6, 3.0, 1.0).make_static())
    r = 2.0 ** -0.5
    b.append(S.make_capsule(len(b), (-3.0, 0.0, -3.0), (r, r, 0.0, 0.0), 0.3, 1.0, 1.0).make_static())
    sens = S.make_box(len(b), (0.0, 0.0, -6.0), (1.0, 0.0, 0.0, 0.0), (2.5, 2.5, 0.04), 1.0).make_static()
    sens.flags |= _abi.FLAG_SENSOR
    b.append(sens)
    lay = S.make_box(len(b), (6.0, 0.0, 0.0), (1.0, 0.0, 0.0, 0.0), (0.04, 2.5, 2.5), 1.0).make_static()
    lay.layer = 6
    b.append(lay)
    if not periodic:
        b.append(S.make_box(len(b), (0.0, -7.5, 0.0), (1.0, 0.0, 0.0, 0.0), (7.5, 0.3, 7.5), 1.0).make_static())
    for k in range(24):
        x = tuple(rng.uniform(-1.5, 1.5, 3))
        kind = k % 5
        q = tuple(S.random_unit_quaternions(rng, 1)[0])
        if kind == 0:
            bd = S.make_sphere(len(b), x, float(rng.uniform(0.1, 0.25)), 1.0)
        elif kind == 1:
            bd = S.make_box(len(b), x, q, tuple(rng.uniform(0.08, 0.2, 3)), 1.0)
        elif kind == 2:
            bd = S.make_capsule(len(b), x, q, 0.1, 0.15, 1.0)
        elif kind == 3:
            bd = S.make_cylinder(len(b), x, q, 0.12, 0.3, 1.0)
        else:
            bd = S.make_convex(len(b), x, q, S.hull12(rng, 0.18), 1.0)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        speed = 700.0 if k % 6 else 2.0   # every sixth is too slow to sweep
        bd.v = tuple(d * speed)
        bd.w = tuple(rng.normal(0, 3.0, 3))
        bd.flags |= _abi.FLAG_CCD
        if k == 7:
            bd.mask = 0xFFFFFFFF & ~(1 << 6)   # passes the layer-6 plate
        if k == 11:
            bd.flags |= _abi.FLAG_SLEEPING
            bd.v, bd.w = (0.0, 0.0, 0.0), (0.0, 0.0, 0.0)
        b.append(bd)
    for k in range(10):   # ordinary bodies in the way
        b.append(S.make_sphere(len(b), tuple(rng.uniform(-2.5, 2.5, 3)), 0.3, 1.0))
    s = S.Scene.from_bodies(b)
    s.params = S.Params(half=(h, h, h), periodic=periodic, dt=1.0 / 60.0,
                        gravity=(0.0, 0.0, 0.0) if periodic else (0.0, -9.81, 0.0), restitution=0.2, friction=0.4)
    return s


def elastic_gas_4096():
    """The plain-method workload as specified: 4096 spheres of radius 0.5 and
    mass 1 uniform in a periodic cube of side 40, rejection-sampled to no
    overlap (seed 0), velocities N(0,1) per axis with the net momentum
    removed, no gravity, restitution 1, friction 0, dt 1/60, 8 iterations."""
    s = S.sphere_gas(4096, L=40.0, seed=0, sigma=1.0)
    assert s.params.dt == 1.0 / 60.0 and s.params.restitution == 1.0 and s.params.friction == 0.0
    return s


def snapshot_scene():
    """The settling pile with island sleeping that the NES1 snapshot fixture
    is written from."""
    return S.sleeping_pile(80, seed=1, periodic=False)


def epa_blowup_pair():
    """A sphere resting on a box face (depth ~0.015) whose EPA in the
    reference never converges: repeated support points leave zero-area faces
    closest to the origin, so all 64 iterations run and the face list grows
    to 99,198 faces (horizon 30,066).  Poses are exact (hex) and the bodies
    are at rest, so the first step's integrate leaves the positions as given."""
    h = float.fromhex
    s = S.Scene(2)
    s.shape[:] = [0, 2]
    s.radius[:] = [h("0x1.ef4dd7f0435fbp-2"), 1.0]
    s.half_extents[1] = [h("0x1.276c43f7bd600p-1"), h("0x1.e7c9c1ac97badp-2"), h("0x1.992e5138addc7p-2")]
    s.x[0] = [h("0x1.c394008719fd0p+1"), h("-0x1.1c5034f14a186p+2"), h("0x1.a7c4b1f09b5b7p-1")]
    s.x[1] = [h("0x1.1e97196826b1bp+2"), h("-0x1.18f2cbc959722p+2"), h("0x1.202bb87e0f37bp+0")]
    s.q[1] = [h("0x1.2805d03e0f562p-1"), h("-0x1.d44633bcb72e1p-5"), h("0x1.f5c9a1d01dbd3p-2"), h("0x1.4cbb46f1f0bb6p-1")]
    s.inv_mass[:] = [h("0x1.0e06a5864bc52p+1"), h("0x1.235c03dc48146p+0")]
    s.inv_inertia[0] = [h("0x1.68abde1878be3p+4")] * 3
    s.inv_inertia[1] = [h("0x1.1aa058a7dba55p+3"), h("0x1.bb9ae8cb9b100p+2"), h("0x1.8653004bf7bddp+2")]
    half = h("0x1.2e9a56266c037p+2")
    s.params = S.Params(half=(half, half, half), periodic=True, dt=1.0 / 60.0, gravity=(0.0, 0.0, 0.0),
                        restitution=h("0x1.869dcfc14a600p-10"), friction=h("0x1.a750454fa41b2p-2"),
                        solver_iterations=8)
    return s


def event_scenes():
    """name -> (scene, steps) for the facade-event fixtures: the feature scene
    (sensors give trigger events; layers, no-collide lists, a plane, a
    kinematic body, cylinders) and a periodic pile with island sleeping and
    a sensor box in it (sleeping bodies, whose stored masses the facade's
    filter does not see)."""
    out = {"features": (features_scene(2), 60)}
    p = S.sleeping_pile(60, seed=3, periodic=True)
    k = int(np.argmax(p.shape == _abi.SHAPE_BOX))
    p.flags[k] |= np.uint32(_abi.FLAG_SENSOR)
    out["sleeping_sensor"] = (p, 240)
    return out


def special_query_scene():
    """Static bodies placed for the ray tests' special cases: a capsule whose
    side exit and cap entry disagree (the reference tries the caps only when
    the side misses within the current best distance), with a sphere of
    lower index between the two on the same ray; a cylinder met along its
    axis; axis-aligned boxes met along a face plane and from inside; nested
    spheres around an origin (distance 0); a rotated box; a hull and a plane
    (never hit by rays, found by overlaps); a sensor and a masked layer."""
    b = []
    b.append(S.make_sphere(0, (0.26, 0.94, 0.0), 0.12, 1.0))                             # 0: sphere on the capsule ray
    b.append(S.make_capsule(1, (0.0, 0.0, 0.0), (1.0, 0.0, 0.0, 0.0), 0.5, 1.0, 1.0))    # 1: capsule, axis y
    b.append(S.make_cylinder(2, (4.0, 0.0, 0.0), (1.0, 0.0, 0.0, 0.0), 0.6, 1.6, 1.0))   # 2: cylinder, axis y
    b.append(S.make_box(3, (-4.0, 0.0, 0.0), (1.0, 0.0, 0.0, 0.0), (0.5, 0.7, 0.9), 1.0))  # 3: axis-aligned box
    b.append(S.make_sphere(4, (0.0, 0.0, 4.0), 0.8, 1.0))                                # 4, 5: nested spheres
    b.append(S.make_sphere(5, (0.1, 0.0, 4.0), 0.5, 1.0))
    r = 2.0 ** -0.5
    b.append(S.make_box(6, (0.0, 0.0, -4.0), (r, 0.0, r, 0.0), (0.4, 0.4, 0.8), 1.0))      # 6: box turned 90 deg about y
    b.append(S.make_convex(7, (4.0, 0.0, 4.0), (1.0, 0.0, 0.0, 0.0), S.hull12(np.random.default_rng(7), 0.6), 1.0))
    b.append(S.make_plane(8, (0.0, -3.0, 0.0), (0.0, 1.0, 0.0)))
    sen = S.make_sphere(9, (-4.0, 0.0, 4.0), 0.7, 1.0)
    sen.flags |= _abi.FLAG_SENSOR
    b.append(sen)
    lay = S.make_capsule(10, (-4.0, 0.0, -4.0), (r, r, 0.0, 0.0), 0.4, 0.6, 1.0)          # 10: capsule along z, layer 5
    lay.layer = 5
    b.append(lay)
    b[0].layer = 4
    for x in b:
        x.make_static()
    s = S.Scene.from_bodies(b)
    s.params = S.Params(half=(8.0, 8.0, 8.0), periodic=False, dt=1.0 / 60.0, gravity=(0.0, 0.0, 0.0))
    return s


def special_rays():
    """(origin, direction, max_dist, mask) rows for special_query_scene."""
    rows = []
    d = np.array([0.3, -1.0, 0.0])
    for md in (10.0, 1.2, 0.5, 0.2):   # capsule: side exit vs cap entry, and the sphere between them
        rows.append(((0.0, 1.8, 0.0), tuple(d), md, 0xFFFFFFFF))
    for md in (10.0, 1.0):               # the sphere's layer masked out: the side exit, then the cap
        rows.append(((0.0, 1.8, 0.0), tuple(d), md, 0xFFFFFFFF & ~(1 << 4)))
    rows.append(((4.1, 5.0, 0.0), (0.0, -1.0, 0.0), 20.0, 0xFFFFFFFF))     # cylinder along its axis: the cap
    rows.append(((4.1, 0.0, 0.0), (0.0, 1.0, 0.0), 20.0, 0xFFFFFFFF))      # from inside the cylinder
    rows.append(((4.0, 0.5, 5.0), (0.0, 0.0, -1.0), 20.0, 0xFFFFFFFF))     # cylinder side, head on
    rows.append(((-6.0, 0.7, 0.0), (1.0, 0.0, 0.0), 20.0, 0xFFFFFFFF))     # box: along its top face plane
    rows.append(((-6.0, 0.2, 0.3), (1.0, 0.0, 0.0), 20.0, 0xFFFFFFFF))     # box: head on
    rows.append(((-4.0, 0.1, 0.2), (0.0, 0.0, 1.0), 20.0, 0xFFFFFFFF))     # box: from inside (no hit)
    rows.append(((0.05, 0.0, 4.0), (1.0, 0.0, 0.0), 20.0, 0xFFFFFFFF))     # inside both spheres: distance 0
    rows.append(((0.0, 0.0, 2.0), (0.0, 0.0, 1.0), 20.0, 0xFFFFFFFF))      # nested spheres from outside
    rows.append(((0.0, 3.0, -4.0), (0.0, -1.0, 0.0), 20.0, 0xFFFFFFFF))    # turned box from above
    rows.append(((2.0, 0.0, -4.0), (-1.0, 0.0, 0.0), 20.0, 0xFFFFFFFF))    # turned box side
    rows.append(((4.0, 3.0, 4.0), (0.0, -1.0, 0.0), 20.0, 0xFFFFFFFF))     # hull, then the plane: no ray hit
    rows.append(((-4.0, 3.0, 4.0), (0.0, -1.0, 0.0), 20.0, 0xFFFFFFFF))    # sensor sphere: hit
    rows.append(((-4.0, 3.0, -4.0), (0.0, -1.0, 0.0), 20.0, 0xFFFFFFFF))   # layer-5 capsule
    rows.append(((-4.0, 3.0, -4.0), (0.0, -1.0, 0.0), 20.0, 0xFFFFFFFF & ~(1 << 5)))
    rows.append(((-4.0, 0.0, -7.0), (0.0, 0.0, 1.0), 20.0, 0xFFFFFFFF))    # that capsule along its axis
    rows.append(((0.0, 6.0, 0.0), (0.0, -1.0, 0.0), 20.0, 0xFFFFFFFF))     # capsule along its axis
    rows.append(((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 0.0, 0xFFFFFFFF))       # zero search distance
    o, dd, m, k = zip(*rows)
    return np.array(o, float), np.array(dd, float), np.array(m, float), np.array(k, np.uint32)


def query_scenes():
    """name -> (scene, steps before the queries) for the scene-query
    fixtures: the feature scene (planes, cylinders, capsules, hulls,
    sensors, layers) after 30 steps, a periodic pile (minimum images) after
    20, and the special-case scene."""
    p = S.mixed_pile(400, half=(6.0, 6.0, 6.0), seed=5, periodic=True, ground=False)
    p.params.gravity = (0.0, 0.0, 0.0)
    rng = np.random.default_rng(5)
    p.v[:] = rng.normal(0, 2.0, p.v.shape)
    p.w[:] = rng.normal(0, 1.0, p.w.shape)
    sel = np.arange(p.n)[::17]
    p.flags[sel] |= np.uint32(_abi.FLAG_SENSOR)
    p.layer[np.arange(p.n)[3::11]] = 2
    return {"features": (features_scene(3), 30), "pile_periodic": (p, 20), "special": (special_query_scene(), 0)}


def query_inputs(s, seed, nrays=1500, nshapes=200):
    """Random rays and query shapes over a scene's box: origins anywhere in
    it (some inside bodies), a tenth of the directions along an axis, search
    distances up to the box's width, a fifth of the masks random; casts of
    radius up to 0.5; overlap spheres and boxes (random orientation), and
    bodies of the scene's own shapes at random places for overlapContacts."""
    from nativeengine_amd import queries as Q
    rng = np.random.default_rng(seed)
    h = np.asarray(s.params.half, float)
    org = rng.uniform(-h, h, (nrays, 3))
    d = rng.normal(size=(nrays, 3))
    ax = rng.integers(0, 3, nrays // 10)
    d[: nrays // 10] = 0.0
    d[np.arange(nrays // 10), ax] = rng.choice([-1.0, 1.0], nrays // 10)
    md = rng.uniform(0.2, 2.0 * h.max(), nrays)
    mask = np.full(nrays, 0xFFFFFFFF, np.uint32)
    pick = rng.random(nrays) < 0.2
    mask[pick] = rng.integers(0, 2 ** 32, int(pick.sum()), dtype=np.uint64).astype(np.uint32)
    rad = rng.uniform(0.0, 0.5, nrays)
    shapes0 = []
    for _ in range(nshapes):
        c = rng.uniform(-h, h)
        if rng.random() < 0.5:
            shapes0.append(Q.sphere(c, rng.uniform(0.05, 1.5)))
        else:
            shapes0.append(Q.box(c, S.random_unit_quaternions(rng, 1)[0], rng.uniform(0.05, 1.0, 3)))
    shapes1 = [Q.body_shape(s, int(i), x=rng.uniform(-h, h), q=S.random_unit_quaternions(rng, 1)[0])
               for i in rng.integers(0, s.n, nshapes)]
    mask0 = mask[:nshapes].copy()
    return dict(org=org, dir=d, max_dist=md, mask=mask, rad=rad, shapes0=shapes0, shapes1=shapes1, mask0=mask0)


def _static(b):
    return b.make_static()


def joint_selftest_scenes():
    """name -> (scene, steps) for the reference's general-joint checks
    (selftest.cpp:1088-1179), built as the facade builds them: a fixed joint
    holding a box beside a static anchor, a ball-joint pendulum, a hinge
    driven by a motor, a breakable ball joint under a heavy box, a slider
    pushed every way, and a bridge of 8 hinged planks."""
    out = {}

    def world(bodies, gravity, dt=1.0 / 240.0):
        s = S.Scene.from_bodies(bodies)
        s.params = S.Params(half=(1000.0, 1000.0, 1000.0), periodic=False, dt=dt, gravity=gravity)
        return s

    q0 = (1.0, 0.0, 0.0, 0.0)
    s = world([_static(S.make_box(0, (0, 5, 0), q0, (0.5, 0.5, 0.5), 1.0)), S.make_box(1, (1, 5, 0), q0, (0.5, 0.5, 0.5), 1.0)],
              (0.0, -10.0, 0.0))
    s.gjoints = np.array([S.general_joint(s, _abi.JOINT_FIXED, 0, 1)], _abi.GJOINT_DTYPE)
    out["fixed"] = (s, 600)
    s = world([_static(S.make_sphere(0, (0, 5, 0), 0.2, 1.0)), S.make_sphere(1, (0, 3, 0), 0.5, 1.0)], (0.0, -10.0, 0.0))
    s.gjoints = np.array([S.general_joint(s, _abi.JOINT_BALL, 0, 1, anchor=(0, 5, 0))], _abi.GJOINT_DTYPE)
    out["ball"] = (s, 800)
    s = world([_static(S.make_box(0, (0, 0, 0), q0, (0.2, 0.2, 0.2), 1.0)), S.make_box(1, (1, 0, 0), q0, (1, 0.1, 0.1), 1.0)],
              (0.0, 0.0, 0.0))
    s.gjoints = np.array([S.general_joint(s, _abi.JOINT_HINGE, 0, 1, anchor=(0, 0, 0), axis=(0, 1, 0), motor=(2.0, 200.0))], _abi.GJOINT_DTYPE)
    out["hinge_motor"] = (s, 240)
    s = world([_static(S.make_sphere(0, (0, 5, 0), 0.2, 1.0)), S.make_box(1, (0, 3, 0), q0, (1, 1, 1), 50.0)],
              (0.0, -30.0, 0.0))
    s.gjoints = np.array([S.general_joint(s, _abi.JOINT_BALL, 0, 1, anchor=(0, 5, 0), break_impulse=5.0)], _abi.GJOINT_DTYPE)
    out["breakable"] = (s, 300)
    s = world([_static(S.make_box(0, (0, 0, 0), q0, (0.2, 0.2, 0.2), 1.0)), S.make_box(1, (0, 0, 0), q0, (0.5, 0.5, 0.5), 1.0)],
              (0.0, 0.0, 0.0))
    s.v[1] = (2.0, 3.0, 1.0)
    s.gjoints = np.array([S.general_joint(s, _abi.JOINT_SLIDER, 0, 1, anchor=(0, 0, 0), axis=(1, 0, 0))], _abi.GJOINT_DTYPE)
    out["slider"] = (s, 240)
    bodies = [_static(S.make_box(0, (-5, 6, 0), q0, (1, 1, 2), 1.0))]
    for i in range(8):
        bodies.append(S.make_box(len(bodies), (-4.5 + i * 1.0, 6, 0), q0, (0.5, 0.12, 2), 1.0))
    bodies.append(_static(S.make_box(len(bodies), (3, 6, 0), q0, (1, 1, 2), 1.0)))
    s = world(bodies, (0.0, -10.0, 0.0))
    js = [S.general_joint(s, _abi.JOINT_HINGE, i, i + 1, anchor=(-4.5 + i * 1.0 - 0.5, 6, 0), axis=(0, 0, 1))
          for i in range(8)]
    js.append(S.general_joint(s, _abi.JOINT_HINGE, 8, 9, anchor=(2.5, 6, 0), axis=(0, 0, 1)))
    s.gjoints = np.array(js, _abi.GJOINT_DTYPE)
    out["bridge"] = (s, 1200)
    return out


def joint_mix_scene(seed=0, periodic=False, limits=False):
    """Chains of boxes, spheres and capsules joined by every joint type at
    random anchors and axes (hinges with motors and angle limits, fixed
    joints, sliders, breakable ball joints), hung from static anchors under
    gravity (or free in a periodic box, chains crossing its boundary), with
    ground contacts and island sleeping on.  Hinge angle limits only with
    `limits` (their angle goes through atan2)."""
    rng = np.random.default_rng(seed)
    bodies = []
    h = 6.0
    anchors = []
    for c in range(5):
        x0 = (-4.0 + 2.0 * c, 4.5, rng.uniform(-1, 1))
        if periodic:
            x0 = (h - 0.6 + 0.3 * c, rng.uniform(-3, 3), rng.uniform(-3, 3))
        a = _static(S.make_box(len(bodies), x0, (1.0, 0.0, 0.0, 0.0), (0.2, 0.2, 0.2), 1.0))
        bodies.append(a)
        anchors.append(len(bodies) - 1)
    chains = []
    for c, ai in enumerate(anchors):
        prev = ai
        links = []
        for k in range(4):
            x = np.array(bodies[ai].x) + np.array([0.0, -0.9 * (k + 1), 0.0]) if not periodic else \
                np.array(bodies[ai].x) + np.array([0.8 * (k + 1), 0.0, 0.0])
            q = tuple(S.random_unit_quaternions(rng, 1)[0]) if k % 2 else (1.0, 0.0, 0.0, 0.0)
            kind = (c + k) % 3
            if kind == 0:
                bd = S.make_box(len(bodies), tuple(x), q, tuple(rng.uniform(0.15, 0.3, 3)), 1.0)
            elif kind == 1:
                bd = S.make_sphere(len(bodies), tuple(x), float(rng.uniform(0.15, 0.3)), 1.0)
            else:
                bd = S.make_capsule(len(bodies), tuple(x), q, 0.12, 0.2, 1.0)
            bd.v = tuple(rng.normal(0, 1.0, 3))
            bodies.append(bd)
            links.append((prev, len(bodies) - 1))
            prev = len(bodies) - 1
        chains.append(links)
    if not periodic:
        bodies.append(_static(S.make_box(len(bodies), (0.0, -5.5, 0.0), (1.0, 0.0, 0.0, 0.0), (5.5, 0.3, 5.5), 1.0)))
        for k in range(6):
            bodies.append(S.make_box(len(bodies), (rng.uniform(-4, 4), -4.5, rng.uniform(-4, 4)),
                                     tuple(S.random_unit_quaternions(rng, 1)[0]), (0.3, 0.3, 0.3), 1.0))
    s = S.Scene.from_bodies(bodies)
    s.params = S.Params(half=(h, h, h), periodic=periodic, dt=1.0 / 120.0,
                        gravity=(0.0, 0.0, 0.0) if periodic else (0.0, -9.81, 0.0), sleep_enabled=not periodic)
    js = []
    for c, links in enumerate(chains):
        for k, (a, b) in enumerate(links):
            mid = (np.array(s.x[a]) + np.array(s.x[b])) * 0.5
            axis = rng.normal(size=3)
            t = (c * 4 + k) % 6
            if t == 0:
                js.append(S.general_joint(s, _abi.JOINT_BALL, a, b, anchor=mid, break_impulse=float(rng.uniform(2, 8))))
            elif t == 1:
                js.append(S.general_joint(s, _abi.JOINT_HINGE, a, b, anchor=mid, axis=axis,
                                          motor=(float(rng.uniform(-3, 3)), float(rng.uniform(0.5, 5)))))
            elif t == 2:
                js.append(S.general_joint(s, _abi.JOINT_HINGE, a, b, anchor=mid, axis=axis,
                                          limit=(-0.3, 0.4) if limits else None))
            elif t == 3:
                js.append(S.general_joint(s, _abi.JOINT_FIXED, a, b))
            elif t == 4:
                js.append(S.general_joint(s, _abi.JOINT_SLIDER, a, b, anchor=mid, axis=axis))
            else:
                js.append(S.general_joint(s, _abi.JOINT_BALL, a, b, anchor=mid))
    s.gjoints = np.array(js, _abi.GJOINT_DTYPE)
    return s
