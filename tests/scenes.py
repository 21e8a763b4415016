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
    return out


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
