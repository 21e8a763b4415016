"""Scenes: bodies as structure-of-arrays numpy buffers plus world parameters.

Mirrors the reference's `ne::World` public state (src/world.hpp:75-136) and
the `ne::Body` fields the step reads (src/body.hpp:41-129).  Factories restate
the reference's mass/inertia formulas (src/body.hpp:206-315) in the same
operation order, so a scene built here has the same bits as one built with the
reference's factories.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _abi

PI = math.pi  # == ne::PI (src/vmath.hpp:16) rounded to double


@dataclass
class Params:
    """`ne::World` scalar fields (src/world.hpp:75-126), with the two extensions
    (per-axis half extents, solver_iterations)."""
    half: tuple = (12.0, 12.0, 12.0)
    periodic: bool = False
    open_boundary: bool = False
    dt: float = 1.0 / 120.0
    gravity: tuple = (0.0, 0.0, 0.0)
    restitution: float = 1.0
    friction: float = 0.0
    contact_beta: float = 0.2
    restitution_combine: int = _abi.COMBINE_MAX
    friction_combine: int = _abi.COMBINE_GEOMETRIC_MEAN
    warm_start: bool = True
    force_brute_force: bool = False
    solver_iterations: int = 8
    sleep_enabled: bool = False
    sleep_lin_vel: float = 0.05
    sleep_ang_vel: float = 0.05
    sleep_time: float = 0.5
    wake_vel: float = 0.5

    def to_c(self) -> _abi.WorldParams:
        p = _abi.WorldParams()
        h = self.half
        if np.isscalar(h):
            h = (h, h, h)
        for a in range(3):
            p.half[a] = float(h[a])
            p.gravity[a] = float(self.gravity[a])
        p.periodic = int(bool(self.periodic))
        p.open_boundary = int(bool(self.open_boundary))
        p.dt = float(self.dt)
        p.restitution = float(self.restitution)
        p.friction = float(self.friction)
        p.contact_beta = float(self.contact_beta)
        p.restitution_combine = int(self.restitution_combine)
        p.friction_combine = int(self.friction_combine)
        p.warm_start = int(bool(self.warm_start))
        p.force_brute_force = int(bool(self.force_brute_force))
        p.solver_iterations = int(self.solver_iterations)
        p.sleep_enabled = int(bool(self.sleep_enabled))
        p.sleep_lin_vel = float(self.sleep_lin_vel)
        p.sleep_ang_vel = float(self.sleep_ang_vel)
        p.sleep_time = float(self.sleep_time)
        p.wake_vel = float(self.wake_vel)
        return p


class Scene:
    """Bodies in SoA form (x: n x 3, q: n x 4 as (w,x,y,z), ...)."""

    def __init__(self, n: int = 0):
        self.n = n
        z3 = lambda: np.zeros((n, 3), np.float64)
        self.x, self.v, self.w = z3(), z3(), z3()
        self.q = np.zeros((n, 4), np.float64)
        self.q[:, 0] = 1.0
        self.force, self.torque = z3(), z3()
        self.inv_mass = np.zeros(n, np.float64)
        self.inv_inertia = z3()
        self.shape = np.zeros(n, np.int32)
        self.radius = np.ones(n, np.float64)
        self.half_height = np.zeros(n, np.float64)
        self.half_extents = z3()
        self.friction = np.full(n, -1.0)
        self.restitution = np.full(n, -1.0)
        self.linear_damping = np.zeros(n)
        self.angular_damping = np.zeros(n)
        self.gravity_scale = np.ones(n)
        self.layer = np.zeros(n, np.uint32)
        self.mask = np.full(n, 0xFFFFFFFF, np.uint32)
        self.flags = np.full(n, _abi.FLAG_DYNAMIC, np.uint32)
        self.hull_offsets = np.zeros(n + 1, np.uint32)
        self.hull_verts = np.zeros((0, 3), np.float64)
        self.nocollide_offsets = np.zeros(n + 1, np.uint32)
        self.nocollide = np.zeros(0, np.uint32)
        self.joints = np.zeros(0, _abi.JOINT_DTYPE)   # World::distanceJoints, in order
        self.gjoints = np.zeros(0, _abi.GJOINT_DTYPE)  # World::joints (general joints), in order
        # static triangle meshes: dicts {body, verts (nv, 3) f64, tris (nt, 3) i32} (Body::mesh)
        self.meshes = []
        # Body::sleepTimer / invMassStore / invIStore (body.hpp); None uploads zeros
        self.sleep_timer = None
        self.inv_mass_store = None
        self.inv_inertia_store = None
        self.params = Params()

    # -- construction -----------------------------------------------------
    @staticmethod
    def from_bodies(bodies, params: Params | None = None) -> "Scene":
        s = Scene(len(bodies))
        hulls, ncs = [], []
        for i, b in enumerate(bodies):
            s.x[i], s.v[i], s.q[i], s.w[i] = b.x, b.v, b.q, b.w
            s.force[i], s.torque[i] = b.force, b.torque
            s.inv_mass[i] = b.inv_mass
            s.inv_inertia[i] = b.inv_inertia
            s.shape[i] = b.shape
            s.radius[i] = b.radius
            s.half_height[i] = b.half_height
            s.half_extents[i] = b.half_extents
            s.friction[i], s.restitution[i] = b.friction, b.restitution
            s.linear_damping[i], s.angular_damping[i] = b.linear_damping, b.angular_damping
            s.gravity_scale[i] = b.gravity_scale
            s.layer[i], s.mask[i] = b.layer, b.mask
            s.flags[i] = b.flags
            hulls.append(np.asarray(b.vertices, np.float64).reshape(-1, 3))
            ncs.append(np.asarray(b.no_collide, np.uint32).reshape(-1))
        s.hull_offsets = np.concatenate([[0], np.cumsum([len(h) for h in hulls])]).astype(np.uint32)
        s.hull_verts = np.concatenate(hulls, axis=0) if hulls else np.zeros((0, 3))
        s.nocollide_offsets = np.concatenate([[0], np.cumsum([len(c) for c in ncs])]).astype(np.uint32)
        s.nocollide = np.concatenate(ncs) if ncs else np.zeros(0, np.uint32)
        if params is not None:
            s.params = params
        return s

    def copy(self) -> "Scene":
        s = Scene(0)
        for k, v in self.__dict__.items():
            setattr(s, k, v.copy() if isinstance(v, np.ndarray) else v)
        s.meshes = [dict(m) for m in self.meshes]
        s.params = Params(**self.params.__dict__)
        return s

    # -- ctypes view ------------------------------------------------------
    def to_c(self):
        """Returns (ne_bodies struct, keepalive list)."""
        keep = []

        def f64(a):
            a = np.ascontiguousarray(a, np.float64)
            keep.append(a)
            return _abi.dptr(a)

        def u32(a):
            a = np.ascontiguousarray(a, np.uint32)
            keep.append(a)
            return _abi.uptr(a)

        b = _abi.Bodies()
        b.n = self.n
        b.x, b.v, b.q, b.w = f64(self.x), f64(self.v), f64(self.q), f64(self.w)
        b.force, b.torque = f64(self.force), f64(self.torque)
        b.inv_mass = f64(self.inv_mass)
        b.inv_inertia = f64(self.inv_inertia)
        sh = np.ascontiguousarray(self.shape, np.int32)
        keep.append(sh)
        b.shape = _abi.iptr(sh)
        b.radius, b.half_height = f64(self.radius), f64(self.half_height)
        b.half_extents = f64(self.half_extents)
        b.friction, b.restitution = f64(self.friction), f64(self.restitution)
        b.linear_damping, b.angular_damping = f64(self.linear_damping), f64(self.angular_damping)
        b.gravity_scale = f64(self.gravity_scale)
        b.layer, b.mask, b.flags = u32(self.layer), u32(self.mask), u32(self.flags)
        b.sleep_timer = None if self.sleep_timer is None else f64(self.sleep_timer)
        b.inv_mass_store = None if self.inv_mass_store is None else f64(self.inv_mass_store)
        b.inv_inertia_store = None if self.inv_inertia_store is None else f64(self.inv_inertia_store)
        b.hull_offsets = u32(self.hull_offsets)
        hv = np.ascontiguousarray(self.hull_verts.reshape(-1, 3) if len(self.hull_verts) else np.zeros((1, 3)), np.float64)
        keep.append(hv)
        b.hull_verts = _abi.dptr(hv)
        b.nocollide_offsets = u32(self.nocollide_offsets)
        nc = np.ascontiguousarray(self.nocollide if len(self.nocollide) else np.zeros(1, np.uint32), np.uint32)
        keep.append(nc)
        b.nocollide = _abi.uptr(nc)
        return b, keep

    def meshes_c(self):
        """(ne_mesh array, keepalive list) for the scene's meshes."""
        keep = []
        arr = (_abi.Mesh * max(1, len(self.meshes)))()
        for k, m in enumerate(self.meshes):
            v = np.ascontiguousarray(m["verts"], np.float64).reshape(-1, 3)
            t = np.ascontiguousarray(m["tris"], np.int32).reshape(-1, 3)
            keep += [v, t]
            arr[k].body = int(m["body"])
            arr[k].nverts = len(v)
            arr[k].verts = _abi.dptr(v)
            arr[k].ntris = len(t)
            arr[k].tris = _abi.iptr(t)
        return arr, keep


# ---------------------------------------------------------------------------
# Body + factories (src/body.hpp:41-315)
# ---------------------------------------------------------------------------

@dataclass
class Body:
    x: tuple = (0.0, 0.0, 0.0)
    v: tuple = (0.0, 0.0, 0.0)
    q: tuple = (1.0, 0.0, 0.0, 0.0)
    w: tuple = (0.0, 0.0, 0.0)
    force: tuple = (0.0, 0.0, 0.0)
    torque: tuple = (0.0, 0.0, 0.0)
    inv_mass: float = 0.0
    inv_inertia: tuple = (0.0, 0.0, 0.0)
    shape: int = _abi.SHAPE_SPHERE
    radius: float = 1.0
    half_height: float = 0.0
    half_extents: tuple = (0.0, 0.0, 0.0)
    vertices: list = field(default_factory=list)
    friction: float = -1.0
    restitution: float = -1.0
    linear_damping: float = 0.0
    angular_damping: float = 0.0
    gravity_scale: float = 1.0
    layer: int = 0
    mask: int = 0xFFFFFFFF
    no_collide: list = field(default_factory=list)
    flags: int = _abi.FLAG_DYNAMIC
    id: int = -1

    def make_static(self):
        """PhysicsWorld::makeStatic (src/physics_world.hpp:112-121)."""
        self.inv_mass = 0.0
        self.inv_inertia = (0.0, 0.0, 0.0)
        self.v = (0.0, 0.0, 0.0)
        self.w = (0.0, 0.0, 0.0)
        self.flags &= ~(_abi.FLAG_DYNAMIC | _abi.FLAG_KINEMATIC)
        return self


def make_sphere(id, x, radius, density) -> Body:            # body.hpp:206-219
    b = Body(shape=_abi.SHAPE_SPHERE, id=id, x=tuple(map(float, x)), radius=float(radius), half_height=0.0)
    m = density * (4.0 / 3.0) * PI * radius * radius * radius
    b.inv_mass = 1.0 / m if m > 0 else 0.0
    I = (2.0 / 5.0) * m * radius * radius
    invI = 1.0 / I if I > 0 else 0.0
    b.inv_inertia = (invI, invI, invI)
    return b


def make_cylinder(id, x, q, radius, height, density) -> Body:    # body.hpp:222-239
    b = Body(shape=_abi.SHAPE_CYLINDER, id=id, x=tuple(map(float, x)), q=tuple(map(float, q)),
             radius=float(radius), half_height=0.5 * height)
    m = density * PI * radius * radius * height
    b.inv_mass = 1.0 / m if m > 0 else 0.0
    Iyy = 0.5 * m * radius * radius
    Ixx = (1.0 / 12.0) * m * (3.0 * radius * radius + height * height)
    b.inv_inertia = (1.0 / Ixx if Ixx > 0 else 0.0, 1.0 / Iyy if Iyy > 0 else 0.0, 1.0 / Ixx if Ixx > 0 else 0.0)
    return b


def make_box(id, x, q, half_extents, density) -> Body:          # body.hpp:243-261
    hx, hy, hz = map(float, half_extents)
    b = Body(shape=_abi.SHAPE_BOX, id=id, x=tuple(map(float, x)), q=tuple(map(float, q)), half_extents=(hx, hy, hz))
    m = density * 8.0 * hx * hy * hz
    b.inv_mass = 1.0 / m if m > 0 else 0.0
    Ix = (1.0 / 3.0) * m * (hy * hy + hz * hz)
    Iy = (1.0 / 3.0) * m * (hx * hx + hz * hz)
    Iz = (1.0 / 3.0) * m * (hx * hx + hy * hy)
    b.inv_inertia = (1.0 / Ix if Ix > 0 else 0.0, 1.0 / Iy if Iy > 0 else 0.0, 1.0 / Iz if Iz > 0 else 0.0)
    return b


def make_capsule(id, x, q, radius, half_height, density) -> Body:   # body.hpp:266-288
    b = Body(shape=_abi.SHAPE_CAPSULE, id=id, x=tuple(map(float, x)), q=tuple(map(float, q)),
             radius=float(radius), half_height=float(half_height))
    r, L = radius, 2.0 * half_height
    mc = density * PI * r * r * L
    mh = density * (4.0 / 3.0) * PI * r * r * r
    m = mc + mh
    b.inv_mass = 1.0 / m if m > 0 else 0.0
    Iyy = 0.5 * mc * r * r + 0.4 * mh * r * r
    Ixx = mc * (0.25 * r * r + L * L / 12.0) + mh * (0.4 * r * r + 0.25 * L * L)
    b.inv_inertia = (1.0 / Ixx if Ixx > 0 else 0.0, 1.0 / Iyy if Iyy > 0 else 0.0, 1.0 / Ixx if Ixx > 0 else 0.0)
    return b


def make_convex(id, x, q, verts, density) -> Body:              # body.hpp:293-315
    verts = [tuple(map(float, v)) for v in verts]
    b = Body(shape=_abi.SHAPE_CONVEX, id=id, x=tuple(map(float, x)), q=tuple(map(float, q)), vertices=verts)
    lo = [1e30, 1e30, 1e30]
    hi = [-1e30, -1e30, -1e30]
    for v in verts:
        lo = [min(lo[a], v[a]) for a in range(3)]
        hi = [max(hi[a], v[a]) for a in range(3)]
    he = [(hi[a] - lo[a]) * 0.5 for a in range(3)]
    hx, hy, hz = max(he[0], 1e-4), max(he[1], 1e-4), max(he[2], 1e-4)
    m = density * 8.0 * hx * hy * hz
    b.inv_mass = 1.0 / m if m > 0 else 0.0
    Ix = (1.0 / 3.0) * m * (hy * hy + hz * hz)
    Iy = (1.0 / 3.0) * m * (hx * hx + hz * hz)
    Iz = (1.0 / 3.0) * m * (hx * hx + hy * hy)
    b.inv_inertia = (1.0 / Ix if Ix > 0 else 0.0, 1.0 / Iy if Iy > 0 else 0.0, 1.0 / Iz if Iz > 0 else 0.0)
    return b


def make_plane(id, point, normal) -> Body:                      # body.hpp:335-345
    n = np.asarray(normal, np.float64)
    nn = math.sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2])
    he = (n[0] / nn, n[1] / nn, n[2] / nn) if nn > 1e-300 else (0.0, 0.0, 0.0)
    b = Body(shape=_abi.SHAPE_PLANE, id=id, x=tuple(map(float, point)), half_extents=tuple(map(float, he)))
    b.inv_mass = 0.0
    b.inv_inertia = (0.0, 0.0, 0.0)
    b.flags = 0
    return b


# ---------------------------------------------------------------------------
# Synthetic workloads
# ---------------------------------------------------------------------------

def _reject_overlaps(pos, radius, L, rng, periodic=True, max_rounds=200):
    """Resample later bodies of overlapping pairs until no two centres are
    closer than 2*radius (minimum image when periodic)."""
    from scipy.spatial import cKDTree
    n = len(pos)
    for _ in range(max_rounds):
        if periodic:
            tree = cKDTree(np.mod(pos + L / 2, L), boxsize=L)
        else:
            tree = cKDTree(pos)
        pairs = tree.query_pairs(2.0 * radius, output_type="ndarray")
        if len(pairs) == 0:
            return pos
        bad = np.unique(pairs.max(axis=1))
        pos[bad] = rng.uniform(-L / 2, L / 2, size=(len(bad), 3))
    raise RuntimeError("could not place non-overlapping bodies")


def sphere_gas(n: int, L: float | None = None, radius: float = 0.5, seed: int = 0, mass: float = 1.0,
               sigma: float = 1.0) -> Scene:
    """Periodic elastic sphere gas: the plain-method workload.

    n spheres of `radius` and `mass` uniform in a periodic cube of side L
    (default: the density of 4096 spheres at L=40), rejection-sampled to no
    overlap; velocities N(0, sigma) per axis with the net momentum removed; no
    gravity, restitution 1, friction 0, dt 1/60, 8 iterations.
    """
    if L is None:
        L = 40.0 * (n / 4096.0) ** (1.0 / 3.0)
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-L / 2, L / 2, size=(n, 3))
    pos = _reject_overlaps(pos, radius, L, rng)
    vel = rng.normal(0.0, sigma, size=(n, 3))
    vel -= vel.mean(axis=0)
    s = Scene(n)
    s.x[:] = pos
    s.v[:] = vel
    density = mass / ((4.0 / 3.0) * PI * radius * radius * radius)
    proto = make_sphere(0, (0, 0, 0), radius, density)
    s.inv_mass[:] = proto.inv_mass
    s.inv_inertia[:] = proto.inv_inertia
    s.shape[:] = _abi.SHAPE_SPHERE
    s.radius[:] = radius
    s.params = Params(half=(L / 2, L / 2, L / 2), periodic=True, dt=1.0 / 60.0, restitution=1.0, friction=0.0,
                      gravity=(0.0, 0.0, 0.0), solver_iterations=8)
    return s


def random_unit_quaternions(rng, n):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return q


def hull12(rng, r):
    """12 points on a jittered sphere of radius r (a random convex hull)."""
    d = rng.normal(size=(12, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return d * r * rng.uniform(0.85, 1.0, size=(12, 1))


def mixed_pile(n: int, half=(8.0, 8.0, 8.0), seed: int = 0, ground: bool = True, spacing: float | None = None,
               periodic: bool = False, iterations: int = 8, sizes: str = "lattice") -> Scene:
    """Mixed-shape pile under gravity on a static ground box: 40% boxes, 30%
    spheres, 20% capsules, 10% 12-vertex hulls, dropped from a jittered
    lattice with random orientations; mu 0.5, e 0.2, g -9.81, dt 1/60.

    sizes="config": box half extents U[0.2,0.6]^3, sphere radii U[0.2,0.6],
    capsules r 0.2 / half-height 0.4, hull radii U[0.2,0.6], on a lattice of
    spacing 2.3 (the largest box's bounding diameter plus the jitter).
    sizes="lattice" (the small test piles): the same distributions scaled to
    a lattice of `spacing` (default 1.3), so that many bodies fit a small box.
    """
    if sizes not in ("lattice", "config"):
        raise ValueError(sizes)
    if spacing is None:
        spacing = 2.3 if sizes == "config" else 1.3
    lat = sizes == "lattice"
    rng = np.random.default_rng(seed)
    bodies = []
    hx, hy, hz = half
    m = int(math.ceil(n ** (1.0 / 3.0)))
    kinds = rng.permutation(np.array([0] * (4 * n // 10) + [1] * (3 * n // 10) + [2] * (2 * n // 10) +
                                     [3] * (n - 4 * n // 10 - 3 * n // 10 - 2 * n // 10)))
    qs = random_unit_quaternions(rng, n)
    floor_y = -hy + 1.0
    idx = 0
    for a in range(m):
        for b in range(m):
            for c in range(m):
                if idx >= n:
                    break
                x = (-0.5 * (m - 1) + a) * spacing + rng.uniform(-0.1, 0.1)
                z = (-0.5 * (m - 1) + c) * spacing + rng.uniform(-0.1, 0.1)
                y = floor_y + 1.0 + b * spacing + rng.uniform(-0.1, 0.1)
                q = tuple(qs[idx])
                k = kinds[idx]
                if k == 0:
                    he = rng.uniform(0.2, 0.6, size=3)
                    he = tuple(he * (spacing / 2.0) / 0.6 * 0.55 if lat else he)
                    bodies.append(make_box(idx, (x, y, z), q, he, 1.0))
                elif k == 1:
                    r = rng.uniform(0.2, 0.6)
                    bodies.append(make_sphere(idx, (x, y, z), r * (spacing / 2.0) / 0.6 * 0.9 if lat else r, 1.0))
                elif k == 2:
                    bodies.append(make_capsule(idx, (x, y, z), q, 0.2, 0.4 * spacing / 1.3 * 0.9 if lat else 0.4, 1.0))
                else:
                    r = rng.uniform(0.2, 0.6)
                    bodies.append(make_convex(idx, (x, y, z), q, hull12(rng, r * spacing / 1.3 if lat else r), 1.0))
                idx += 1
    if ground:
        g = make_box(idx, (0.0, -hy + 0.5, 0.0), (1.0, 0.0, 0.0, 0.0), (hx - 0.5, 0.5, hz - 0.5), 1.0).make_static()
        bodies.append(g)
    s = Scene.from_bodies(bodies)
    s.params = Params(half=tuple(half), periodic=periodic, dt=1.0 / 60.0, gravity=(0.0, -9.81, 0.0),
                      restitution=0.2, friction=0.5, warm_start=True, solver_iterations=iterations)
    return s


# ---------------------------------------------------------------------------
# Batched worlds: many independent periodic worlds in one body array
# ---------------------------------------------------------------------------

def concat_scenes(scenes) -> Scene:
    """One body array holding the scenes one after another (hull and
    no-collide indices shifted); params are the first scene's."""
    n = sum(s.n for s in scenes)
    out = Scene(n)
    per_body = ["x", "v", "q", "w", "force", "torque", "inv_mass", "inv_inertia", "shape", "radius", "half_height",
                "half_extents", "friction", "restitution", "linear_damping", "angular_damping", "gravity_scale",
                "layer", "mask", "flags"]
    for k in per_body:
        setattr(out, k, np.concatenate([getattr(s, k) for s in scenes], axis=0))
    hv, ho, nc, no = [], [0], [], [0]
    base = 0
    for s in scenes:
        hv.append(s.hull_verts.reshape(-1, 3))
        ho.extend((s.hull_offsets[1:] + ho[-1]).tolist())
        nc.append(s.nocollide.astype(np.uint32) + np.uint32(base))
        no.extend((s.nocollide_offsets[1:] + no[-1]).tolist())
        base += s.n
    out.hull_verts = np.concatenate(hv, axis=0) if hv else np.zeros((0, 3))
    out.hull_offsets = np.asarray(ho, np.uint32)
    out.nocollide = np.concatenate(nc) if nc else np.zeros(0, np.uint32)
    out.nocollide_offsets = np.asarray(no, np.uint32)
    js, base = [], 0
    for s in scenes:
        j = s.joints.copy()
        j["a"] += np.uint32(base)
        j["b"] += np.uint32(base)
        js.append(j)
        base += s.n
    out.joints = np.concatenate(js) if js else np.zeros(0, _abi.JOINT_DTYPE)
    out.params = Params(**scenes[0].params.__dict__) if scenes else Params()
    return out


def slice_scene(s: Scene, begin: int, end: int) -> Scene:
    """Bodies [begin, end) of a concatenated scene as a scene of their own."""
    out = Scene(end - begin)
    for k in ["x", "v", "q", "w", "force", "torque", "inv_mass", "inv_inertia", "shape", "radius", "half_height",
              "half_extents", "friction", "restitution", "linear_damping", "angular_damping", "gravity_scale",
              "layer", "mask", "flags"]:
        setattr(out, k, getattr(s, k)[begin:end].copy())
    h0, h1 = int(s.hull_offsets[begin]), int(s.hull_offsets[end])
    out.hull_verts = s.hull_verts.reshape(-1, 3)[h0:h1].copy()
    out.hull_offsets = (s.hull_offsets[begin:end + 1] - s.hull_offsets[begin]).astype(np.uint32)
    c0, c1 = int(s.nocollide_offsets[begin]), int(s.nocollide_offsets[end])
    out.nocollide = (s.nocollide[c0:c1] - begin).astype(np.uint32)
    out.nocollide_offsets = (s.nocollide_offsets[begin:end + 1] - s.nocollide_offsets[begin]).astype(np.uint32)
    sel = (s.joints["a"] >= begin) & (s.joints["a"] < end) & (s.joints["b"] >= begin) & (s.joints["b"] < end)
    out.joints = s.joints[sel].copy()
    out.joints["a"] -= np.uint32(begin)
    out.joints["b"] -= np.uint32(begin)
    out.params = Params(**s.params.__dict__)
    return out


@dataclass
class Batch:
    """Concatenated worlds: world w is bodies [offsets[w], offsets[w+1]) with
    box half extents half[w] and default restitution / friction."""
    scene: Scene
    offsets: np.ndarray
    half: np.ndarray
    restitution: np.ndarray
    friction: np.ndarray

    @property
    def nworlds(self) -> int:
        return len(self.offsets) - 1

    def world(self, w: int) -> Scene:
        """World w alone, with its own box and material defaults."""
        s = slice_scene(self.scene, int(self.offsets[w]), int(self.offsets[w + 1]))
        s.params.half = tuple(float(h) for h in self.half[w])
        s.params.restitution = float(self.restitution[w])
        s.params.friction = float(self.friction[w])
        return s


def _lattice(rng, n, L, jitter=0.1):
    m = int(math.ceil(n ** (1.0 / 3.0)))
    sp = L / m
    g = np.stack(np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij"), -1).reshape(-1, 3)
    g = g[rng.permutation(len(g))[:n]]
    return (-L / 2 + (g + 0.5) * sp + rng.uniform(-jitter, jitter, size=(n, 3)) * sp).astype(np.float64)


def _place_periodic(rng, rho, L, max_rounds=200, tries=8, k=32):
    """Uniform positions in a periodic cube of side L, resampled until no two
    bodies are closer than rho_i + rho_j (minimum image).  Each round moves one
    body of every conflicting pair to the first of `tries` random spots that
    is free of the other bodies (checked against the k nearest; a spot with k
    bodies in reach counts as taken)."""
    from scipy.spatial import cKDTree
    n = len(rho)
    pos = rng.uniform(-L / 2, L / 2, size=(n, 3))
    reach = 2.0 * float(rho.max())
    k = min(k, n)
    for _ in range(max_rounds):
        tree = cKDTree(np.mod(pos + L / 2, L), boxsize=L)
        pairs = tree.query_pairs(reach, output_type="ndarray")
        if len(pairs):
            d = pos[pairs[:, 0]] - pos[pairs[:, 1]]
            d -= L * np.floor(d / L + 0.5)
            pairs = pairs[np.einsum("ij,ij->i", d, d) < (rho[pairs[:, 0]] + rho[pairs[:, 1]]) ** 2]
        if len(pairs) == 0:
            return pos
        bad = np.unique(pairs.max(axis=1))
        cand = rng.uniform(-L / 2, L / 2, size=(len(bad), tries, 3))
        dist, idx = tree.query(np.mod(cand.reshape(-1, 3) + L / 2, L), k=k, distance_upper_bound=reach)
        dist = dist.reshape(len(bad), tries, k)
        idx = idx.reshape(len(bad), tries, k)
        present = idx < n
        other = np.where(present & (idx != bad[:, None, None]), idx, n)
        rr = np.concatenate([rho, [0.0]])[other]
        hit = (other < n) & (dist < rho[bad][:, None, None] + rr)
        full = present[:, :, -1]          # k bodies in reach: cannot tell, call it taken
        free = ~(hit.any(axis=2) | full)
        pick = np.argmax(free, axis=1)
        pos[bad] = cand[np.arange(len(bad)), pick]
    raise RuntimeError("could not place the bodies")


def _sphere_mass(radius, density):
    """make_sphere's inv_mass and inverse inertia (body.hpp:206-219), vectorised
    in the same operation order."""
    m = density * (4.0 / 3.0) * PI * radius * radius * radius
    I = (2.0 / 5.0) * m * radius * radius
    return np.where(m > 0, 1.0 / m, 0.0), np.where(I > 0, 1.0 / I, 0.0)


def _box_mass(he, density):
    """make_box's inv_mass and inverse inertia (body.hpp:243-261), vectorised."""
    hx, hy, hz = he[:, 0], he[:, 1], he[:, 2]
    m = density * 8.0 * hx * hy * hz
    Ix = (1.0 / 3.0) * m * (hy * hy + hz * hz)
    Iy = (1.0 / 3.0) * m * (hx * hx + hz * hz)
    Iz = (1.0 / 3.0) * m * (hx * hx + hy * hy)
    inv = lambda a: np.where(a > 0, 1.0 / a, 0.0)
    return inv(m), np.stack([inv(Ix), inv(Iy), inv(Iz)], axis=1)


def batched_worlds(nworlds: int = 1024, seed: int = 1, nmin: int = 64, nmax: int = 4096, phi: float = 0.2,
                   iterations: int = 8) -> Batch:
    """Independent periodic worlds of uneven size: N_w log-uniform in
    [nmin, nmax]; half spheres (r U[0.2,0.6]) and half boxes (half extents
    U[0.2,0.6]^3), density 1, random orientations, in a cube of side
    (sum of volumes / phi)^(1/3); placed uniformly with no sphere overlap and
    no two box centres closer than half their bounding radii (box corners
    may touch at start); velocities N(0,1); per-world friction U[0,0.8] and
    restitution U[0,1]; no gravity, dt 1/60."""
    rng = np.random.default_rng(seed)
    sizes = np.floor(np.exp(rng.uniform(math.log(nmin), math.log(nmax + 1), size=nworlds))).astype(int)
    sizes = np.clip(sizes, nmin, nmax)
    scenes, halves = [], []
    mu = rng.uniform(0.0, 0.8, size=nworlds)
    e = rng.uniform(0.0, 1.0, size=nworlds)
    for w in range(nworlds):
        n = int(sizes[w])
        ns = n // 2
        r = rng.uniform(0.2, 0.6, size=ns)
        he = rng.uniform(0.2, 0.6, size=(n - ns, 3))
        vol = float(np.sum(4.0 / 3.0 * PI * r ** 3) + np.sum(8.0 * np.prod(he, axis=1)))
        L = (vol / phi) ** (1.0 / 3.0)
        kinds = rng.permutation(n) < ns          # True: sphere
        rho = np.empty(n)
        rho[kinds] = r
        rho[~kinds] = 0.5 * np.linalg.norm(he, axis=1)
        pos = _place_periodic(rng, rho, L)
        qs = random_unit_quaternions(rng, n)
        vel = rng.normal(0.0, 1.0, size=(n, 3))
        s = Scene(n)
        s.x[:] = pos
        s.v[:] = vel
        sph, box = np.flatnonzero(kinds), np.flatnonzero(~kinds)
        s.shape[sph] = _abi.SHAPE_SPHERE
        s.radius[sph] = r
        s.inv_mass[sph], ii = _sphere_mass(r, 1.0)
        s.inv_inertia[sph] = ii[:, None]
        s.shape[box] = _abi.SHAPE_BOX
        s.half_extents[box] = he
        s.q[box] = qs[box]
        s.inv_mass[box], s.inv_inertia[box] = _box_mass(he, 1.0)
        s.params = Params(half=(L / 2, L / 2, L / 2), periodic=True, dt=1.0 / 60.0, gravity=(0.0, 0.0, 0.0),
                          restitution=float(e[w]), friction=float(mu[w]), solver_iterations=iterations)
        scenes.append(s)
        halves.append((L / 2, L / 2, L / 2))
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint32)
    return Batch(concat_scenes(scenes), off, np.asarray(halves, np.float64), e, mu)


def adjoint_worlds(nworlds: int = 256, n_spheres: int = 256, n_boxes: int = 256, side: float = 16.0, seed: int = 2,
                   friction: float = 0.4, restitution: float = 0.3, ground: bool = True) -> Batch:
    """Differentiable-rollout worlds: each a periodic cube of `side` holding a
    static ground box (first body, spanning the cube in x and z) and
    `n_spheres` spheres r=0.4 plus `n_boxes` boxes of half extent 0.3 on a
    jittered lattice above it; velocities N(0,1); gravity -9.81, dt 1/60,
    8 iterations."""
    rng = np.random.default_rng(seed)
    h = side / 2.0
    scenes = []
    for w in range(nworlds):
        bodies = []
        if ground:
            bodies.append(make_box(0, (0.0, -h + 0.5, 0.0), (1.0, 0.0, 0.0, 0.0), (h, 0.5, h), 1.0).make_static())
        n = n_spheres + n_boxes
        m = int(math.ceil((n / 2.0) ** (1.0 / 3.0)))
        sp = side / max(m, 1)
        cells = np.stack(np.meshgrid(np.arange(m), np.arange(2 * m), np.arange(m), indexing="ij"), -1).reshape(-1, 3)
        cells = cells[rng.permutation(len(cells))[:n]]
        qs = random_unit_quaternions(rng, n)
        kinds = rng.permutation(n) < n_spheres
        for t in range(n):
            c = cells[t]
            x = -h + (c[0] + 0.5) * sp + rng.uniform(-0.1, 0.1)
            y = -h + 1.5 + (c[1] + 0.5) * (side - 2.0) / (2 * m) + rng.uniform(-0.05, 0.05)
            z = -h + (c[2] + 0.5) * sp + rng.uniform(-0.1, 0.1)
            k = len(bodies)
            if kinds[t]:
                b = make_sphere(k, (x, y, z), 0.4, 1.0)
            else:
                b = make_box(k, (x, y, z), tuple(qs[t]), (0.3, 0.3, 0.3), 1.0)
            b.v = tuple(rng.normal(0.0, 1.0, size=3))
            bodies.append(b)
        s = Scene.from_bodies(bodies)
        s.params = Params(half=(h, h, h), periodic=True, dt=1.0 / 60.0, gravity=(0.0, -9.81, 0.0),
                          restitution=restitution, friction=friction, solver_iterations=8)
        scenes.append(s)
    off = np.concatenate([[0], np.cumsum([s.n for s in scenes])]).astype(np.uint32)
    return Batch(concat_scenes(scenes), off, np.full((nworlds, 3), h), np.full(nworlds, restitution),
                 np.full(nworlds, friction))


def rope_scene(nropes: int = 8, links: int = 16, radius: float = 0.2, half=(6.0, 6.0, 6.0), seed: int = 0,
               periodic: bool = True, anchored: bool = True, springs: int = 0, spacing: float = 1.5) -> Scene:
    """Ropes of `links` spheres joined surface to surface by rigid distance
    joints (rest = the gap between surfaces, so the joints act at off-centre
    anchors and turn the links), each optionally hung from a static anchor
    sphere; `springs` extra spring-damper joints tie random pairs of links.
    A few loose boxes fall among the ropes.  Gravity -9.81, friction 0.4,
    restitution 0.1, dt 1/60."""
    rng = np.random.default_rng(seed)
    bodies, joints = [], []
    top = half[1] - 1.0
    for r in range(nropes):
        x0 = -half[0] + 1.0 + (r % 4) * spacing + rng.uniform(-0.1, 0.1)
        z0 = -half[2] + 1.0 + (r // 4) * spacing + rng.uniform(-0.1, 0.1)
        prev = None
        if anchored:
            bodies.append(make_sphere(len(bodies), (x0, top, z0), 0.1, 1.0).make_static())
            prev = len(bodies) - 1
        for k in range(links):
            # links step sideways as well as down so the rope starts bent and swings
            x = (x0 + 0.6 * (k + 1) * 2 * radius, top - (k + 1) * 2.2 * radius, z0)
            bodies.append(make_sphere(len(bodies), x, radius, 1.0))
            cur = len(bodies) - 1
            if prev is not None:
                ra = 0.1 if (anchored and k == 0) else radius
                joints.append((prev, cur, (0.0, -ra, 0.0), (0.0, radius, 0.0), 0.2 * radius, 0.0, 0.0))
            prev = cur
    for b in range(max(2, nropes // 2)):
        q = tuple(random_unit_quaternions(rng, 1)[0])
        bodies.append(make_box(len(bodies), (rng.uniform(-2, 2), top - 0.5, rng.uniform(-2, 2)), q, (0.3, 0.2, 0.25), 1.0))
    dyn = [i for i, bd in enumerate(bodies) if bd.inv_mass > 0 and bd.shape == _abi.SHAPE_SPHERE]
    for _ in range(springs):
        a, b = rng.choice(dyn, size=2, replace=False)
        joints.append((int(min(a, b)), int(max(a, b)), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), float(rng.uniform(0.5, 2.0)),
                       float(rng.uniform(20.0, 200.0)), float(rng.uniform(0.5, 5.0))))
    s = Scene.from_bodies(bodies)
    s.joints = np.array([(a, b, la, lb, rest, k, d) for a, b, la, lb, rest, k, d in joints], _abi.JOINT_DTYPE)
    s.params = Params(half=tuple(half), periodic=periodic, dt=1.0 / 60.0, gravity=(0.0, -9.81, 0.0),
                      restitution=0.1, friction=0.4, solver_iterations=8)
    return s


def sleeping_pile(n: int = 120, seed: int = 0, half=(6.0, 6.0, 6.0), periodic: bool = False,
                  projectile_delay: float = 2.5) -> Scene:
    """A pile of boxes and spheres settling on a static ground box with island
    sleeping on, and a fast sphere that reaches the pile after about
    `projectile_delay` seconds and wakes part of it."""
    s = mixed_pile(n, half=half, seed=seed, periodic=periodic, spacing=1.2)
    bodies = []
    for i in range(s.n):
        b = Body(shape=int(s.shape[i]), id=i, x=tuple(s.x[i]), q=tuple(s.q[i]), radius=float(s.radius[i]),
                 half_height=float(s.half_height[i]), half_extents=tuple(s.half_extents[i]))
        b.inv_mass = float(s.inv_mass[i])
        b.inv_inertia = tuple(s.inv_inertia[i])
        b.flags = int(s.flags[i])
        b.vertices = s.hull_verts[s.hull_offsets[i]:s.hull_offsets[i + 1]].tolist()
        bodies.append(b)
    speed = 8.0
    start = (-half[0] + 0.6, -half[1] + 2.0, 0.0)
    # flies along +x from behind a wall of empty space: distance = speed * delay (clamped to the box)
    x0 = max(-half[0] + 0.6, -speed * projectile_delay * 0.1)
    proj = make_sphere(len(bodies), (x0, start[1], start[2]), 0.35, 4.0)
    proj.v = (speed, 0.0, 0.0)
    bodies.append(proj)
    out = Scene.from_bodies(bodies)
    out.params = s.params
    out.params.sleep_enabled = True
    return out


def wake_chain(chain: int = 6, rows: int = 3, seed: int = 0, drift: float = 1.0) -> Scene:
    """Rows of touching spheres set asleep with a velocity above wakeVel, each
    row led by an awake sphere moving into it.  The reference wakes along a
    row within one step, since each woken sphere keeps its velocity and the
    pairs come in index order (world.cpp:527-531).  Every sleeper carries its
    stored inverse mass and inertia and a nonzero timer."""
    rng = np.random.default_rng(seed)
    bodies = []
    r = 0.5
    for row in range(rows):
        y = -4.0 + 2.5 * row
        for k in range(chain + 1):
            b = make_sphere(len(bodies), (-5.0 + 0.98 * k, y, rng.uniform(-0.01, 0.01)), r, 1.0)
            b.v = (drift + 0.5 * k, 0.0, 0.0) if k else (3.0, 0.0, 0.0)
            bodies.append(b)
    s = Scene.from_bodies(bodies)
    n = s.n
    s.sleep_timer = np.zeros(n)
    s.inv_mass_store = np.zeros(n)
    s.inv_inertia_store = np.zeros((n, 3))
    for i in range(n):
        if i % (chain + 1) == 0:
            continue
        s.inv_mass_store[i] = s.inv_mass[i]
        s.inv_inertia_store[i] = s.inv_inertia[i]
        s.inv_mass[i] = 0.0
        s.inv_inertia[i] = 0.0
        s.flags[i] |= _abi.FLAG_SLEEPING
        s.sleep_timer[i] = 0.75
    s.params.gravity = (0.0, 0.0, 0.0)
    s.params.half = (8.0, 8.0, 8.0)
    s.params.sleep_enabled = True
    return s


# ---------------------------------------------------------------------------
# Heightfields (physics_world.hpp:96-109) and scenes on them
# ---------------------------------------------------------------------------

def value_noise(rows: int, cols: int, seed: int = 0, octaves: int = 4, amplitude: float = 1.0,
                base_cells: int = 4) -> np.ndarray:
    """Seeded value noise on a rows x cols grid: `octaves` layers of smoothly
    interpolated random lattice values, each of twice the frequency and half
    the amplitude of the last; the sum is scaled to [-amplitude, amplitude]."""
    rng = np.random.default_rng(seed)
    out = np.zeros((rows, cols))
    amp, cells, total = 1.0, base_cells, 0.0
    for _ in range(octaves):
        lat = rng.uniform(-1.0, 1.0, (cells + 2, cells + 2))
        u = np.linspace(0.0, cells, rows)
        v = np.linspace(0.0, cells, cols)
        iu, iv = np.minimum(u.astype(int), cells - 1), np.minimum(v.astype(int), cells - 1)
        fu, fv = u - iu, v - iv
        su, sv = fu * fu * (3 - 2 * fu), fv * fv * (3 - 2 * fv)
        a = lat[iu][:, iv]
        b = lat[iu + 1][:, iv]
        c = lat[iu][:, iv + 1]
        d = lat[iu + 1][:, iv + 1]
        top = a + (b - a) * su[:, None]
        bot = c + (d - c) * su[:, None]
        out += amp * (top + (bot - top) * sv[None, :])
        total += amp
        amp *= 0.5
        cells *= 2
    return out * (amplitude / total)


def heightfield_mesh(origin, spacing: float, heights: np.ndarray):
    """Vertices and triangles of PhysicsWorld::addHeightfield
    (physics_world.hpp:96-109): vertex (r, c) at origin + (c*spacing,
    heights[r, c], r*spacing), two triangles per cell, (i00, i10, i11) and
    (i00, i11, i01)."""
    rows, cols = heights.shape
    o = np.asarray(origin, np.float64)
    r, c = np.meshgrid(np.arange(rows), np.arange(cols), indexing="ij")
    verts = np.empty((rows * cols, 3))
    verts[:, 0] = o[0] + (c.ravel() * spacing)
    verts[:, 1] = o[1] + heights.ravel()
    verts[:, 2] = o[2] + (r.ravel() * spacing)
    rr, cc = np.meshgrid(np.arange(rows - 1), np.arange(cols - 1), indexing="ij")
    i00 = (rr * cols + cc).ravel()
    i01, i10 = i00 + 1, i00 + cols
    i11 = i10 + 1
    tris = np.stack([i00, i10, i11, i00, i11, i01], axis=1).reshape(-1, 3).astype(np.int32)
    return verts, tris


def mesh_body(index: int) -> Body:
    """makeMesh (body.hpp:319-330): a static mesh body at the origin."""
    b = Body(shape=_abi.SHAPE_MESH, id=index, x=(0.0, 0.0, 0.0))
    b.inv_mass = 0.0
    b.inv_inertia = (0.0, 0.0, 0.0)
    b.flags = 0
    return b


def terrain_pile(n: int = 120, seed: int = 0, rows: int = 24, cols: int = 24, spacing: float = 0.5,
                 amplitude: float = 1.0, half: float = 8.0, periodic: bool = False, shapes: str = "mixed",
                 drop: float = 1.2, spread: float | None = None) -> Scene:
    """Bodies dropped onto a value-noise heightfield (the mesh is body 0, so
    every pair with it is (0, j)): spheres, boxes, capsules and hulls, or one
    shape kind, placed at random over the terrain within `spread` above it
    (default 4 + 0.05 n).  The terrain is centred under the bodies and lies
    within the box."""
    rng = np.random.default_rng(seed)
    h = value_noise(rows, cols, seed=seed + 101, amplitude=amplitude)
    origin = (-(cols - 1) * spacing / 2.0, -half + amplitude + 1.0, -(rows - 1) * spacing / 2.0)
    verts, tris = heightfield_mesh(origin, spacing, h)
    bodies = [mesh_body(0)]
    xs = (cols - 1) * spacing / 2.0 - 0.6
    zs = (rows - 1) * spacing / 2.0 - 0.6
    top = origin[1] + amplitude + drop
    kinds = {"mixed": ["sphere", "box", "capsule", "hull"], "sphere": ["sphere"], "box": ["box"],
             "capsule": ["capsule"], "hull": ["hull"]}[shapes]
    qs = random_unit_quaternions(rng, n)
    placed = []
    tries = 0
    while len(bodies) < n + 1 and tries < 200 * n:
        tries += 1
        p = np.array([rng.uniform(-xs, xs), top + rng.uniform(0.0, 4.0 + 0.05 * n if spread is None else spread),
                      rng.uniform(-zs, zs)])
        if any(np.linalg.norm(p - q) < 0.95 for q in placed):
            continue
        placed.append(p)
        i = len(bodies)
        k = kinds[i % len(kinds)]
        q = tuple(qs[i - 1])
        if k == "sphere":
            b = make_sphere(i, tuple(p), rng.uniform(0.2, 0.4), 1.0)
        elif k == "box":
            b = make_box(i, tuple(p), q, tuple(rng.uniform(0.15, 0.35, 3)), 1.0)
        elif k == "capsule":
            b = make_capsule(i, tuple(p), q, 0.15, 0.2, 1.0)
        else:
            b = make_convex(i, tuple(p), q, hull12(rng, 0.3), 1.0)
        b.v = tuple(rng.uniform(-0.5, 0.5, 3))
        bodies.append(b)
    s = Scene.from_bodies(bodies)
    s.meshes = [dict(body=0, verts=verts, tris=tris)]
    s.params.half = (half, half, half)
    s.params.periodic = periodic
    s.params.gravity = (0.0, -9.81, 0.0)
    s.params.friction = 0.6
    s.params.restitution = 0.1
    return s


def triangle_soup(ntris: int = 300, seed: int = 0, lattice: bool = True, degenerate: int = 5):
    """Vertices and triangles for tree checks: on a coarse integer lattice
    (many equal split keys) or continuous, plus a few degenerate
    (collinear) triangles."""
    rng = np.random.default_rng(seed)
    nv = max(3, ntris // 2)
    if lattice:
        verts = rng.integers(-4, 5, (nv, 3)).astype(np.float64)
    else:
        verts = rng.uniform(-4.0, 4.0, (nv, 3))
    tris = np.stack([rng.choice(nv, 3, replace=False) for _ in range(ntris)]).astype(np.int32)
    if degenerate:
        extra = []
        for _ in range(degenerate):
            a, b = rng.uniform(-3, 3, 3), rng.uniform(-3, 3, 3)
            base = len(verts) + 3 * len(extra)
            extra.append([a, b, 0.5 * (a + b)])
        verts = np.concatenate([verts, np.asarray(extra).reshape(-1, 3)])
        dt = np.arange(len(extra) * 3, dtype=np.int32).reshape(-1, 3) + nv
        tris = np.concatenate([tris, dt])
    return verts, tris


def tilted_mesh_pile(n: int = 40, seed: int = 0, half: float = 8.0, periodic: bool = False) -> Scene:
    """Bodies on a heightfield mesh whose body is moved off the origin and
    tilted (PhysicsWorld::addMesh with a pose, physics_world.hpp:83-93), so the
    queries run in the mesh frame and the triangles are transformed."""
    s = terrain_pile(n, seed=seed, rows=16, cols=16, spacing=0.6, amplitude=0.6, half=half, periodic=periodic)
    ang = 0.15
    s.x[0] = (0.3, 0.4, -0.2)
    s.q[0] = (math.cos(ang / 2), 0.0, 0.0, math.sin(ang / 2))
    return s


def large_scene(n: int = 1 << 20, seed: int = 5, height: float = 100.0, amplitude: float = 8.0,
                iterations: int = 10, sleep: bool = True) -> Scene:
    """The large terrain scene: `n` bodies (the heightfield mesh is body 0) in
    a periodic side x height x side domain over a static heightfield.

    At n = 2^20 the domain is 400 x 100 x 400 and the heightfield 1024 x 1024
    vertices; smaller n shrink side and vertex count with sqrt(n) so the
    density stays the same.  Height is 4 octaves of seeded value noise of
    amplitude `amplitude`.  Of the n - 1 dynamic bodies, 70% are boxes in
    stacked columns of 8 (half extents U[0.25,0.5] across, U[0.15,0.35]
    tall, each column on the terrain), 5% are spheres of radius 0.2 in
    16-link ropes joined by rigid distance joints (surface anchors, rest
    0.1), and the rest are spheres of radius U[0.2,0.5] on a lattice above
    the columns with velocities N(0, 0.5).  Density 1, friction 0.6,
    restitution 0.1, gravity -9.81, dt 1/60, island sleeping on,
    `iterations` solver iterations."""
    from scipy.ndimage import maximum_filter
    rng = np.random.default_rng(seed)
    scale = math.sqrt(n / float(1 << 20))
    side = 400.0 * scale
    nv = max(8, int(round(1024 * scale)))
    ndyn = n - 1
    n_rope = (int(0.05 * ndyn) // 16) * 16
    n_box = (int(0.70 * ndyn) // 8) * 8
    n_sph = ndyn - n_rope - n_box
    hy = height / 2.0

    # terrain: vertex (r, c) at (-side/2 + c*sp, base + h, -side/2 + r*sp)
    sp_t = side / (nv - 1)
    h = value_noise(nv, nv, seed=seed + 1000, amplitude=amplitude)
    base = -hy + amplitude + 2.0
    verts, tris = heightfield_mesh((-side / 2.0, base, -side / 2.0), sp_t, h)
    reach = int(math.ceil(0.8 / sp_t))
    hmax = maximum_filter(h, size=2 * reach + 1, mode="nearest")

    s = Scene(n)
    s.shape[0] = _abi.SHAPE_MESH
    s.inv_mass[0] = 0.0
    s.flags[0] = 0
    s.radius[0] = 1.0

    # box columns on a lattice of sites, each sitting on the terrain
    ncol = n_box // 8
    m = int(math.ceil(math.sqrt(ncol)))
    sp_c = side / m
    sites = rng.permutation(m * m)[:ncol]
    cx = -side / 2.0 + (sites % m + 0.5) * sp_c + rng.uniform(-0.1, 0.1, ncol)
    cz = -side / 2.0 + (sites // m + 0.5) * sp_c + rng.uniform(-0.1, 0.1, ncol)
    ci = np.clip(np.rint((cz + side / 2.0) / sp_t).astype(int), 0, nv - 1)
    cj = np.clip(np.rint((cx + side / 2.0) / sp_t).astype(int), 0, nv - 1)
    ground = base + hmax[ci, cj]
    he = np.empty((ncol, 8, 3))
    he[:, :, 0] = rng.uniform(0.25, 0.5, (ncol, 8))
    he[:, :, 1] = rng.uniform(0.15, 0.35, (ncol, 8))
    he[:, :, 2] = rng.uniform(0.25, 0.5, (ncol, 8))
    tops = np.cumsum(2.0 * he[:, :, 1] + 0.02, axis=1)
    cy = ground[:, None] + 0.05 + tops - he[:, :, 1] - 0.02
    b0, b1 = 1, 1 + n_box
    s.shape[b0:b1] = _abi.SHAPE_BOX
    s.x[b0:b1, 0] = np.repeat(cx, 8)
    s.x[b0:b1, 1] = cy.ravel()
    s.x[b0:b1, 2] = np.repeat(cz, 8)
    s.half_extents[b0:b1] = he.reshape(-1, 3)
    s.inv_mass[b0:b1], s.inv_inertia[b0:b1] = _box_mass(he.reshape(-1, 3), 1.0)

    # loose spheres on a lattice above the columns
    col_top = float(ground.max() + tops[:, -1].max()) + 1.5
    ms = int(side / 1.2)
    per_layer = ms * ms
    layers = -(-n_sph // per_layer)
    slots = rng.permutation(per_layer * layers)[:n_sph]
    lay, rem = slots // per_layer, slots % per_layer
    s0, s1 = b1, b1 + n_sph
    r = rng.uniform(0.2, 0.5, n_sph)
    s.shape[s0:s1] = _abi.SHAPE_SPHERE
    s.radius[s0:s1] = r
    s.x[s0:s1, 0] = -side / 2.0 + (rem % ms + 0.5) * (side / ms) + rng.uniform(-0.05, 0.05, n_sph)
    s.x[s0:s1, 1] = col_top + 1.2 * lay + rng.uniform(-0.05, 0.05, n_sph)
    s.x[s0:s1, 2] = -side / 2.0 + (rem // ms + 0.5) * (side / ms) + rng.uniform(-0.05, 0.05, n_sph)
    s.v[s0:s1] = rng.normal(0.0, 0.5, (n_sph, 3))
    im, ii = _sphere_mass(r, 1.0)
    s.inv_mass[s0:s1] = im
    s.inv_inertia[s0:s1] = ii[:, None]

    # ropes: 16 spheres of radius 0.2 along x, 0.5 apart, above the spheres
    nr = n_rope // 16
    rx = max(1, int(side // 9.0))
    rz = max(1, int(side / 1.2))
    rope_y = col_top + 1.2 * layers + 1.5
    rlayers = -(-nr // (rx * rz))
    rslots = rng.permutation(rx * rz * rlayers)[:nr]
    rl, rrem = rslots // (rx * rz), rslots % (rx * rz)
    x0 = -side / 2.0 + (rrem % rx) * (side / rx) + 0.5
    z0 = -side / 2.0 + (rrem // rx + 0.5) * (side / rz)
    r0 = 1 + n_box + n_sph
    rr = 0.2
    k = np.arange(16)
    s.shape[r0:] = _abi.SHAPE_SPHERE
    s.radius[r0:] = rr
    s.x[r0:, 0] = (x0[:, None] + 0.5 * k[None, :]).ravel()
    s.x[r0:, 1] = np.repeat(rope_y + 1.2 * rl, 16)
    s.x[r0:, 2] = np.repeat(z0, 16)
    im, ii = _sphere_mass(np.full(n_rope, rr), 1.0)
    s.inv_mass[r0:] = im
    s.inv_inertia[r0:] = ii[:, None]
    a = (r0 + np.arange(nr)[:, None] * 16 + np.arange(15)[None, :]).ravel()
    j = np.zeros(len(a), _abi.JOINT_DTYPE)
    j["a"], j["b"] = a, a + 1
    j["local_a"] = (rr, 0.0, 0.0)
    j["local_b"] = (-rr, 0.0, 0.0)
    j["rest"] = 0.1
    s.joints = j

    s.meshes = [dict(body=0, verts=verts, tris=tris)]
    s.params = Params(half=(side / 2.0, hy, side / 2.0), periodic=True, dt=1.0 / 60.0, gravity=(0.0, -9.81, 0.0),
                      restitution=0.1, friction=0.6, solver_iterations=iterations, sleep_enabled=sleep)
    return s


# ---- general joints as PhysicsWorld::addGeneralJoint builds them
# (src/physics_world.hpp:191-203, 408-423), with the reference's V3 / Q
# arithmetic (src/vmath.hpp:24-77) in its operation order.

def _v_cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def _v_normalized(a):
    n = math.sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2])
    return (a[0] / n, a[1] / n, a[2] / n) if n > 1e-300 else (0.0, 0.0, 0.0)


def _q_rotate(q, v):
    w, qv = q[0], (q[1], q[2], q[3])
    c = _v_cross(qv, v)
    t = (c[0] * 2.0, c[1] * 2.0, c[2] * 2.0)
    c2 = _v_cross(qv, t)
    return ((v[0] + t[0] * w) + c2[0], (v[1] + t[1] * w) + c2[1], (v[2] + t[2] * w) + c2[2])


def _q_conj(q):
    return (q[0], -q[1], -q[2], -q[3])


def _q_mul(a, o):
    return (a[0] * o[0] - a[1] * o[1] - a[2] * o[2] - a[3] * o[3],
            a[0] * o[1] + a[1] * o[0] + a[2] * o[3] - a[3] * o[2],
            a[0] * o[2] - a[1] * o[3] + a[2] * o[0] + a[3] * o[1],
            a[0] * o[3] + a[1] * o[2] - a[2] * o[1] + a[3] * o[0])


def general_joint(s: Scene, kind: int, a: int, b: int, anchor=None, axis=(0.0, 1.0, 0.0), motor=None, limit=None,
                  break_impulse=None):
    """One ne_joint record between bodies a and b of scene s at their
    current poses: a world anchor (a fixed joint's defaults to the midpoint
    of the centres, addFixedJoint) and axis, converted to each body's frame,
    with ref_rel = qA * conj(qB); motor (speed, max impulse), limit
    (lower, upper), break_impulse as setJointMotor / setJointLimit /
    setJointBreakable."""
    xa, xb = tuple(float(v) for v in s.x[a]), tuple(float(v) for v in s.x[b])
    qa, qb = tuple(float(v) for v in s.q[a]), tuple(float(v) for v in s.q[b])
    if anchor is None:
        anchor = ((xa[0] + xb[0]) * 0.5, (xa[1] + xb[1]) * 0.5, (xa[2] + xb[2]) * 0.5)
    anchor = tuple(float(v) for v in anchor)
    j = np.zeros((), _abi.GJOINT_DTYPE)
    j["type"], j["a"], j["b"] = kind, a, b
    j["local_a"] = _q_rotate(_q_conj(qa), (anchor[0] - xa[0], anchor[1] - xa[1], anchor[2] - xa[2]))
    j["local_b"] = _q_rotate(_q_conj(qb), (anchor[0] - xb[0], anchor[1] - xb[1], anchor[2] - xb[2]))
    ax = _v_normalized(tuple(float(v) for v in axis))
    j["axis_a"] = _q_rotate(_q_conj(qa), ax)
    j["axis_b"] = _q_rotate(_q_conj(qb), ax)
    j["ref_rel"] = _q_mul(qa, _q_conj(qb))
    if motor is not None:
        j["use_motor"], j["motor_speed"], j["max_motor_impulse"] = 1, float(motor[0]), float(motor[1])
    if limit is not None:
        j["use_limit"], j["lower"], j["upper"] = 1, float(limit[0]), float(limit[1])
    if break_impulse is not None:
        j["breakable"], j["break_impulse"] = 1, float(break_impulse)
    return j
