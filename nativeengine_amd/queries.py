"""Scene queries, batched: PhysicsWorld::raycast, raycastAll, sphereCast,
overlapSphere, overlapBox and overlapContacts (src/physics_world.cpp:146-251).

These functions pack numpy arguments for an engine with the C signatures of
ne_gpu_raycast / ne_gpu_raycast_all / ne_gpu_overlap (include/ne_gpu.h) and
unpack its results; GpuWorld's query methods call them with the HIP
library's entry points.  Bodies are reported by index (the facade's handle
order); no hit is body -1.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _abi


@dataclass
class Shape:
    """A query body: a shape at a pose (the fields of ne_query_shape)."""
    shape: int
    x: tuple
    q: tuple = (1.0, 0.0, 0.0, 0.0)
    half_extents: tuple = (0.0, 0.0, 0.0)
    radius: float = 0.0
    half_height: float = 0.0
    verts: np.ndarray | None = field(default=None, repr=False)


def sphere(center, radius) -> Shape:
    """overlapSphere's query: makeSphere(-1, center, radius, 1.0)
    (src/physics_world.cpp:193)."""
    return Shape(_abi.SHAPE_SPHERE, tuple(float(v) for v in center), radius=float(radius))


def box(center, rot, half) -> Shape:
    """overlapBox's query: makeBox(-1, center, rot, half, 1.0)
    (src/physics_world.cpp:205); rot is (w, x, y, z), used as given."""
    return Shape(_abi.SHAPE_BOX, tuple(float(v) for v in center), q=tuple(float(v) for v in rot),
                 half_extents=tuple(float(v) for v in half))


def body_shape(scene, i: int, x=None, q=None) -> Shape:
    """Body i of a Scene as a query shape (overlapContacts' `query` Body),
    at its own pose unless x / q are given."""
    verts = None
    if scene.shape[i] == _abi.SHAPE_CONVEX:
        verts = np.ascontiguousarray(scene.hull_verts[scene.hull_offsets[i]:scene.hull_offsets[i + 1]], np.float64)
    return Shape(int(scene.shape[i]), tuple(float(v) for v in (scene.x[i] if x is None else x)),
                 q=tuple(float(v) for v in (scene.q[i] if q is None else q)),
                 half_extents=tuple(float(v) for v in scene.half_extents[i]), radius=float(scene.radius[i]),
                 half_height=float(scene.half_height[i]), verts=verts)


def _rays(origin, direction, max_dist, *extra):
    o = np.ascontiguousarray(np.atleast_2d(np.asarray(origin, np.float64)))
    d = np.ascontiguousarray(np.atleast_2d(np.asarray(direction, np.float64)))
    n = max(len(o), len(d))
    o = np.ascontiguousarray(np.broadcast_to(o, (n, 3)))
    d = np.ascontiguousarray(np.broadcast_to(d, (n, 3)))
    m = np.ascontiguousarray(np.broadcast_to(np.asarray(max_dist, np.float64), (n,)))
    out = [o, d, m]
    for a, dt in extra:
        out.append(None if a is None else np.ascontiguousarray(np.broadcast_to(np.asarray(a, dt), (n,))))
    return n, out


def raycast(fn, h, check, origin, direction, max_dist, mask=None, radius=None) -> dict:
    """Nearest hit per ray (sphere casts when radius is given): dict of
    body (int32, -1: none), distance, point (n, 3), normal (n, 3)."""
    n, (o, d, m, mk, rad) = _rays(origin, direction, max_dist, (mask, np.uint32), (radius, np.float64))
    body = np.zeros(n, np.int32)
    dist = np.zeros(n)
    pt, nrm = np.zeros((n, 3)), np.zeros((n, 3))
    check(fn(h, n, _abi.dptr(o), _abi.dptr(d), _abi.dptr(m), _abi.uptr(mk), _abi.dptr(rad), _abi.iptr(body),
             _abi.dptr(dist), _abi.dptr(pt), _abi.dptr(nrm)))
    return dict(body=body, distance=dist, point=pt, normal=nrm)


def raycast_all(fn, h, check, origin, direction, max_dist, mask=None, cap_guess=None) -> dict:
    """Every hit per ray within max_dist, ray k's at offsets[k]:offsets[k+1],
    sorted by distance: dict of offsets, body, distance, point, normal."""
    n, (o, d, m, mk) = _rays(origin, direction, max_dist, (mask, np.uint32))
    off = np.zeros(n + 1, np.uint32)
    cap = int(cap_guess if cap_guess is not None else 4 * n + 256)
    while True:
        body = np.zeros(cap, np.int32)
        dist = np.zeros(cap)
        pt, nrm = np.zeros((cap, 3)), np.zeros((cap, 3))
        rc = fn(h, n, _abi.dptr(o), _abi.dptr(d), _abi.dptr(m), _abi.uptr(mk), cap, _abi.uptr(off), _abi.iptr(body),
                _abi.dptr(dist), _abi.dptr(pt), _abi.dptr(nrm))
        if rc == _abi.NE_ERR_CAPACITY and int(off[n]) > cap:
            cap = int(off[n])
            continue
        check(rc)
        t = int(off[n])
        return dict(offsets=off, body=body[:t], distance=dist[:t], point=pt[:t], normal=nrm[:t])


SHAPE_DTYPE = np.dtype({"names": ["shape", "nverts", "x", "q", "half_extents", "radius", "half_height", "verts"],
                        "formats": ["<i4", "<u4", ("<f8", (3,)), ("<f8", (4,)), ("<f8", (3,)), "<f8", "<f8", "<u8"],
                        "offsets": [_abi.QueryShape.shape.offset, _abi.QueryShape.nverts.offset,
                                    _abi.QueryShape.x.offset, _abi.QueryShape.q.offset,
                                    _abi.QueryShape.half_extents.offset, _abi.QueryShape.radius.offset,
                                    _abi.QueryShape.half_height.offset, _abi.QueryShape.verts.offset],
                        "itemsize": C.sizeof(_abi.QueryShape)})


def spheres(centers, radii) -> np.ndarray:
    """Many overlapSphere queries at once, packed as ne_query_shape records."""
    c = np.atleast_2d(np.asarray(centers, np.float64))
    a = np.zeros(len(c), SHAPE_DTYPE)
    a["shape"] = _abi.SHAPE_SPHERE
    a["x"] = c
    a["q"] = (1.0, 0.0, 0.0, 0.0)
    a["radius"] = np.broadcast_to(np.asarray(radii, np.float64), (len(c),))
    return a


def boxes(centers, rots, halves) -> np.ndarray:
    """Many overlapBox queries at once, packed as ne_query_shape records."""
    c = np.atleast_2d(np.asarray(centers, np.float64))
    a = np.zeros(len(c), SHAPE_DTYPE)
    a["shape"] = _abi.SHAPE_BOX
    a["x"] = c
    a["q"] = np.broadcast_to(np.asarray(rots, np.float64), (len(c), 4))
    a["half_extents"] = np.broadcast_to(np.asarray(halves, np.float64), (len(c), 3))
    return a


def _shape_array(shapes):
    if isinstance(shapes, np.ndarray):   # packed records (spheres / boxes): no vertices to keep alive
        assert shapes.dtype == SHAPE_DTYPE and not np.any(shapes["nverts"])
        a = np.ascontiguousarray(shapes)
        return a.ctypes.data_as(C.POINTER(_abi.QueryShape)), [a]
    arr = (_abi.QueryShape * max(1, len(shapes)))()
    keep = []
    for k, s in enumerate(shapes):
        e = arr[k]
        e.shape = int(s.shape)
        e.x[:] = [float(v) for v in s.x]
        e.q[:] = [float(v) for v in s.q]
        e.half_extents[:] = [float(v) for v in s.half_extents]
        e.radius = float(s.radius)
        e.half_height = float(s.half_height)
        if s.verts is not None:
            v = np.ascontiguousarray(s.verts, np.float64).reshape(-1, 3)
            keep.append(v)
            e.nverts = len(v)
            e.verts = _abi.dptr(v)
    return arr, keep


def overlap(fn, h, check, shapes, mask=None, contacts=False, cap_guess=None) -> dict:
    """Bodies overlapping each query shape (a list of Shape, or records from
    spheres() / boxes()), query k's at
    offsets[k]:offsets[k+1] in body order (overlapSphere / overlapBox; with
    contacts, overlapContacts: sensors skipped, normal body -> query and
    depth): dict of offsets, body, normal, depth."""
    if not isinstance(shapes, np.ndarray):
        shapes = list(shapes)
    n = len(shapes)
    arr, keep = _shape_array(shapes)
    mk = None if mask is None else np.ascontiguousarray(np.broadcast_to(np.asarray(mask, np.uint32), (n,)))
    off = np.zeros(n + 1, np.uint32)
    cap = int(cap_guess if cap_guess is not None else 16 * n + 256)
    while True:
        body = np.zeros(cap, np.int32)
        nrm, dep = np.zeros((cap, 3)), np.zeros(cap)
        rc = fn(h, n, arr, _abi.uptr(mk), 1 if contacts else 0, cap, _abi.uptr(off), _abi.iptr(body),
                _abi.dptr(nrm), _abi.dptr(dep))
        if rc == _abi.NE_ERR_CAPACITY and int(off[n]) > cap:
            cap = int(off[n])
            continue
        check(rc)
        del keep
        t = int(off[n])
        return dict(offsets=off, body=body[:t], normal=nrm[:t], depth=dep[:t])


def split(res: dict, key: str = "body") -> list:
    """A CSR result as one array per ray / query."""
    off = res["offsets"]
    return [res[key][off[k]:off[k + 1]] for k in range(len(off) - 1)]


def bind(lib, prefix: str):
    """argtypes of the three query entry points of a library (prefix
    'ne_gpu', or a test engine's with the same signatures)."""
    vp, dp, u32p, i32p = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_uint32), C.POINTER(C.c_int32)
    getattr(lib, prefix + "_raycast").argtypes = [vp, C.c_uint32, dp, dp, dp, u32p, dp, i32p, dp, dp, dp]
    getattr(lib, prefix + "_raycast_all").argtypes = [vp, C.c_uint32, dp, dp, dp, u32p, C.c_uint32, u32p, i32p, dp,
                                                      dp, dp]
    getattr(lib, prefix + "_overlap").argtypes = [vp, C.c_uint32, C.POINTER(_abi.QueryShape), u32p, C.c_int32,
                                                  C.c_uint32, u32p, i32p, dp, dp]
