"""Python binding of the HIP step (include/ne_gpu.h) over ctypes.

`GpuWorld` is a resident world: upload once, step many times, download when
needed.  There is no CPU fallback: if the HIP library is missing or no device
is visible, constructing a GpuWorld raises.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import sys
import warnings
import weakref

import numpy as np

from . import _abi
from . import queries
from .scene import Params, Scene

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnegpu.so")

_lib = None

# every entry point declared in include/ne_gpu.h
EXPORTS = [
    "ne_gpu_create", "ne_gpu_destroy", "ne_gpu_set_params", "ne_gpu_set_bodies", "ne_gpu_set_state",
    "ne_gpu_step", "ne_gpu_get_state", "ne_gpu_get_body_aux", "ne_gpu_reset_warm_cache", "ne_gpu_get_contacts",
    "ne_gpu_get_pairs", "ne_gpu_diagnostics", "ne_gpu_get_counters", "ne_gpu_get_stream", "ne_gpu_last_error",
    "ne_gpu_fp_probe", "ne_gpu_time_kernel", "ne_gpu_kernel_time", "ne_gpu_export_warm_cache",
    "ne_gpu_import_warm_cache", "ne_gpu_set_worlds", "ne_gpu_use_graph", "ne_gpu_autotune", "ne_gpu_epa_stats",
    "ne_gpu_rollout", "ne_gpu_rollout_backward", "ne_gpu_set_joints", "ne_gpu_set_meshes", "ne_gpu_mesh_bvh",
    "ne_gpu_get_sleep_store", "ne_gpu_save_state", "ne_gpu_load_state", "ne_gpu_save_state_noalloc",
    "ne_gpu_set_events", "ne_gpu_get_events", "ne_gpu_raycast", "ne_gpu_raycast_all", "ne_gpu_overlap",
    "ne_gpu_set_general_joints", "ne_gpu_get_general_joints", "ne_gpu_epa_queues",
]

KERNELS = {"integrate": 1, "cell_count": 2, "cell_fill": 3, "cand_count": 4, "cand_write": 5, "narrow": 6,
           "cons_build": 7, "solve": 8, "boundary": 9, "step": 10, "ray": 11}


class NeGpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ne_gpu error {code}: {msg}")
        self.code = code


class EpaCappedWarning(RuntimeWarning):
    """A step completed, but an EPA run stopped at the workspace's face cap
    (NE_WARN_EPA_CAPPED): its contact can differ from the reference's."""


# Contexts still open at interpreter exit are destroyed by an exit handler,
# while the HIP runtime is still up; finalizers that run later skip the call.
_live = weakref.WeakSet()


def _close_all():
    for w in list(_live):
        try:
            w.close()
        except Exception:
            pass


atexit.register(_close_all)


def load_library(path: str = LIB_PATH):
    """Loads libnegpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"HIP library not built: {path} (run __graft_entry__.build())")
    lib = C.CDLL(path)
    vp = C.c_void_p
    u32p = C.POINTER(C.c_uint32)
    dp = C.POINTER(C.c_double)
    lib.ne_gpu_create.argtypes = [C.POINTER(_abi.GpuConfig), C.POINTER(vp)]
    lib.ne_gpu_destroy.argtypes = [vp]
    lib.ne_gpu_set_params.argtypes = [vp, C.POINTER(_abi.WorldParams)]
    lib.ne_gpu_set_bodies.argtypes = [vp, C.POINTER(_abi.Bodies)]
    lib.ne_gpu_set_state.argtypes = [vp, dp, dp, dp, dp]
    lib.ne_gpu_step.argtypes = [vp, C.c_int32]
    lib.ne_gpu_get_state.argtypes = [vp, dp, dp, dp, dp]
    lib.ne_gpu_get_body_aux.argtypes = [vp, u32p, dp, dp, dp]
    lib.ne_gpu_get_sleep_store.argtypes = [vp, dp, dp]
    lib.ne_gpu_set_meshes.argtypes = [vp, C.c_uint32, C.POINTER(_abi.Mesh)]
    lib.ne_gpu_mesh_bvh.argtypes = [vp, C.c_uint32, C.POINTER(C.c_int32), C.c_uint32, u32p, dp, C.POINTER(C.c_int32)]
    lib.ne_gpu_reset_warm_cache.argtypes = [vp]
    lib.ne_gpu_save_state.argtypes = [vp, vp, C.c_uint64, C.POINTER(C.c_uint64)]
    lib.ne_gpu_load_state.argtypes = [vp, vp, C.c_uint64]
    lib.ne_gpu_save_state_noalloc.argtypes = [vp, vp, C.c_uint64, C.POINTER(C.c_uint64)]
    lib.ne_gpu_set_events.argtypes = [vp, C.c_int32]
    lib.ne_gpu_get_events.argtypes = [vp, C.c_int32, C.c_uint32, u32p, u32p, u32p, dp, dp, dp]
    lib.ne_gpu_get_contacts.argtypes = [vp, C.c_uint32, u32p, u32p, u32p, dp, dp, dp, dp]
    lib.ne_gpu_get_pairs.argtypes = [vp, C.c_uint32, u32p, u32p, u32p]
    lib.ne_gpu_diagnostics.argtypes = [vp, dp, dp, dp]
    lib.ne_gpu_get_counters.argtypes = [vp, u32p]
    lib.ne_gpu_get_stream.argtypes = [vp, C.POINTER(vp)]
    lib.ne_gpu_last_error.argtypes = [vp]
    lib.ne_gpu_last_error.restype = C.c_char_p
    lib.ne_gpu_fp_probe.argtypes = [vp, C.c_int32, C.c_uint32, dp, dp, dp]
    lib.ne_gpu_time_kernel.argtypes = [vp, C.c_int32]
    u64p = C.POINTER(C.c_uint64)
    lib.ne_gpu_export_warm_cache.argtypes = [vp, C.c_uint32, u32p, u64p, dp]
    lib.ne_gpu_import_warm_cache.argtypes = [vp, C.c_uint32, u64p, dp]
    lib.ne_gpu_kernel_time.argtypes = [vp, C.POINTER(C.c_double), u32p]
    lib.ne_gpu_set_worlds.argtypes = [vp, C.c_uint32, u32p, dp, dp, dp]
    lib.ne_gpu_use_graph.argtypes = [vp, C.c_int32]
    lib.ne_gpu_epa_stats.argtypes = [vp, u32p, u32p]
    lib.ne_gpu_epa_queues.argtypes = [vp, u32p]
    lib.ne_gpu_set_joints.argtypes = [vp, C.c_uint32, vp]
    lib.ne_gpu_rollout.argtypes = [vp, vp, dp, dp, C.c_int32, C.c_int32, vp]
    lib.ne_gpu_rollout_backward.argtypes = [vp, vp, vp, dp, dp]
    lib.ne_gpu_autotune.argtypes = [vp, C.c_int32, C.POINTER(_abi.TuneResult)]
    queries.bind(lib, "ne_gpu")
    lib.ne_gpu_set_general_joints.argtypes = [vp, C.c_uint32, vp]
    lib.ne_gpu_get_general_joints.argtypes = [vp, C.c_uint32, C.POINTER(C.c_int32), dp, dp]
    for name in EXPORTS:
        if name != "ne_gpu_last_error":
            getattr(lib, name).restype = C.c_int
    _lib = lib
    return lib


class GpuWorld:
    """A world resident on one MI355X."""

    def __init__(self, device: int = 0, pair_capacity: int = 0, contact_capacity: int = 0, cell_scale: float = 0.0,
                 broadphase: int = 0, cells_per_body: int = 0, epa_face_cap: int = 0, strict: bool = True):
        """strict: a step whose EPA run reached the face cap (NE_WARN_EPA_CAPPED)
        raises NeGpuError; otherwise it warns (EpaCappedWarning) and counts
        the call in `epa_capped_calls`."""
        self.lib = load_library()
        cfg = _abi.GpuConfig(device, pair_capacity, contact_capacity, cell_scale, broadphase, cells_per_body,
                             epa_face_cap)
        h = C.c_void_p()
        rc = self.lib.ne_gpu_create(C.byref(cfg), C.byref(h))
        if rc != 0:
            raise NeGpuError(rc, "ne_gpu_create failed (no HIP device?)")
        self.h = h
        self.n = 0
        self.params = Params()
        self.strict = strict
        self.epa_capped_calls = 0
        _live.add(self)

    def _check(self, rc):
        if rc != 0:
            msg = self.lib.ne_gpu_last_error(self.h).decode(errors="replace")
            raise NeGpuError(rc, msg)

    def close(self):
        if getattr(self, "h", None):
            self.lib.ne_gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        if sys.is_finalizing():   # the exit handler has closed it, or the runtime is going away
            return
        try:
            self.close()
        except Exception:
            pass

    # -- upload -----------------------------------------------------------
    def set_params(self, p: Params):
        self.params = p
        cp = p.to_c()
        self._check(self.lib.ne_gpu_set_params(self.h, C.byref(cp)))

    def set_scene(self, s: Scene):
        self.set_params(s.params)
        b, keep = s.to_c()
        self._check(self.lib.ne_gpu_set_bodies(self.h, C.byref(b)))
        self.n = s.n
        self._ngj = 0
        del keep
        if len(s.joints):
            self.set_joints(s.joints)
        if len(getattr(s, "gjoints", ())):
            self.set_general_joints(s.gjoints)
        if s.meshes:
            arr, keep = s.meshes_c()
            self._check(self.lib.ne_gpu_set_meshes(self.h, len(s.meshes), arr))
            del keep

    def mesh_bvh(self, m, ntris):
        """(order, bounds (nn, 6), links (nn, 4)) of mesh m's tree, as built here."""
        nn = C.c_uint32(0)
        self._check(self.lib.ne_gpu_mesh_bvh(self.h, m, None, 0, C.byref(nn), None, None))
        order = np.zeros(ntris, np.int32)
        bounds = np.zeros((nn.value, 6))
        links = np.zeros((nn.value, 4), np.int32)
        ip = C.POINTER(C.c_int32)
        self._check(self.lib.ne_gpu_mesh_bvh(self.h, m, order.ctypes.data_as(ip), nn.value, C.byref(nn),
                                             _abi.dptr(bounds), links.ctypes.data_as(ip)))
        return order, bounds, links

    def set_general_joints(self, joints):
        """World::joints: ball / hinge / fixed / slider joints (records of
        _abi.GJOINT_DTYPE, e.g. from scene.general_joint), solved in order."""
        j = np.ascontiguousarray(joints, _abi.GJOINT_DTYPE)
        self._check(self.lib.ne_gpu_set_general_joints(self.h, len(j), j.ctypes.data if len(j) else None))
        self._ngj = len(j)

    def general_joints(self):
        """(broken, applied impulse, motor accumulator) of the general joints
        after the last step."""
        n = getattr(self, "_ngj", 0)
        br, ap, mo = np.zeros(n, np.int32), np.zeros(n), np.zeros(n)
        self._check(self.lib.ne_gpu_get_general_joints(self.h, n, br.ctypes.data_as(C.POINTER(C.c_int32)),
                                                       _abi.dptr(ap), _abi.dptr(mo)))
        return br, ap, mo

    def set_joints(self, joints):
        """World::distanceJoints: a JOINT_DTYPE record array, in order."""
        j = np.ascontiguousarray(joints, _abi.JOINT_DTYPE)
        self._check(self.lib.ne_gpu_set_joints(self.h, len(j), j.ctypes.data if len(j) else None))

    def set_worlds(self, offsets, half, restitution=None, friction=None):
        """Batched worlds: bodies [offsets[w], offsets[w+1]) form periodic world w
        with box half extents half[w] (3) and default restitution / friction."""
        off = np.ascontiguousarray(offsets, np.uint32)
        nw = len(off) - 1
        hh = np.ascontiguousarray(np.broadcast_to(np.asarray(half, np.float64).reshape(-1, 3) if np.ndim(half) > 1
                                                  else np.asarray(half, np.float64).reshape(-1, 1), (nw, 3)))
        e = None if restitution is None else np.ascontiguousarray(restitution, np.float64)
        m = None if friction is None else np.ascontiguousarray(friction, np.float64)
        self._check(self.lib.ne_gpu_set_worlds(self.h, nw, _abi.uptr(off), _abi.dptr(hh), _abi.dptr(e), _abi.dptr(m)))

    def epa_stats(self):
        """(large-workspace EPA reruns, runs stopped at the largest workspace)
        during the last step call."""
        a, b = C.c_uint32(), C.c_uint32()
        self._check(self.lib.ne_gpu_epa_stats(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    EPA_QUEUES = ("overlapping_pairs", "past_tier_a", "past_tier_b", "past_tier_c", "block_tier", "mesh_overlapping",
                  "mesh_past_tier_a", "mesh_past_group_tier")

    def epa_queues(self):
        """The last step's EPA tier queue lengths (ne_gpu_epa_queues), by name."""
        out = np.zeros(8, np.uint32)
        self._check(self.lib.ne_gpu_epa_queues(self.h, _abi.uptr(out)))
        return dict(zip(self.EPA_QUEUES, map(int, out)))

    def rollout(self, v0=None, mu=None, e=None, nsteps: int = 120, every: int = 16):
        """Differentiable rollout from the current state (velocities v0, n x 3,
        when given; mu, e: one value per world or None).  Returns the final
        positions (n x 3)."""
        v = None if v0 is None else np.ascontiguousarray(v0, np.float64)
        m = None if mu is None else np.ascontiguousarray(mu, np.float64)
        r = None if e is None else np.ascontiguousarray(e, np.float64)
        x = np.empty((self.n, 3))
        self._check_step(self.lib.ne_gpu_rollout(self.h, None if v is None else v.ctypes.data, _abi.dptr(m),
                                                 _abi.dptr(r), int(nsteps), int(every), x.ctypes.data))
        return x

    def rollout_backward(self, gx, nworlds: int):
        """Adjoint of the last rollout for d loss / d x_final = gx (n x 3):
        returns (d loss / d v0, d loss / d mu per world, d loss / d e per world)."""
        g = np.ascontiguousarray(gx, np.float64)
        gv = np.empty((self.n, 3))
        gmu = np.zeros(nworlds)
        ge = np.zeros(nworlds)
        self._check(self.lib.ne_gpu_rollout_backward(self.h, g.ctypes.data, gv.ctypes.data, _abi.dptr(gmu),
                                                     _abi.dptr(ge)))
        return gv, gmu, ge

    def use_graph(self, enable: bool = True):
        self._check(self.lib.ne_gpu_use_graph(self.h, 1 if enable else 0))

    def autotune(self, steps: int = 3):
        r = _abi.TuneResult()
        self._check(self.lib.ne_gpu_autotune(self.h, int(steps), C.byref(r)))
        return dict(cell_scale=r.cell_scale, broadphase=r.broadphase, graph=r.graph, ms_per_step=r.ms_per_step,
                    trials=r.trials)

    def set_state(self, x=None, v=None, q=None, w=None):
        arrs = [None if a is None else np.ascontiguousarray(a, np.float64) for a in (x, v, q, w)]
        self._check(self.lib.ne_gpu_set_state(self.h, *[None if a is None else _abi.dptr(a) for a in arrs]))

    # -- step -------------------------------------------------------------
    def _check_step(self, rc):
        """A step or rollout status: NE_WARN_EPA_CAPPED raises in strict mode,
        else warns and is counted."""
        if rc == _abi.NE_WARN_EPA_CAPPED:
            self.epa_capped_calls += 1
            msg = self.lib.ne_gpu_last_error(self.h).decode(errors="replace")
            if self.strict:
                raise NeGpuError(rc, msg)
            warnings.warn(msg, EpaCappedWarning, stacklevel=3)
            return
        self._check(rc)

    def step(self, nsteps: int = 1):
        self._check_step(self.lib.ne_gpu_step(self.h, int(nsteps)))

    def reset_warm_cache(self):
        self._check(self.lib.ne_gpu_reset_warm_cache(self.h))

    def export_warm_cache(self):
        cnt = C.c_uint32(0)
        self._check(self.lib.ne_gpu_export_warm_cache(self.h, 0, C.byref(cnt), None, None))
        keys = np.zeros(cnt.value, np.uint64)
        imp = np.zeros(cnt.value)
        if cnt.value:
            self._check(self.lib.ne_gpu_export_warm_cache(self.h, cnt.value, C.byref(cnt),
                                                          keys.ctypes.data_as(C.POINTER(C.c_uint64)), _abi.dptr(imp)))
        return keys, imp

    def import_warm_cache(self, keys, imp):
        keys = np.ascontiguousarray(keys, np.uint64)
        imp = np.ascontiguousarray(imp, np.float64)
        self._check(self.lib.ne_gpu_import_warm_cache(self.h, len(keys), keys.ctypes.data_as(C.POINTER(C.c_uint64)),
                                                      _abi.dptr(imp)))

    # -- download ---------------------------------------------------------
    def state(self):
        n = self.n
        x, v, w = (np.zeros((n, 3)) for _ in range(3))
        q = np.zeros((n, 4))
        self._check(self.lib.ne_gpu_get_state(self.h, _abi.dptr(x), _abi.dptr(v), _abi.dptr(q), _abi.dptr(w)))
        return x, v, q, w

    def body_aux(self):
        """(flags, sleep timer, stored inverse mass, stored inverse inertia (n, 3))
        of every body: Body::flags, sleepTimer, invMassStore, invIStore."""
        n = self.n
        fl, tm, im, ii = np.zeros(n, np.uint32), np.zeros(n), np.zeros(n), np.zeros((n, 3))
        self._check(self.lib.ne_gpu_get_body_aux(self.h, _abi.uptr(fl), _abi.dptr(tm), _abi.dptr(im), _abi.dptr(ii)))
        return fl, tm, im, ii

    def save_state(self) -> bytes:
        """The bodies as the reference's NES1 snapshot (PhysicsWorld::saveState)."""
        size = C.c_uint64(0)
        self._check(self.lib.ne_gpu_save_state(self.h, None, 0, C.byref(size)))
        buf = C.create_string_buffer(size.value)
        self._check(self.lib.ne_gpu_save_state(self.h, buf, size.value, C.byref(size)))
        return buf.raw[: size.value]

    def save_state_noalloc(self) -> bytes:
        """save_state through ne_gpu_save_state_noalloc: the same bytes, staged
        in the library's static device buffer (no device allocation)."""
        size = C.c_uint64(0)
        self._check(self.lib.ne_gpu_save_state_noalloc(self.h, None, 0, C.byref(size)))
        buf = C.create_string_buffer(size.value)
        self._check(self.lib.ne_gpu_save_state_noalloc(self.h, buf, size.value, C.byref(size)))
        return buf.raw[: size.value]

    def load_state(self, blob: bytes):
        """Restores an NES1 snapshot (PhysicsWorld::loadState); resets the warm cache."""
        buf = C.create_string_buffer(bytes(blob), len(blob))
        self._check(self.lib.ne_gpu_load_state(self.h, buf, len(blob)))

    def set_events(self, enable: bool = True):
        """Facade contact / trigger events after every step
        (PhysicsWorld::processContacts); one step per call while on."""
        self._check(self.lib.ne_gpu_set_events(self.h, 1 if enable else 0))

    def events(self, kind: int):
        """The last step's events of one kind (_abi.NE_EVENT_*): dict of i, j,
        point, normal, depth, in (i, j) order."""
        cnt = C.c_uint32(0)
        self._check(self.lib.ne_gpu_get_events(self.h, int(kind), 0, C.byref(cnt), None, None, None, None, None))
        m = cnt.value
        i, j = np.zeros(m, np.uint32), np.zeros(m, np.uint32)
        pt, nrm, dep = np.zeros((m, 3)), np.zeros((m, 3)), np.zeros(m)
        if m:
            self._check(self.lib.ne_gpu_get_events(self.h, int(kind), m, C.byref(cnt), _abi.uptr(i), _abi.uptr(j),
                                                   _abi.dptr(pt), _abi.dptr(nrm), _abi.dptr(dep)))
        return dict(i=i, j=j, point=pt, normal=nrm, depth=dep)

    # -- scene queries (PhysicsWorld, src/physics_world.cpp:146-251) --------
    def raycast(self, origin, direction, max_dist, mask=None) -> dict:
        """Nearest hit per ray (PhysicsWorld::raycast): dict of body (-1: no
        hit), distance, point, normal; origin / direction (n, 3), max_dist
        and mask per ray or scalar."""
        return queries.raycast(self.lib.ne_gpu_raycast, self.h, self._check, origin, direction, max_dist, mask)

    def sphere_cast(self, origin, radius, direction, max_dist, mask=None) -> dict:
        """Nearest hit per swept sphere (PhysicsWorld::sphereCast)."""
        return queries.raycast(self.lib.ne_gpu_raycast, self.h, self._check, origin, direction, max_dist, mask,
                               radius=radius)

    def raycast_all(self, origin, direction, max_dist, mask=None) -> dict:
        """Every hit per ray, sorted by distance (PhysicsWorld::raycastAll):
        CSR dict (offsets, body, distance, point, normal)."""
        return queries.raycast_all(self.lib.ne_gpu_raycast_all, self.h, self._check, origin, direction, max_dist,
                                   mask)

    def overlap(self, shapes, mask=None, contacts=False) -> dict:
        """Bodies overlapping each query shape (queries.Shape), in body order:
        CSR dict (offsets, body, normal, depth).  Sphere / box shapes without
        contacts are overlapSphere / overlapBox; with contacts,
        overlapContacts (sensors skipped, normal body -> query, depth)."""
        return queries.overlap(self.lib.ne_gpu_overlap, self.h, self._check_step, shapes, mask, contacts)

    def overlap_sphere(self, centers, radii, mask=None) -> list:
        """PhysicsWorld::overlapSphere per (center, radius): one index array each."""
        return queries.split(self.overlap(queries.spheres(centers, radii), mask))

    def overlap_box(self, centers, rots, halves, mask=None) -> list:
        """PhysicsWorld::overlapBox per (center, rot, half): one index array each."""
        return queries.split(self.overlap(queries.boxes(centers, rots, halves), mask))

    def sleep_store(self):
        """(Body::invMassStore (n), Body::invInertiaStore (n, 3)) of every body."""
        im, ii = np.zeros(self.n), np.zeros((self.n, 3))
        self._check(self.lib.ne_gpu_get_sleep_store(self.h, _abi.dptr(im), _abi.dptr(ii)))
        return im, ii

    def contacts(self):
        """Last step's constraint list: dict of i, j, normal, point, depth, impulse."""
        cnt = C.c_uint32(0)
        self._check(self.lib.ne_gpu_get_contacts(self.h, 0, C.byref(cnt), None, None, None, None, None, None))
        m = cnt.value
        i = np.zeros(m, np.uint32)
        j = np.zeros(m, np.uint32)
        nrm = np.zeros((m, 3))
        pt = np.zeros((m, 3))
        dep = np.zeros(m)
        imp = np.zeros(m)
        if m:
            self._check(self.lib.ne_gpu_get_contacts(self.h, m, C.byref(cnt), _abi.uptr(i), _abi.uptr(j), _abi.dptr(nrm),
                                                     _abi.dptr(pt), _abi.dptr(dep), _abi.dptr(imp)))
        return dict(i=i, j=j, normal=nrm, point=pt, depth=dep, impulse=imp)

    def pairs(self):
        cnt = C.c_uint32(0)
        self._check(self.lib.ne_gpu_get_pairs(self.h, 0, C.byref(cnt), None, None))
        m = cnt.value
        i = np.zeros(m, np.uint32)
        j = np.zeros(m, np.uint32)
        if m:
            self._check(self.lib.ne_gpu_get_pairs(self.h, m, C.byref(cnt), _abi.uptr(i), _abi.uptr(j)))
        return i, j

    def diagnostics(self):
        p = np.zeros(3)
        L = np.zeros(3)
        ke = C.c_double(0.0)
        self._check(self.lib.ne_gpu_diagnostics(self.h, _abi.dptr(p), _abi.dptr(L), C.byref(ke)))
        return p, L, ke.value

    def counters(self):
        out = np.zeros(8, np.uint32)
        self._check(self.lib.ne_gpu_get_counters(self.h, _abi.uptr(out)))
        keys = ["candidate_pairs", "mesh_triangles", "constraints", "islands", "max_island", "gridded", "cells", "fault"]
        return dict(zip(keys, map(int, out)))

    def stream(self) -> int:
        s = C.c_void_p()
        self._check(self.lib.ne_gpu_get_stream(self.h, C.byref(s)))
        return s.value or 0

    def time_kernel(self, name: str | None):
        """Bracket every launch of one step kernel with HIP events (None: off)."""
        self._check(self.lib.ne_gpu_time_kernel(self.h, KERNELS[name] if name else 0))

    def kernel_time(self):
        """(total ms, launches) of the timed kernel since time_kernel()."""
        ms = C.c_double(0.0)
        k = C.c_uint32(0)
        self._check(self.lib.ne_gpu_kernel_time(self.h, C.byref(ms), C.byref(k)))
        return ms.value, k.value

    def fp_probe(self, op: int, a, b=None):
        a = np.ascontiguousarray(a, np.float64)
        bb = None if b is None else np.ascontiguousarray(b, np.float64)
        out = np.zeros_like(a)
        self._check(self.lib.ne_gpu_fp_probe(self.h, int(op), len(a), _abi.dptr(a), _abi.dptr(bb), _abi.dptr(out)))
        return out
