"""Test-side engines with the GpuWorld interface, for parity checks.

  OracleWorld  the C restatement (oracle/libne_oracle.so)
  RefWorld     the unmodified reference behind a public-API harness
               (oracle/_ref/libne_ref.so), present where it was built

Both are checkers only: the product never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from nativeengine_amd import _abi
from nativeengine_amd import queries
from nativeengine_amd.scene import Scene

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_LIB = os.path.join(ROOT, "oracle", "libne_oracle.so")
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libne_ref.so")

_dp = C.POINTER(C.c_double)
_u32p = C.POINTER(C.c_uint32)
_vp = C.c_void_p


def _bind(lib, prefix):
    g = lambda n: getattr(lib, prefix + n)
    g("create").argtypes = [C.POINTER(_vp)]
    g("destroy").argtypes = [_vp]
    g("set_params").argtypes = [_vp, C.POINTER(_abi.WorldParams)]
    g("set_bodies").argtypes = [_vp, C.POINTER(_abi.Bodies)]
    g("step").argtypes = [_vp, C.c_int32]
    g("get_state").argtypes = [_vp, _dp, _dp, _dp, _dp]
    g("get_body_aux").argtypes = [_vp, _u32p, _dp, _dp, _dp]
    g("reset_warm_cache").argtypes = [_vp]
    g("pair_contacts").argtypes = [_vp, C.c_uint32, C.c_uint32, C.c_uint32, _u32p, _dp, _dp, _dp]
    g("diagnostics").argtypes = [_vp, _dp, _dp, _dp]
    g("min_image").argtypes = [_vp, _dp, _dp]
    g("wrap").argtypes = [_vp, _dp, _dp]
    g("set_joints").argtypes = [_vp, C.c_uint32, _vp]
    g("set_meshes").argtypes = [_vp, C.c_uint32, C.POINTER(_abi.Mesh)]
    g("mesh_bvh").argtypes = [_vp, C.c_uint32, C.POINTER(C.c_int32), C.c_uint32, _u32p, _dp, C.POINTER(C.c_int32)]
    queries.bind(lib, prefix.rstrip("_"))
    g("set_general_joints").argtypes = [_vp, C.c_uint32, _vp]
    g("get_general_joints").argtypes = [_vp, C.c_uint32, C.POINTER(C.c_int32), _dp, _dp]
    return lib


_libs = {}


def mesh_bvh_of(fn, h, m, ntris):
    nn = C.c_uint32(0)
    rc = fn(h, m, None, 0, C.byref(nn), None, None)
    if rc != 0:
        raise RuntimeError(f"mesh_bvh failed: {rc}")
    order = np.zeros(ntris, np.int32)
    bounds = np.zeros((nn.value, 6))
    links = np.zeros((nn.value, 4), np.int32)
    ip = C.POINTER(C.c_int32)
    rc = fn(h, m, order.ctypes.data_as(ip), nn.value, C.byref(nn), _abi.dptr(bounds), links.ctypes.data_as(ip))
    if rc != 0:
        raise RuntimeError(f"mesh_bvh failed: {rc}")
    return order, bounds, links


def oracle_available():
    return os.path.exists(ORACLE_LIB)


def ref_available():
    return os.path.exists(REF_LIB)


def _load(kind):
    if kind in _libs:
        return _libs[kind]
    if kind == "oracle":
        lib = _bind(C.CDLL(ORACLE_LIB), "ne_oracle_")
        lib.ne_oracle_get_contacts.argtypes = [_vp, C.c_uint32, _u32p, _u32p, _u32p, _dp, _dp, _dp, _dp]
        lib.ne_oracle_get_pairs.argtypes = [_vp, C.c_uint32, _u32p, _u32p, _u32p]
        lib.ne_oracle_set_state.argtypes = [_vp, _dp, _dp, _dp, _dp]
        lib.ne_oracle_candidate_count.argtypes = [_vp, _u32p]
        lib.ne_oracle_tape.argtypes = [_vp, C.c_int32]
        lib.ne_oracle_set_events.argtypes = [_vp, C.c_int32]
        lib.ne_oracle_get_events.argtypes = [_vp, C.c_int32, C.c_uint32, _u32p, _u32p, _u32p, _dp, _dp, _dp]
    else:
        lib = _bind(C.CDLL(REF_LIB), "ne_ref_")
        lib.ne_ref_get_contacts.argtypes = [_vp, C.c_uint32, _u32p, _u32p, _u32p, _dp, _dp, _dp]
        lib.ne_ref_set_capture.argtypes = [_vp, C.c_int]
        lib.ne_ref_save_state.argtypes = [_vp, _vp, C.c_uint64, C.POINTER(C.c_uint64)]
        lib.ne_ref_load_state.argtypes = [_vp, _vp, C.c_uint64]
        lib.ne_ref_facade_begin.argtypes = [_vp]
        lib.ne_ref_facade_step.argtypes = [_vp, C.c_int32]
        lib.ne_ref_facade_events.argtypes = [_vp, C.c_int32, C.c_uint32, _u32p, _u32p, _u32p, _dp, _dp, _dp]
    _libs[kind] = lib
    return lib


EVENT_KINDS = ["contact", "begin", "stay", "end", "trigger_enter", "trigger_stay", "trigger_exit"]


def events_of(fn, h, kind):
    """One kind of facade event as a dict of i, j, point, normal, depth."""
    cnt = C.c_uint32(0)
    rc = fn(h, kind, 0, C.byref(cnt), None, None, None, None, None)
    if rc != 0:
        raise RuntimeError(f"events failed: {rc}")
    m = cnt.value
    i, j = np.zeros(m, np.uint32), np.zeros(m, np.uint32)
    pt, nrm, dep = np.zeros((m, 3)), np.zeros((m, 3)), np.zeros(m)
    if m:
        rc = fn(h, kind, m, C.byref(cnt), _abi.uptr(i), _abi.uptr(j), _abi.dptr(pt), _abi.dptr(nrm), _abi.dptr(dep))
        if rc != 0:
            raise RuntimeError(f"events failed: {rc}")
    return dict(i=i, j=j, point=pt, normal=nrm, depth=dep)


def body_aux(w, n):
    """(flags, sleep timer, inverse mass, inverse inertia (n, 3)) of an oracle
    or reference engine's bodies."""
    fl, tm, im, ii = np.zeros(n, np.uint32), np.zeros(n), np.zeros(n), np.zeros((n, 3))
    w._f("get_body_aux")(w.h, fl.ctypes.data_as(C.POINTER(C.c_uint32)), _abi.dptr(tm), _abi.dptr(im), _abi.dptr(ii))
    return fl, tm, im, ii


class _CpuWorld:
    prefix = ""
    kind = ""

    def __init__(self):
        self.lib = _load(self.kind)
        h = _vp()
        assert self._f("create")(C.byref(h)) == 0
        self.h = h
        self.n = 0

    def _f(self, name):
        return getattr(self.lib, self.prefix + name)

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{self.kind} {what} failed: {rc}")

    def close(self):
        if getattr(self, "h", None):
            self._f("destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # scene queries, through the same packing as GpuWorld's
    def _qcheck(self, rc):
        self._check(rc, "query")

    def raycast(self, origin, direction, max_dist, mask=None):
        return queries.raycast(self._f("raycast"), self.h, self._qcheck, origin, direction, max_dist, mask)

    def sphere_cast(self, origin, radius, direction, max_dist, mask=None):
        return queries.raycast(self._f("raycast"), self.h, self._qcheck, origin, direction, max_dist, mask,
                               radius=radius)

    def raycast_all(self, origin, direction, max_dist, mask=None):
        return queries.raycast_all(self._f("raycast_all"), self.h, self._qcheck, origin, direction, max_dist, mask)

    def overlap(self, shapes, mask=None, contacts=False):
        return queries.overlap(self._f("overlap"), self.h, self._qcheck, shapes, mask, contacts)

    def set_params(self, p):
        cp = p.to_c()
        self._check(self._f("set_params")(self.h, C.byref(cp)), "set_params")

    def set_scene(self, s: Scene):
        self.set_params(s.params)
        b, keep = s.to_c()
        self._check(self._f("set_bodies")(self.h, C.byref(b)), "set_bodies")
        self.n = s.n
        self._ngj = 0
        del keep
        j = np.ascontiguousarray(s.joints, _abi.JOINT_DTYPE)
        if len(j) or getattr(self, "_had_joints", False):
            self._check(self._f("set_joints")(self.h, len(j), j.ctypes.data if len(j) else None), "set_joints")
            self._had_joints = len(j) > 0
        if s.meshes:
            arr, keep = s.meshes_c()
            self._check(self._f("set_meshes")(self.h, len(s.meshes), arr), "set_meshes")
            del keep
        gj = np.ascontiguousarray(getattr(s, "gjoints", np.zeros(0, _abi.GJOINT_DTYPE)), _abi.GJOINT_DTYPE)
        if len(gj):
            self.set_general_joints(gj)

    def set_general_joints(self, gj):
        gj = np.ascontiguousarray(gj, _abi.GJOINT_DTYPE)
        self._check(self._f("set_general_joints")(self.h, len(gj), gj.ctypes.data if len(gj) else None),
                    "set_general_joints")
        self._ngj = len(gj)

    def general_joints(self):
        """(broken, applied impulse, motor accumulator) of the general joints."""
        n = getattr(self, "_ngj", 0)
        br, ap, mo = np.zeros(n, np.int32), np.zeros(n), np.zeros(n)
        self._check(self._f("get_general_joints")(self.h, n, br.ctypes.data_as(C.POINTER(C.c_int32)), _abi.dptr(ap),
                                                  _abi.dptr(mo)), "get_general_joints")
        return br, ap, mo

    def mesh_bvh(self, m, ntris):
        """(order, bounds (nn, 6), links (nn, 4)) of mesh m's tree."""
        return mesh_bvh_of(self._f("mesh_bvh"), self.h, m, ntris)

    def step(self, nsteps=1):
        self._check(self._f("step")(self.h, int(nsteps)), "step")

    def reset_warm_cache(self):
        self._check(self._f("reset_warm_cache")(self.h), "reset_warm_cache")

    def state(self):
        n = self.n
        x, v, w = (np.zeros((n, 3)) for _ in range(3))
        q = np.zeros((n, 4))
        self._check(self._f("get_state")(self.h, _abi.dptr(x), _abi.dptr(v), _abi.dptr(q), _abi.dptr(w)), "get_state")
        return x, v, q, w

    def pair_contacts(self, a, b):
        cnt = C.c_uint32(0)
        nrm = np.zeros((4, 3))
        pt = np.zeros((4, 3))
        dep = np.zeros(4)
        self._check(self._f("pair_contacts")(self.h, a, b, 4, C.byref(cnt), _abi.dptr(nrm), _abi.dptr(pt), _abi.dptr(dep)),
                    "pair_contacts")
        m = cnt.value
        return nrm[:m].copy(), pt[:m].copy(), dep[:m].copy()

    def diagnostics(self):
        p = np.zeros(3)
        L = np.zeros(3)
        ke = C.c_double(0.0)
        self._check(self._f("diagnostics")(self.h, _abi.dptr(p), _abi.dptr(L), C.byref(ke)), "diagnostics")
        return p, L, ke.value

    def min_image(self, d):
        d = np.ascontiguousarray(d, np.float64)
        out = np.zeros(3)
        self._f("min_image")(self.h, _abi.dptr(d), _abi.dptr(out))
        return out

    def wrap(self, p):
        p = np.ascontiguousarray(p, np.float64)
        out = np.zeros(3)
        self._f("wrap")(self.h, _abi.dptr(p), _abi.dptr(out))
        return out

    def pairs(self):
        c = self.contacts()
        if len(c["i"]) == 0:
            return c["i"], c["j"]
        keep = np.ones(len(c["i"]), bool)
        keep[1:] = (c["i"][1:] != c["i"][:-1]) | (c["j"][1:] != c["j"][:-1])
        return c["i"][keep], c["j"][keep]


class OracleWorld(_CpuWorld):
    prefix = "ne_oracle_"
    kind = "oracle"

    def contacts(self):
        cnt = C.c_uint32(0)
        self._check(self.lib.ne_oracle_get_contacts(self.h, 0, C.byref(cnt), None, None, None, None, None, None), "contacts")
        m = cnt.value
        i = np.zeros(m, np.uint32)
        j = np.zeros(m, np.uint32)
        nrm, pt = np.zeros((m, 3)), np.zeros((m, 3))
        dep, imp = np.zeros(m), np.zeros(m)
        if m:
            self._check(self.lib.ne_oracle_get_contacts(self.h, m, C.byref(cnt), _abi.uptr(i), _abi.uptr(j), _abi.dptr(nrm),
                                                        _abi.dptr(pt), _abi.dptr(dep), _abi.dptr(imp)), "contacts")
        return dict(i=i, j=j, normal=nrm, point=pt, depth=dep, impulse=imp)

    def candidate_count(self):
        c = C.c_uint32(0)
        self.lib.ne_oracle_candidate_count(self.h, C.byref(c))
        return c.value

    def set_state(self, x=None, v=None, q=None, w=None):
        arrs = [None if a is None else np.ascontiguousarray(a, np.float64) for a in (x, v, q, w)]
        self._check(self.lib.ne_oracle_set_state(self.h, *[None if a is None else _abi.dptr(a) for a in arrs]), "set_state")

    def set_events(self, on=True):
        self._check(self.lib.ne_oracle_set_events(self.h, 1 if on else 0), "set_events")

    def events(self, kind):
        """The last step's facade events of one kind (NE_EVENT_*)."""
        return events_of(self.lib.ne_oracle_get_events, self.h, kind)

    def tape(self, mode):
        """Contact-set tape: 1 records each following step's contact lists, 2
        replays them from the first recorded step (fixed contact set), 0 off."""
        self._check(self.lib.ne_oracle_tape(self.h, int(mode)), "tape")


class RefWorld(_CpuWorld):
    prefix = "ne_ref_"
    kind = "ref"

    def __init__(self, capture=True):
        super().__init__()
        self.lib.ne_ref_set_capture(self.h, 1 if capture else 0)

    def save_state(self) -> bytes:
        """The reference's own PhysicsWorld::saveState bytes for this world."""
        size = C.c_uint64(0)
        self._check(self.lib.ne_ref_save_state(self.h, None, 0, C.byref(size)), "save_state")
        buf = C.create_string_buffer(size.value)
        self._check(self.lib.ne_ref_save_state(self.h, buf, size.value, C.byref(size)), "save_state")
        return buf.raw[: size.value]

    def load_state(self, blob: bytes):
        buf = C.create_string_buffer(bytes(blob), len(blob))
        self._check(self.lib.ne_ref_load_state(self.h, buf, len(blob)), "load_state")

    def facade_begin(self):
        """Step through PhysicsWorld from now on, recording its contact and
        trigger callbacks."""
        self._check(self.lib.ne_ref_facade_begin(self.h), "facade_begin")

    def facade_step(self, nsteps=1):
        self._check(self.lib.ne_ref_facade_step(self.h, int(nsteps)), "facade_step")

    def events(self, kind):
        """The last facade step's callbacks of one kind, as body indices."""
        return events_of(self.lib.ne_ref_facade_events, self.h, kind)

    def contacts(self):
        cnt = C.c_uint32(0)
        self._check(self.lib.ne_ref_get_contacts(self.h, 0, C.byref(cnt), None, None, None, None, None), "contacts")
        m = cnt.value
        i = np.zeros(m, np.uint32)
        j = np.zeros(m, np.uint32)
        nrm, pt = np.zeros((m, 3)), np.zeros((m, 3))
        dep = np.zeros(m)
        if m:
            self._check(self.lib.ne_ref_get_contacts(self.h, m, C.byref(cnt), _abi.uptr(i), _abi.uptr(j), _abi.dptr(nrm),
                                                     _abi.dptr(pt), _abi.dptr(dep)), "contacts")
        return dict(i=i, j=j, normal=nrm, point=pt, depth=dep)
