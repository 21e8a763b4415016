"""The NES1 snapshot format on the CPU (no device): the reference's own
PhysicsWorld::saveState bytes through the harness (src/physics_world.cpp:22-64),
and the device calls' argument checks."""
import ctypes as C
import struct

import pytest

import engines
from nativeengine_amd import _abi, gpu, scene


def test_snapshot_calls_reject_a_null_context():
    lib = gpu.load_library()
    size = C.c_uint64(0)
    assert lib.ne_gpu_save_state(None, None, 0, C.byref(size)) == _abi.NE_ERR_INVALID
    assert lib.ne_gpu_load_state(None, None, 0) == _abi.NE_ERR_INVALID


def test_reference_snapshot_layout_and_round_trip(oracle_lib):
    """Magic and body count, 177 bytes per body in saveState's field order
    (x, q, v, w first), the facade joint count last; loadState then saveState
    gives the same bytes."""
    if not engines.ref_available():
        pytest.skip("oracle/_ref/libne_ref.so (the reference build) is not present")
    s = scene.sleeping_pile(30, seed=2)
    r = engines.RefWorld(capture=False)
    r.set_scene(s)
    r.step(30)
    b = r.save_state()
    assert len(b) == 8 + 177 * s.n + 4
    assert struct.unpack_from("<II", b, 0) == (0x4E455331, s.n)
    assert struct.unpack_from("<I", b, 8 + 177 * s.n) == (0,)
    x, v, q, w = r.state()
    for i in (0, s.n // 2, s.n - 1):
        o = 8 + 177 * i
        assert struct.unpack_from("<3d", b, o) == tuple(x[i])
        assert struct.unpack_from("<4d", b, o + 24) == tuple(q[i])
        assert struct.unpack_from("<3d", b, o + 56) == tuple(v[i])
        assert struct.unpack_from("<3d", b, o + 80) == tuple(w[i])
        assert b[o + 176] in (0, 1)
    r2 = engines.RefWorld(capture=False)
    r2.set_scene(s)
    r2.load_state(b)
    assert r2.save_state() == b
    r.close()
    r2.close()


def _nes1_record(blob, i):
    """Body i of an NES1 snapshot: x, q, v, w, invMass, invInertiaBody,
    invMassStore, invInertiaStore, sleepTimer, sleeping (physics_world.cpp:25-39)."""
    o = 8 + 177 * i
    d = struct.unpack_from("<22d", blob, o)
    return dict(x=d[0:3], q=d[3:7], v=d[7:10], w=d[10:13], im=d[13], ii=d[14:17], ims=d[17], iis=d[18:21],
                timer=d[21], sleeping=blob[o + 176])


def test_snapshot_fixture_matches_the_oracle_state(oracle_lib):
    """The committed NES1 fixture (the reference's own saveState bytes) holds,
    field by field, the oracle's state after the same steps: pose, velocity,
    live and stored masses, sleep timer and flag -- so the GPU snapshot tests
    can compare against it with no reference build present."""
    import fixtures
    import numpy as np
    s, saves, every, load_blob, load_at, after, after_steps = fixtures.load_snapshots()
    o = engines.OracleWorld()
    o.set_scene(s)
    asleep = 0
    for k, blob in enumerate(saves):
        o.step(every)
        assert len(blob) == 8 + 177 * s.n + 4
        assert struct.unpack_from("<II", blob, 0) == (0x4E455331, s.n)
        assert struct.unpack_from("<I", blob, 8 + 177 * s.n) == (0,)
        x, v, q, w = o.state()
        fl, tm, im, ii = engines.body_aux(o, s.n)
        for i in range(s.n):
            r = _nes1_record(blob, i)
            assert r["x"] == tuple(x[i]) and r["q"] == tuple(q[i]) and r["v"] == tuple(v[i]) and r["w"] == tuple(w[i])
            assert r["im"] == im[i] and r["ii"] == tuple(ii[i]) and r["timer"] == tm[i]
            assert r["sleeping"] == (1 if fl[i] & _abi.FLAG_SLEEPING else 0)
        asleep = max(asleep, sum(blob[8 + 177 * i + 176] for i in range(s.n)))
    assert asleep > 5
    o.close()
