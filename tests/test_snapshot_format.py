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
