"""ne_gpu_save_state_noalloc: the snapshot entry point that allocates no
device memory (its staging is the library's static device buffer).  Its
bytes must equal ne_gpu_save_state's, on two calls in a row -- the second
with fewer bodies, so stale bytes of the first call left in the shared
buffer would show -- and a world larger than the buffer is refused with a
clear error."""
import ctypes as C
import sys

import numpy as np
import pytest

from nativeengine_amd import _abi, scene

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _raw_noalloc(g):
    """The C ABI call itself: size query, then the bytes."""
    size = C.c_uint64(0)
    assert g.lib.ne_gpu_save_state_noalloc(g.h, None, 0, C.byref(size)) == _abi.NE_OK
    buf = C.create_string_buffer(size.value)
    assert g.lib.ne_gpu_save_state_noalloc(g.h, buf, size.value, C.byref(size)) == _abi.NE_OK
    return buf.raw[: size.value]


def test_noalloc_snapshot_equals_the_allocating_one(gpu_available):
    from nativeengine_amd.gpu import GpuWorld
    big = scene.sleeping_pile(300, seed=3, half=(8.0, 8.0, 8.0))
    small = scene.sleeping_pile(60, seed=5)
    out = []
    for s, steps in ((big, 90), (small, 40)):
        _say(f"noalloc snapshot: {s.n} bodies after {steps} steps")
        g = GpuWorld()
        g.set_scene(s)
        g.step(steps)
        a = _raw_noalloc(g)
        b = g.save_state()
        assert len(a) == 8 + 177 * s.n + 4
        assert a == b, "the no-allocation snapshot differs from ne_gpu_save_state's"
        g.step(3)
        assert _raw_noalloc(g) == g.save_state()
        out.append(a)
        g.close()
    assert len(out[1]) < len(out[0])


def test_noalloc_snapshot_refuses_more_than_its_buffer(gpu_available):
    from nativeengine_amd.gpu import GpuWorld
    cap_bodies = (8 << 20) // 177
    s = scene.sphere_gas(cap_bodies + 1, seed=4)
    g = GpuWorld()
    g.set_scene(s)
    size = C.c_uint64(0)
    rc = g.lib.ne_gpu_save_state_noalloc(g.h, None, 0, C.byref(size))
    assert rc == _abi.NE_ERR_INVALID
    assert size.value == 8 + 177 * s.n + 4
    msg = g.lib.ne_gpu_last_error(g.h).decode()
    assert "ne_gpu_save_state_noalloc" in msg and str(cap_bodies) in msg.replace(",", ""), msg
    buf = C.create_string_buffer(size.value)
    assert g.lib.ne_gpu_save_state_noalloc(g.h, buf, size.value, C.byref(size)) == _abi.NE_ERR_INVALID
    # one body fewer fits, and equals the allocating call
    t = scene.sphere_gas(cap_bodies, seed=4)
    h = GpuWorld()
    h.set_scene(t)
    h.step(2)
    assert _raw_noalloc(h) == h.save_state()
    g.close()
    h.close()
