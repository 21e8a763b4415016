"""Snapshots in the reference's "NES1" bytes (PhysicsWorld::saveState /
loadState, src/physics_world.cpp:22-64) written and read on the device.

The expected bytes come from the reference itself: oracle/_ref builds
src/physics_world.cpp unmodified, and its harness runs saveState / loadState
on a copy of the reference world stepped beside the GPU one.
"""
import struct
import sys

import numpy as np
import pytest

import engines
from nativeengine_amd import scene

pytestmark = pytest.mark.gpu

RECORD = 177   # 22 doubles + the sleeping byte


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _gpu():
    from nativeengine_amd.gpu import GpuWorld
    return GpuWorld()


def _need_ref():
    if not engines.ref_available():
        pytest.skip("oracle/_ref/libne_ref.so (the reference build) is not present")


def _sleeping(blob, n):
    return sum(blob[8 + RECORD * i + 176] for i in range(n))


def test_save_state_matches_the_reference_bytes(gpu_available):
    """A settling pile with island sleeping: after every 60 steps the device's
    snapshot equals the reference's saveState byte for byte, including bodies
    asleep (zeroed masses, stores, timers, the sleeping byte)."""
    _need_ref()
    s = scene.sleeping_pile(80, seed=1, periodic=False)
    _say(f"snapshot: sleeping pile of {s.n} bodies")
    g = _gpu()
    g.set_scene(s)
    r = engines.RefWorld(capture=False)
    r.set_scene(s)
    asleep = 0
    for t in range(4):
        g.step(60)
        r.step(60)
        bg, br = g.save_state(), r.save_state()
        assert len(br) == 8 + RECORD * s.n + 4
        assert struct.unpack_from("<II", br, 0) == (0x4E455331, s.n)
        if bg != br:
            d = next(k for k in range(len(br)) if bg[k] != br[k])
            raise AssertionError(f"after {60 * (t + 1)} steps: first differing byte {d} "
                                 f"(body {(d - 8) // RECORD}, offset {(d - 8) % RECORD})")
        asleep = max(asleep, _sleeping(bg, s.n))
    assert asleep > 5, asleep
    g.close()
    r.close()


def test_load_state_then_step_matches_the_reference(gpu_available):
    """loadState restores every field and resets the warm cache: a world loaded
    from a snapshot and stepped on, on the device and in the reference, stays
    byte-identical; the reference's own bytes load on the device unchanged."""
    _need_ref()
    s = scene.sleeping_pile(80, seed=1, periodic=False)
    src = _gpu()
    src.set_scene(s)
    src.step(150)
    blob = src.save_state()
    assert _sleeping(blob, s.n) > 0
    g = _gpu()
    g.set_scene(s)
    g.step(7)                       # a different state and warm cache before the load
    g.load_state(blob)
    assert g.save_state() == blob
    r = engines.RefWorld(capture=False)
    r.set_scene(s)
    r.step(3)
    r.load_state(blob)
    assert r.save_state() == blob
    for t in range(3):
        g.step(30)
        r.step(30)
        assert g.save_state() == r.save_state(), f"{30 * (t + 1)} steps after the load"
    # the reference's bytes, loaded on the device, read back unchanged
    g2 = _gpu()
    g2.set_scene(s)
    rb = r.save_state()
    g2.load_state(rb)
    assert g2.save_state() == rb
    for w in (src, g, g2, r):
        w.close()


def test_load_state_rejects_bad_snapshots_and_keeps_the_state(gpu_available):
    """Wrong magic, another body count, facade joints, truncation, and a static
    ground the grid leaves out made movable are refused, with the world as it
    was (the reference returns false, possibly after loading some bodies)."""
    s = scene.sleeping_pile(40, seed=2, periodic=False)
    g = _gpu()
    g.set_scene(s)
    g.step(20)
    good = g.save_state()
    n = s.n
    bad = []
    b = bytearray(good); b[0] ^= 1; bad.append(("magic", bytes(b)))
    b = bytearray(good); struct.pack_into("<I", b, 4, n + 1); bad.append(("count", bytes(b)))
    b = bytearray(good); struct.pack_into("<I", b, 8 + RECORD * n, 1); bad.append(("joints", bytes(b + b"\x00")))
    bad.append(("truncated", good[:-5]))
    ground = int(np.argmax(np.where(s.inv_mass == 0, np.max(s.half_extents, axis=1), -1.0)))
    assert s.inv_mass[ground] == 0
    b = bytearray(good); struct.pack_into("<d", b, 8 + RECORD * ground + 8 * 13, 1.0)
    bad.append(("static ground made movable", bytes(b)))
    for what, blob in bad:
        with pytest.raises(RuntimeError):
            g.load_state(blob)
        assert g.save_state() == good, f"state changed by the refused load ({what})"
    # a good snapshot still loads, and extra trailing bytes are ignored as the reference ignores them
    g.step(5)
    g.load_state(good + b"\x00\x01")
    assert g.save_state() == good
    g.close()
