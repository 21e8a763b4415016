"""General joints on the device (ne_gpu_set_general_joints:
World::solveJoints, src/world.cpp:206-315), through the C ABI: bit for bit
the oracle's state and joint states (the golden joint scenes run with every
scene fixture in test_gpu_golden.py); larger chains, eager and
graph-replayed; the NES1 snapshot's joint flags; the device atan2 of the
hinge limit against the host's."""
import sys

import numpy as np
import pytest

import engines
import scenes
from nativeengine_amd import _abi
from nativeengine_amd import scene as S

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _gpu():
    from nativeengine_amd.gpu import GpuWorld
    return GpuWorld()


def _chains(nscenes, limits=False):
    """Copies of the chain scene (different seeds) side by side on a 5-wide
    grid 13 apart, in one closed box."""
    parts = []
    for k in range(nscenes):
        p = scenes.joint_mix_scene(10 + k, limits=limits)
        p.x[:, 0] += 13.0 * (k % 5 - 2)
        p.x[:, 2] += 13.0 * (k // 5 - (nscenes // 5 - 1) / 2.0)
        g = np.array(p.gjoints, copy=True)   # frames are relative: a shift keeps them
        p.gjoints = g
        parts.append(p)
    big = S.concat_scenes(parts)
    off = 0
    js = []
    for p in parts:
        g = np.array(p.gjoints, copy=True)
        g["a"] += np.uint32(off)
        g["b"] += np.uint32(off)
        js.append(g)
        off += p.n
    big.gjoints = np.concatenate(js)
    big.params = parts[0].params
    big.params.half = (36.0, 6.0, 6.5 * ((nscenes + 4) // 5) + 1.0)
    return big


@pytest.mark.parametrize("graph", [False, True])
def test_gpu_joint_chains_match_the_oracle(gpu_available, oracle_lib, graph):
    """20 overlaid copies of the chain scene (every joint type without
    limits, sleeping on): 150 steps bitwise, joint states included."""
    s = _chains(20)
    _say(f"joint chains: {s.n} bodies, {len(s.gjoints)} joints, graph={graph}")
    g = _gpu()
    g.set_scene(s)
    g.use_graph(graph)
    o = engines.OracleWorld()
    o.set_scene(s)
    for t in range(150):
        g.step(1)
        o.step(1)
        if t % 25 == 24:
            for name, a, b in zip("xvqw", g.state(), o.state()):
                assert np.array_equal(a, b), f"step {t}: {name}"
    for a, b in zip(g.general_joints(), o.general_joints()):
        assert np.array_equal(a, b)
    g.close()


def test_gpu_snapshot_keeps_broken_joints(gpu_available, oracle_lib):
    """The breakable joint scene: after the joint breaks, the NES1 snapshot
    ends with the joint count and its broken byte; loading it into a fresh
    world restores the flag, and the run continues as the uninterrupted one."""
    import fixtures
    s, steps, out = fixtures.load_scene("joints_breakable")
    g = _gpu()
    g.set_scene(s)
    g.step(200)
    assert g.general_joints()[0][0] == 1
    blob = g.save_state()
    assert blob[-5:-1] == (1).to_bytes(4, "little") and blob[-1] == 1
    h = _gpu()
    h.set_scene(s)
    assert h.general_joints()[0][0] == 0
    h.load_state(blob)
    assert h.general_joints()[0][0] == 1
    g.step(100)
    h.step(100)
    for a, b in zip(g.state(), h.state()):
        assert np.array_equal(a, b)
    for a, k in zip(g.state(), "xvqw"):
        assert np.array_equal(a, out[k])
    g.close()
    h.close()


def test_device_atan2_against_the_host(gpu_available, tmp_path):
    """The hinge limit's atan2 (rounded to nearest, ne_atan2.hpp) gives the
    same bits on the device as the same source built for the host, on a
    million arguments like the limit's (small rotations and the full
    circle); the device library's atan2, for comparison, differs from the
    host C library's on a large share of them (printed)."""
    import ctypes as C
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    so = str(tmp_path / "libatan2rn.so")
    _say("building the host atan2 probe")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-shared", "-fPIC", "-DNE_ATAN2_LIB",
                    os.path.join(root, "tests", "cpp", "atan2rn.cpp"), "-o", so], check=True, timeout=120)
    lib = C.CDLL(so)
    dp = C.POINTER(C.c_double)
    lib.atan2rn_batch.argtypes = [dp, dp, dp, dp, C.c_size_t]
    rng = np.random.default_rng(0)
    n = 1 << 20
    y = np.concatenate([rng.uniform(-1, 1, n // 2), rng.normal(0, 1e-2, n // 2)])
    x = np.concatenate([rng.uniform(-1, 1, n // 2), rng.uniform(0.9, 1.0, n // 2)])
    host_rn, host_lib = np.zeros(n), np.zeros(n)
    lib.atan2rn_batch(_abi.dptr(y), _abi.dptr(x), _abi.dptr(host_rn), _abi.dptr(host_lib), n)
    _say("host atan2s done")
    g = _gpu()
    dev_rn = g.fp_probe(4, y, x)
    dev_lib = g.fp_probe(5, y, x)
    g.close()
    assert dev_rn.tobytes() == host_rn.tobytes()
    _say(f"atan2 rounded to nearest: differs from the host library on {np.mean(dev_rn != host_lib):.4%}; "
         f"the device library's differs on {np.mean(dev_lib != host_lib):.3%} "
         f"(max {np.max(np.abs(dev_lib.view(np.int64) - host_lib.view(np.int64)))} ulp)")
    assert np.mean(dev_rn != host_lib) < 2e-3


def test_gpu_general_joints_refuse_batches_and_bad_records(gpu_available):
    from nativeengine_amd.gpu import NeGpuError
    s = scenes.joint_mix_scene(3)
    g = _gpu()
    g.set_scene(s)
    bad = np.array(s.gjoints, copy=True)
    bad[0]["b"] = bad[0]["a"]
    with pytest.raises(NeGpuError) as err:
        g.set_general_joints(bad)
    assert err.value.code == _abi.NE_ERR_INVALID
    bad = np.array(s.gjoints, copy=True)
    bad[1]["type"] = 9
    with pytest.raises(NeGpuError):
        g.set_general_joints(bad)
    g.close()
