"""Continuous collision on the device (World::ccdPass, src/world.cpp:
658-696), through the C ABI: the candidate pass over (CCD body, body tile)
and the ordered fold give the oracle's state bit for bit -- the golden CCD
scenes run with every scene fixture in test_gpu_golden.py; here the
reference's own check, larger scenes (closed with tile skips, periodic with
minimum images, eager and graph-replayed), and a sweep whose candidates
outgrow their buffer (the fold then visits every body)."""
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


def test_gpu_ccd_selftest(gpu_available, oracle_lib):
    for ccd in (False, True):
        s = scenes.ccd_selftest_scene(ccd)
        g = _gpu()
        g.set_scene(s)
        g.step(20)
        o = engines.OracleWorld()
        o.set_scene(s)
        o.step(20)
        for a, b in zip(g.state(), o.state()):
            assert np.array_equal(a, b)
        x = g.state()[0][1][0]
        assert (x < 0.3) if ccd else (x > 1.0)
        g.close()


def _ccd_pile(n, seed, periodic):
    s = S.mixed_pile(n, half=(10.0, 10.0, 10.0), seed=seed, periodic=periodic, ground=not periodic)
    rng = np.random.default_rng(seed)
    if periodic:
        s.params.gravity = (0.0, 0.0, 0.0)
    dyn = np.arange(s.n)[s.inv_mass > 0]
    fast = dyn[::9]
    d = rng.normal(size=(len(fast), 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    s.v[fast] = d * rng.uniform(200.0, 900.0, (len(fast), 1))
    s.flags[fast] |= np.uint32(_abi.FLAG_CCD)
    s.flags[dyn[4::31]] |= np.uint32(_abi.FLAG_CCD)   # slow CCD bodies
    s.flags[dyn[5::37]] |= np.uint32(_abi.FLAG_SENSOR)
    return s


@pytest.mark.parametrize("periodic,graph", [(False, False), (True, False), (False, True)])
def test_gpu_ccd_piles_match_the_oracle(gpu_available, oracle_lib, periodic, graph):
    """Piles of 3,000 bodies with every ninth dynamic body a fast CCD body
    (200-900 units/s), more slow CCD bodies and sensors: 40 steps bitwise.
    (The periodic pile's seed is one whose EPA runs stay below the face cap:
    seed 3 there grows one to 1.67M faces in the oracle.)"""
    s = _ccd_pile(3000, 4 if periodic else 3, periodic)
    _say(f"ccd pile: {s.n} bodies, {int(((s.flags & _abi.FLAG_CCD) != 0).sum())} CCD, periodic={periodic}, graph={graph}")
    g = _gpu()
    g.set_scene(s)
    g.use_graph(graph)
    o = engines.OracleWorld()
    o.set_scene(s)
    for t in range(40):
        g.step(1)
        o.step(1)
        if t % 10 == 9:
            for name, a, b in zip("xvqw", g.state(), o.state()):
                assert np.array_equal(a, b), f"step {t}: {name}"
    g.close()


def test_gpu_ccd_candidate_overflow_falls_back_to_every_body(gpu_available, oracle_lib):
    """One CCD sphere swept 100 units in a step along a row of 6,000 small
    static spheres: the full sweep hits more bodies than the candidate
    buffer holds (4,096), so the fold visits every body; the stop equals the
    oracle's."""
    bodies = [S.make_sphere(k, (0.05 * k - 150.0, 0.0, 0.0), 0.02, 1.0).make_static() for k in range(6000)]
    ball = S.make_sphere(6000, (-160.0, 0.0, 0.0), 0.1, 1.0)
    ball.v = (6000.0, 0.0, 0.0)
    ball.flags |= _abi.FLAG_CCD
    s = S.Scene.from_bodies(bodies + [ball])
    s.params = S.Params(half=(400.0, 400.0, 400.0), periodic=False, dt=1.0 / 60.0, gravity=(0.0, 0.0, 0.0))
    g = _gpu()
    g.set_scene(s)
    o = engines.OracleWorld()
    o.set_scene(s)
    g.step(2)
    o.step(2)
    for a, b in zip(g.state(), o.state()):
        assert np.array_equal(a, b)
    assert g.state()[0][6000][0] < -149.0
    g.close()


def test_gpu_ccd_refuses_batches_and_rollouts(gpu_available):
    from nativeengine_amd.gpu import NeGpuError
    batch = S.batched_worlds(2, seed=1, nmin=40, nmax=60)
    batch.scene.flags[3] |= np.uint32(_abi.FLAG_CCD)
    g = _gpu()
    g.set_scene(batch.scene)
    with pytest.raises(NeGpuError) as err:
        g.set_worlds(batch.offsets, batch.half, batch.restitution, batch.friction)
    assert err.value.code == _abi.NE_ERR_UNSUPPORTED
    g.close()
