"""Batched worlds on the GPU: many independent periodic worlds stepped in
one set of launches (ne_gpu_set_worlds), eagerly and as HIP-graph replays,
against the oracle run on each world alone.

Bar: every world's state (x, v, q, w) and constraint list bit-identical to
the oracle's for that world, with pair indices offset by the world's first
body; and the same bits whatever the autotuner chose.
"""
import sys

import numpy as np
import pytest

import engines
from nativeengine_amd import scene as S

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _gpu_batch(batch, graph):
    from nativeengine_amd.gpu import GpuWorld
    g = GpuWorld()
    g.set_scene(batch.scene)
    g.set_worlds(batch.offsets, batch.half, batch.restitution, batch.friction)
    g.use_graph(graph)
    return g


def _oracle_worlds(batch, steps):
    out = []
    for w in range(batch.nworlds):
        o = engines.OracleWorld()
        o.set_scene(batch.world(w))
        o.step(steps)
        out.append((o.state(), o.contacts()))
    return out


def _check(batch, g, want, what):
    x, v, q, w = g.state()
    cg = g.contacts()
    off = batch.offsets
    for k in range(batch.nworlds):
        a, b = int(off[k]), int(off[k + 1])
        (ox, ov, oq, ow), co = want[k]
        for name, u, t in (("x", x[a:b], ox), ("v", v[a:b], ov), ("q", q[a:b], oq), ("w", w[a:b], ow)):
            if not np.array_equal(u, t):
                i = int(np.argwhere(u != t)[0][0])
                raise AssertionError(f"{what}: world {k} {name} differs at body {i}: {u[i]} vs {t[i]}")
        sel = (cg["i"] >= a) & (cg["i"] < b)
        assert np.all((cg["j"][sel] >= a) & (cg["j"][sel] < b)), f"{what}: world {k} has a cross-world contact"
        assert np.array_equal(cg["i"][sel] - a, co["i"]) and np.array_equal(cg["j"][sel] - a, co["j"]), \
            f"{what}: world {k} pair list differs"
        for f in ("normal", "point", "depth", "impulse"):
            assert np.array_equal(cg[f][sel], co[f]), f"{what}: world {k} contact {f} differs"


@pytest.mark.parametrize("graph", [False, True])
def test_batched_gas_worlds_match_each_world_alone(gpu_available, oracle_lib, graph):
    """Uneven worlds of spheres and boxes, per-world box, restitution and
    friction (a small version of the 1024-world batch)."""
    batch = S.batched_worlds(9, seed=3, nmin=64, nmax=600)
    steps = 12
    _say(f"batched gas: {batch.nworlds} worlds, {batch.scene.n} bodies, graph={graph}")
    want = _oracle_worlds(batch, steps)
    g = _gpu_batch(batch, graph)
    for t in range(steps):
        g.step(1)
        assert g.epa_stats()[1] == 0
    _check(batch, g, want, f"batched gas graph={graph}")
    assert len(g.contacts()["i"]) > 0
    g.close()


def test_epa_blowup_pair_matches_through_the_large_workspace(gpu_available, oracle_lib):
    """A sphere-box contact whose EPA grows to 99,198 faces in the reference
    (tests/scenes.py): the GPU reruns it in the wave-cooperative large
    workspace and gives the oracle's contact and state bit for bit."""
    import scenes
    from nativeengine_amd.gpu import GpuWorld
    s = scenes.epa_blowup_pair()
    o = engines.OracleWorld()
    o.set_scene(s)
    o.step(1)
    g = GpuWorld()
    g.set_scene(s)
    g.step(1)
    reruns, trunc = g.epa_stats()
    assert reruns >= 1 and trunc == 0
    assert g.epa_queues()["block_tier"] >= 1   # the run went all the way to the one-block tier
    cg, co = g.contacts(), o.contacts()
    for f in ("i", "j", "normal", "point", "depth", "impulse"):
        assert np.array_equal(cg[f], co[f]), f
    for u, v in zip(g.state(), o.state()):
        assert np.array_equal(u, v)
    g.close()


def test_epa_face_cap_is_reported(gpu_available, oracle_lib):
    """The same pair with the large workspace capped at 16,384 faces: the run
    stops at its closest face, so the step reports NE_WARN_EPA_CAPPED (raised
    in strict mode, a warning otherwise) and counts the truncation; with the
    default cap the same step returns NE_OK and equals the oracle."""
    import warnings
    import scenes
    from nativeengine_amd import _abi
    from nativeengine_amd.gpu import EpaCappedWarning, GpuWorld, NeGpuError
    s = scenes.epa_blowup_pair()
    g = GpuWorld(epa_face_cap=16384)
    g.set_scene(s)
    with pytest.raises(NeGpuError) as err:
        g.step(1)
    assert err.value.code == _abi.NE_WARN_EPA_CAPPED
    assert g.epa_stats()[1] >= 1
    g.close()
    g = GpuWorld(epa_face_cap=16384, strict=False)
    g.set_scene(s)
    with warnings.catch_warnings(record=True) as caught:
        warnings.simplefilter("always")
        g.step(1)
    assert any(issubclass(w.category, EpaCappedWarning) for w in caught)
    assert g.epa_capped_calls == 1 and g.epa_stats()[1] >= 1
    assert all(np.all(np.isfinite(a)) for a in g.state())
    g.close()
    g = GpuWorld()
    g.set_scene(s)
    g.step(1)
    assert g.epa_stats()[1] == 0 and g.epa_capped_calls == 0
    g.close()


@pytest.mark.parametrize("graph", [False, True])
def test_batched_ground_worlds_with_excluded_ground(gpu_available, oracle_lib, graph):
    """Gravity worlds whose first body is a static ground box spanning the
    periodic cube: it is kept out of each world's grid and paired with that
    world's bodies only.  Multi-step calls exercise the deferred wrap."""
    batch = S.adjoint_worlds(5, n_spheres=20, n_boxes=20, side=8.0, seed=4)
    _say(f"ground worlds: {batch.nworlds} worlds, {batch.scene.n} bodies, graph={graph}")
    steps = 40
    want = _oracle_worlds(batch, steps)
    g = _gpu_batch(batch, graph)
    g.step(15)
    g.step(25)
    _check(batch, g, want, f"ground worlds graph={graph}")
    c = g.contacts()
    assert np.any(np.isin(c["i"], batch.offsets[:-1])), "no contact with a ground box"
    g.close()


def test_graph_steps_sweep_large_islands_in_order(gpu_available, oracle_lib):
    """Dense piles on a static ground, stepped from a HIP graph: the dataflow
    solve cannot run there (no read-back), so one wave sweeps each island
    above 4 constraints in ordered windows.  Every world equals the oracle
    run alone, bit for bit, and the step did have islands of more than 16."""
    batch = S.adjoint_worlds(2, n_spheres=100, n_boxes=100, side=8.0, seed=7)
    _say(f"dense ground worlds: {batch.nworlds} worlds, {batch.scene.n} bodies, graph")
    steps = 60
    want = _oracle_worlds(batch, steps)
    g = _gpu_batch(batch, True)
    g.step(steps)
    big = g.counters()["max_island"]
    _say(f"largest island: {big} constraints")
    assert big > 16, big
    _check(batch, g, want, "dense ground worlds graph")
    g.close()


def test_autotune_keeps_the_bits(gpu_available, oracle_lib):
    """The tuner runs trial steps from a snapshot and restores it: stepping
    after tuning gives the oracle's bits, for whatever configuration won."""
    batch = S.batched_worlds(6, seed=5, nmin=64, nmax=300)
    g = _gpu_batch(batch, False)
    g.step(3)
    res = g.autotune(steps=2)
    _say("autotune:", res)
    assert res["trials"] == 12 and res["ms_per_step"] > 0
    g.step(7)
    want = _oracle_worlds(batch, 10)
    _check(batch, g, want, "after autotune")
    g.close()


def test_set_worlds_rejects_bad_layouts(gpu_available):
    from nativeengine_amd.gpu import GpuWorld, NeGpuError
    batch = S.batched_worlds(3, seed=6, nmin=64, nmax=100)
    g = GpuWorld()
    g.set_scene(batch.scene)
    bad = batch.offsets.copy()
    bad[-1] -= 1
    with pytest.raises(NeGpuError):
        g.set_worlds(bad, batch.half)
    with pytest.raises(NeGpuError):
        g.set_worlds(batch.offsets, -batch.half)
    p = batch.scene.params
    p.periodic = False
    g.set_params(p)
    with pytest.raises(NeGpuError):
        g.set_worlds(batch.offsets, batch.half)
    g.close()
