"""Every EPA tier is on a bit-exact path: scenes whose runs reach each tier
(ne_gpu_epa_queues reports the tiers' queue lengths every step) stepped on
the GPU against the C restatement, state every step and the constraint list
every 10 steps bit for bit.

Tiers (src/epa.hpp:71-150 restated in each): A one lane per run (4
iterations), B 4 lanes a run (20), C 8 lanes a run (32), one wave, one block
of 8 waves.  Mesh triangle candidates: A, then the 8-lane group tier, one
wave, one block."""
import sys

import numpy as np
import pytest

import engines
from nativeengine_amd import scene

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _step_both(s, steps, what):
    from nativeengine_amd.gpu import GpuWorld
    g = GpuWorld()
    g.set_scene(s)
    o = engines.OracleWorld()
    o.set_scene(s)
    seen = {}
    for t in range(steps):
        g.step(1)
        o.step(1)
        for k, v in g.epa_queues().items():
            seen[k] = max(seen.get(k, 0), v)
        for name, a, b in zip("xvqw", g.state(), o.state()):
            assert np.array_equal(a, b), f"{what} step {t}: {name} differs"
        if t % 10 == 9:
            cg, co = g.contacts(), o.contacts()
            for f in ("i", "j", "normal", "point", "depth", "impulse"):
                assert np.array_equal(cg[f], co[f]), f"{what} step {t}: contact {f} differs"
    g.close()
    o.close()
    _say(f"{what}: largest queue per tier over {steps} steps {seen}")
    return seen


def test_pile_runs_reach_tiers_a_b_c_and_the_wave_tier(gpu_available, oracle_lib):
    """The pile at the specified sizes (boxes U[0.2,0.6], spheres, capsules,
    hulls): sphere-box and capsule-box runs outlast tiers A, B and C.  (The
    one-block tier: test_gpu_batched.py's 99,198-face pair.)"""
    s = scene.mixed_pile(300, half=(5.0, 8.0, 5.0), seed=2, sizes="config")
    _say(f"case: config-size pile, {s.n} bodies, 120 steps")
    seen = _step_both(s, 120, "pile300")
    assert seen["overlapping_pairs"] > 0
    assert seen["past_tier_a"] > 0, "no run reached tier B"
    assert seen["past_tier_b"] > 0, "no run reached tier C"
    assert seen["past_tier_c"] > 0, "no run reached the one-wave tier"


def test_terrain_runs_reach_the_mesh_group_tier(gpu_available, oracle_lib):
    """Mixed bodies settling on a heightfield: prism runs outlast the mesh tier A."""
    s = scene.terrain_pile(60, seed=7, shapes="mixed")
    _say(f"case: terrain pile, {s.n} bodies, 150 steps")
    seen = _step_both(s, 150, "terrain60")
    assert seen["mesh_overlapping"] > 0
    assert seen["mesh_past_tier_a"] > 0, "no prism run reached the group tier"
