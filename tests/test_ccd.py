"""Continuous collision (World::ccdPass, src/world.cpp:658-696) in the C
restatement, pinned against the reference: the golden CCD scenes (fast CCD
spheres, boxes, capsules, cylinders and hulls fired at thin plates and at
each other, slow and sleeping CCD bodies, a sensor and a masked plate, closed
and periodic; tests/golden/scene_ccd*.npz, checked with every scene fixture
in test_oracle.py), the live reference where built, and the reference's own
check (selftest.cpp:1241-1257): without CCD the fast ball tunnels through
the wall, with it the ball stops at the wall.  What the sweep fixes without
the reference: a stopped body lies on the inflated obstacle's surface, at
the sweep's first contact (its bounding sphere touches the wall)."""
import numpy as np
import pytest

import engines
import scenes
from nativeengine_amd import _abi


def test_oracle_ccd_selftest(oracle_lib):
    xs = {}
    for ccd in (False, True):
        o = engines.OracleWorld()
        o.set_scene(scenes.ccd_selftest_scene(ccd))
        o.step(20)
        xs[ccd] = o.state()[0][1][0]
    assert xs[False] > 1.0 and xs[True] < 0.3
    # stopped with its bounding sphere (r 0.2) touching the wall face at x = -0.05
    assert abs(xs[True] - (-0.25)) < 1e-9


def test_oracle_ccd_stops_at_the_first_contact(oracle_lib):
    """One step of the selftest ball: its stop is the sweep of its bounding
    sphere from the step-start position to the wall's inflated face, and the
    velocity into the wall is gone (restitution and friction aside)."""
    s = scenes.ccd_selftest_scene(True)
    o = engines.OracleWorld()
    o.set_scene(s)
    o.step(1)
    x, v, _, _ = o.state()
    assert x[1][0] == -0.25 and v[1][0] <= 0.0


@pytest.mark.skipif(not engines.ref_available(), reason="oracle/_ref not built (reference sources absent)")
@pytest.mark.parametrize("periodic", [False, True])
def test_oracle_ccd_matches_the_reference_live(periodic):
    s = scenes.ccd_scene(7, periodic=periodic)
    o = engines.OracleWorld()
    o.set_scene(s)
    r = engines.RefWorld()
    r.set_scene(s)
    for t in range(40):
        o.step(1)
        r.step(1)
        for a, b in zip(o.state(), r.state()):
            assert np.array_equal(a, b), t
