"""The HIP step against the golden fixtures written from the unmodified
reference (tests/golden/*.npz, by tests/golden/make_fixtures.py), through the
C ABI.  Nothing here needs the reference or its build: the fixtures hold the
inputs and what the reference produced.

  - every scene fixture, stepped on the device: final state, the last step's
    constraint list and every step's constraint count, bit for bit.  The
    feature scenes carry the paths no other scene reaches: planes
    (detect.hpp:19-31), cylinders (narrowphase.hpp:34-81), layers and masks,
    no-collide lists and sensors (world.cpp:513-516, world.hpp:94-103), the
    Average and Min combine modes (world.cpp:42-51), a kinematic body,
    damping, gravity scale, applied forces (world.cpp:373-411), an open
    boundary and a cold start;
  - the elastic gas of 4096 spheres as specified (side 40, N(0,1), 1000
    steps), with the diagnostics of world.cpp:704-729 every 100 steps equal
    to the reference's;
  - the narrow-phase vectors, each pair stepped as a world of its own, so
    every shape pair goes through the device's detectContact /
    boxBoxManifold / GJK / EPA in the order gatherContacts calls them.
"""
import sys

import numpy as np
import pytest

import fixtures
import scenes
from nativeengine_amd import scene

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _gpu():
    from nativeengine_amd.gpu import GpuWorld
    return GpuWorld()


def _diag_row(g):
    p, L, ke = g.diagnostics()
    return np.concatenate([p, L, [ke]])


@pytest.mark.parametrize("name", [n for n in fixtures.scene_names() if n not in scenes.ATAN2_SCENES])
def test_gpu_matches_golden_scene(gpu_available, name):
    s, steps, out = fixtures.load_scene(name)
    _say(f"golden scene {name}: {s.n} bodies, {steps} steps")
    g = _gpu()
    g.set_scene(s)
    every = out.get("diag_every", 0)
    diag = [_diag_row(g)] if every else []
    counts = []
    for t in range(steps):
        g.step(1)
        counts.append(g.counters()["constraints"])
        if every and (t + 1) % every == 0:
            diag.append(_diag_row(g))
    assert g.epa_stats()[1] == 0
    for k, a in zip("xvqw", g.state()):
        if not np.array_equal(a, out[k]):
            bad = np.argwhere(a != out[k])
            raise AssertionError(f"{name}: {k} differs from the reference at body {bad[0][0]} "
                                 f"({len(np.unique(bad[:, 0]))} bodies, max {np.max(np.abs(a - out[k])):.3e})")
    c = g.contacts()
    assert np.array_equal(np.array(counts), out["counts"]), f"{name}: per-step constraint counts differ"
    assert np.array_equal(c["i"], out["ci"]) and np.array_equal(c["j"], out["cj"]), f"{name}: pair list differs"
    for k in ["normal", "point", "depth"]:
        assert np.array_equal(c[k], out[k]), f"{name}: contact {k} differs"
    if every:
        assert np.array_equal(np.array(diag), out["diag"]), f"{name}: diagnostics differ from the reference's"
    if "gj_broken" in out:   # general joints: broken flags and the last step's accumulators
        for k, a in zip(("gj_broken", "gj_applied", "gj_motor"), g.general_joints()):
            assert np.array_equal(a, out[k]), f"{name}: {k} differs"
    g.close()


@pytest.mark.parametrize("name", sorted(scenes.ATAN2_SCENES))
def test_gpu_matches_golden_hinge_limit_scene_to_tolerance(gpu_available, name):
    """Hinge limits take atan2 rounded to nearest, which the host C library's
    atan2 is not for about 3 arguments in 10,000; the scene (chains with
    limited hinges among every other joint type, 120 steps) stays within
    1e-7 of the reference's state (bit for bit unless one of those arguments
    comes up; a last-bit angle difference then grows over the run's
    contacts and joints) and breaks the same joints."""
    s, steps, out = fixtures.load_scene(name)
    g = _gpu()
    g.set_scene(s)
    g.step(steps)
    for k, a in zip("xvqw", g.state()):
        err = np.max(np.abs(a - out[k]))
        _say(f"{name}: {k} max |GPU - reference| {err:.3e}")
        assert err < 1e-7, f"{name}: {k} off by {err:.3e}"
    assert np.array_equal(g.general_joints()[0], out["gj_broken"])
    g.close()


def test_elastic_gas_conservation_as_the_reference(gpu_available):
    """The 4096-sphere elastic gas over 1000 steps: momentum stays at rounding
    level, and the kinetic energy drifts exactly as the reference's does
    (the reference loses energy at approach speeds under its restitution slop
    of 0.5, world.cpp:40 and :544, so its own drift is the target, bit for
    bit, not 1e-10)."""
    s, steps, out = fixtures.load_scene("gas4096")
    g = _gpu()
    g.set_scene(s)
    d0 = _diag_row(g)
    g.step(steps)
    d1 = _diag_row(g)
    ref0, ref1 = out["diag"][0], out["diag"][-1]
    assert np.array_equal(d0, ref0) and np.array_equal(d1, ref1)
    dp = np.abs(d1[:3] - d0[:3]).max()
    de = (d1[6] - d0[6]) / d0[6]
    _say(f"gas4096: |dp| {dp:.3e}, dE/E {de:.6e} (the reference's own drift, bit for bit)")
    assert dp < 1e-10
    assert -1e-2 < de < 0.0
    g.close()


def test_isolated_elastic_collisions_conserve_energy_to_1e10(gpu_available):
    """Where the restitution slop does not act (every approach faster than
    0.5), the elastic frictionless collision conserves momentum and energy to
    1e-10: 256 two-sphere worlds, each pair approaching at 1..4 along a random
    line, offset sideways by at most 0.24 (so the normal approach speed at
    impact is at least 0.95 of the closing speed), stepped through the
    collision and on."""
    rng = np.random.default_rng(12)
    worlds = []
    for w in range(256):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        off = rng.normal(size=3)
        off -= d * off.dot(d)
        off *= rng.uniform(0.0, 0.24) / np.linalg.norm(off)
        speed = rng.uniform(1.0, 4.0)
        a = scene.make_sphere(0, tuple(-0.8 * d), 0.5, 1.9)
        b = scene.make_sphere(1, tuple(0.8 * d + off), rng.uniform(0.3, 0.6), 1.0)
        a.v = tuple(0.5 * speed * d)
        b.v = tuple(-0.5 * speed * d)
        ws = scene.Scene.from_bodies([a, b])
        worlds.append(ws)
    s = scene.concat_scenes(worlds)
    s.params = scene.Params(half=(4.0, 4.0, 4.0), periodic=True, dt=1.0 / 60.0, restitution=1.0, friction=0.0,
                            gravity=(0.0, 0.0, 0.0))
    g = _gpu()
    g.set_scene(s)
    g.set_worlds(np.arange(0, 2 * len(worlds) + 1, 2), np.full((len(worlds), 3), 4.0))
    m = 1.0 / s.inv_mass

    def energy_momentum():
        x, v, q, w = g.state()
        ke = 0.5 * m * np.sum(v * v, axis=1)
        p = m[:, None] * v
        return ke.reshape(-1, 2).sum(axis=1), p.reshape(-1, 2, 3).sum(axis=1)

    e0, p0 = energy_momentum()
    hits = 0
    for _ in range(60):
        g.step(1)
        hits += len(g.contacts()["i"])
    e1, p1 = energy_momentum()
    assert hits >= len(worlds), "every pair collided"
    assert np.max(np.abs(e1 - e0) / e0) < 1e-10
    assert np.max(np.abs(p1 - p0)) < 1e-10
    g.close()


def test_gpu_narrow_pairs_match_golden(gpu_available):
    """1100 reference narrow-phase vectors (every pair of sphere, box,
    capsule, cylinder, 12-vertex hull, and plane against each), each pair
    stepped as a periodic world of its own from rest poses that the step
    leaves unchanged: its constraint list is the reference's detectContact
    (+ box-box manifold) on the pair, bit for bit (contacts within 1e-9
    relative is the stated tolerance)."""
    s, pairs, count, off, nrm, pt, dep = fixtures.load_narrow()
    assert np.all(pairs[:, 1] == pairs[:, 0] + 1) and np.all(pairs[:, 0] == 2 * np.arange(len(pairs)))
    s.params.gravity = (0.0, 0.0, 0.0)
    _say(f"narrow pairs: {len(pairs)} two-body worlds")
    g = _gpu()
    g.set_scene(s)
    g.set_worlds(np.arange(0, s.n + 1, 2), np.tile(np.asarray(s.params.half, np.float64), (len(pairs), 1)))
    g.step(1)   # integrate leaves the rest poses as they are; the contacts are found on them
    c = g.contacts()
    assert np.all(c["j"] == c["i"] + 1)
    world = c["i"] // 2
    start = np.searchsorted(world, np.arange(len(pairs)))
    end = np.searchsorted(world, np.arange(len(pairs)), side="right")
    kinds = {}
    for t in range(len(pairs)):
        sl = slice(off[t], off[t + 1])
        got = slice(start[t], end[t])
        a, b = int(pairs[t][0]), int(pairs[t][1])
        what = f"pair {t} (shapes {s.shape[a]}, {s.shape[b]})"
        assert end[t] - start[t] == count[t], f"{what}: {end[t] - start[t]} contacts vs {count[t]}"
        assert np.array_equal(c["normal"][got], nrm[sl]), f"{what}: normal"
        assert np.array_equal(c["point"][got], pt[sl]), f"{what}: point"
        assert np.array_equal(c["depth"][got], dep[sl]), f"{what}: depth"
        if count[t]:
            key = tuple(sorted((int(s.shape[a]), int(s.shape[b]))))
            kinds[key] = kinds.get(key, 0) + 1
    _say(f"narrow pairs in contact by shape pair: {sorted(kinds.items())}")
    assert (0, 5) in kinds and (1, 5) in kinds and (0, 1) in kinds and (1, 1) in kinds
    g.close()
