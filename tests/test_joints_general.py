"""General joints (World::solveJoints, src/world.cpp:206-315: ball, hinge with
motor and angle limit, fixed, slider, breakable) in the C restatement,
pinned against the reference: the golden joint scenes (the reference's own
joint checks built as the facade builds them, chains of every joint type,
closed with sleeping and periodic; tests/golden/scene_joints_*.npz, checked
with every scene fixture in test_oracle.py, joint states included), the live
reference where built, and the reference's checks themselves
(selftest.cpp:1088-1179) on the oracle's runs."""
import ctypes as C
import math

import numpy as np
import pytest

import engines
import fixtures
import scenes
from nativeengine_amd import _abi
from nativeengine_amd import scene as S


def _run(name):
    s, steps, _ = fixtures.load_scene(f"joints_{name}")
    o = engines.OracleWorld()
    o.set_scene(s)
    o.step(steps)
    return o


def test_reference_joint_checks_on_the_oracle(oracle_lib):
    x = _run("fixed").state()[0][1]
    assert abs(x[0] - 1.0) < 0.08 and abs(x[1] - 5.0) < 0.15 and abs(x[2]) < 0.08
    x = _run("ball").state()[0][1]
    assert abs(np.linalg.norm(x - np.array([0.0, 5.0, 0.0])) - 2.0) < 0.05 and x[1] < 5.0
    o = _run("hinge_motor")
    assert abs(o.state()[3][1][1]) > 1.0 and abs(o.state()[0][1][2]) > 0.2
    o = _run("breakable")
    assert o.general_joints()[0][0] == 1 and o.state()[0][1][1] < 2.0
    x = _run("slider").state()[0][1]
    assert abs(x[0]) > 0.5 and abs(x[1]) < 0.05 and abs(x[2]) < 0.05
    o = _run("bridge")
    q = o.state()[2]
    tilt = 0.0
    for k in range(1, 9):
        w, a, b, c = q[k]
        lz_z = 1 - 2 * (a * a + b * b)   # z of the plank's local z axis in the world
        tilt = max(tilt, math.degrees(math.acos(min(1.0, abs(lz_z)))))
    assert tilt < 5.0


@pytest.mark.skipif(not engines.ref_available(), reason="oracle/_ref not built (reference sources absent)")
def test_joint_frames_equal_the_facades():
    """scene.general_joint's frames (local anchors and axes, ref_rel) are the
    bits PhysicsWorld::addGeneralJoint computes, on random poses."""
    lib = engines._load("ref")
    dp = C.POINTER(C.c_double)
    lib.ne_ref_joint_frames.argtypes = [dp, dp, dp, dp, dp, dp, C.POINTER(_abi.Joint)]
    rng = np.random.default_rng(3)
    for k in range(200):
        s = S.Scene(2)
        s.x[:] = rng.uniform(-5, 5, (2, 3))
        s.q[:] = S.random_unit_quaternions(rng, 2)
        anchor = rng.uniform(-5, 5, 3) if k % 4 else None
        axis = rng.normal(size=3)
        j = S.general_joint(s, _abi.JOINT_HINGE, 0, 1, anchor=anchor, axis=axis)
        out = _abi.Joint()
        arr = lambda a: np.ascontiguousarray(a, np.float64)
        xa, qa, xb, qb, ax = arr(s.x[0]), arr(s.q[0]), arr(s.x[1]), arr(s.q[1]), arr(axis)
        an = None if anchor is None else arr(anchor)
        assert lib.ne_ref_joint_frames(_abi.dptr(xa), _abi.dptr(qa), _abi.dptr(xb), _abi.dptr(qb),
                                       None if an is None else _abi.dptr(an), _abi.dptr(ax), C.byref(out)) == 0
        for f in ("local_a", "local_b", "axis_a", "axis_b", "ref_rel"):
            assert np.array_equal(np.array(list(getattr(out, f))), j[f]), (k, f)


@pytest.mark.skipif(not engines.ref_available(), reason="oracle/_ref not built (reference sources absent)")
@pytest.mark.parametrize("seed,periodic,limits", [(5, False, True), (6, True, True), (7, False, False)])
def test_oracle_joints_match_the_reference_live(seed, periodic, limits):
    s = scenes.joint_mix_scene(seed, periodic=periodic, limits=limits)
    o = engines.OracleWorld()
    o.set_scene(s)
    r = engines.RefWorld()
    r.set_scene(s)
    for t in range(150):
        o.step(1)
        r.step(1)
        if t % 10 == 9:
            for a, b in zip(o.state(), r.state()):
                assert np.array_equal(a, b), t
    for a, b in zip(o.general_joints(), r.general_joints()):
        assert np.array_equal(a, b)


def test_limits_scene_has_active_limits(oracle_lib):
    """The hinge-limit fixture drives its limited hinges past their limits
    (the angle 2 atan2(.) leaves [-0.3, 0.4]), so the limit branch is run."""
    s, steps, _ = fixtures.load_scene("joints_limits")
    o = engines.OracleWorld()
    o.set_scene(s)
    lim = [k for k in range(len(s.gjoints)) if s.gjoints[k]["use_limit"]]
    assert lim
    seen = 0
    for _ in range(steps):
        o.step(1)
        _, _, q, _ = o.state()
        for k in lim:
            j = s.gjoints[k]
            a, b = int(j["a"]), int(j["b"])
            qa, qb, rr = q[a], q[b], j["ref_rel"]
            dq = S._q_mul(S._q_mul(tuple(qa), S._q_conj(tuple(qb))), S._q_conj(tuple(rr)))
            aa = S._v_normalized(S._q_rotate(tuple(qa), tuple(j["axis_a"])))
            ang = 2.0 * math.atan2(dq[1] * aa[0] + dq[2] * aa[1] + dq[3] * aa[2], dq[0])
            seen += ang <= j["lower"] or ang >= j["upper"]
    assert seen > 0
