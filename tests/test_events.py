"""Facade contact and trigger events (PhysicsWorld::processContacts,
src/physics_world.cpp:253-302) in the C restatement, pinned against the
reference's own callbacks: the golden event fixtures (written from
PhysicsWorld::step with every callback set, tests/golden/make_fixtures.py)
and, where it is built, the live reference.

The events of a step: every pair i < j of the post-step state that passes
the facade's filters (a sensor pair always, else not two immovable bodies;
layers; no-collide lists) and overlaps under detectContact -- one contact
per pair, no box-box manifold -- split into solid pairs (contact, begin,
stay, end) and pairs with a sensor (trigger enter, stay, exit), diffed
against the last step's lists.  End and exit carry the last step's contact.
"""
import numpy as np
import pytest

import engines
import fixtures
import scenes
from nativeengine_amd import _abi

NAMES = ["features", "sleeping_sensor"]


def _same(a, b, what):
    for f in ("i", "j", "point", "normal", "depth"):
        assert np.array_equal(a[f], b[f]), f"{what}: {f} differs ({len(a['i'])} vs {len(b['i'])} events)"


@pytest.mark.parametrize("name", NAMES)
def test_oracle_events_match_the_reference_fixture(oracle_lib, name):
    s, steps, ev, out = fixtures.load_events(name)
    o = engines.OracleWorld()
    o.set_scene(s)
    o.set_events(True)
    totals = np.zeros(len(engines.EVENT_KINDS), int)
    asleep = 0
    for t in range(steps):
        o.step(1)
        for k, kind in enumerate(engines.EVENT_KINDS):
            got = o.events(k)
            _same(got, ev[t][k], f"{name} step {t} {kind}")
            totals[k] += len(got["i"])
        asleep = max(asleep, int(np.sum((engines.body_aux(o, s.n)[0] & _abi.FLAG_SLEEPING) != 0)))
    for k, a in zip("xvqw", o.state()):
        assert np.array_equal(a, out[k])
    assert totals[_abi.NE_EVENT_BEGIN] > 0 and totals[_abi.NE_EVENT_END] > 0
    if name == "features":
        assert totals[_abi.NE_EVENT_TRIGGER_ENTER] > 0 and totals[_abi.NE_EVENT_TRIGGER_STAY] > 0
        assert totals[_abi.NE_EVENT_TRIGGER_EXIT] > 0, totals
    else:
        assert asleep > 10, asleep
        assert totals[_abi.NE_EVENT_TRIGGER_ENTER] > 0, totals


@pytest.mark.skipif(not engines.ref_available(), reason="oracle/_ref not built (reference sources absent)")
def test_oracle_events_match_the_reference_live():
    s = scenes.features_scene(5)
    o = engines.OracleWorld()
    o.set_scene(s)
    o.set_events(True)
    r = engines.RefWorld()
    r.set_scene(s)
    r.facade_begin()
    for t in range(40):
        o.step(1)
        r.facade_step(1)
        for k, kind in enumerate(engines.EVENT_KINDS):
            _same(o.events(k), r.events(k), f"step {t} {kind}")


def test_events_start_empty_and_reset_with_the_world(oracle_lib):
    """Enabling events, or uploading a new world, empties the last-step
    lists: the first step's pairs all begin (or enter) and nothing ends."""
    s = scenes.features_scene(0)
    o = engines.OracleWorld()
    o.set_scene(s)
    o.step(10)
    o.set_events(True)
    o.step(1)
    assert len(o.events(_abi.NE_EVENT_END)["i"]) == 0 and len(o.events(_abi.NE_EVENT_STAY)["i"]) == 0
    assert len(o.events(_abi.NE_EVENT_BEGIN)["i"]) == len(o.events(_abi.NE_EVENT_CONTACT)["i"]) > 0
    o.set_scene(s)
    o.step(1)
    assert len(o.events(_abi.NE_EVENT_END)["i"]) == 0 and len(o.events(_abi.NE_EVENT_TRIGGER_EXIT)["i"]) == 0
