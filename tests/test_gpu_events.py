"""Contact events (enter / stay / exit) on the GPU against the CPU oracle: the reference's own
event test (physxcoopa/tests/contact_events_test.cpp), lockstep event lists on scenes with solid
and trigger contacts, a removed body, queue accumulation and clearing (world.h:744-748, 1230),
queue overflow, and the set identities at the full 16k-body pile size."""
import numpy as np
import pytest

from physxcoopa_amd import scenes, _abi
from physxcoopa_amd._abi import PXG_DYNAMIC
from parity import lockstep
from test_oracle_events import (ENTER, STAY, EXIT, trigger_probe_scene, count_pair_events, events_scene,
                                expected_events)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def World():
    from physxcoopa_amd.world import World
    return World


def rows(ev):
    return np.stack([ev["a"], ev["b"], ev["type"], ev["is_trigger"]], axis=1) if len(ev) else np.zeros((0, 4), np.uint32)


def test_trigger_enter_exit_once_on_gpu(World, oracle):
    # contact_events_test.cpp:23-61, both sides step by step
    b, s, t, p = trigger_probe_scene()
    cfg = _abi.default_config(gravity=(0, 0, 0))
    gw, ow = World(cfg, b, s), oracle.OracleWorld(cfg, b, s)
    total = {ENTER: 0, STAY: 0, EXIT: 0}
    for step in range(200):
        gw.clear_events()
        ow.clear_events()
        gw.step_fixed()
        ow.step_fixed()
        g = rows(gw.events())
        np.testing.assert_array_equal(g, ow.events(), err_msg="step %d" % step)
        for k, v in count_pair_events(g, t, p, True).items():
            total[k] += v
    print("trigger events:", total, flush=True)
    assert total[ENTER] == 1 and total[EXIT] == 1
    assert 60 <= total[STAY] <= 70


def test_lockstep_events_with_triggers(World, oracle):
    b, s = events_scene()
    print("lockstep events: %d bodies" % len(b), flush=True)
    _, _, rep = lockstep(World, oracle.OracleWorld, _abi.default_config(), b, s, 90)
    print(rep, flush=True)
    assert rep["events"] > 0


def test_lockstep_events_config1(World, oracle):
    sc = scenes.config1_scene(seed=1)
    _, _, rep = lockstep(World, oracle.OracleWorld, _abi.default_config(**sc.config), sc.bodies, sc.shapes, 120)
    print(rep, flush=True)
    assert rep["events"] > 0


def test_events_accumulate_and_removed_body_drops_its_exits(World, oracle):
    b, s = events_scene(seed=12)
    cfg = _abi.default_config()
    gw, ow = World(cfg, b, s), oracle.OracleWorld(cfg, b, s)
    for _ in range(60):
        gw.step_fixed()
        ow.step_fixed()
    for w in (gw, ow):
        w.clear_events()
    for _ in range(3):  # no clear in between: the queue keeps all three steps' events
        gw.step_fixed()
        ow.step_fixed()
    g3 = rows(gw.events())
    np.testing.assert_array_equal(g3, ow.events())
    assert len(g3) > len(gw.manifolds())
    # remove a body in contact: body slot dead, its shape disabled, on both sides
    m = gw.manifolds()
    victim = int(m["a"][np.nonzero(b["type"][m["a"]] == PXG_DYNAMIC)[0][0]])
    bodies = gw.bodies()
    bodies["alive"][victim] = 0
    s2 = s.copy()
    s2["enabled"][s2["owner"] == victim] = 0
    gw.set_bodies(bodies)
    ow.set_bodies(bodies)
    gw.set_shapes(s2)
    ow.set_shapes(s2)
    gw.set_fat_aabbs(ow.fat_aabbs())
    for w in (gw, ow):
        w.clear_events()
    gw.step_fixed()
    ow.step_fixed()
    g = rows(gw.events())
    np.testing.assert_array_equal(g, ow.events())
    np.testing.assert_array_equal(g, expected_events(m, gw.manifolds(), s2, bodies))
    assert not np.any((g[:, 0] == victim) | (g[:, 1] == victim))


def test_step_clears_the_queue_and_overflow_is_reported(World, oracle):
    from physxcoopa_amd._lib import PxgError
    sc = scenes.config1_scene(seed=1)
    cfg = _abi.default_config(**sc.config)
    gw, ow = World(cfg, sc.bodies, sc.shapes), oracle.OracleWorld(cfg, sc.bodies, sc.shapes)
    for _ in range(80):
        gw.step_fixed()
        ow.step_fixed()
    # step() clears the queue, then runs its substeps (world.h:1229-1240): two here
    assert gw.step(2.0 * float(cfg.fixed_dt) + 1e-5) == 2
    ow.clear_events()
    ow.step_fixed()
    ow.step_fixed()
    g = rows(gw.events())
    assert len(g) > 0
    np.testing.assert_array_equal(g, ow.events())
    # a queue of 16 events overflows in one step: reported, the events up to the capacity kept
    gw.set_events(True, 16)
    gw.step_fixed()
    with pytest.raises(PxgError, match="dropped"):
        gw.events()
    gw.clear_events()
    assert len(gw.events()) == 0
    # switched off: no events; switched on again: an empty snapshot, so every contact enters
    gw.set_events(False)
    gw.step_fixed()
    assert len(gw.events()) == 0
    gw.set_events(True, 1 << 16)
    gw.step_fixed()
    ev = gw.events()
    assert len(ev) == len(gw.manifolds()) and np.all(ev["type"] == ENTER)


def test_full_pile_events_are_the_manifold_set_difference(World):
    """At the 16k-body pile size (no oracle run): a step's events are exactly the set difference
    of the GPU's own consecutive manifold sets, in the reference's order; checked while the pile
    falls onto the ground (step 100) and once it has landed (step 200)."""
    sc = scenes.config2_scene(seed=2, n=16384, lattice=(64, 64, 40))
    gw = World(_abi.default_config(**sc.config), sc.bodies, sc.shapes)
    for step in range(201):
        if step in (100, 200):
            prev = gw.manifolds()
            gw.clear_events()
            gw.step_fixed()
            got = rows(gw.events())
            exp = expected_events(prev, gw.manifolds(), sc.shapes, sc.bodies)
            print("pile step %d: %d events, %d manifolds" % (step, len(got), len(prev)), flush=True)
            np.testing.assert_array_equal(got, exp, err_msg="step %d" % step)
            assert np.count_nonzero(got[:, 2] == ENTER) > 0
        else:
            gw.step_fixed()
