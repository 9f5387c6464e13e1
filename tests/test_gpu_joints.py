"""Joints on the GPU against the CPU oracle: the reference's joint tests (physxcoopa/tests/joints_test.cpp,
sleep_test.cpp:134-175) run through the C ABI with the reference's bars, ragdoll scenes in lockstep
and free running, the eager launch path, and the joint / ignored-pair contact filters.

The GPU evaluates atan2 in double and rounds once; the oracle does the same in its rounded mode,
which is what the bit-exact checks compare against. Against the reference's float std::atan2 the
results agree to the lockstep tolerances (a one-ulp difference in an angle per call)."""
import math

import numpy as np
import pytest

from physxcoopa_amd import scenes, _abi, joints
from physxcoopa_amd._abi import PXG_DYNAMIC, PXG_STATIC
from parity import lockstep

pytestmark = pytest.mark.gpu

F = np.float32
H = F(1.0) / F(60.0)


@pytest.fixture(scope="module")
def sides(oracle):
    from physxcoopa_amd.world import World
    return World, oracle.OracleWorld


def _pair(World, OracleWorld, sb, jts, gravity=(0, 0, -9.81), ignored=None):
    b, s = sb.arrays()
    cfg = _abi.default_config(gravity=gravity)
    gw, ow = World(cfg, b, s), OracleWorld(cfg, b, s)
    ow.set_atan2_mode(True)
    for w in (gw, ow):
        w.set_joints(joints.table(jts))
        if ignored is not None:
            w.set_ignored_pairs(ignored)
    return gw, ow


def _run_both(gw, ow, steps, h=H, each=None):
    for i in range(steps):
        gw.step_fixed(h)
        ow.step_fixed(h)
        assert gw.state_hash() == ow.state_hash(), "step %d: GPU and oracle differ" % (i + 1)
        if each:
            each(gw.bodies())


def _anchor(body, local):
    return body["position"].astype(np.float64) + joints.quat_rotate(body["orientation"], local).astype(np.float64)


def _hinge_angle(oracle, qa, qb, axis, rest):
    import ctypes
    from physxcoopa_amd._abi import ptr
    arrs = [np.ascontiguousarray(np.asarray(x, np.float32)) for x in (qa, qb, axis, rest)]
    return float(oracle.lib().pxo_hinge_angle(*[ptr(a) for a in arrs], 1))


def test_door_stops_at_its_limit(sides, oracle):
    # joints_test.cpp:23-62
    World, OW = sides
    sb = scenes.SceneBuilder()
    fr = sb.add_body(PXG_STATIC, (0, 0, 0))
    door = sb.add_body(PXG_DYNAMIC, (1, 0, 0), angular_velocity=(0, 0, 5))
    b, _ = sb.arrays()
    j = joints.hinge(b, fr, door, (0, 0, 0), (-1, 0, 0), (0, 0, 1), (0, 0, 1), True, 0.0, F(math.radians(90.0)))
    gw, ow = _pair(World, OW, sb, [j], gravity=(0, 0, 0))
    _run_both(gw, ow, 300)
    bb = gw.bodies()
    assert np.linalg.norm(_anchor(bb[door], j["local_anchor_b"]) - _anchor(bb[fr], j["local_anchor_a"])) < 0.02
    ang = _hinge_angle(oracle, bb[fr]["orientation"], bb[door]["orientation"], j["local_axis_a"], j["rest_relative_rotation"])
    assert abs(ang - math.radians(90)) < math.radians(3)
    assert np.linalg.norm(bb[door]["angular_velocity"]) < 0.05
    gj, oj = gw.joints(), ow.joints()
    assert np.array_equal(gj.view(np.uint8), oj.view(np.uint8))


def test_pendulum_and_zero_range_hinge(sides, oracle):
    # joints_test.cpp:70-153
    World, OW = sides
    sb = scenes.SceneBuilder()
    ceil = sb.add_body(PXG_STATIC, (0, 0, 0))
    pend = sb.add_body(PXG_DYNAMIC, (1, 0, 0))
    frame = sb.add_body(PXG_STATIC, (0, 5, 0))
    panel = sb.add_body(PXG_DYNAMIC, (1, 5, 0), angular_velocity=(0, 0, 5))
    b, _ = sb.arrays()
    jts = [joints.hinge(b, ceil, pend, (0, 0, 0), (-1, 0, 0), (0, 1, 0), (0, 1, 0)),
           joints.hinge(b, frame, panel, (0, 0, 0), (-1, 0, 0), (0, 0, 1), (0, 0, 1), True, 0.0, 0.0)]
    gw, ow = _pair(World, OW, sb, jts)
    drift = []
    _run_both(gw, ow, 120, each=lambda bb: drift.append(
        np.linalg.norm(_anchor(bb[pend], jts[0]["local_anchor_b"]) - _anchor(bb[ceil], jts[0]["local_anchor_a"]))))
    bb = gw.bodies()
    assert max(drift) < 0.05 and bb[pend]["position"][2] < -0.3
    ang = _hinge_angle(oracle, bb[frame]["orientation"], bb[panel]["orientation"], jts[1]["local_axis_a"],
                       jts[1]["rest_relative_rotation"])
    assert abs(ang) < math.radians(3)


def test_ball_joint_holds_and_releases(sides):
    # joints_test.cpp:160-196
    World, OW = sides
    sb = scenes.SceneBuilder()
    ceil = sb.add_body(PXG_STATIC, (0, 0, 0))
    bob = sb.dynamic_capsule((0, 0, -0.5), 0.08, 0.42, mass=2.0, velocity=(2.0, 1.5, 0.0))
    b, _ = sb.arrays()
    j = joints.ball(b, ceil, bob, (0, 0, 0), (0, 0, 0.5))
    gw, ow = _pair(World, OW, sb, [j])
    drift = []
    _run_both(gw, ow, 180, each=lambda bb: drift.append(np.linalg.norm(_anchor(bb[bob], j["local_anchor_b"]))))
    assert max(drift) < 0.02
    assert abs(np.linalg.norm(gw.bodies()[bob]["position"]) - 0.5) < 0.03
    for w in (gw, ow):
        jt = w.joints()
        jt["valid"] = 0
        w.set_joints(jt)
    _run_both(gw, ow, 30)
    assert np.linalg.norm(gw.bodies()[bob]["position"]) > 0.8


def test_cone_twist_limits(sides, oracle):
    # joints_test.cpp:204-254
    World, OW = sides
    sb = scenes.SceneBuilder()
    sh = sb.add_body(PXG_STATIC, (0, 0, 0))
    arm = sb.dynamic_capsule((0, 0, -0.3), 0.06, 0.24, mass=2.0, angular_velocity=(6.0, 0.0, 9.0))
    b, _ = sb.arrays()
    swing, tmin, tmax = F(math.radians(40.0)), F(math.radians(-20.0)), F(math.radians(25.0))
    j = joints.cone_twist(b, sh, arm, (0, 0, 0), (0, 0, 0.3), (0, 0, -1), swing, tmin, tmax)
    gw, ow = _pair(World, OW, sb, [j], gravity=(0, 0, 0))
    tw = []
    _run_both(gw, ow, 240, each=lambda bb: tw.append(
        _hinge_angle(oracle, bb[sh]["orientation"], bb[arm]["orientation"], j["local_axis_a"], j["rest_relative_rotation"])))
    tol = math.radians(4.0)
    assert tmin - tol <= min(tw) <= tmin + tol and max(tw) <= tmax + tol


def test_jointed_island_sleep_and_wake(sides):
    # sleep_test.cpp:134-175, and wake_jointed_islands (solver.h:887-901)
    World, OW = sides
    sb = scenes.SceneBuilder()
    a = sb.add_body(PXG_DYNAMIC, (0, 0, 0))
    c = sb.add_body(PXG_DYNAMIC, (1, 0, 0), angular_velocity=(0, 0, 2))
    s1 = sb.add_body(PXG_DYNAMIC, (0, 3, 0), velocity=(1, 0, 0))
    s2 = sb.add_body(PXG_DYNAMIC, (1, 3, 0), awake=False)
    s3 = sb.add_body(PXG_DYNAMIC, (2, 3, 0), awake=False)
    b, _ = sb.arrays()
    jts = [joints.hinge(b, a, c, (0, 0, 0), (-1, 0, 0), (0, 0, 1), (0, 0, 1)),
           joints.ball(b, s2, s3, (0.5, 0, 0), (-0.5, 0, 0)), joints.ball(b, s1, s2, (0.5, 0, 0), (-0.5, 0, 0))]
    gw, ow = _pair(World, OW, sb, jts, gravity=(0, 0, 0))
    _run_both(gw, ow, 90)
    bb = gw.bodies()
    assert bb[a]["awake"] == 1 and np.linalg.norm(bb[c]["angular_velocity"]) > 0.5
    assert bb[s3]["awake"] == 1


def test_contact_filters(sides):
    # world.h:1616-1620
    World, OW = sides
    cfg = _abi.default_config(gravity=(0, 0, 0))
    for cc, ign, expect in ((False, None, 0), (True, None, 1), (True, [[1, 0]], 0)):
        sb = scenes.SceneBuilder()
        p = sb.dynamic_box((0, 0, 0), (0.5, 0.5, 0.5))
        q = sb.dynamic_box((0.8, 0, 0), (0.5, 0.5, 0.5))
        b, s = sb.arrays()
        j = joints.ball(b, p, q, (0.4, 0, 0), (-0.4, 0, 0), collide_connected=cc)
        gw = World(cfg, b, s)
        gw.set_joints(joints.table([j]))
        if ign is not None:
            gw.set_ignored_pairs(ign)
        gw.step_fixed(H)
        assert len(gw.manifolds()) == expect


def _ragdolls(n, seed):
    sc, recs, ig = scenes.ragdoll_scene(seed=seed, n_ragdolls=n)
    return sc, joints.table(recs), ig


def test_ragdoll_lockstep_bit_exact(sides):
    """Four ragdolls and 32 boxes: every step the GPU's bodies and joint accumulators equal the
    oracle's bit for bit (oracle fed the GPU's state and contact order)."""
    World, OW = sides
    sc, jt, ig = _ragdolls(4, 11)
    c = _abi.default_config(**sc.config)
    gw, ow, rep = lockstep(World, OW, c, sc.bodies, sc.shapes, 200, joints=jt, ignored=ig)
    print(rep, flush=True)
    assert rep["manifolds"] > 0
    assert rep["bitexact_bodies"] == rep["steps"] and rep["bitexact_joints"] == rep["steps"]


def test_ragdoll_lockstep_against_libm_atan2(sides):
    """The same against the reference's float std::atan2: within the lockstep tolerances."""
    World, OW = sides
    sc, jt, ig = _ragdolls(2, 12)
    c = _abi.default_config(**sc.config)
    gw, ow, rep = lockstep(World, OW, c, sc.bodies, sc.shapes, 120, joints=jt, ignored=ig, rounded_atan2=False)
    print(rep, flush=True)
    assert rep["steps"] == 120


def test_ragdoll_free_running_equals_oracle(sides):
    """One ragdoll (the per-scene content of the batched ragdoll workload), 600 free-running steps
    in the reference's own solve order: the state hash equals the oracle's at every step."""
    World, OW = sides
    sc, jt, ig = _ragdolls(1, 5)
    c = _abi.default_config(**sc.config)
    gw, ow = World(c, sc.bodies, sc.shapes), OW(c, sc.bodies, sc.shapes)
    ow.set_atan2_mode(True)
    for w in (gw, ow):
        w.set_joints(jt)
        w.set_ignored_pairs(ig)
    for i in range(600):
        gw.step_fixed()
        ow.step_fixed()
        assert gw.state_hash() == ow.state_hash(), "step %d" % (i + 1)
    gj, oj = gw.joints(), ow.joints()
    assert np.array_equal(gj.view(np.uint8), oj.view(np.uint8))


def test_ragdoll_eager_path(sides, monkeypatch):
    """The one-launch-per-layer fallback gives the same results."""
    monkeypatch.setenv("PXG_EAGER", "1")
    World, OW = sides
    sc, jt, ig = _ragdolls(3, 13)
    c = _abi.default_config(**sc.config)
    gw, ow, rep = lockstep(World, OW, c, sc.bodies, sc.shapes, 60, joints=jt, ignored=ig, feed_order=False)
    assert rep["bitexact_bodies"] == rep["steps"]


def test_ragdoll_residual_mode(sides):
    """Residual mode with joints (a joint pass after every contact pass, as in the fixed mode)."""
    World, OW = sides
    sc, jt, ig = _ragdolls(2, 14)
    c = _abi.default_config(residual_mode=1, residual_tolerance=1e-3, residual_max_iterations=48, **sc.config)
    gw, ow, rep = lockstep(World, OW, c, sc.bodies, sc.shapes, 60, joints=jt, ignored=ig)
    assert rep["bitexact_bodies"] == rep["steps"]
    assert gw.stats()["velocity_passes"] == ow.stats()["velocity_passes"]
