"""GPU step vs the CPU oracle, step by step (see tests/parity.py for the bar).

Runs through the C ABI (libpxg.so via physxcoopa_amd.world.World)."""
import numpy as np
import pytest

from physxcoopa_amd import scenes, _abi
from parity import lockstep, compare_step

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sides(oracle):
    from physxcoopa_amd.world import World
    return World, oracle.OracleWorld


def _run(sides, scene_arrays, steps, resync=True, feed_order=True, **cfg):
    World, OracleWorld = sides
    b, s = scene_arrays
    c = _abi.default_config(**cfg)
    print("lockstep: %d bodies, %d shapes, %d steps" % (len(b), len(s), steps), flush=True)
    gw, ow, rep = lockstep(World, OracleWorld, c, b, s, steps, resync=resync, feed_order=feed_order)
    print(rep, flush=True)
    return gw, ow, rep


def test_parity_config1_boxes(sides):
    sc = scenes.config1_scene(seed=1)
    _, _, rep = _run(sides, (sc.bodies, sc.shapes), 120, **sc.config)
    assert rep["manifolds"] > 0


def test_parity_sphere_rain(sides):
    _, _, rep = _run(sides, scenes.sphere_rain_scene(), 150)
    assert rep["manifolds"] > 0


def test_parity_mixed_pile(sides):
    _, _, rep = _run(sides, scenes.mixed_pile_scene(n=300, seed=7), 120, velocity_iterations=8, position_iterations=3)
    assert rep["manifolds"] > 0


def test_parity_large_pile_few_steps(sides):
    sc = scenes.config2_scene(seed=2, n=2048, lattice=(32, 32, 8))
    _, _, rep = _run(sides, (sc.bodies, sc.shapes), 60, **sc.config)
    assert rep["manifolds"] > 0


def test_free_running_hash_matches_oracle_config1(sides):
    """No resync: 300 steps of config 1, the GPU and the oracle (fed the GPU's colour order each
    step) must agree on the final world_state_hash."""
    sc = scenes.config1_scene(seed=1)
    gw, ow, rep = _run(sides, (sc.bodies, sc.shapes), 300, resync=False, **sc.config)
    assert gw.state_hash() == ow.state_hash()


# In the default (reference) solve order the GPU's colour-by-colour solve is the reference's own
# sequential order up to commuting manifolds: the oracle solving in its own order, with no hint
# from the GPU, must give the same bits.
@pytest.mark.parametrize("name", ["config1", "mixed_pile", "sphere_rain", "pile_2048"])
def test_reference_order_equals_oracle_natural_order(sides, name):
    if name == "config1":
        sc = scenes.config1_scene(seed=5)
        arrays, cfg, steps = (sc.bodies, sc.shapes), sc.config, 200
    elif name == "mixed_pile":
        arrays, cfg, steps = scenes.mixed_pile_scene(n=300, seed=21), dict(velocity_iterations=8, position_iterations=3), 150
    elif name == "sphere_rain":
        arrays, cfg, steps = scenes.sphere_rain_scene(seed=99), {}, 200
    else:
        sc = scenes.config2_scene(seed=4, n=2048, lattice=(32, 32, 8))
        arrays, cfg, steps = (sc.bodies, sc.shapes), sc.config, 80
    _, _, rep = _run(sides, arrays, steps, feed_order=False, **cfg)
    assert rep["manifolds"] > 0
    assert rep["bitexact_bodies"] == rep["steps"]


def test_reference_order_free_running_equals_oracle(sides):
    """No resync and no order hint: 300 free-running steps of config 1 (the reference's scene) on
    the GPU end in the same world_state_hash as the oracle's own sequential solve."""
    sc = scenes.config1_scene(seed=1)
    gw, ow, rep = _run(sides, (sc.bodies, sc.shapes), 300, resync=False, feed_order=False, **sc.config)
    assert gw.state_hash() == ow.state_hash()


def test_coloured_order_parity(sides):
    """The faster greedy colouring: a different sequential order, checked against the oracle fed
    that order."""
    _, _, rep = _run(sides, scenes.mixed_pile_scene(n=300, seed=8), 100, velocity_iterations=8, position_iterations=3,
                     solve_order=_abi.PXG_ORDER_COLOURED)
    assert rep["manifolds"] > 0 and rep["bitexact_bodies"] == rep["steps"]


def test_gpu_deterministic_across_runs(sides):
    World, _ = sides
    b, s = scenes.mixed_pile_scene(n=300, seed=3)
    c = _abi.default_config()
    hashes = []
    for _ in range(2):
        w = World(c, b, s)
        for _ in range(100):
            w.step_fixed()
        hashes.append(w.state_hash())
        w.close()
    assert hashes[0] == hashes[1]


def test_parity_eager_launch_path(sides, monkeypatch):
    """The one-launch-per-colour fallback gives the same results as the persistent kernels."""
    monkeypatch.setenv("PXG_EAGER", "1")
    _, _, rep = _run(sides, scenes.mixed_pile_scene(n=200, seed=11), 60, velocity_iterations=8, position_iterations=3)
    assert rep["manifolds"] > 0


@pytest.mark.parametrize("every,eager", [(1, False), (3, False), (3, True)], ids=["every_pass", "every_3rd", "every_3rd_eager"])
def test_parity_residual_mode(sides, every, eager, monkeypatch):
    """Residual mode: velocity passes repeat until the residual <= 1e-4 (cap 64 here), the residual
    evaluated after every `every`-th pass and after the 64th (64 is not a multiple of 3); the GPU
    must run the same number of passes as the oracle and reach the same state, on the persistent
    and on the one-launch-per-colour path."""
    World, OracleWorld = sides
    if eager:
        monkeypatch.setenv("PXG_EAGER", "1")
    sc = scenes.config1_scene(seed=1)
    c = _abi.default_config(residual_mode=1, residual_tolerance=1e-4, residual_max_iterations=64,
                            residual_check_interval=every, **sc.config)
    gw, ow, rep = lockstep(World, OracleWorld, c, sc.bodies, sc.shapes, 80, resync=True)
    print(rep, gw.stats(), ow.stats(), flush=True)
    assert gw.stats()["velocity_passes"] == ow.stats()["velocity_passes"]
    assert abs(gw.stats()["residual"] - ow.stats()["residual"]) <= 1e-6 + 1e-4 * abs(ow.stats()["residual"])


@pytest.mark.parametrize("n,lattice,warm,steps,every", [(2048, (32, 32, 8), 90, 6, 1), (2048, (32, 32, 8), 90, 4, 8),
                                                        (16384, (64, 64, 40), 240, 3, 1)],
                         ids=["pile_2048", "pile_2048_every_8th", "pile_16384"])
def test_parity_residual_mode_coloured_piles(sides, n, lattice, warm, steps, every):
    """Residual mode on piles large enough for a multi-block solve with a tail of short colours
    (the coloured order): the residual of each manifold is computed during the colour phase after
    the last colour writing its bodies, and the GPU must run the oracle's pass count and reach its
    state bit for bit (with the residual after every pass, or after every 8th and the last). The
    GPU state after `warm` fixed-iteration steps is handed to the oracle; the oracle then solves in
    the GPU's order."""
    World, OracleWorld = sides
    sc = scenes.config2_scene(seed=2, n=n, lattice=lattice)
    c = _abi.default_config(**sc.config)
    c.solve_order = _abi.PXG_ORDER_COLOURED
    gw = World(c, sc.bodies, sc.shapes)
    for _ in range(warm):
        gw.step_fixed()
    rc = _abi.default_config(residual_mode=1, residual_tolerance=1e-4, residual_max_iterations=256,
                             residual_check_interval=every, **sc.config)
    rc.solve_order = _abi.PXG_ORDER_COLOURED
    gw.set_config(rc)
    ow = OracleWorld(rc, gw.bodies(), sc.shapes)
    report = dict(steps=0, manifolds=0, bitexact_manifolds=0, bitexact_bodies=0, bitexact_joints=0, max_body_diff=0.0,
                  colours=0)
    for i in range(steps):
        ow.set_bodies(gw.bodies())
        ow.set_fat_aabbs(gw.fat_aabbs())
        ow.set_warm_cache(*gw.warm_cache())
        pre = gw.bodies()
        gw.clear_events()
        ow.clear_events()
        gw.step_fixed()
        gm = gw.manifolds()
        order, _ = gw.solve_order()
        ow.set_order(np.stack([gm["shape_a"][order], gm["shape_b"][order]], axis=1))
        ow.step_fixed()
        # the oracle's first step has no previous pair set to compare events against
        compare_step(gw, ow, pre, report, check_events=i > 0)
        gs, os_ = gw.stats(), ow.stats()
        print("step %d: passes %d / %d, residual %.3g / %.3g, colours %d" %
              (i, gs["velocity_passes"], os_["velocity_passes"], gs["residual"], os_["residual"], gs["colours"]), flush=True)
        assert gs["velocity_passes"] == os_["velocity_passes"]
        assert gs["residual"] == os_["residual"]
    assert report["bitexact_bodies"] == steps


def test_full_pile_free_running_equals_oracle(sides):
    """The 16384-body pile (64x64x40 lattice) at full size, free-running for 300 steps with no
    resync: every step's world_state_hash and awake count equal the oracle's. Islands start to
    fall asleep after about 100 steps, so this covers the union-find and sleep kernels at the
    concurrency of the full workload."""
    World, OracleWorld = sides
    sc = scenes.config2_scene(seed=2, n=16384, lattice=(64, 64, 40))
    c = _abi.default_config(**sc.config)
    gw = World(c, sc.bodies, sc.shapes)
    ow = OracleWorld(c, sc.bodies, sc.shapes)
    asleep = 0
    for i in range(300):
        gw.step_fixed()
        ow.step_fixed()
        if gw.state_hash() != ow.state_hash():
            gb, ob = gw.bodies(), ow.bodies()
            raise AssertionError("step %d: hashes differ (awake %d vs %d)" % (i + 1, gb["awake"].sum(), ob["awake"].sum()))
        if i % 50 == 49:
            asleep = int(len(sc.bodies) - gw.bodies()["awake"].sum())
            print("step", i + 1, "asleep", asleep, flush=True)
    assert asleep > 0


def test_gpu_state_hash_equals_fnv_over_body_bytes(sides):
    """pxg_state_hash equals the numpy FNV-1a restatement of world.h:1334-1359 over the bodies
    downloaded from the GPU (config-1 scene, 30 steps), and a dead body is left out."""
    from test_oracle_residual import state_hash_numpy
    World, _ = sides
    sc = scenes.config1_scene(seed=1)
    gw = World(_abi.default_config(**sc.config), sc.bodies, sc.shapes)
    for _ in range(30):
        gw.step_fixed()
    b = gw.bodies()
    assert gw.state_hash() == state_hash_numpy(b)
    b["alive"][5] = 0
    gw.set_bodies(b)
    assert gw.state_hash() == state_hash_numpy(b)
    gw.close()
