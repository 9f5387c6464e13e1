"""The tiled solve order (PXG_ORDER_TILED) against the CPU oracle fed the GPU's order.

Each tile's colours run in one workgroup (its bodies and impulses in LDS where they fit), the
manifolds crossing tiles in two later phases; the order is colour-major and every colour is a
valid colour of the whole world, so the oracle's sequential solve of it must give the GPU's
state bit for bit. Covered: fixed iterations and residual mode on piles of 2,048 and 16,384 bodies
(tiles of 64 and 512 bodies), a small pile where most manifolds cross tiles (the later phases
carry the solve), and the global-memory form of the same kernel (PXG_TILED_LDS=0)."""
import numpy as np
import pytest

from physxcoopa_amd import scenes, _abi
from parity import lockstep, compare_step

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sides(oracle):
    from physxcoopa_amd.world import World
    return World, oracle.OracleWorld


def _pile_lockstep(sides, n, lattice, warm, steps, residual=False, every=1, cap=256):
    """`warm` fixed-iteration steps on the GPU in the tiled order, then `steps` steps in lockstep:
    each step the oracle gets the GPU's input state and the GPU's solve order."""
    World, OracleWorld = sides
    sc = scenes.config2_scene(seed=2, n=n, lattice=lattice)
    c = _abi.default_config(**sc.config)
    c.solve_order = _abi.PXG_ORDER_TILED
    gw = World(c, sc.bodies, sc.shapes)
    for _ in range(warm):
        gw.step_fixed()
    rc = _abi.default_config(**sc.config)
    rc.solve_order = _abi.PXG_ORDER_TILED
    if residual:
        rc.residual_mode, rc.residual_tolerance, rc.residual_max_iterations = 1, 1e-4, cap
        rc.residual_check_interval = every
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
        order, offs = gw.solve_order()
        ow.set_order(np.stack([gm["shape_a"][order], gm["shape_b"][order]], axis=1))
        ow.step_fixed()
        compare_step(gw, ow, pre, report, check_events=i > 0)
        gs, os_ = gw.stats(), ow.stats()
        print("step %d: passes %d / %d, residual %.3g / %.3g, colours %d, manifolds %d" %
              (i, gs["velocity_passes"], os_["velocity_passes"], gs["residual"], os_["residual"], gs["colours"],
               len(gm)), flush=True)
        assert gs["velocity_passes"] == os_["velocity_passes"]
        if residual:
            assert gs["residual"] == os_["residual"]
    assert report["bitexact_bodies"] == steps
    return gw, report


@pytest.mark.parametrize("lds", ["1", "0"], ids=["lds", "global"])
def test_tiled_order_fixed_iterations_pile_2048(sides, lds, monkeypatch):
    monkeypatch.setenv("PXG_TILED_LDS", lds)
    _pile_lockstep(sides, 2048, (32, 32, 8), 90, 5)


def test_tiled_order_residual_mode_pile_2048(sides):
    _pile_lockstep(sides, 2048, (32, 32, 8), 90, 4, residual=True, every=1)


def test_tiled_order_residual_every_8th_pile_2048(sides):
    _pile_lockstep(sides, 2048, (32, 32, 8), 90, 4, residual=True, every=8)


def test_tiled_order_residual_mode_pile_16384(sides):
    _pile_lockstep(sides, 16384, (64, 64, 40), 240, 2, residual=True, every=8, cap=512)


def test_tiled_order_small_pile_crossing_tiles(sides):
    """300 bodies over 32 tiles: most manifolds cross tiles, so the later phases (and the one
    group of the third) carry most of the solve."""
    World, OracleWorld = sides
    b, s = scenes.mixed_pile_scene(n=300, seed=8)
    c = _abi.default_config(velocity_iterations=8, position_iterations=3)
    c.solve_order = _abi.PXG_ORDER_TILED
    gw, ow, rep = lockstep(World, OracleWorld, c, b, s, 100, resync=True)
    print(rep, flush=True)
    assert rep["manifolds"] > 0 and rep["bitexact_bodies"] == rep["steps"]


def test_tiled_lds_and_global_forms_agree(sides, monkeypatch):
    """The LDS form and the global-memory form of the tiled solve reach the same state hash over
    60 free-running steps of a 2,048-body pile in residual mode (every 4th pass)."""
    World, _ = sides
    sc = scenes.config2_scene(seed=2, n=2048, lattice=(32, 32, 8))
    hashes = []
    for lds in ("1", "0"):
        monkeypatch.setenv("PXG_TILED_LDS", lds)
        c = _abi.default_config(residual_mode=1, residual_tolerance=1e-4, residual_max_iterations=128,
                                residual_check_interval=4, **sc.config)
        c.solve_order = _abi.PXG_ORDER_TILED
        w = World(c, sc.bodies, sc.shapes)
        for _ in range(60):
            w.step_fixed()
        hashes.append(w.state_hash())
        w.close()
    assert hashes[0] == hashes[1]
