"""The tiled solve order (PXG_ORDER_TILED) against the CPU oracle fed the GPU's order.

Each tile's colours run in one workgroup (its bodies and impulses in LDS where they fit), the
manifolds crossing tiles in two later phases; the order is colour-major and every colour is a
valid colour of the whole world, so the oracle's sequential solve of it must give the GPU's
state bit for bit. Covered: fixed iterations and residual mode on piles of 2,048 and 16,384 bodies
(tiles of 64 and 512 bodies), a small pile where most manifolds cross tiles (the later phases
carry the solve), and the global-memory form of the same kernel (PXG_TILED_LDS=0). Stacks of
16,384 boxes (four-point face contacts in nearly every manifold: colours with more multi-point
manifolds than eight lanes each can hold) and a pile of spheres (one-point manifolds only) take
the lane plans the mixed pile uses little."""
import numpy as np
import pytest

from physxcoopa_amd import scenes, _abi
from parity import lockstep, compare_step

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sides(oracle):
    from physxcoopa_amd.world import World
    return World, oracle.OracleWorld


def _box_stacks(nx, ny, nz, seed=12):
    """nx * ny columns of nz unit boxes (1 kg, friction 0.6), each turned by a small random yaw,
    1.05 m apart, on a static ground."""
    rng = np.random.default_rng(seed)
    sb = scenes.SceneBuilder()
    sb.static_box((0, 0, -0.5), (100, 100, 0.5), friction=0.6)
    for ix in range(nx):
        for iy in range(ny):
            for iz in range(nz):
                q = scenes.quat_axis_angle((0.0, 0.0, 1.0), float(rng.uniform(-0.05, 0.05)))
                sb.dynamic_box(((ix - nx / 2) * 1.05, (iy - ny / 2) * 1.05, 0.5 + iz * 1.001), (0.5, 0.5, 0.5), q,
                               mass=1.0, friction=0.6)
    b, s = sb.arrays()
    return scenes.Scene("box_stacks", b, s, dict(velocity_iterations=8, position_iterations=3))


def _sphere_pile(nx, ny, nz, seed=13):
    """nx * ny * nz spheres (radius 0.35-0.5 m, 1 kg) on a jittered 1 m lattice over a ground."""
    rng = np.random.default_rng(seed)
    sb = scenes.SceneBuilder()
    sb.static_box((0, 0, -0.5), (100, 100, 0.5), friction=0.5)
    for ix in range(nx):
        for iy in range(ny):
            for iz in range(nz):
                pos = (ix - nx / 2 + float(rng.uniform(-0.05, 0.05)), iy - ny / 2 + float(rng.uniform(-0.05, 0.05)),
                       1.0 + iz)
                sb.dynamic_sphere(pos, float(rng.uniform(0.35, 0.5)), mass=1.0, friction=0.5)
    b, s = sb.arrays()
    return scenes.Scene("sphere_pile", b, s, dict(velocity_iterations=8, position_iterations=3))


def _pile_lockstep(sides, n, lattice, warm, steps, residual=False, every=1, cap=256, sc=None):
    """`warm` fixed-iteration steps on the GPU in the tiled order, then `steps` steps in lockstep:
    each step the oracle gets the GPU's input state and the GPU's solve order."""
    World, OracleWorld = sides
    if sc is None:
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


def test_tiled_order_box_stacks_16384(sides):
    gw, rep = _pile_lockstep(sides, 0, None, 20, 2, sc=_box_stacks(32, 32, 16))
    gm = gw.manifolds()
    assert (gm["count"] == 4).sum() > len(gm) // 2


def test_tiled_order_sphere_pile_16384(sides):
    gw, rep = _pile_lockstep(sides, 0, None, 60, 2, sc=_sphere_pile(32, 32, 16))
    assert gw.manifolds()["count"].max() == 1


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
