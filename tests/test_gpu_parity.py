"""GPU step vs the CPU oracle, step by step (see tests/parity.py for the bar).

Runs through the C ABI (libpxg.so via physxcoopa_amd.world.World)."""
import numpy as np
import pytest

from physxcoopa_amd import scenes, _abi
from parity import lockstep

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sides(oracle):
    from physxcoopa_amd.world import World
    return World, oracle.OracleWorld


def _run(sides, scene_arrays, steps, resync=True, **cfg):
    World, OracleWorld = sides
    b, s = scene_arrays
    c = _abi.default_config(**cfg)
    print("lockstep: %d bodies, %d shapes, %d steps" % (len(b), len(s), steps), flush=True)
    gw, ow, rep = lockstep(World, OracleWorld, c, b, s, steps, resync=resync)
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


def test_parity_barrier_solve_path(sides, monkeypatch):
    """The solve with a grid barrier between colours (the residual-mode kernel, here in fixed
    iteration mode) gives the same results as the default barrier-free dataflow solve."""
    monkeypatch.setenv("PXG_SOLVE", "barrier")
    _, _, rep = _run(sides, scenes.mixed_pile_scene(n=200, seed=12), 60, velocity_iterations=8, position_iterations=3)
    assert rep["manifolds"] > 0


def test_dataflow_and_barrier_solves_agree_bit_for_bit(monkeypatch):
    """Same scene, same steps: the two persistent solves must produce identical bodies."""
    from physxcoopa_amd.world import World
    sc = scenes.config2_scene(seed=3, n=2048, lattice=(32, 32, 8))
    cfg = _abi.default_config(**sc.config)
    out = []
    for mode in ("dataflow", "barrier"):
        monkeypatch.setenv("PXG_SOLVE", mode)
        w = World(cfg, sc.bodies, sc.shapes)
        for _ in range(150):
            w.step_fixed()
        out.append(w.bodies())
        w.close()
    assert out[0].tobytes() == out[1].tobytes()


def test_parity_residual_mode(sides):
    """Residual mode: velocity passes repeat until the residual <= 1e-4 (cap 64 here); the GPU
    must run the same number of passes as the oracle and reach the same state."""
    World, OracleWorld = sides
    sc = scenes.config1_scene(seed=1)
    c = _abi.default_config(residual_mode=1, residual_tolerance=1e-4, residual_max_iterations=64, **sc.config)
    gw, ow, rep = lockstep(World, OracleWorld, c, sc.bodies, sc.shapes, 80, resync=True)
    print(rep, gw.stats(), ow.stats(), flush=True)
    assert gw.stats()["velocity_passes"] == ow.stats()["velocity_passes"]
    assert abs(gw.stats()["residual"] - ow.stats()["residual"]) <= 1e-6 + 1e-4 * abs(ow.stats()["residual"])
