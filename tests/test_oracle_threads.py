"""The oracle's thread-pool mode (bounds, pair discovery and narrowphase on several host threads,
the stages the reference runs on its job system, world.h:1489-1499, 1545-1577, 1598-1645) gives the
serial run's results bit for bit: every chunk writes only its own indices or its own pair bucket,
and the buckets are merged into the sorted pair cache."""
import pytest

from physxcoopa_amd import scenes, _abi


@pytest.mark.parametrize("scene", ["config1", "mixed_pile", "pile_2048"])
def test_threaded_oracle_equals_serial(oracle, scene):
    if scene == "config1":
        sc = scenes.config1_scene(seed=1)
        b, s, settings = sc.bodies, sc.shapes, sc.config
    elif scene == "mixed_pile":
        b, s = scenes.mixed_pile_scene(n=200, seed=7)
        settings = {}
    else:
        sc = scenes.config2_scene(seed=2, n=2048, lattice=(32, 32, 8))
        b, s, settings = sc.bodies, sc.shapes, sc.config
    cfg = _abi.default_config(**settings)
    serial = oracle.OracleWorld(cfg, b, s)
    threaded = oracle.OracleWorld(cfg, b, s)
    threaded.set_threads(4)
    for i in range(60):
        serial.step_fixed()
        threaded.step_fixed()
        assert serial.state_hash() == threaded.state_hash(), "step %d" % (i + 1)
    assert len(serial.manifolds()) > 0
    assert serial.pairs().tobytes() == threaded.pairs().tobytes()
