"""Convex-hull shapes on the GPU (pxg_set_hulls, collision/sat_hull.h) against the CPU oracle: a
pile of cube, tetrahedron and hexagonal-prism hulls among spheres and capsules over a box ground,
in lockstep (every step's pairs, manifolds with feature ids, impulses and bodies) and free-running
(every step's world_state_hash); and the upload's validation."""
import math

import numpy as np
import pytest

from physxcoopa_amd import scenes, _abi
from parity import lockstep

pytestmark = pytest.mark.gpu


def cube(half):
    return [(sx * half, sy * half, sz * half) for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)]


@pytest.fixture(scope="module")
def hulls(oracle):
    return [oracle.build_hull(cube(0.35)),
            oracle.build_hull([(0.5, 0, -0.3), (-0.3, 0.45, -0.3), (-0.3, -0.45, -0.3), (0, 0, 0.5)]),
            oracle.build_hull([(0.4 * math.cos(a), 0.4 * math.sin(a), z) for a in np.arange(6) * math.pi / 3
                               for z in (-0.25, 0.25)])]


def families(manifolds):
    f = manifolds["points"]["feature_id"]
    return {int(x) >> 28 for m, row in zip(manifolds, f) for x in row[:m["count"]]}


@pytest.mark.parametrize("order", [_abi.PXG_ORDER_REFERENCE, _abi.PXG_ORDER_COLOURED], ids=["reference", "coloured"])
def test_hull_pile_lockstep(oracle, hulls, order):
    from physxcoopa_amd.world import World
    b, s = scenes.hull_pile_scene(hulls, n=40, seed=4)
    cfg = _abi.default_config(velocity_iterations=8, position_iterations=3)
    cfg.solve_order = order
    gw, ow, rep = lockstep(World, oracle.OracleWorld, cfg, b, s, 200, resync=True, hulls=hulls)
    fam = families(gw.manifolds())
    print("hull pile lockstep:", rep, "feature id families", sorted(fam), flush=True)
    assert rep["bitexact_bodies"] == rep["steps"] == 200
    assert {4, 8} <= fam  # hull faces and hull-capsule samples in the last step's contacts


def test_hull_pile_free_running_hash(oracle, hulls):
    """No resync: 300 steps in the reference order, every step's state hash equal to the oracle's."""
    from physxcoopa_amd.world import World
    b, s = scenes.hull_pile_scene(hulls, n=40, seed=9)
    cfg = _abi.default_config(velocity_iterations=8, position_iterations=3)
    gw = World(cfg, b, s, hulls=hulls)
    ow = oracle.OracleWorld(cfg, b, s, hulls=hulls)
    seen = set()
    for i in range(300):
        gw.step_fixed()
        ow.step_fixed()
        assert gw.state_hash() == ow.state_hash(), "step %d" % (i + 1)
        if i % 25 == 0:
            seen |= families(gw.manifolds())
    print("feature id families seen:", sorted(seen), flush=True)
    assert {4, 5, 7, 8} <= seen


def test_hull_uploads_are_validated(oracle, hulls):
    """A face of more than 8 vertices or a shape pointing past the hull table is refused."""
    from physxcoopa_amd.world import World
    from physxcoopa_amd._lib import lib
    nine = oracle.build_hull([(math.cos(a), math.sin(a), z) for a in np.arange(9) * 2 * math.pi / 9 for z in (-0.2, 0.2)])
    assert max(len(nine.face(f)) for f in range(nine.n_faces)) == 9
    b, s = scenes.hull_pile_scene(hulls[:1], n=3, seed=1, others=False)
    w = World(_abi.default_config(), b, s, hulls=hulls[:1])
    tab = _abi.hull_table([nine])
    import ctypes
    assert lib().pxg_set_hulls(w.h, ctypes.byref(tab), 1) == _abi.PXG_ERR_INVALID
    bad = s.copy()
    bad["capsule_axis"][1] = 5
    with pytest.raises(RuntimeError):
        World(_abi.default_config(), b, bad, hulls=hulls[:1])
    w.close()


def test_queries_skip_hull_targets(oracle, hulls):
    """Scene queries test sphere, box and capsule targets: a ray straight down onto a hull resting on
    the ground hits the ground under it, as if the hull were not there."""
    from physxcoopa_amd.world import World
    b, s = scenes.hull_pile_scene(hulls[:1], n=1, seed=1, others=False)
    w = World(_abi.default_config(), b, s, hulls=hulls[:1])
    for _ in range(120):
        w.step_fixed()
    p = w.bodies()["position"][1]
    hits = w.raycast(_abi.rays([p[0], p[1], 5.0], [0, 0, -1], 20.0))
    assert hits["hit"][0] == 1 and hits["body"][0] == 0 and abs(hits["distance"][0] - 5.0) < 1e-4
    w.close()
