"""Triangle-mesh shapes on the GPU (pxg_set_meshes, collision/mesh_contact.h) against the CPU
oracle: spheres, boxes, capsules and convex hulls dropped on a bumpy terrain mesh, in lockstep
(every step's pairs, manifolds with feature ids, impulses and bodies) and free-running (every
step's world_state_hash), under a turned, scaled and offset mesh too; the reference's mesh world
tests (narrowphase_test.cpp:255-305, 360-455) on the GPU; and the upload's validation."""
import ctypes
import math

import numpy as np
import pytest

from physxcoopa_amd import scenes, _abi
from parity import lockstep

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hulls(oracle):
    return [oracle.build_hull([(sx * 0.25, sy * 0.2, sz * 0.15) for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)]),
            oracle.build_hull([(0.3, 0, -0.2), (-0.2, 0.28, -0.2), (-0.2, -0.28, -0.2), (0, 0, 0.3)])]


def families(manifolds):
    out = set()
    for m in manifolds:
        for k in range(int(m["count"])):
            f = int(m["points"]["feature_id"][k])
            out.add("recovery" if f & 0x80000000 else "low" if f < 8 else "tri")
    return out


@pytest.mark.parametrize("order", [_abi.PXG_ORDER_REFERENCE, _abi.PXG_ORDER_COLOURED], ids=["reference", "coloured"])
def test_terrain_pile_lockstep(oracle, hulls, order):
    b, s = scenes.terrain_pile_scene(hulls, n=40, seed=11)
    cfg = _abi.default_config(velocity_iterations=8, position_iterations=3)
    cfg.solve_order = order
    gw, ow, rep = lockstep(World(), oracle.OracleWorld, cfg, b, s, 240, resync=True, hulls=hulls,
                           meshes=[scenes.bumpy_terrain()])
    print("terrain pile lockstep:", rep, flush=True)
    assert rep["bitexact_bodies"] == rep["steps"] == 240
    assert rep["manifolds"] > 0 and rep["bitexact_manifolds"] == rep["manifolds"]


def test_turned_scaled_terrain_free_running_hash(oracle, hulls):
    """No resync: 300 steps in the reference order under a mesh whose body is turned and offset,
    with a local rotation, local centre and scale 1.2; every step's state hash equals the oracle's."""
    b, s = scenes.terrain_pile_scene(hulls, n=36, seed=5, mesh_position=(-4.5, -4.0, 0.2),
                                     mesh_orientation=(0.0, 0.0, math.sin(0.1), math.cos(0.1)))
    mesh_rows = s["type"] == _abi.PXG_TRIANGLE_MESH
    s["radius"][mesh_rows] = 1.2
    s["local_center"][mesh_rows] = (0.3, -0.2, 0.0)
    s["local_rotation"][mesh_rows] = (0.0, 0.0, math.sin(-0.05), math.cos(-0.05))
    cfg = _abi.default_config(velocity_iterations=8, position_iterations=3)
    meshes = [scenes.bumpy_terrain()]
    gw = World()(cfg, b, s, hulls=hulls, meshes=meshes)
    ow = oracle.OracleWorld(cfg, b, s, hulls=hulls, meshes=meshes)
    seen = set()
    for i in range(300):
        gw.step_fixed()
        ow.step_fixed()
        assert gw.state_hash() == ow.state_hash(), "step %d" % (i + 1)
        if i % 20 == 0:
            seen |= families(gw.manifolds())
    print("feature id kinds seen:", sorted(seen), flush=True)
    assert {"low", "tri"} <= seen
    gw.close()


def World():
    from physxcoopa_amd.world import World as W
    return W


def run_world(oracle, bodies, shapes, meshes, steps, config=None):
    cfg = config or _abi.default_config()
    gw = World()(cfg, bodies, shapes, meshes=meshes)
    ow = oracle.OracleWorld(cfg, bodies, shapes, meshes=meshes)
    for i in range(steps):
        gw.step_fixed()
        ow.step_fixed()
        assert gw.state_hash() == ow.state_hash(), "step %d" % (i + 1)
    return gw


def test_box_slides_across_internal_mesh_edge_without_a_bump(oracle):
    sb = scenes.SceneBuilder()
    scenes.mesh_shape_on(sb, 0, friction=0.05)
    b = sb.dynamic_box((-1.5, 0.0, 0.27), (0.25, 0.25, 0.25), mass=1.0, friction=0.05)
    bodies, shapes = sb.arrays()
    meshes = [scenes.two_triangle_floor()]
    gw = run_world(oracle, bodies, shapes, meshes, 60)
    bb = gw.bodies()
    bb["awake"][b] = 1
    bb["linear_velocity"][b] = (3.0, 0.0, 0.0)
    gw.set_bodies(bb)
    prev, jump = float(bb["linear_velocity"][b][2]), 0.0
    for _ in range(40):
        gw.step_fixed()
        vz = float(gw.bodies()["linear_velocity"][b][2])
        jump = max(jump, abs(vz - prev))
        prev = vz
    p = gw.bodies()["position"][b]
    assert p[0] > 0.0 and abs(p[2] - 0.25) < 0.05 and jump < 1.0
    gw.close()


def test_box_settles_flat_on_a_finely_tessellated_floor(oracle):
    from test_oracle_mesh import quat_from_euler, best_up_alignment
    sb = scenes.SceneBuilder()
    scenes.mesh_shape_on(sb, 0)
    b = sb.dynamic_box((0.1, -0.2, 1.5), (0.6,) * 3, orientation=quat_from_euler((0.17, 0.35, 0.52)), mass=40.0)
    bodies, shapes = sb.arrays()
    gw = run_world(oracle, bodies, shapes, [scenes.tiled_floor(16)], 300)
    r = gw.bodies()[b]
    assert best_up_alignment(r["orientation"]) > 0.999 and abs(r["position"][2] - 0.6) < 0.02
    assert np.linalg.norm(r["linear_velocity"]) < 0.05 and np.linalg.norm(r["angular_velocity"]) < 0.05
    gw.close()


def test_an_embedded_box_climbs_back_onto_the_floor(oracle):
    sb = scenes.SceneBuilder()
    scenes.mesh_shape_on(sb, 0)
    b = sb.dynamic_box((0.3, -0.4, -0.1), (0.11,) * 3, mass=1.2)
    bodies, shapes = sb.arrays()
    gw = run_world(oracle, bodies, shapes, [scenes.tiled_floor(8)], 240)
    r = gw.bodies()[b]
    assert abs(r["position"][2] - 0.11) < 0.02 and np.linalg.norm(r["linear_velocity"]) < 0.05
    gw.close()


def test_mesh_uploads_are_validated(oracle):
    from physxcoopa_amd._lib import lib
    sb = scenes.SceneBuilder()
    scenes.mesh_shape_on(sb, 0)
    sb.dynamic_sphere((0, 0, 1), 0.3)
    bodies, shapes = sb.arrays()
    floor = scenes.tiled_floor(4)
    with pytest.raises(RuntimeError):  # shape index past the table
        World()(_abi.default_config(), bodies, shapes)
    w = World()(_abi.default_config(), bodies, shapes, meshes=[floor])
    empty = _abi.mesh_table([_abi.TriangleMesh([(0, 0, 0), (1, 0, 0), (2, 0, 0)], [0, 1, 2])])
    assert lib().pxg_set_meshes(w.h, ctypes.byref(empty), 1) == _abi.PXG_ERR_INVALID  # nothing left to collide with
    assert lib().pxg_set_meshes(w.h, None, 0) == _abi.PXG_ERR_STATE  # the floor shape still uses mesh 0
    bad = _abi.mesh_table([_abi.TriangleMesh([(0, 0, 0), (1, 0, 0), (0, 1, 0)], [0, 1, 7])])
    assert lib().pxg_set_meshes(w.h, ctypes.byref(bad), 1) == _abi.PXG_ERR_INVALID
    w.step_fixed()  # the floor is still there and the world still steps
    w.close()


def test_too_many_mesh_vertices_inside_one_box_is_reported(oracle):
    """A 1 m box sunk into a floor of 5 cm tiles has hundreds of mesh vertices inside it: more than
    the 32 a box keeps, which the step reports as a capacity error instead of dropping them."""
    sb = scenes.SceneBuilder()
    scenes.mesh_shape_on(sb, 0)
    sb.dynamic_box((0.0, 0.0, 0.45), (0.5, 0.5, 0.5), mass=1.0)
    bodies, shapes = sb.arrays()
    w = World()(_abi.default_config(), bodies, shapes, meshes=[scenes.tiled_floor(80)])
    with pytest.raises(RuntimeError, match="mesh vertices"):
        for _ in range(3):
            w.step_fixed()
    w.close()
