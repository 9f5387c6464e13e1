"""XPBD cloth and hair on the GPU (pxg_set_cloths / pxg_set_strand_sets, stepped by pxg_step once
per frame) against the CPU oracle's restatement of world.h:1749-1957, cloth/cloth_solver.h and
hair/strand_solver.h. The oracle runs with sin, cos, pow and acos evaluated in double and rounded
once, as the GPU evaluates them, and visits colliders in shape-slot order within the static and
dynamic groups, as the GPU gathers them; with both, every frame's particles, multipliers, anchors,
roots, bounds and sleep state are compared bit for bit, and the world state hash (which includes
the cloths' particles) every frame."""
import math

import numpy as np
import pytest

from physxcoopa_amd import scenes, _abi

pytestmark = pytest.mark.gpu

H = np.float32(1.0 / 60.0)
PFIELDS = ("position", "inv_mass", "prev_position", "velocity")


def World():
    from physxcoopa_amd.world import World as W
    return W


def pair(oracle, sb, cloths=(), strands=(), hulls=None, meshes=None, **cfg):
    """The same scene on the GPU and in the oracle (rounded trig, slot-order colliders)."""
    if sb is None:
        sb = scenes.SceneBuilder()
    far = sb.add_body(_abi.PXG_STATIC, (0, 0, -100), mass=1.0)  # keeps the body and shape tables non-empty
    sb.add_shape(far, _abi.PXG_SPHERE, radius=0.5)
    bodies, shapes = sb.arrays()
    config = _abi.default_config(**cfg)
    gw = World()(config, bodies, shapes, hulls=hulls, meshes=meshes, cloths=[c.copy() for c in cloths],
                 strands=[s.copy() for s in strands])
    ow = oracle.OracleWorld(config, bodies, shapes, hulls=hulls, meshes=meshes, cloths=[c.copy() for c in cloths],
                            strands=[s.copy() for s in strands])
    ow.set_atan2_mode(True)
    ow.set_collider_order(False)
    return gw, ow


def bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


def same_particles(g, o, what):
    for f in PFIELDS:
        if not np.array_equal(bits(g.particles[f]), bits(o.particles[f])):
            bad = np.nonzero(np.any(g.particles[f] != o.particles[f], axis=tuple(range(1, g.particles[f].ndim))))[0]
            raise AssertionError("%s: particles' %s differ at %s (gpu %s, oracle %s)" % (
                what, f, bad[:8], g.particles[f][bad[:2]], o.particles[f][bad[:2]]))


def same_cloth(g, o, what):
    same_particles(g, o, what)
    for f in ("stretch", "bend"):
        assert np.array_equal(bits(getattr(g, f)["lambda"]), bits(getattr(o, f)["lambda"])), what + ": " + f + " multipliers"
    for f in ("world_position", "prev_world_position"):
        assert np.array_equal(bits(g.anchors[f]), bits(o.anchors[f])), what + ": anchors' " + f
    assert (g.awake, g.sleep_timer, g.time) == (o.awake, o.sleep_timer, o.time), what + ": sleep state"
    assert np.array_equal(bits(g.bounds), bits(o.bounds)), what + ": bounds"


def same_groom(g, o, what):
    same_particles(g, o, what)
    assert np.array_equal(bits(g.stretch["lambda"]), bits(o.stretch["lambda"])), what + ": multipliers"
    assert np.array_equal(bits(g.strands), bits(o.strands)), what + ": roots"
    assert (g.awake, g.sleep_timer, g.time) == (o.awake, o.sleep_timer, o.time), what + ": sleep state"
    assert np.array_equal(bits(g.bounds), bits(o.bounds)), what + ": bounds"


def run_cloths(gw, ow, frames, n_cloths, dts=None, each=None, check_every=1):
    for f in range(frames):
        dt = H if dts is None else dts(f)
        if each:
            each(f)
        assert gw.step(dt) == ow.step(dt), "frame %d substeps" % f
        if (f + 1) % check_every == 0 or f + 1 == frames:
            assert gw.state_hash() == ow.state_hash(), "frame %d hash" % f
            for k in range(n_cloths):
                same_cloth(gw.cloth(k), ow.cloth(k), "frame %d cloth %d" % (f, k))


def sheet(oracle, res, center, params=None, **kw):
    return oracle.grid_cloth(res, res, 4.0, 4.0, center, params=params, **kw)


def test_drape_over_sphere_with_wind_self_collision_and_tethers(oracle):
    """A pinned, tethered 13 x 13 sheet with self-collision in turbulent wind drapes over a static
    sphere: bit-exact particles, multipliers, anchors and bounds every frame for 180 frames."""
    sb = scenes.SceneBuilder()
    b = sb.add_body(_abi.PXG_STATIC, (0, 0, 3), mass=1.0)
    sb.add_shape(b, _abi.PXG_SPHERE, radius=1.0)
    p = _abi.ClothParams(self_collision=1, wind=(0.4, 0.0, 0.0), wind_turbulence=0.6)
    c = sheet(oracle, 13, (0, 0, 4.05), p, pins=[((0, 0, 4.05), 0.4)], tethers=True)
    assert len(c.tethers) and len(c.anchors)
    gw, ow = pair(oracle, sb, cloths=[c])
    print("drape: %d particles, %d stretch, %d bend, %d tethers" % (len(c.particles), len(c.stretch), len(c.bend),
                                                                    len(c.tethers)), flush=True)
    run_cloths(gw, ow, 180, 1)
    pos = gw.cloth(0).particles["position"]
    assert np.min(np.linalg.norm(pos - np.array([0, 0, 3.0], np.float32), axis=1)) >= 1.0
    gw.close()


def test_anchor_follows_a_kinematic_body_over_uneven_frames(oracle):
    """Anchors on a moving kinematic ball; frames of 0.9 and 1.1 fixed steps, so some frames run
    no rigid substep (the broadphase refit then runs before the cloth)."""
    sb = scenes.SceneBuilder()
    ball = sb.add_body(_abi.PXG_KINEMATIC, (0, 0, 3), mass=1.0, inv_mass=0.0, use_gravity=False)
    sb.add_shape(ball, _abi.PXG_SPHERE, radius=1.0)
    c = sheet(oracle, 15, (0, 0, 4.05), pins=[((0, 0, 4.0), 0.4, (ball, 0), ((0, 0, 3), (0, 0, 0, 1)))], tethers=True)
    gw, ow = pair(oracle, sb, cloths=[c])
    zero = []

    def move(f, dt):
        for w in (gw, ow):
            b = w.bodies()
            if f >= 30:
                b["position"][ball][0] += np.float32(4.0) * dt
            b["linear_velocity"][ball] = (4.0, 0.0, 0.0) if f >= 30 else (0.0, 0.0, 0.0)
            w.set_bodies(b)

    def dts(f):
        return np.float32(H * np.float32(0.9)) if f % 2 == 0 else np.float32(H * np.float32(1.1))

    for f in range(150):
        dt = dts(f)
        move(f, dt)
        n = gw.step(dt)
        assert n == ow.step(dt)
        if n == 0:
            zero.append(f)
        assert gw.state_hash() == ow.state_hash(), "frame %d" % f
        if f % 10 == 9:
            same_cloth(gw.cloth(0), ow.cloth(0), "frame %d" % f)
    assert zero, "no frame ran without a rigid substep"
    assert gw.bodies()["position"][ball][0] > 1.0
    gw.close()


def test_stale_anchor_body_turns_into_a_static_pin(oracle):
    """The followed body's generation changes (a new body in the slot): the anchor stays put."""
    sb = scenes.SceneBuilder()
    post = sb.add_body(_abi.PXG_KINEMATIC, (0, 0, 5), mass=1.0, inv_mass=0.0, use_gravity=False)
    c = sheet(oracle, 9, (0, 0, 5), pins=[((0, 0, 5), 0.6, (post, 0), ((0, 0, 5), (0, 0, 0, 1)))])
    gw, ow = pair(oracle, sb, cloths=[c])
    for f in range(40):
        if f == 10:
            for w in (gw, ow):
                b = w.bodies()
                b["generation"][post] = 1
                b["position"][post] = (0.5, 0.0, 5.0)
                w.set_bodies(b)
        assert gw.step(H) == ow.step(H)
        same_cloth(gw.cloth(0), ow.cloth(0), "frame %d" % f)
    a = gw.cloth(0).anchors
    np.testing.assert_allclose(a["world_position"], c.anchors["world_position"], atol=1e-5)
    gw.close()


@pytest.mark.parametrize("kind", ["box", "capsule", "hull", "mesh"])
def test_sheet_on_each_collider_kind(oracle, kind):
    """A sheet dropped on a box, a lying capsule, a cube hull, or (fast, to take the swept
    crossing test) on a two-triangle mesh floor."""
    sb = scenes.SceneBuilder()
    hulls = meshes = None
    b = sb.add_body(_abi.PXG_STATIC, (0, 0, 0), mass=1.0)
    if kind == "box":
        sb.add_shape(b, _abi.PXG_BOX, half_extents=(1.5, 1.5, 1.0))
    elif kind == "capsule":
        sb.add_shape(b, _abi.PXG_CAPSULE, capsule_radius=1.0, capsule_half_height=1.0, capsule_axis=0)
    elif kind == "hull":
        hulls = [oracle.build_hull([(sx * 1.5, sy * 1.5, sz * 1.0) for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)])]
        sb.add_shape(b, _abi.PXG_CONVEX_HULL, capsule_axis=0, radius=1.0)
    else:
        meshes = [scenes.two_triangle_floor()]
        sb.add_shape(b, _abi.PXG_TRIANGLE_MESH, radius=1.0, capsule_axis=0)
    c = oracle.grid_cloth(7, 7, 0.6, 0.6, (0, 0, 1.3), params=_abi.ClothParams(thickness=0.03))
    if kind == "mesh":
        c.particles["velocity"] = (0.0, 0.0, -60.0)
    gw, ow = pair(oracle, sb, cloths=[c], hulls=hulls, meshes=meshes)
    run_cloths(gw, ow, 150, 1, check_every=5)
    gw.close()


def test_many_cloths_sleep_and_wake(oracle):
    """Six cloths of different sizes and parameters in one world (one workgroup each); one is
    pinned to a post that moves after it has gone to sleep, which wakes it."""
    sb = scenes.SceneBuilder()
    post = sb.add_body(_abi.PXG_KINEMATIC, (6, 0, 5), mass=1.0, inv_mass=0.0, use_gravity=False)
    floor = sb.add_body(_abi.PXG_STATIC, (0, 0, -0.5), mass=1.0)
    sb.add_shape(floor, _abi.PXG_BOX, half_extents=(20.0, 20.0, 0.5))
    cloths = [
        oracle.grid_cloth(5, 5, 1.0, 1.0, (-6, 0, 2), params=_abi.ClothParams(air_drag=0.1)),
        oracle.grid_cloth(9, 7, 2.0, 1.5, (-3, 0, 1), shear=False, params=_abi.ClothParams(bend_compliance=1e-3)),
        oracle.grid_cloth(21, 21, 2.0, 2.0, (0, 0, 3), params=_abi.ClothParams(substeps=2, iterations=3)),
        oracle.grid_cloth(11, 11, 1.5, 1.5, (3, 0, 2), params=_abi.ClothParams(self_collision=1, friction=0.9),
                          pins=[((3, 0.75, 2), 0.3)], tethers=True),
        oracle.grid_cloth(9, 9, 2.0, 2.0, (6, 0, 5), params=_abi.ClothParams(damping=4.0, sleep_threshold=0.25,
                                                                              sleep_time=0.25),
                          pins=[((6, 0, 5), 3.0, (post, 0), ((6, 0, 5), (0, 0, 0, 1)))]),
        oracle.grid_cloth(4, 30, 0.3, 3.0, (9, 0, 4), params=_abi.ClothParams(max_velocity=3.0, stretch_compliance=1e-4)),
    ]
    gw, ow = pair(oracle, sb, cloths=cloths)
    run_cloths(gw, ow, 300, len(cloths), check_every=25)
    assert not gw.cloth(4).awake
    for w in (gw, ow):
        b = w.bodies()
        b["position"][post][0] += 0.5
        w.set_bodies(b)
    run_cloths(gw, ow, 5, len(cloths))
    assert gw.cloth(4).awake
    gw.close()


def test_serial_tethers_when_an_anchor_is_another_tethers_particle(oracle):
    """Hand-made tethers chained through each other (particle 2's anchor is tether particle 1):
    the GPU runs them in order on one lane and agrees with the oracle."""
    c = oracle.grid_cloth(6, 6, 1.0, 1.0, (0, 0, 3), params=_abi.ClothParams(external_acceleration=(30.0, 0.0, 0.0)),
                          pins=[((-0.5, -0.5, 3), 0.05)])
    t = np.zeros(4, _abi.TETHER_DTYPE)
    t["particle"], t["anchor"], t["max_length"] = (1, 2, 3, 7), (0, 1, 2, 0), (0.21, 0.21, 0.21, 0.3)
    c.tethers = t
    gw, ow = pair(oracle, None, cloths=[c])
    run_cloths(gw, ow, 60, 1)
    gw.close()


def test_cloth_upload_validation(oracle):
    c = oracle.grid_cloth(5, 5, 1.0, 1.0, (0, 0, 3))
    gw, _ = pair(oracle, None, cloths=[c])
    bad = c.copy()
    bad.stretch_batches = np.array([[0, len(bad.stretch)]], np.uint32)  # every stretch edge in one batch
    with pytest.raises(RuntimeError, match="share a particle"):
        gw.set_cloths([bad])
    bad = c.copy()
    bad.triangles = bad.triangles.copy()
    bad.triangles[4] = 25
    with pytest.raises(RuntimeError, match="out of range"):
        gw.set_cloths([bad])
    big = oracle.grid_cloth(91, 91, 1.0, 1.0, (0, 0, 3))
    with pytest.raises(RuntimeError, match="8192"):
        gw.set_cloths([big])
    dead = c.copy()
    dead.alive = False
    gw.set_cloths([dead, c])  # an empty slot, then a live cloth
    gw.step(H)
    assert gw.cloth(1).particles["position"][0][2] < 3.0
    gw.close()


# ---- hair ----

def straight(root, count=8, spacing=0.05):
    return [(root[0] + spacing * k, root[1], root[2]) for k in range(count)]


def bent(root, spacing=0.05):
    return [(root[0] + spacing * k, root[1], root[2]) for k in range(4)] + \
           [(root[0] + spacing * 3, root[1] + spacing * k, root[2]) for k in range(1, 5)]


def run_grooms(gw, ow, frames, n_sets, each=None, check_every=1):
    for f in range(frames):
        if each:
            each(f)
        assert gw.step(H) == ow.step(H)
        if (f + 1) % check_every == 0 or f + 1 == frames:
            for k in range(n_sets):
                same_groom(gw.strand_set(k), ow.strand_set(k), "frame %d set %d" % (f, k))


def test_groom_with_moving_turning_roots_wind_and_a_collider(oracle):
    """Two curled and three straight strands in turbulent wind, roots moved and turned every frame,
    over a static sphere: bit-exact against the oracle every frame for 150 frames."""
    sb = scenes.SceneBuilder()
    b = sb.add_body(_abi.PXG_STATIC, (0.1, 0, -0.35), mass=1.0)
    sb.add_shape(b, _abi.PXG_SPHERE, radius=0.3)
    strands = [bent((0, 0, 0)), straight((0.1, 0, 0), 12), bent((0, 0.2, 0)), straight((-0.1, 0, 0), 5),
               straight((0, -0.2, 0), 9)]
    s = oracle.make_strands(strands, _abi.StrandParams(wind=(1.0, 0.3, 0.0), wind_turbulence=0.5, thickness=0.02))
    gw, ow = pair(oracle, sb, strands=[s], cloth_iterations=4)
    roots = np.array([st[0] for st in strands], np.float32)

    def move(f):
        pos = roots + np.float32(0.002) * np.float32(f)
        q = [(0, 0, math.sin(0.01 * f), math.cos(0.01 * f))] * len(strands)
        gw.set_strand_roots(0, pos, q)
        ow.set_strand_roots(0, pos, q)

    run_grooms(gw, ow, 150, 1, each=move)
    gw.close()


def test_grooms_collide_bodies_sleep_and_stale_ids(oracle):
    """Three grooms: one colliding with everything, one only with the listed body (a stale id in
    its list is skipped), one that settles and goes to sleep."""
    sb = scenes.SceneBuilder()
    near = sb.add_body(_abi.PXG_STATIC, (0, 0, -0.4), mass=1.0)
    sb.add_shape(near, _abi.PXG_SPHERE, radius=0.3)
    other = sb.add_body(_abi.PXG_STATIC, (1.0, 0, -0.4), mass=1.0)
    sb.add_shape(other, _abi.PXG_BOX, half_extents=(0.3, 0.3, 0.3))
    a = oracle.make_strands([straight((-0.1, 0, 0), 12), straight((0.05, 0, 0), 12)], _abi.StrandParams(global_shape=0.0))
    b = oracle.make_strands([straight((0.9, 0, 0), 12), straight((1.05, 0, 0), 12)], _abi.StrandParams(local_shape=0.6))
    b.collide_bodies = np.array([[near, 0], [other, 7]], np.uint32)  # `other` is listed with a stale generation
    c = oracle.make_strands([straight((3, 0, 0))], _abi.StrandParams(sleep_time=0.2, sleep_threshold=0.05))
    gw, ow = pair(oracle, sb, strands=[a, b, c], cloth_iterations=2)
    run_grooms(gw, ow, 420, 3, check_every=20)
    print("groom sleep states:", [gw.strand_set(k).awake for k in range(3)], flush=True)
    gw.close()


def test_groom_upload_validation(oracle):
    s = oracle.make_strands([straight((0, 0, 0)), straight((0.1, 0, 0))])
    gw, _ = pair(oracle, None, strands=[s])
    bad = s.copy()
    bad.strands = bad.strands.copy()
    bad.strands["first"][1] = 4  # overlaps the first strand
    with pytest.raises(RuntimeError, match="share a particle"):
        gw.set_strand_sets([bad])
    bad = s.copy()
    bad.strands = bad.strands.copy()
    bad.strands["count"][1] = 9
    with pytest.raises(RuntimeError, match="out of range"):
        gw.set_strand_sets([bad])
    gw.close()
