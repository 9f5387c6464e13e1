"""Pins of the oracle's velocity residual and of world_state_hash, by constructed cases whose value
follows from the definitions, not from the oracle's own formula.

The residual (residual mode's stopping test): over the active contact rows, the normal row's
complementarity violation of (lambda_n, v_n - target) and each friction row's violation of the
Coulomb conditions (stick: v_t = 0; slip at |lambda_t| = mu lambda_n: the friction impulse, +lambda_t
on body B along the tangent, opposes v_t), with v the velocity of B's contact point relative to
A's (solver.h:86-95 applies +J to B and -J to A). Each case below sets body velocities and
accumulated impulses by hand and states the residual a reader can check with pencil and paper;
a flipped slip sign, a dropped tangent row, a cross product in the wrong order, a wrong A/B
sign or a wrong speculative target fails at least one of them.

world_state_hash (world.h:1334-1359): FNV-1a 64 over each alive body's position, orientation
(x, y, z, w, glm's storage order), linear and angular velocity bytes, in index order; restated
here in numpy, its byte loop checked against the published FNV-1a test vectors first. The
reference starts from 1469598103934665603 (world.h:1335), which is the published offset basis
14695981039346656037 with its last digit missing; the hash keeps the reference's value, since it
must equal the reference's.
"""
import numpy as np
import pytest

from physxcoopa_amd import scenes, _abi
from physxcoopa_amd._abi import MANIFOLD_DTYPE

H = np.float32(1.0) / np.float32(60.0)


def ground_and_box(box_v=(0, 0, 0), box_w=(0, 0, 0), box_awake=True, box_first=False, box_static=False):
    """A static ground (top face z = 0) and a 1 kg unit box resting on it at z = 0.5; returns the
    oracle world's inputs and the body indices (ground, box)."""
    sb = scenes.SceneBuilder()
    def add_box():
        return sb.add_body(_abi.PXG_STATIC if box_static else _abi.PXG_DYNAMIC, position=(0, 0, 0.5), mass=1.0,
                           inv_inertia=(6, 6, 6), velocity=box_v, angular_velocity=box_w, awake=box_awake)
    if box_first:
        ib = add_box()
        ig = sb.add_body(_abi.PXG_STATIC, position=(0, 0, -0.5))
    else:
        ig = sb.add_body(_abi.PXG_STATIC, position=(0, 0, -0.5))
        ib = add_box()
    sb.add_shape(ig, _abi.PXG_BOX, half_extents=(10, 10, 0.5))
    sb.add_shape(ib, _abi.PXG_BOX)
    b, s = sb.arrays()
    return b, s, ig, ib


def contact(a, b, normal, points, mu=0.5, trigger=False):
    """One manifold: points = [(position, lambda_n, lambda_t1, lambda_t2, separation), ...]."""
    m = np.zeros(1, MANIFOLD_DTYPE)
    m["a"], m["b"], m["normal"], m["friction"], m["valid"] = a, b, normal, mu, 1
    m["is_trigger"] = 1 if trigger else 0
    m["count"] = len(points)
    for i, (pos, ln, lt1, lt2, sep) in enumerate(points):
        p = m["points"][0][i]
        p["position"], p["normal_impulse"], p["tangent_impulse"], p["separation"] = pos, ln, (lt1, lt2), sep
    return m


def residual(oracle, b, s, ms, **kw):
    ow = oracle.OracleWorld(_abi.default_config(), b, s)
    return ow.eval_residual(np.concatenate(ms) if isinstance(ms, list) else ms, **kw), ow


UP = (0.0, 0.0, 1.0)  # orthonormal_basis((0, 0, 1)) = t1 (1, 0, 0), t2 (0, 1, 0)


def test_tangent_basis_of_up_normal(oracle):
    t1, t2 = oracle.orthonormal_basis(np.array(UP, np.float32))
    assert np.array_equal(t1, [1, 0, 0]) and np.array_equal(t2, [0, 1, 0])


def test_sliding_box_friction_opposing_slip_is_converged(oracle):
    """The box slides at +0.3 m/s along x; friction at the cone's edge pushes it along -x (the box
    is B: lambda_t1 = -mu lambda_n): every condition holds, residual 0."""
    b, s, ig, ib = ground_and_box(box_v=(0.3, 0, 0))
    r, _ = residual(oracle, b, s, contact(ig, ib, UP, [((0, 0, 0), 10.0, -5.0, 0.0, 0.0)]))
    assert r == 0.0


def test_sliding_box_friction_along_slip_violates_by_slip_speed(oracle):
    """Same, friction flipped to push along +x with the slip: the residual is the slip speed 0.3."""
    b, s, ig, ib = ground_and_box(box_v=(0.3, 0, 0))
    r, ow = residual(oracle, b, s, contact(ig, ib, UP, [((0, 0, 0), 10.0, 5.0, 0.0, 0.0)]))
    assert r == pytest.approx(0.3, rel=1e-6)
    assert ow.residual_detail()["rt1"][0] == pytest.approx(0.3, rel=1e-6)


def test_body_a_side_sign(oracle):
    """The dynamic box as body A over the ground as B: v_t = v_B - v_A = -0.3 along x. Friction that
    opposes A's slip pushes A along -x, i.e. +lambda on B: residual 0; the opposite sign 0.3."""
    b, s, ig, ib = ground_and_box(box_v=(0.3, 0, 0), box_first=True)
    down = (0.0, 0.0, -1.0)  # normal from A (box) to B (ground)
    t1, t2 = oracle.orthonormal_basis(np.array(down, np.float32))
    # impulse along t1 * lambda on B, -t1 * lambda on A; pick lambda so that A is pushed along -x
    lam = 5.0 * np.sign(t1[0])
    r_ok, _ = residual(oracle, b, s, contact(ib, ig, down, [((0, 0, 0), 10.0, lam, 0.0, 0.0)]))
    r_bad, _ = residual(oracle, b, s, contact(ib, ig, down, [((0, 0, 0), 10.0, -lam, 0.0, 0.0)]))
    assert r_ok == 0.0
    assert r_bad == pytest.approx(0.3, rel=1e-6)


@pytest.mark.parametrize("v,expected", [((0.3, 0, 0), 0.3), ((0, -0.2, 0), 0.2), ((0.3, -0.4, 0), 0.4)])
def test_stick_inside_the_cone_needs_zero_tangent_speed(oracle, v, expected):
    """|lambda_t| < mu lambda_n (sticking) with the contact still moving: each tangent row's residual
    is its speed, the second tangent row included."""
    b, s, ig, ib = ground_and_box(box_v=v)
    r, ow = residual(oracle, b, s, contact(ig, ib, UP, [((0, 0, 0), 10.0, 2.0, -1.0, 0.0)]))
    assert r == pytest.approx(expected, rel=1e-6)
    d = ow.residual_detail()
    assert d["rt1"][0] == pytest.approx(abs(v[0]), abs=1e-7) and d["rt2"][0] == pytest.approx(abs(v[1]), abs=1e-7)


def test_rotation_enters_through_omega_cross_r(oracle):
    """The box spins at 2 rad/s about z; its contact point at the bottom corner (0.5, 0, 0) moves at
    omega x r = (0, 0, 2) x (0.5, 0, -0.5) = (0, 1, 0). Friction at the cone's edge along +y (with the
    motion) violates by 1.0; along -y it is converged."""
    b, s, ig, ib = ground_and_box(box_w=(0, 0, 2))
    pt = (0.5, 0.0, 0.0)
    r_with, _ = residual(oracle, b, s, contact(ig, ib, UP, [(pt, 10.0, 0.0, 5.0, 0.0)]))
    r_against, _ = residual(oracle, b, s, contact(ig, ib, UP, [(pt, 10.0, 0.0, -5.0, 0.0)]))
    assert r_with == pytest.approx(1.0, rel=1e-6)
    assert r_against == 0.0


@pytest.mark.parametrize("vz,ln,expected", [(0.25, 0.0, 0.0), (-0.4, 0.0, 0.4), (0.25, 3.0, 0.25), (0.0, 3.0, 0.0)])
def test_normal_row_complementarity(oracle, vz, ln, expected):
    """Normal row, target 0: separating with no impulse is fine (0); approaching with no impulse
    violates by the approach speed; an impulse while the contact separates violates by the
    separation speed (it should have been released); resting with an impulse is converged."""
    b, s, ig, ib = ground_and_box(box_v=(0, 0, vz))
    r, _ = residual(oracle, b, s, contact(ig, ib, UP, [((0, 0, 0), ln, 0.0, 0.0, 0.0)]))
    assert r == pytest.approx(expected, abs=1e-7)


@pytest.mark.parametrize("vz,ln,expected", [(-1.0, 3.0, 2.0), (-1.0, 0.0, 0.0), (-4.0, 0.0, 1.0)])
def test_speculative_row_targets_the_gap(oracle, vz, ln, expected):
    """A speculative point 5 cm apart (separation 0.05, h = 1/60): the row allows an approach of up to
    0.05 / h = 3 m/s. With an impulse the closing speed must equal 3 (|-1 + 3| = 2 otherwise);
    without one, approaching at 1 m/s is allowed (0) and at 4 m/s violates by 1."""
    b, s, ig, ib = ground_and_box(box_v=(0, 0, vz))
    r, ow = residual(oracle, b, s, contact(ig, ib, UP, [((0, 0, 0), ln, 0.0, 0.0, 0.05)]), h=H)
    assert r == pytest.approx(expected, rel=1e-5)
    assert ow.residual_detail()["target"][0] == pytest.approx(-0.05 / float(H), rel=1e-6)


def test_restitution_target_only_in_relax_passes(oracle):
    """A restitution bias of 0.2 m/s is the normal target only when restitution is in use: a contact
    separating at 0.2 m/s with an impulse is converged in a relax pass and violates by 0.2 in a
    velocity pass."""
    b, s, ig, ib = ground_and_box(box_v=(0, 0, 0.2))
    m = contact(ig, ib, UP, [((0, 0, 0), 3.0, 0.0, 0.0, 0.0)])
    m["points"][0][0]["restitution_bias"] = 0.2
    r_vel, _ = residual(oracle, b, s, m, use_restitution=False)
    r_rel, _ = residual(oracle, b, s, m, use_restitution=True)
    assert r_vel == pytest.approx(0.2, rel=1e-6)
    assert r_rel == pytest.approx(0.0, abs=1e-7)


def test_inactive_and_trigger_contacts_are_not_counted(oracle):
    """A sleeping box (both sides inactive) and a trigger manifold add nothing, however wrong their
    velocities and impulses are; the maximum runs over every point of every counted manifold."""
    b, s, ig, ib = ground_and_box(box_v=(0, 0, -5.0), box_awake=False)
    r, _ = residual(oracle, b, s, contact(ig, ib, UP, [((0, 0, 0), 0.0, 0.0, 0.0, 0.0)]))
    assert r == 0.0
    b, s, ig, ib = ground_and_box(box_v=(0, 0, -5.0))
    r, _ = residual(oracle, b, s, contact(ig, ib, UP, [((0, 0, 0), 0.0, 0.0, 0.0, 0.0)], trigger=True))
    assert r == 0.0
    ms = [contact(ig, ib, UP, [((0, 0, 0), 1.0, 0.0, 0.0, 0.0), ((0.5, 0.5, 0), 1.0, 0.0, 0.0, 0.0)]),
          contact(ig, ib, UP, [((0, 0, 0), 0.0, 0.0, 0.0, 0.0)], trigger=True)]
    b, s, ig, ib = ground_and_box(box_v=(0, 0, -0.7))
    r, ow = residual(oracle, b, s, ms)
    assert r == pytest.approx(0.7, rel=1e-6)  # |v_n - 0| with an impulse, both points
    assert len(ow.residual_detail()["rn"]) == 2


def test_residual_mode_stops_at_the_tolerance(oracle):
    """Run to the tolerance, the oracle's stopping test agrees with the constructed-case residual:
    a box dropped onto the ground with a residual-mode solve ends with residual <= 1e-4 and every
    recorded point term below it."""
    sb = scenes.SceneBuilder()
    sb.static_box((0, 0, -0.5), (10, 10, 0.5), friction=0.5)
    sb.dynamic_box((0, 0, 0.49), (0.5, 0.5, 0.5), mass=1.0, friction=0.5, velocity=(0.4, 0.1, -0.3),
                   angular_velocity=(0.0, 0.0, 0.5))
    b, s = sb.arrays()
    c = _abi.default_config(residual_mode=1, residual_tolerance=1e-4, residual_max_iterations=512)
    ow = oracle.OracleWorld(c, b, s)
    ow.want_residual_detail(True)
    ow.step_fixed()
    st = ow.stats()
    d = ow.residual_detail()
    assert 1 <= st["velocity_passes"] < 512 and st["residual"] <= 1e-4
    assert len(d["rn"]) == 4
    assert max(d["rn"].max(), d["rt1"].max(), d["rt2"].max()) == pytest.approx(st["residual"], abs=0)


# ---- world_state_hash ----

FNV_OFFSET = np.uint64(1469598103934665603)       # world.h:1335
FNV_OFFSET_PUBLISHED = np.uint64(14695981039346656037)
FNV_PRIME = np.uint64(1099511628211)


def fnv1a64(data, h=FNV_OFFSET):
    h = np.uint64(h)
    with np.errstate(over="ignore"):
        for byte in bytes(data):
            h = np.uint64(h ^ np.uint64(byte)) * FNV_PRIME
    return int(h)


def state_hash_numpy(bodies):
    """world.h:1334-1359 restated: alive bodies in index order, each position (3 x f32),
    orientation (x, y, z, w as 4 x f32), linear and angular velocity (3 x f32 each)."""
    h = FNV_OFFSET
    for b in bodies:
        if not b["alive"]:
            continue
        data = b"".join(np.ascontiguousarray(b[f], np.float32).tobytes()
                        for f in ("position", "orientation", "linear_velocity", "angular_velocity"))
        h = np.uint64(fnv1a64(data, h))
    return int(h)


def test_fnv1a64_published_vectors():
    """The FNV-1a 64-bit test vectors of the algorithm's authors (empty, "a", "foobar")."""
    assert fnv1a64(b"", FNV_OFFSET_PUBLISHED) == 0xcbf29ce484222325
    assert fnv1a64(b"a", FNV_OFFSET_PUBLISHED) == 0xaf63dc4c8601ec8c
    assert fnv1a64(b"foobar", FNV_OFFSET_PUBLISHED) == 0x85944171f73967e8
    assert fnv1a64(b"") == 1469598103934665603  # the reference's basis: a world with no bodies


def test_state_hash_equals_fnv_over_body_bytes(oracle):
    """The oracle's world_state_hash equals the numpy FNV-1a over the downloaded body bytes after
    20 steps of the config-1 scene, and a dead body is left out."""
    sc = scenes.config1_scene(seed=1)
    ow = oracle.OracleWorld(_abi.default_config(**sc.config), sc.bodies, sc.shapes)
    for _ in range(20):
        ow.step_fixed()
    bodies = ow.bodies()
    assert ow.state_hash() == state_hash_numpy(bodies)
    b = bodies.copy()
    b["alive"][7] = 0
    ow2 = oracle.OracleWorld(_abi.default_config(**sc.config), b, sc.shapes)
    assert ow2.state_hash() == state_hash_numpy(b) != state_hash_numpy(bodies)


def test_state_hash_depends_on_field_order(oracle):
    """Swapping two bodies, or a quaternion's w with x, changes the hash (the order is pinned)."""
    sc = scenes.config1_scene(seed=1)
    b = sc.bodies.copy()
    ref = oracle.OracleWorld(_abi.default_config(), b, sc.shapes).state_hash()
    b2 = b.copy()
    b2[[1, 2]] = b2[[2, 1]]
    assert oracle.OracleWorld(_abi.default_config(), b2, sc.shapes).state_hash() != ref
    b3 = b.copy()
    b3["orientation"][:, [0, 3]] = b3["orientation"][:, [3, 0]]
    assert oracle.OracleWorld(_abi.default_config(), b3, sc.shapes).state_hash() == state_hash_numpy(b3) != ref
