"""The CPU oracle against the reference's own known answers, closed forms and brute force.

Each test names the reference test or formula it pins. None of these retypes the oracle's own
expressions: they check outputs against values the mathematics or the reference's tests fix.
"""
import math

import numpy as np
import pytest

from physxcoopa_amd import scenes, _abi
from physxcoopa_amd._abi import PXG_BOX, PXG_SPHERE, PXG_CAPSULE, PXG_DYNAMIC, PXG_STATIC, PXG_KINEMATIC

ID = (0.0, 0.0, 0.0, 1.0)


def quat_wxyz(w, x, y, z):
    return (x, y, z, w)


def rotmat(q):
    x, y, z, w = [float(c) for c in q]
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def rand_quat(rng):
    return scenes.random_quaternion(rng)


# ---------------------------------------------------------------- narrowphase KATs

def test_box_box_face_contact_four_points_depth_02(oracle):
    # physxcoopa/tests/narrowphase_test.cpp:26-43
    hit, m = oracle.box_box((0, 0, 0), (0.5, 0.5, 0.5), ID, (0.8, 0, 0), (0.5, 0.5, 0.5), ID)
    assert hit
    assert m["count"] == 4
    np.testing.assert_allclose(m["normal"], [1, 0, 0], atol=1e-6)
    for i in range(4):
        assert abs(m["points"][i]["penetration"] - 0.2) < 1e-6
        # reference face: A's +x (id 0); incident face: B's -x (axis 0, negative -> id 1)
        assert m["points"][i]["feature_id"] == (0 << 8) | (1 << 4) | i
        p = m["points"][i]["position"]
        assert abs(p[0] - 0.3) < 1e-6 and abs(abs(p[1]) - 0.5) < 1e-6 and abs(abs(p[2]) - 0.5) < 1e-6


def test_box_box_edge_contact_single_edge_feature(oracle):
    # physxcoopa/tests/narrowphase_test.cpp:79-97 (orientation and centre found there by search)
    q = quat_wxyz(-0.996603, -0.0278176, -0.0606705, -0.0482471)
    hit, m = oracle.box_box((0, 0, 0), (0.5, 0.5, 0.5), ID, (0.864856, 0.686217, 0.79309), (0.5, 0.5, 0.5), q)
    assert hit
    assert m["count"] == 1
    assert m["points"][0]["feature_id"] >= 0x1000


def _sat_brute(ca, hea, qa, cb, heb, qb):
    """Minimum biased overlap over the 15 SAT axes in float64 (the rule of sat.h:75-160)."""
    Ra, Rb = rotmat(qa), rotmat(qb)
    d = np.array(cb, float) - np.array(ca, float)
    best = (math.inf, None, None)
    sep = False
    axes = [(Ra[:, i], i, 0.0) for i in range(3)] + [(Rb[:, i], 3 + i, 0.0) for i in range(3)]
    axes += [(np.cross(Ra[:, i], Rb[:, j]), 6 + 3 * i + j, 0.01) for i in range(3) for j in range(3)]
    for raw, aid, bias in axes:
        l2 = raw @ raw
        if l2 < 1e-8:
            continue
        ax = raw / math.sqrt(l2)
        ra = sum(hea[k] * abs(ax @ Ra[:, k]) for k in range(3))
        rb = sum(heb[k] * abs(ax @ Rb[:, k]) for k in range(3))
        ov = ra + rb - abs(d @ ax)
        if ov < 0:
            sep = True
            continue
        if ov + bias < best[0]:
            best = (ov + bias, aid, ov)
    return sep, best


def test_sat_axis_matches_brute_force_on_random_boxes(oracle):
    rng = np.random.default_rng(11)
    checked = 0
    for _ in range(400):
        hea = rng.uniform(0.2, 0.8, 3)
        heb = rng.uniform(0.2, 0.8, 3)
        qa, qb = rand_quat(rng), rand_quat(rng)
        cb = rng.uniform(-1.2, 1.2, 3)
        hit, axis, depth, normal, separated = oracle.sat((0, 0, 0), hea, qa, cb, heb, qb)
        sep, (score, aid, ov) = _sat_brute((0, 0, 0), hea, qa, cb, heb, qb)
        assert hit == (not sep)
        if not hit:
            continue
        # float32 vs float64: the chosen axis's biased score must be (near) the minimum
        bias = 0.01 if axis >= 6 else 0.0
        assert depth + bias <= score + 2e-4, (axis, aid, depth, score)
        assert abs(np.linalg.norm(normal) - 1) < 1e-5
        checked += 1
    assert checked > 50


def _inside_box(p, c, he, q, tol):
    local = rotmat(q).T @ (np.array(p, float) - np.array(c, float))
    return np.all(np.abs(local) <= np.array(he) + tol)


def test_box_box_points_lie_on_both_boxes(oracle):
    rng = np.random.default_rng(5)
    n_face = 0
    for _ in range(300):
        hea, heb = rng.uniform(0.2, 0.7, 3), rng.uniform(0.2, 0.7, 3)
        qa, qb = rand_quat(rng), rand_quat(rng)
        cb = rng.uniform(-1.0, 1.0, 3)
        hit, m = oracle.box_box((0, 0, 0), hea, qa, cb, heb, qb)
        if not hit:
            continue
        n = np.array(m["normal"], float)
        assert abs(np.linalg.norm(n) - 1) < 1e-5
        for i in range(m["count"]):
            pt = m["points"][i]
            fid = int(pt["feature_id"])
            pen = float(pt["penetration"])
            assert pen >= 0
            if fid < 0x1000:
                n_face += 1
                # clipped incident points sit on the incident box's surface and within penetration of the reference box
                assert _inside_box(pt["position"], (0, 0, 0), hea, qa, pen + 1e-4) or \
                    _inside_box(pt["position"], cb, heb, qb, pen + 1e-4)
            else:
                assert fid & 0xF < 3 and (fid >> 4) & 0xF < 3
        # normal points A -> B
        assert n @ np.array(cb) >= -1e-4
    assert n_face > 50


def _shape(kind, **kw):
    sb = scenes.SceneBuilder()
    i = sb.add_body(PXG_DYNAMIC)
    sb.add_shape(i, kind, **kw)
    return sb.arrays()[1][0]


def test_sphere_sphere_closed_form(oracle):
    a, b = _shape(PXG_SPHERE, radius=0.5), _shape(PXG_SPHERE, radius=0.3)
    hit, m = oracle.generate_contacts(a, (0, 0, 0), ID, b, (0.6, 0.0, 0.0), ID)
    assert hit and m["count"] == 1
    assert abs(m["points"][0]["penetration"] - 0.2) < 1e-6
    np.testing.assert_allclose(m["normal"], [1, 0, 0], atol=1e-7)
    np.testing.assert_allclose(m["points"][0]["position"], [0.5, 0, 0], atol=1e-7)
    hit, _ = oracle.generate_contacts(a, (0, 0, 0), ID, b, (0.81, 0.0, 0.0), ID)
    assert not hit
    # speculative: separated by 0.2 along +y
    hit, m = oracle.generate_contacts(a, (0, 0, 0), ID, b, (0.0, 1.0, 0.0), ID, speculative=True)
    assert hit and abs(m["points"][0]["separation"] - 0.2) < 1e-6 and m["points"][0]["penetration"] == 0


def test_sphere_box_outside_and_inside(oracle):
    box, sph = _shape(PXG_BOX, half_extents=(1, 1, 1)), _shape(PXG_SPHERE, radius=0.5)
    # outside over the +z face: distance 0.3 from the face -> depth 0.2, normal from box to sphere
    hit, m = oracle.generate_contacts(box, (0, 0, 0), ID, sph, (0.2, -0.1, 1.3), ID)
    assert hit and abs(m["points"][0]["penetration"] - 0.2) < 1e-6
    np.testing.assert_allclose(m["normal"], [0, 0, 1], atol=1e-6)
    np.testing.assert_allclose(m["points"][0]["position"], [0.2, -0.1, 1.0], atol=1e-6)
    # sphere first: normal flips (narrowphase.h:221-223)
    hit, m2 = oracle.generate_contacts(sph, (0.2, -0.1, 1.3), ID, box, (0, 0, 0), ID)
    np.testing.assert_allclose(m2["normal"], [0, 0, -1], atol=1e-6)
    # centre inside, nearest face +x at distance 0.1: depth = 0.1 + r
    hit, m = oracle.generate_contacts(box, (0, 0, 0), ID, sph, (0.9, 0.2, -0.3), ID)
    assert hit and abs(m["points"][0]["penetration"] - 0.6) < 1e-6
    np.testing.assert_allclose(m["normal"], [1, 0, 0], atol=1e-6)


def _seg_dist_brute(p1, q1, p2, q2, n=801):
    t = np.linspace(0, 1, n)
    A = np.array(p1)[None] + t[:, None] * (np.array(q1) - np.array(p1))[None]
    B = np.array(p2)[None] + t[:, None] * (np.array(q2) - np.array(p2))[None]
    d = np.linalg.norm(A[:, None, :] - B[None, :, :], axis=2)
    return d.min()


def test_segment_segment_known_answers_and_brute_force(oracle):
    # physxcoopa/tests/geometry_test.cpp:40-57
    c1, c2 = oracle.seg_seg((0, 0, 0), (0, 0, 1), (1, 0, 0), (1, 0, 1))
    assert abs(np.linalg.norm(c2 - c1) - 1) < 1e-4 and abs(c1[2] - c2[2]) < 1e-4
    c1, c2 = oracle.seg_seg((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0))
    np.testing.assert_allclose(c1, 0, atol=1e-6)
    np.testing.assert_allclose(c2, 0, atol=1e-6)
    rng = np.random.default_rng(3)
    for _ in range(60):
        p1, q1, p2, q2 = (rng.uniform(-1, 1, 3) for _ in range(4))
        c1, c2 = oracle.seg_seg(p1, q1, p2, q2)
        d = np.linalg.norm(c2.astype(float) - c1)
        assert d <= _seg_dist_brute(p1, q1, p2, q2) + 1e-4


def test_capsule_pairs_against_segment_distance(oracle):
    rng = np.random.default_rng(9)
    for _ in range(80):
        a = _shape(PXG_CAPSULE, capsule_radius=0.3, capsule_half_height=0.5, capsule_axis=2)
        b = _shape(PXG_CAPSULE, capsule_radius=0.2, capsule_half_height=0.4, capsule_axis=0)
        qa, qb = rand_quat(rng), rand_quat(rng)
        pb = rng.uniform(-1, 1, 3)
        hit, m = oracle.generate_contacts(a, (0, 0, 0), qa, b, pb, qb)
        Ra, Rb = rotmat(qa), rotmat(qb)
        sa = (-0.5 * Ra[:, 2], 0.5 * Ra[:, 2])
        sb = (pb - 0.4 * Rb[:, 0], pb + 0.4 * Rb[:, 0])
        dist = _seg_dist_brute(sa[0], sa[1], sb[0], sb[1])
        if dist < 0.5 - 2e-3:
            assert hit and abs(m["points"][0]["penetration"] - (0.5 - dist)) < 2e-3
        elif dist > 0.5 + 2e-3:
            assert not hit


def test_capsule_box_flat_gets_two_points(oracle):
    # segment.h:91-136: a capsule lying flat on a box face gets fid 0 and 1
    box = _shape(PXG_BOX, half_extents=(2, 2, 0.5))
    cap = _shape(PXG_CAPSULE, capsule_radius=0.3, capsule_half_height=0.5, capsule_axis=0)
    hit, m = oracle.generate_contacts(cap, (0, 0, 0.79), ID, box, (0, 0, 0), ID)
    assert hit and m["count"] == 2
    assert [int(p["feature_id"]) for p in m["points"][:2]] == [0, 1]
    for p in m["points"][:2]:
        assert abs(p["penetration"] - 0.01) < 1e-5
    np.testing.assert_allclose(m["normal"], [0, 0, -1], atol=1e-5)  # capsule (A) -> box (B)


def test_orthonormal_basis_is_orthonormal(oracle):
    rng = np.random.default_rng(1)
    for _ in range(200):
        n = rng.normal(size=3)
        n = (n / np.linalg.norm(n)).astype(np.float32)
        t1, t2 = oracle.orthonormal_basis(n)
        M = np.stack([n, t1, t2]).astype(float)
        np.testing.assert_allclose(M @ M.T, np.eye(3), atol=1e-5)


def test_quaternion_rotation_and_mat3_cast(oracle):
    rng = np.random.default_rng(2)
    for _ in range(50):
        q = rand_quat(rng)
        v = rng.normal(size=3)
        np.testing.assert_allclose(oracle.quat_rotate(q, v), rotmat(q) @ v, atol=2e-5)
        m = oracle.mat3_cast(q)  # [col][row]
        np.testing.assert_allclose(m.T, rotmat(q), atol=1e-6)


def test_material_combine_priority(oracle):
    # physxcoopa/tests/physics_config_test.cpp:21-39
    assert abs(oracle.combine(0.2, 0.8, 0, 0) - 0.5) < 1e-5
    assert abs(oracle.combine(0.2, 0.8, 1, 0) - 0.2) < 1e-5
    assert abs(oracle.combine(0.1, 0.9, 2, 0) - 0.9) < 1e-5
    assert abs(oracle.combine(0.1, 0.9, 3, 2) - 0.09) < 1e-5


def test_inertia_closed_forms():
    # physxcoopa/tests/geometry_test.cpp:70-85 and the textbook formulas
    s = scenes.sphere_inverse_inertia(1.0, 2.0)
    assert abs(s[0] - 1.0 / (0.4 * 2.0)) < 1e-5
    b = scenes.box_inverse_inertia((0.5, 0.25, 1.0), 3.0)
    ex, ey, ez = 1.0, 0.5, 2.0
    np.testing.assert_allclose(b, [12 / (3 * (ey ** 2 + ez ** 2)), 12 / (3 * (ex ** 2 + ez ** 2)),
                                   12 / (3 * (ex ** 2 + ey ** 2))], rtol=1e-6)
    c = scenes.capsule_inverse_inertia(0.5, 1.0, 2.0)
    assert c[2] >= c[0]
    # capsule with zero cylinder is a sphere
    c0 = scenes.capsule_inverse_inertia(0.5, 0.0, 2.0)
    np.testing.assert_allclose(c0, scenes.sphere_inverse_inertia(0.5, 2.0), rtol=1e-6)


def test_warm_start_key_packing(oracle):
    assert oracle.pack_key(3, 9, 0x12) == (3 << 40) ^ (9 << 12) ^ 0x12
    assert oracle.pack_key(9, 3, 0x12) == oracle.pack_key(3, 9, 0x12)
    # the documented aliasing: edge fid bit 12 aliases the high body's bit 0
    assert oracle.pack_key(5, 10, 0x1021) == oracle.pack_key(5, 11, 0x021)


def test_world_bounds(oracle):
    box = _shape(PXG_BOX, half_extents=(1, 2, 3))
    q = scenes.quat_axis_angle((0, 0, 1), math.pi / 2)
    b = oracle.world_bounds(box, (1, 1, 1), q)
    np.testing.assert_allclose(b, [-1, 0, -2, 3, 2, 4], atol=1e-6)
    cap = _shape(PXG_CAPSULE, capsule_radius=0.2, capsule_half_height=0.4, capsule_axis=2)
    np.testing.assert_allclose(oracle.world_bounds(cap, (0, 0, 1), ID), [-0.2, -0.2, 0.4, 0.2, 0.2, 1.6], atol=1e-6)
