"""Scene raycasts in the CPU oracle: the reference's own query tests
(physxcoopa/tests/queries_test.cpp) with their known answers, the ray-vs-shape functions against
an independent method (the first root of the shape's signed distance along the ray, found by
sampling and bisection in float64), the inside-start rules of geometry/ray.h, the slab test the
trees cull with, and the QueryFilter fields."""
import math

import numpy as np
import pytest

from physxcoopa_amd import scenes, _abi
from physxcoopa_amd._abi import PXG_STATIC, PXG_SPHERE, PXG_BOX, PXG_CAPSULE


def static_box_world(oracle, centres, half=0.5, **shape_kw):
    sb = scenes.SceneBuilder()
    ids = []
    for c in centres:
        i = sb.add_body(PXG_STATIC, c)
        sb.add_shape(i, PXG_BOX, half_extents=(half, half, half), **shape_kw)
        ids.append(i)
    b, s = sb.arrays()
    return oracle.OracleWorld(_abi.default_config(), b, s), ids


def x_ray(maxd=100.0, direction=(1.0, 0.0, 0.0)):
    return _abi.rays([(0.0, 0.0, 0.0)], [direction], maxd)


def test_raycast_hits_box_with_exact_distance_point_and_normal(oracle):
    # queries_test.cpp:22-37
    w, (box,) = static_box_world(oracle, [(5.0, 0.0, 0.0)])
    h = w.raycast(x_ray())[0]
    assert h["hit"] == 1 and h["body"] == box
    assert abs(h["distance"] - 4.5) < 1e-3
    np.testing.assert_allclose(h["point"], [4.5, 0, 0], atol=1e-3)
    np.testing.assert_allclose(h["normal"], [-1, 0, 0], atol=1e-3)


def test_raycast_all_sorts_hits_and_any_agrees(oracle):
    # queries_test.cpp:73-97
    w, _ = static_box_world(oracle, [(9.0, 0, 0), (3.0, 0, 0), (6.0, 0, 0)])
    hits = w.raycast_all(x_ray()[0])
    assert len(hits) == 3
    np.testing.assert_allclose(hits["distance"], [2.5, 5.5, 8.5], atol=1e-3)
    assert w.raycast(x_ray())[0]["hit"] == 1
    assert w.raycast(x_ray(direction=(-1.0, 0, 0)))[0]["hit"] == 0  # nothing behind the origin


def test_query_filter_ignore_layer_and_triggers(oracle):
    # queries_test.cpp:270-293: `ignore` skips the near body; a layer mask and include_triggers
    # hide shapes the same way
    sb = scenes.SceneBuilder()
    near = sb.add_body(PXG_STATIC, (3.0, 0, 0))
    sb.add_shape(near, PXG_BOX, half_extents=(0.5, 0.5, 0.5), layer=3, is_trigger=True)
    far = sb.add_body(PXG_STATIC, (6.0, 0, 0))
    sb.add_shape(far, PXG_BOX, half_extents=(0.5, 0.5, 0.5))
    b, s = sb.arrays()
    w = oracle.OracleWorld(_abi.default_config(), b, s)
    r = x_ray()
    assert w.raycast(r)[0]["body"] == near
    for f in (_abi.query_filter(ignore_body=near), _abi.query_filter(layer_mask=~np.uint32(1 << 3) & 0xFFFFFFFF),
              _abi.query_filter(include_triggers=False)):
        h = w.raycast(r, f)[0]
        assert h["body"] == far and abs(h["distance"] - 5.5) < 1e-4
        assert len(w.raycast_all(r[0], f)) == 1
    assert len(w.raycast_all(r[0])) == 2


def test_inside_start_rules(oracle):
    # geometry/ray.h: from inside a sphere the hit is where the ray leaves it (t1); from inside a
    # box the hit is at t = 0; a ray whose max distance ends short of a shape misses it
    sph = np.zeros((), _abi.SHAPE_DTYPE)
    sph["type"], sph["radius"], sph["enabled"] = PXG_SPHERE, 2.0, 1
    sph["local_rotation"] = (0, 0, 0, 1)
    r = _abi.rays([(0.5, 0, 0)], [(1.0, 0, 0)], 100.0)[0]
    hit, t, n = oracle.ray_shape(sph, (0, 0, 0), (0, 0, 0, 1), r)
    assert hit and abs(t - 1.5) < 1e-6 and np.allclose(n, [1, 0, 0])
    box = np.zeros((), _abi.SHAPE_DTYPE)
    box["type"], box["half_extents"], box["enabled"] = PXG_BOX, (1, 1, 1), 1
    box["local_rotation"] = (0, 0, 0, 1)
    hit, t, n = oracle.ray_shape(box, (0, 0, 0), (0, 0, 0, 1), r)
    assert hit and t == 0.0
    short = _abi.rays([(-5.0, 0, 0)], [(1.0, 0, 0)], 2.9)[0]
    assert not oracle.ray_shape(sph, (0, 0, 0), (0, 0, 0, 1), short)[0]
    assert oracle.ray_shape(sph, (0, 0, 0), (0, 0, 0, 1), _abi.rays([(-5.0, 0, 0)], [(1.0, 0, 0)], 3.0)[0])[0]


# --- independent method: first root of the signed distance along the ray, float64 ---

def quat_matrix(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def sdf(kind, p, geo):
    if kind == PXG_SPHERE:
        c, r = geo
        return np.linalg.norm(p - c, axis=-1) - r
    if kind == PXG_BOX:
        c, R, he = geo
        lp = (p - c) @ R
        q = np.abs(lp) - he
        return np.linalg.norm(np.maximum(q, 0.0), axis=-1) + np.minimum(np.max(q, axis=-1), 0.0)
    a, b, r = geo
    ab = b - a
    t = np.clip(((p - a) @ ab) / (ab @ ab), 0.0, 1.0)
    return np.linalg.norm(p - (a + t[..., None] * ab), axis=-1) - r


def first_root(kind, geo, o, d, maxd, n=20001):
    ts = np.linspace(0.0, maxd, n)
    f = sdf(kind, o + ts[:, None] * d, geo)
    if f[0] <= 0:
        return None  # starts inside: covered by the inside-start test
    idx = np.nonzero(f <= 0)[0]
    if len(idx) == 0:
        return np.inf
    lo, hi = ts[idx[0] - 1], ts[idx[0]]
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        if sdf(kind, (o + mid * d)[None], geo)[0] > 0:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def sdf_normal(kind, geo, p, h=1e-6):
    g = np.array([(sdf(kind, (p + h * e)[None], geo)[0] - sdf(kind, (p - h * e)[None], geo)[0]) / (2 * h)
                  for e in np.eye(3)])
    return g / np.linalg.norm(g)


@pytest.mark.parametrize("kind", [PXG_SPHERE, PXG_BOX, PXG_CAPSULE])
def test_ray_vs_shape_equals_first_root_of_signed_distance(oracle, kind):
    rng = np.random.default_rng(100 + kind)
    checked = hits = 0
    while checked < 120:
        s = np.zeros((), _abi.SHAPE_DTYPE)
        s["type"], s["enabled"], s["capsule_axis"] = kind, 1, int(rng.integers(0, 3))
        s["local_rotation"] = (0, 0, 0, 1)
        s["radius"] = rng.uniform(0.2, 1.0)
        s["half_extents"] = rng.uniform(0.2, 1.0, 3)
        s["capsule_radius"], s["capsule_half_height"] = rng.uniform(0.1, 0.5), rng.uniform(0.1, 0.8)
        pos = rng.uniform(-1, 1, 3)
        q = scenes.random_quaternion(rng)
        # aim near the shape so that many of the rays hit
        o = pos + rng.uniform(-4, 4, 3)
        tgt = pos + rng.uniform(-1.2, 1.2, 3) * (0.4 if kind == PXG_CAPSULE else 1.0)
        d = (tgt - o) / np.linalg.norm(tgt - o)
        o32, d32 = o.astype(np.float32), d.astype(np.float32)
        o, d = o32.astype(np.float64), d32.astype(np.float64)
        R = quat_matrix(np.asarray(q, np.float64))
        if kind == PXG_SPHERE:
            geo = (pos, float(s["radius"]))
        elif kind == PXG_BOX:
            geo = (pos, R, s["half_extents"].astype(np.float64))
        else:
            ax = R[:, int(s["capsule_axis"])]
            hh = float(s["capsule_half_height"])
            geo = (pos - ax * hh, pos + ax * hh, float(s["capsule_radius"]))
        maxd = 12.0
        ref = first_root(kind, geo, o, d, maxd)
        if ref is None:
            continue
        ray = _abi.rays([o32], [d32], maxd)[0]
        hit, t, n = oracle.ray_shape(s, pos.astype(np.float32), np.asarray(q, np.float32), ray)
        checked += 1
        if not np.isfinite(ref):
            # grazing rays within a hair of the surface are left out
            assert not hit or sdf(kind, (o + t * d)[None], geo)[0] < 1e-4
            continue
        assert hit, (kind, ref)
        assert abs(t - ref) <= 2e-4 * max(1.0, ref), (kind, t, ref)
        p = o + t * d
        n_ref = sdf_normal(kind, geo, p)
        edge = kind == PXG_BOX and np.sort(np.abs(((p - geo[0]) @ geo[1]) / geo[2]))[-2] > 0.999
        if not edge:
            assert np.allclose(n, n_ref, atol=2e-3), (kind, n, n_ref)
        hits += 1
    assert hits > 30


def test_slab_test_agrees_with_interval_overlap(oracle):
    """Ray::intersect (geometry/ray.h:47-60): the ray segment [0, max] meets the box iff the
    per-axis entry/exit intervals overlap (float64 reference, boxes clearly hit or missed)."""
    rng = np.random.default_rng(7)
    agree = 0
    for _ in range(400):
        mn = rng.uniform(-2, 1, 3)
        mx = mn + rng.uniform(0.1, 2, 3)
        o = rng.uniform(-4, 4, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        maxd = rng.uniform(0.5, 8)
        o32, d32 = o.astype(np.float32), d.astype(np.float32)
        o, d = o32.astype(np.float64), d32.astype(np.float64)
        with np.errstate(divide="ignore"):
            t0, t1 = (mn - o) / d, (mx - o) / d
        near = max(np.max(np.minimum(t0, t1)), 0.0)
        far = min(np.min(np.maximum(t0, t1)), maxd)
        if abs(far - near) < 1e-4:
            continue
        got = oracle.ray_aabb(_abi.rays([o32], [d32], maxd)[0], mn.astype(np.float32), mx.astype(np.float32))
        assert got == (near <= far)
        agree += 1
    assert agree > 300


def test_world_raycast_is_the_nearest_of_all_hits(oracle):
    """On a settled pile: raycast() is raycast_all()'s first hit, every hit lies on the shape
    surface it names, and raycast_all() is sorted."""
    b, s = scenes.mixed_pile_scene(n=150, seed=3)
    w = oracle.OracleWorld(_abi.default_config(), b, s)
    for _ in range(90):
        w.step_fixed()
    rng = np.random.default_rng(5)
    o = np.stack([rng.uniform(-4, 4, 64), rng.uniform(-4, 4, 64), np.full(64, 12.0)], axis=1)
    d = np.tile([0.0, 0.0, -1.0], (64, 1)) + rng.normal(scale=0.2, size=(64, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = _abi.rays(o, d, 40.0)
    closest = w.raycast(r)
    n_hit = 0
    for i in range(64):
        allh = w.raycast_all(r[i])
        assert np.all(np.diff(allh["distance"]) >= 0)
        if len(allh) == 0:
            assert closest[i]["hit"] == 0
            continue
        n_hit += 1
        assert closest[i].tobytes() == allh[0].tobytes()
        np.testing.assert_allclose(allh["point"], r[i]["origin"] + allh["distance"][:, None] * r[i]["direction"],
                                   atol=1e-5)
    assert n_hit > 40


def test_overlaps_find_only_touched_bodies(oracle):
    # queries_test.cpp:100-123
    w, (near, far) = static_box_world(oracle, [(0.0, 0.0, 0.0), (20.0, 0.0, 0.0)])
    assert list(w.overlap(_abi.overlap_sphere((0.2, 0, 0), 0.5))) == [near]
    assert list(w.overlap(_abi.overlap_box((0, 0, 0), (1, 1, 1)))) == [near]
    assert list(w.overlap(_abi.overlap_capsule((0, 0, -0.4), (0, 0, 0.4), 0.3))) == [near]
    assert list(w.overlap(_abi.overlap_sphere((10.0, 0, 0), 1.0))) == []
    assert list(w.overlap(_abi.overlap_sphere((0.2, 0, 0), 0.5), _abi.query_filter(ignore_body=near))) == []


def test_sphere_overlaps_match_distance_closed_forms(oracle):
    """A sphere query against spheres and boxes: overlap iff the centre distance to the shape is
    at most the query radius (float64 closed forms: |c - c_i| <= r + r_i for spheres, the clamped
    box distance for boxes), away from the boundary."""
    rng = np.random.default_rng(21)
    sb = scenes.SceneBuilder()
    geo = []
    for k in range(60):
        c = rng.uniform(-6, 6, 3)
        q = scenes.random_quaternion(rng)
        i = sb.add_body(PXG_STATIC, c, q)
        if k % 2:
            r = rng.uniform(0.2, 0.8)
            sb.add_shape(i, PXG_SPHERE, radius=r)
            geo.append(("s", c, r))
        else:
            he = rng.uniform(0.2, 0.8, 3)
            sb.add_shape(i, PXG_BOX, half_extents=he)
            geo.append(("b", c, quat_matrix(np.asarray(q, np.float64)), he))
    b, s = sb.arrays()
    w = oracle.OracleWorld(_abi.default_config(), b, s)
    checked = 0
    for _ in range(200):
        qc, qr = rng.uniform(-6, 6, 3), rng.uniform(0.3, 2.0)
        qc32, qr32 = qc.astype(np.float32), np.float32(qr)
        got = set(int(x) for x in w.overlap(_abi.overlap_sphere(qc32, qr32)))
        for i, g in enumerate(geo):
            if g[0] == "s":
                dist = np.linalg.norm(qc32.astype(np.float64) - g[1]) - g[2]
            else:
                lp = (qc32.astype(np.float64) - g[1]) @ g[2]
                dist = np.linalg.norm(lp - np.clip(lp, -g[3], g[3]))
            if abs(dist - qr) < 1e-4:
                continue
            assert (i in got) == (dist <= qr), (i, dist, qr)
            checked += 1
    assert checked > 10000


# --- shape casts ---

def test_sweeps_are_exact_against_box_corner_and_capsule(oracle):
    # queries_test.cpp:132-174
    w, (box,) = static_box_world(oracle, [(5.0, 0.0, 0.0)])
    h = w.shape_cast(_abi.sphere_cast((0.0, 0.8, 0.8), 0.5, (1, 0, 0), 100.0))
    e = 4.5 - math.sqrt(0.07)
    assert h["hit"] == 1 and h["body"] == box and h["shape_index"] == box and not h["started_inside"]
    assert abs(h["distance"] - e) < 1e-3
    np.testing.assert_allclose(h["point"], [4.5, 0.5, 0.5], atol=1e-3)
    np.testing.assert_allclose(h["normal"], np.array([-math.sqrt(0.07), 0.3, 0.3]) / 0.5, atol=1e-2)
    # a box yawed 45 degrees: its leading edge meets the face x = 4.5 at 4.5 - 0.5 sqrt(2)
    yaw = (0.0, 0.0, math.sin(math.pi / 8), math.cos(math.pi / 8))
    h = w.shape_cast(_abi.box_cast((0, 0, 0), (0.5, 0.5, 0.5), yaw, (1, 0, 0), 100.0))
    assert h["hit"] == 1 and abs(h["distance"] - (4.5 - 0.5 * math.sqrt(2.0))) < 1e-3
    np.testing.assert_allclose(h["normal"], [-1, 0, 0], atol=1e-3)
    np.testing.assert_allclose(h["point"], [4.5, 0, 0], atol=1e-2)
    sb = scenes.SceneBuilder()
    t = sb.add_body(PXG_STATIC, (3.0, 0.0, 0.8))
    sb.add_shape(t, PXG_CAPSULE, capsule_radius=0.2, capsule_half_height=1.0, capsule_axis=1)
    b, s = sb.arrays()
    cw = oracle.OracleWorld(_abi.default_config(), b, s)
    h = cw.shape_cast(_abi.capsule_cast((0, 0, 0), 0.3, 0.5, 2, (0, 0, 0, 1), (1, 0, 0), 100.0))
    assert h["hit"] == 1 and abs(h["distance"] - 2.6) < 1e-3
    np.testing.assert_allclose(h["normal"], [-0.8, 0, -0.6], atol=1e-2)
    assert cw.shape_cast(_abi.capsule_cast((0, 0, -0.65), 0.3, 0.5, 2, (0, 0, 0, 1), (1, 0, 0), 100.0))["hit"] == 0


def test_sweep_reports_started_inside_and_ignores_resting_contact(oracle):
    # queries_test.cpp:224-246
    w, (box,) = static_box_world(oracle, [(5.0, 0.0, 0.0)])
    h = w.shape_cast(_abi.sphere_cast((4.8, 0, 0), 0.25, (1, 0, 0), 2.0))
    assert h["hit"] == 1 and h["started_inside"] == 1 and abs(h["distance"]) < 1e-6 and h["body"] == box
    np.testing.assert_allclose(h["normal"], [-1, 0, 0], atol=1e-3)
    assert w.shape_cast(_abi.sphere_cast((4.8, 0, 0), 0.25, (-1, 0, 0), 2.0))["hit"] == 0
    h = w.shape_cast(_abi.box_cast((4.1, 0, 0), (0.5, 0.5, 0.5), (0, 0, 0, 1), (1, 0, 0), 2.0))
    assert h["hit"] == 1 and h["started_inside"] == 1
    np.testing.assert_allclose(h["normal"], [-1, 0, 0], atol=1e-3)
    assert w.shape_cast(_abi.box_cast((4.1, 0, 0), (0.5, 0.5, 0.5), (0, 0, 0, 1), (-1, 0, 0), 2.0))["hit"] == 0
    sb = scenes.SceneBuilder()
    sb.static_box((0, 0, -0.5), (10, 10, 0.5))
    b, s = sb.arrays()
    fw = oracle.OracleWorld(_abi.default_config(), b, s)
    assert fw.shape_cast(_abi.sphere_cast((0, 0, 0.5), 0.5, (1, 0, 0), 3.0))["hit"] == 0
    assert fw.shape_cast(_abi.capsule_cast((0, 0, 0.9), 0.4, 0.5, 2, (0, 0, 0, 1), (0, 1, 0), 3.0))["hit"] == 0
    assert fw.shape_cast(_abi.box_cast((0, 0, 0.25), (0.25, 0.25, 0.25), (0, 0, 0, 1), (1, 0, 0), 3.0))["hit"] == 0
    h = fw.shape_cast(_abi.sphere_cast((0, 0, 0.5), 0.5, (0, 0, -1), 3.0))
    assert h["hit"] == 1 and abs(h["distance"]) < 1e-3 and h["started_inside"] == 0


def cast_first_root(kind, geo, core_a, core_b, r, d, maxd, n=4001):
    """First t at which the cast core (segment a-b, radius r) moved by d*t touches the shape,
    by sampling the gap min over the core of sdf - r, then bisection (float64)."""
    us = np.linspace(0.0, 1.0, 41)

    def gap(t):
        pts = core_a[None] + us[:, None] * (core_b - core_a)[None] + t * d[None]
        return np.min(sdf(kind, pts, geo)) - r
    ts = np.linspace(0.0, maxd, n)
    g = np.array([gap(t) for t in ts])
    if g[0] <= 0:
        return None
    idx = np.nonzero(g <= 0)[0]
    if len(idx) == 0:
        return np.inf
    lo, hi = ts[idx[0] - 1], ts[idx[0]]
    for _ in range(50):
        mid = 0.5 * (lo + hi)
        if gap(mid) > 0:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


@pytest.mark.parametrize("kind", [PXG_SPHERE, PXG_BOX, PXG_CAPSULE])
@pytest.mark.parametrize("cast_kind", [PXG_SPHERE, PXG_CAPSULE])
def test_cast_distance_equals_first_touch(oracle, kind, cast_kind):
    """The cast distance of a sphere or capsule against one target is where the swept core first
    comes within its radius of the target (float64, sampled along the capsule's core); the
    conservative advancement stops within 1e-4 of contact."""
    rng = np.random.default_rng(40 + 3 * kind + cast_kind)
    hits = 0
    for trial in range(40):
        sb = scenes.SceneBuilder()
        pos = rng.uniform(-1, 1, 3)
        q = scenes.random_quaternion(rng)
        i = sb.add_body(PXG_STATIC, pos, q)
        he = rng.uniform(0.3, 0.9, 3)
        rad, cr, chh = rng.uniform(0.3, 0.9), rng.uniform(0.2, 0.5), rng.uniform(0.2, 0.7)
        sb.add_shape(i, kind, radius=rad, half_extents=he, capsule_radius=cr, capsule_half_height=chh, capsule_axis=2)
        b, s = sb.arrays()
        w = oracle.OracleWorld(_abi.default_config(), b, s)
        R = quat_matrix(np.asarray(q, np.float64))
        if kind == PXG_SPHERE:
            geo = (pos, rad)
        elif kind == PXG_BOX:
            geo = (pos, R, he)
        else:
            geo = (pos - R[:, 2] * chh, pos + R[:, 2] * chh, cr)
        o = (pos + rng.uniform(-5, 5, 3)).astype(np.float32).astype(np.float64)
        tgt = pos + rng.uniform(-0.8, 0.8, 3)
        d = ((tgt - o) / np.linalg.norm(tgt - o)).astype(np.float32).astype(np.float64)
        cr_cast = 0.25
        if cast_kind == PXG_SPHERE:
            c = _abi.sphere_cast(o, cr_cast, d, 12.0)
            ca = cb = o
        else:
            cq = scenes.random_quaternion(rng)
            Rc = quat_matrix(np.asarray(cq, np.float64))
            hh = 0.4
            c = _abi.capsule_cast(o, cr_cast, hh, 2, cq, d, 12.0)
            ca, cb = o - Rc[:, 2] * hh, o + Rc[:, 2] * hh
        ref = cast_first_root(kind, geo, ca, cb, cr_cast, d, 12.0)
        if ref is None:
            continue
        h = w.shape_cast(c)
        if not np.isfinite(ref):
            assert h["hit"] == 0 or h["distance"] > 11.0
            continue
        assert h["hit"] == 1, (kind, cast_kind, ref)
        assert -2e-3 <= ref - h["distance"] <= 2e-3, (kind, cast_kind, float(h["distance"]), ref)
        hits += 1
    assert hits > 8


def boxes_overlap(ca, Ra, ha, cb, Rb, hb):
    """Static separating-axis test of two OBBs in float64 (15 axes)."""
    axes = [Ra[:, i] for i in range(3)] + [Rb[:, i] for i in range(3)]
    axes += [np.cross(Ra[:, i], Rb[:, j]) for i in range(3) for j in range(3)]
    d = cb - ca
    for L in axes:
        n = np.linalg.norm(L)
        if n < 1e-9:
            continue
        L = L / n
        ra = np.sum(ha * np.abs(Ra.T @ L))
        rb = np.sum(hb * np.abs(Rb.T @ L))
        if abs(d @ L) > ra + rb:
            return False
    return True


@pytest.mark.parametrize("kind", [PXG_SPHERE, PXG_BOX, PXG_CAPSULE])
def test_box_cast_distance_equals_first_touch(oracle, kind):
    """A box cast against one target: where the moving box first touches it (float64: bisection on
    the static 15-axis SAT for a box target, the signed distance of the moving box at the round
    target's core for spheres and capsules)."""
    rng = np.random.default_rng(70 + kind)
    hits = 0
    for trial in range(40):
        sb = scenes.SceneBuilder()
        pos = rng.uniform(-1, 1, 3)
        q = scenes.random_quaternion(rng)
        i = sb.add_body(PXG_STATIC, pos, q)
        he = rng.uniform(0.3, 0.9, 3)
        rad, cr, chh = rng.uniform(0.3, 0.9), rng.uniform(0.2, 0.5), rng.uniform(0.2, 0.7)
        sb.add_shape(i, kind, radius=rad, half_extents=he, capsule_radius=cr, capsule_half_height=chh, capsule_axis=2)
        b, s = sb.arrays()
        w = oracle.OracleWorld(_abi.default_config(), b, s)
        R = quat_matrix(np.asarray(q, np.float64))
        o = (pos + rng.uniform(-5, 5, 3)).astype(np.float32).astype(np.float64)
        tgt = pos + rng.uniform(-0.8, 0.8, 3)
        d = ((tgt - o) / np.linalg.norm(tgt - o)).astype(np.float32).astype(np.float64)
        cq = scenes.random_quaternion(rng)
        Rc = quat_matrix(np.asarray(cq, np.float64))
        chex = rng.uniform(0.15, 0.45, 3).astype(np.float32).astype(np.float64)
        c = _abi.box_cast(o, chex, cq, d, 12.0)
        if kind == PXG_BOX:
            def touching(t):
                return boxes_overlap(o + d * t, Rc, chex, pos, R, he)
        else:
            core = (pos, pos, rad) if kind == PXG_SPHERE else (pos - R[:, 2] * chh, pos + R[:, 2] * chh, cr)
            us = np.linspace(0.0, 1.0, 41)
            pts = core[0][None] + us[:, None] * (core[1] - core[0])[None]

            def touching(t):
                return np.min(sdf(PXG_BOX, pts, (o + d * t, Rc, chex))) <= core[2]
        if touching(0.0):
            continue
        ts = np.linspace(0.0, 12.0, 2401)
        first = next((t for t in ts if touching(t)), None)
        h = w.shape_cast(c)
        if first is None:
            assert h["hit"] == 0 or h["distance"] > 11.9
            continue
        lo, hi = first - ts[1], first
        for _ in range(40):
            mid = 0.5 * (lo + hi)
            if touching(mid):
                hi = mid
            else:
                lo = mid
        ref = 0.5 * (lo + hi)
        assert h["hit"] == 1, (kind, ref)
        assert -2e-3 <= ref - h["distance"] <= 2e-3, (kind, float(h["distance"]), ref)
        hits += 1
    assert hits > 8
