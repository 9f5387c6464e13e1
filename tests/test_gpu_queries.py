"""Scene queries on the GPU (pxg_raycast / pxg_raycast_any / pxg_raycast_all / pxg_overlap)
against the CPU oracle: the reference's query tests (physxcoopa/tests/queries_test.cpp) through
the C ABI, bit-identical raycast hits (distance, point, normal, body, shape) on settled piles up
to the full 16k-body pile, and identical overlap body lists, with the oracle given the GPU's poses
and fat AABBs."""
import time

import numpy as np
import pytest

from physxcoopa_amd import scenes, _abi
from physxcoopa_amd._abi import PXG_STATIC, PXG_BOX

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def World():
    from physxcoopa_amd.world import World
    return World


def test_reference_query_tests_on_gpu(World):
    # queries_test.cpp:22-37, 73-97, 270-293
    sb = scenes.SceneBuilder()
    ids = []
    for x in (9.0, 3.0, 6.0):
        i = sb.add_body(PXG_STATIC, (x, 0, 0))
        sb.add_shape(i, PXG_BOX, half_extents=(0.5, 0.5, 0.5), layer=3 if x == 3.0 else 0, is_trigger=x == 3.0)
        ids.append(i)
    b, s = sb.arrays()
    w = World(_abi.default_config(), b, s)
    w.step_fixed()  # statics do not move; the step fills the fat AABBs as the reference's trees hold them
    r = _abi.rays([(0, 0, 0)], [(1, 0, 0)], 100.0)
    h = w.raycast(r)[0]
    assert h["hit"] == 1 and h["body"] == ids[1] and abs(h["distance"] - 2.5) < 1e-3
    np.testing.assert_allclose(h["point"], [2.5, 0, 0], atol=1e-3)
    np.testing.assert_allclose(h["normal"], [-1, 0, 0], atol=1e-3)
    allh = w.raycast_all(r[0])
    np.testing.assert_allclose(allh["distance"], [2.5, 5.5, 8.5], atol=1e-3)
    assert list(w.raycast_any(_abi.rays([(0, 0, 0)] * 2, [(1, 0, 0), (-1, 0, 0)], 100.0))) == [True, False]
    for f in (_abi.query_filter(ignore_body=ids[1]), _abi.query_filter(layer_mask=0xFFFFFFFF & ~(1 << 3)),
              _abi.query_filter(include_triggers=False)):
        h = w.raycast(r, f)[0]
        assert h["body"] == ids[2] and abs(h["distance"] - 5.5) < 1e-4
        assert len(w.raycast_all(r[0], f)) == 2


def random_rays(rng, n, lo, hi, height, spread=0.3, maxd=60.0):
    o = np.stack([rng.uniform(lo, hi, n), rng.uniform(lo, hi, n), np.full(n, height)], axis=1)
    d = np.tile([0.0, 0.0, -1.0], (n, 1)) + rng.normal(scale=spread, size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = _abi.rays(o, d, maxd)
    r["max_distance"][::7] = np.finfo(np.float32).max  # some unbounded rays
    return r


def oracle_from(oracle, gw, cfg, shapes):
    ow = oracle.OracleWorld(cfg, gw.bodies(), shapes)
    ow.set_fat_aabbs(gw.fat_aabbs())
    return ow


@pytest.mark.parametrize("seed", [3, 4])
def test_raycast_equals_oracle_on_settled_pile(World, oracle, seed):
    b, s = scenes.mixed_pile_scene(n=300, seed=seed)
    s = s.copy()
    s["is_trigger"][5::23] = 1
    s["layer"][3::17] = 2
    cfg = _abi.default_config()
    gw = World(cfg, b, s)
    for _ in range(80):
        gw.step_fixed()
    ow = oracle_from(oracle, gw, cfg, s)
    rng = np.random.default_rng(seed)
    r = random_rays(rng, 3000, -6, 6, 15.0)
    for f in (None, _abi.query_filter(include_triggers=False, layer_mask=0xFFFFFFFF & ~(1 << 2)),
              _abi.query_filter(ignore_body=int(s["owner"][10]))):
        g, o = gw.raycast(r, f), ow.raycast(r, f)
        print("seed %d: %d of %d rays hit" % (seed, int(g["hit"].sum()), len(r)), flush=True)
        assert g.tobytes() == o.tobytes()
        assert np.array_equal(gw.raycast_any(r, f), o["hit"] == 1)
        assert g["hit"].sum() > len(r) // 3
    for i in range(0, 3000, 150):
        ga, oa = gw.raycast_all(r[i]), ow.raycast_all(r[i])
        assert ga.tobytes() == oa.tobytes()


def test_raycast_full_pile(World, oracle):
    """The 16k-body pile after 200 steps: 65,536 rays on the GPU, every 32nd checked against the
    oracle bit for bit."""
    sc = scenes.config2_scene(seed=2, n=16384, lattice=(64, 64, 40))
    cfg = _abi.default_config(**sc.config)
    gw = World(cfg, sc.bodies, sc.shapes)
    for _ in range(200):
        gw.step_fixed()
    rng = np.random.default_rng(9)
    r = random_rays(rng, 65536, -34, 34, 45.0, maxd=200.0)
    gw.raycast(r[:256])  # buffers and code warm
    t0 = time.perf_counter()
    g = gw.raycast(r)
    dt = time.perf_counter() - t0
    print("full pile: %d rays in %.2f ms (host to host), %d hit" % (len(r), 1e3 * dt, int(g["hit"].sum())), flush=True)
    ow = oracle_from(oracle, gw, cfg, sc.shapes)
    sub = np.arange(0, len(r), 32)
    o = ow.raycast(r[sub])
    assert g[sub].tobytes() == o.tobytes()
    assert g["hit"].mean() > 0.9


def test_overlaps_equal_oracle(World, oracle):
    """overlap_sphere / overlap_box / overlap_capsule on a settled pile: the GPU's body lists equal
    the oracle's (queries_test.cpp:100-123 semantics, slot order)."""
    b, s = scenes.mixed_pile_scene(n=300, seed=6)
    cfg = _abi.default_config()
    gw = World(cfg, b, s)
    for _ in range(80):
        gw.step_fixed()
    ow = oracle_from(oracle, gw, cfg, s)
    rng = np.random.default_rng(6)
    total = 0
    for k in range(300):
        c = rng.uniform([-5, -5, 0], [5, 5, 3])
        if k % 3 == 0:
            q = _abi.overlap_sphere(c, rng.uniform(0.2, 1.5))
        elif k % 3 == 1:
            q = _abi.overlap_box(c, rng.uniform(0.2, 1.2, 3), scenes.random_quaternion(rng))
        else:
            q = _abi.overlap_capsule(c, c + rng.uniform(-1, 1, 3), rng.uniform(0.1, 0.6))
        f = _abi.query_filter(include_triggers=bool(k % 2))
        g, o = gw.overlap(q, f), ow.overlap(q, f)
        assert np.array_equal(g, o), (k, g, o)
        total += len(g)
    print("overlaps: %d bodies over 300 queries" % total, flush=True)
    assert total > 300


def test_reference_sweep_tests_on_gpu(World):
    # queries_test.cpp:132-174, 224-246
    sb = scenes.SceneBuilder()
    box = sb.add_body(PXG_STATIC, (5.0, 0, 0))
    sb.add_shape(box, PXG_BOX, half_extents=(0.5, 0.5, 0.5))
    b, s = sb.arrays()
    w = World(_abi.default_config(), b, s)
    h = w.shape_cast(_abi.sphere_cast((0.0, 0.8, 0.8), 0.5, (1, 0, 0), 100.0))
    assert h["hit"] == 1 and h["body"] == box and abs(h["distance"] - (4.5 - 0.07 ** 0.5)) < 1e-3
    np.testing.assert_allclose(h["point"], [4.5, 0.5, 0.5], atol=1e-3)
    h = w.shape_cast(_abi.sphere_cast((4.8, 0, 0), 0.25, (1, 0, 0), 2.0))
    assert h["hit"] == 1 and h["started_inside"] == 1 and abs(h["distance"]) < 1e-6
    assert w.shape_cast(_abi.sphere_cast((4.8, 0, 0), 0.25, (-1, 0, 0), 2.0))["hit"] == 0
    yaw = (0.0, 0.0, float(np.sin(np.pi / 8)), float(np.cos(np.pi / 8)))
    h = w.shape_cast(_abi.box_cast((0, 0, 0), (0.5, 0.5, 0.5), yaw, (1, 0, 0), 100.0))
    assert h["hit"] == 1 and abs(h["distance"] - (4.5 - 0.5 * 2 ** 0.5)) < 1e-3
    h = w.shape_cast(_abi.box_cast((4.1, 0, 0), (0.5, 0.5, 0.5), (0, 0, 0, 1), (1, 0, 0), 2.0))
    assert h["hit"] == 1 and h["started_inside"] == 1


def test_shape_casts_equal_oracle(World, oracle):
    """Sphere, box and capsule casts from above and beside a settled pile: the GPU's hit records
    equal the oracle's bit for bit (conservative advancement, swept SAT, start-overlap rules,
    nearest target)."""
    b, s = scenes.mixed_pile_scene(n=300, seed=8)
    cfg = _abi.default_config()
    gw = World(cfg, b, s)
    for _ in range(80):
        gw.step_fixed()
    ow = oracle_from(oracle, gw, cfg, s)
    rng = np.random.default_rng(8)
    hits = started = 0
    for k in range(400):
        o = rng.uniform([-5, -5, 0.2], [5, 5, 6])
        d = rng.normal(size=3)
        d[2] -= 1.0
        if k % 3 == 0:
            c = _abi.sphere_cast(o, rng.uniform(0.1, 0.6), d, rng.uniform(1, 12))
        elif k % 3 == 1:
            c = _abi.box_cast(o, rng.uniform(0.1, 0.5, 3), scenes.random_quaternion(rng), d, rng.uniform(1, 12))
        else:
            c = _abi.capsule_cast(o, rng.uniform(0.1, 0.4), rng.uniform(0.1, 0.6), int(rng.integers(0, 3)),
                                  scenes.random_quaternion(rng), d, rng.uniform(1, 12))
        g, oo = gw.shape_cast(c), ow.shape_cast(c)
        assert g.tobytes() == oo.tobytes(), (k, g, oo)
        hits += int(g["hit"])
        started += int(g["started_inside"])
    print("casts: %d of 400 hit, %d started inside" % (hits, started), flush=True)
    assert hits > 150 and started > 20


def test_reference_penetration_tests_on_gpu(World):
    # queries_test.cpp:335-360 (the convex part)
    sb = scenes.SceneBuilder()
    floor = sb.add_body(PXG_STATIC, (0.0, 0.0, -0.5))
    sb.add_shape(floor, PXG_BOX, half_extents=(2.0, 2.0, 0.5))
    b, s = sb.arrays()
    w = World(_abi.default_config(), b, s)
    ident = (0, 0, 0, 1)
    for probe, depth in ((_abi.sphere_cast((0, 0, 0.4), 0.5, (0, 0, 1), 0.0), 0.1),
                         (_abi.box_cast((0.5, 0, 0.1), (0.25, 0.25, 0.25), ident, (0, 0, 1), 0.0), 0.15),
                         (_abi.capsule_cast((0, 0, -0.3), 0.2, 0.5, 0, ident, (0, 0, 1), 0.0), 0.5)):
        p = w.compute_penetration(probe)
        assert len(p) == 1 and p[0]["body"] == floor and abs(p[0]["depth"] - depth) < 1e-4
        np.testing.assert_allclose(p[0]["normal"], [0, 0, 1], atol=1e-4)
    assert len(w.compute_penetration(_abi.sphere_cast((0, 0, 0.6), 0.5, (0, 0, 1), 0.0))) == 0
    assert len(w.compute_penetration(_abi.sphere_cast((0, 0, 0.4), 0.5, (0, 0, 1), 0.0),
                                     _abi.query_filter(ignore_body=floor))) == 0


def test_penetrations_equal_oracle(World, oracle):
    """Sphere, box and capsule probes placed inside a settled pile: the GPU's penetration lists
    (normal, depth, point, body, shape, slot order) equal the oracle's bit for bit."""
    b, s = scenes.mixed_pile_scene(n=300, seed=9)
    s = s.copy()
    s["is_trigger"][4::19] = 1
    cfg = _abi.default_config()
    gw = World(cfg, b, s)
    for _ in range(80):
        gw.step_fixed()
    ow = oracle_from(oracle, gw, cfg, s)
    rng = np.random.default_rng(9)
    total = 0
    for k in range(300):
        c = rng.uniform([-5, -5, 0], [5, 5, 3])
        if k % 3 == 0:
            p = _abi.sphere_cast(c, rng.uniform(0.2, 1.0), (0, 0, 1), 0.0)
        elif k % 3 == 1:
            p = _abi.box_cast(c, rng.uniform(0.1, 0.8, 3), scenes.random_quaternion(rng), (0, 0, 1), 0.0)
        else:
            p = _abi.capsule_cast(c, rng.uniform(0.1, 0.5), rng.uniform(0.1, 0.7), int(rng.integers(0, 3)),
                                  scenes.random_quaternion(rng), (0, 0, 1), 0.0)
        f = _abi.query_filter(include_triggers=bool(k % 2))
        g, o = gw.compute_penetration(p, f), ow.compute_penetration(p, f)
        assert g.tobytes() == o.tobytes(), (k, g, o)
        total += len(g)
    print("penetrations: %d over 300 probes" % total, flush=True)
    assert total > 300


def random_casts(rng, n, lo, hi, zlo, zhi, scale=1.0):
    casts = np.zeros(n, _abi.CAST_DTYPE)
    for k in range(n):
        o = rng.uniform([lo, lo, zlo], [hi, hi, zhi])
        d = rng.normal(size=3)
        d[2] -= 1.0
        maxd = rng.uniform(1, 12) * scale
        if k % 3 == 0:
            casts[k] = _abi.sphere_cast(o, rng.uniform(0.1, 0.6), d, maxd)[0]
        elif k % 3 == 1:
            casts[k] = _abi.box_cast(o, rng.uniform(0.1, 0.5, 3), scenes.random_quaternion(rng), d, maxd)[0]
        else:
            casts[k] = _abi.capsule_cast(o, rng.uniform(0.1, 0.4), rng.uniform(0.1, 0.6), int(rng.integers(0, 3)),
                                         scenes.random_quaternion(rng), d, maxd)[0]
    casts["max_distance"][7::50] = -1.0  # a negative maximum: no hit
    casts["direction"][11::50] = 0.0  # no direction: no hit
    return casts


def test_shape_cast_batch_equals_single_casts_and_oracle(World, oracle):
    """600 sphere, box and capsule casts in one pass over a settled pile: every record equals the
    single-cast entry point's and the oracle's bit for bit, with and without a filter."""
    b, s = scenes.mixed_pile_scene(n=300, seed=10)
    s = s.copy()
    s["is_trigger"][3::13] = 1
    cfg = _abi.default_config()
    gw = World(cfg, b, s)
    for _ in range(80):
        gw.step_fixed()
    ow = oracle_from(oracle, gw, cfg, s)
    casts = random_casts(np.random.default_rng(10), 600, -5, 5, 0.2, 6)
    for f in (None, _abi.query_filter(include_triggers=False, ignore_body=int(s["owner"][20]))):
        g = gw.shape_cast_batch(casts, f)
        hits = int(g["hit"].sum())
        print("batched casts: %d of %d hit, %d started inside" % (hits, len(casts), int(g["started_inside"].sum())),
              flush=True)
        assert hits > 250
        for k in range(len(casts)):
            one = gw.shape_cast(casts[k], f)
            assert g[k].tobytes() == one.tobytes(), (k, g[k], one)
            if k % 3 == 0:
                assert g[k].tobytes() == ow.shape_cast(casts[k], f).tobytes(), k


def test_shape_cast_batch_full_pile(World, oracle):
    """The 16k-body pile after 200 steps: 16,384 casts in one pass, every 64th checked against the
    oracle bit for bit."""
    sc = scenes.config2_scene(seed=2, n=16384, lattice=(64, 64, 40))
    cfg = _abi.default_config(**sc.config)
    gw = World(cfg, sc.bodies, sc.shapes)
    for _ in range(200):
        gw.step_fixed()
    casts = random_casts(np.random.default_rng(12), 16384, -34, 34, 5, 45, scale=3.0)
    gw.shape_cast_batch(casts[:256])
    t0 = time.perf_counter()
    g = gw.shape_cast_batch(casts)
    dt = time.perf_counter() - t0
    print("full pile: %d casts in %.2f ms (host to host), %d hit" % (len(casts), 1e3 * dt, int(g["hit"].sum())),
          flush=True)
    ow = oracle_from(oracle, gw, cfg, sc.shapes)
    for k in range(0, len(casts), 64):
        assert g[k].tobytes() == ow.shape_cast(casts[k]).tobytes(), k
    assert g["hit"].mean() > 0.15


def random_overlaps(rng, n, lo, hi, zlo, zhi):
    q = np.zeros(n, _abi.OVERLAP_DTYPE)
    for k in range(n):
        c = rng.uniform([lo, lo, zlo], [hi, hi, zhi])
        if k % 3 == 0:
            q[k] = _abi.overlap_sphere(c, rng.uniform(0.2, 1.5))[0]
        elif k % 3 == 1:
            q[k] = _abi.overlap_box(c, rng.uniform(0.2, 1.2, 3), scenes.random_quaternion(rng))[0]
        else:
            q[k] = _abi.overlap_capsule(c, c + rng.uniform(-1, 1, 3), rng.uniform(0.1, 0.6))[0]
    return q


def test_overlap_batch_equals_single_queries_and_oracle(World, oracle):
    """500 overlap queries in one pass over a settled pile: each list equals the single-query entry
    point's (and the oracle's), in slot order; a small capacity truncates the list but not the count."""
    b, s = scenes.mixed_pile_scene(n=300, seed=11)
    cfg = _abi.default_config()
    gw = World(cfg, b, s)
    for _ in range(80):
        gw.step_fixed()
    ow = oracle_from(oracle, gw, cfg, s)
    q = random_overlaps(np.random.default_rng(11), 500, -5, 5, 0, 3)
    for f in (None, _abi.query_filter(include_triggers=False, ignore_body=int(s["owner"][30]))):
        counts, bodies = gw.overlap_batch(q, f, cap=64)
        small_counts, small = gw.overlap_batch(q, f, cap=3)
        assert np.array_equal(counts, small_counts)
        print("batched overlaps: %d bodies over %d queries, max %d" % (int(counts.sum()), len(q), int(counts.max())),
              flush=True)
        assert counts.sum() > 500 and counts.max() <= 64
        for k in range(len(q)):
            one = gw.overlap(q[k], f)
            assert counts[k] == len(one) and np.array_equal(bodies[k, :len(one)], one), (k, one, bodies[k])
            assert np.array_equal(small[k, :min(3, len(one))], one[:3])
            if k % 5 == 0:
                assert np.array_equal(ow.overlap(q[k], f), one), k


def test_overlap_batch_full_pile(World, oracle):
    """The 16k-body pile after 200 steps: 16,384 overlap queries in one pass, every 64th checked
    against the oracle."""
    sc = scenes.config2_scene(seed=2, n=16384, lattice=(64, 64, 40))
    cfg = _abi.default_config(**sc.config)
    gw = World(cfg, sc.bodies, sc.shapes)
    for _ in range(200):
        gw.step_fixed()
    q = random_overlaps(np.random.default_rng(13), 16384, -34, 34, 0, 20)
    gw.overlap_batch(q[:256])
    t0 = time.perf_counter()
    counts, bodies = gw.overlap_batch(q, cap=128)
    dt = time.perf_counter() - t0
    print("full pile: %d overlap queries in %.2f ms (host to host), %d bodies" % (len(q), 1e3 * dt, int(counts.sum())),
          flush=True)
    assert counts.max() <= 128
    ow = oracle_from(oracle, gw, cfg, sc.shapes)
    for k in range(0, len(q), 64):
        o = ow.overlap(q[k])
        assert counts[k] == len(o) and np.array_equal(bodies[k, :len(o)], o), k
    assert (counts > 0).mean() > 0.1


def test_penetration_batch_equals_single_queries(World, oracle):
    """300 posed probes in one pass over a settled pile: each record list equals the single-query
    entry point's (and the oracle's for every 4th), in slot order; a capacity of 2 truncates the
    list but not the count."""
    b, s = scenes.mixed_pile_scene(n=300, seed=12)
    cfg = _abi.default_config()
    gw = World(cfg, b, s)
    for _ in range(80):
        gw.step_fixed()
    ow = oracle_from(oracle, gw, cfg, s)
    probes = random_casts(np.random.default_rng(12), 300, -5, 5, 0, 3)
    for f in (None, _abi.query_filter(include_triggers=False, ignore_body=int(s["owner"][40]))):
        counts, recs = gw.compute_penetration_batch(probes, f, cap=16)
        small_counts, small = gw.compute_penetration_batch(probes, f, cap=2)
        assert np.array_equal(counts, small_counts) and counts.max() <= 16
        print("batched penetrations: %d over %d probes" % (int(counts.sum()), len(probes)), flush=True)
        assert counts.sum() > 300
        for k in range(len(probes)):
            one = gw.compute_penetration(probes[k], f)
            assert counts[k] == len(one) and recs[k, :len(one)].tobytes() == one.tobytes(), k
            assert small[k, :min(2, len(one))].tobytes() == one[:2].tobytes(), k
            if k % 4 == 0:
                assert ow.compute_penetration(probes[k], f).tobytes() == one.tobytes(), k


def test_penetration_batch_full_pile(World, oracle):
    """The 16k-body pile after 200 steps: 16,384 posed probes in one pass (more than 64 chunks of
    64 shapes, so every probe's wavefront loops over several ballot rounds of chunk boxes). Every
    64th probe's records equal the single-query entry point's and the oracle's; a capacity of 2
    keeps every count and truncates the lists to their first two records."""
    sc = scenes.config2_scene(seed=2, n=16384, lattice=(64, 64, 40))
    cfg = _abi.default_config(**sc.config)
    gw = World(cfg, sc.bodies, sc.shapes)
    for _ in range(200):
        gw.step_fixed()
    probes = random_casts(np.random.default_rng(17), 16384, -34, 34, 0.0, 2.5)
    counts, recs = gw.compute_penetration_batch(probes, None, cap=32)
    small_counts, small = gw.compute_penetration_batch(probes, None, cap=2)
    print("full pile: %d penetrations over %d probes, max %d per probe" % (int(counts.sum()), len(probes),
                                                                          int(counts.max())), flush=True)
    assert np.array_equal(counts, small_counts)
    assert counts.max() <= 32 and (counts > 0).mean() > 0.2
    ow = oracle_from(oracle, gw, cfg, sc.shapes)
    for k in range(0, len(probes), 64):
        one = gw.compute_penetration(probes[k], None)
        assert counts[k] == len(one) and recs[k, :len(one)].tobytes() == one.tobytes(), k
        assert small[k, :min(2, len(one))].tobytes() == one[:2].tobytes(), k
        assert ow.compute_penetration(probes[k], None).tobytes() == one.tobytes(), k


def test_batch_queries_ignore_a_nan_fat_box(World):
    """A shape whose fat AABB is NaN is accepted by no box test; as the first slot of its 64-slot
    chunk it must not hide the chunk's other shapes from the batched overlaps and penetrations
    (the chunk box reduction drops NaN operands). The world is only queried after the boxes are
    set, never stepped."""
    b, s = scenes.mixed_pile_scene(n=300, seed=11)
    cfg = _abi.default_config()
    gw = World(cfg, b, s)
    for _ in range(80):
        gw.step_fixed()
    fat = gw.fat_aabbs()
    fat[64] = np.nan
    fat[128, 0] = np.nan
    gw.set_fat_aabbs(fat)
    rng = np.random.default_rng(21)
    q = random_overlaps(rng, 300, -5, 5, 0, 3)
    counts, bodies = gw.overlap_batch(q, cap=64)
    hits_in_chunks = 0
    for k in range(len(q)):
        one = gw.overlap(q[k])
        assert counts[k] == len(one) and np.array_equal(bodies[k, :len(one)], one), k
        hits_in_chunks += int(np.any((s["owner"][64:192][:, None] == one[None, :])))
    assert hits_in_chunks > 0
    probes = random_casts(rng, 200, -5, 5, 0, 3)
    pc, recs = gw.compute_penetration_batch(probes, None, cap=16)
    for k in range(len(probes)):
        one = gw.compute_penetration(probes[k], None)
        assert pc[k] == len(one) and recs[k, :len(one)].tobytes() == one.tobytes(), k
