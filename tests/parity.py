"""Lockstep comparison of the GPU step (libpxg.so) with the CPU oracle.

Each step the oracle is given the GPU's input state (bodies, fat AABBs, warm-start cache) and the
GPU's solve order (colour by colour), then both step and every output is compared:
  * broadphase pair sets: identical;
  * manifolds: identical key sets, counts and feature ids; points, normals, depths within 1e-5 m;
  * solver impulses within 1e-4 relative (atol 1e-6);
  * body state after integration within 1e-5;
  * the colouring is valid (no two manifolds of one colour share an awake dynamic body);
  * contact events (enter / stay / exit): identical lists, in the same order.
Bit-exact agreement is also counted and reported.
"""
import numpy as np

from physxcoopa_amd._abi import PXG_DYNAMIC

TOL_GEOM = 1e-5
TOL_IMP_REL = 1e-4
TOL_IMP_ABS = 1e-6
TOL_BODY = 1e-5


def sorted_pairs(p):
    p = np.asarray(p, np.uint64).reshape(-1, 2)
    key = (p[:, 0] << np.uint64(32)) | p[:, 1]
    return np.sort(key)


def check_colouring(gm, order, offs, bodies):
    active = (bodies["type"] == PXG_DYNAMIC) & (bodies["awake"] != 0) & (bodies["alive"] != 0)
    for c in range(len(offs) - 1):
        seen = set()
        for s in range(offs[c], offs[c + 1]):
            m = gm[order[s]]
            for b in (int(m["a"]), int(m["b"])):
                if active[b]:
                    assert b not in seen, "colour %d uses body %d twice" % (c, b)
                    seen.add(b)


def compare_step(gw, ow, pre_bodies, report, check_events=True):
    gp, op = sorted_pairs(gw.pairs()), sorted_pairs(ow.pairs())
    assert len(gp) == len(op) and np.array_equal(gp, op), "pair sets differ: gpu %d oracle %d" % (len(gp), len(op))
    gm, om = gw.manifolds(), ow.manifolds()
    assert len(gm) == len(om), "manifold count %d vs %d" % (len(gm), len(om))
    gk = {(int(m["shape_a"]), int(m["shape_b"])): i for i, m in enumerate(gm)}
    ok_ = {(int(m["shape_a"]), int(m["shape_b"])): i for i, m in enumerate(om)}
    assert set(gk) == set(ok_), "manifold key sets differ"
    for key, gi in gk.items():
        g, o = gm[gi], om[ok_[key]]
        assert g["count"] == o["count"], "count differs at %s" % (key,)
        assert g["a"] == o["a"] and g["b"] == o["b"]
        assert g["is_trigger"] == o["is_trigger"]
        n = int(g["count"])
        gf = g["points"]["feature_id"][:n]
        of = o["points"]["feature_id"][:n]
        assert np.array_equal(gf, of), "feature ids differ at %s: %s vs %s" % (key, gf, of)
        dn = np.abs(g["normal"] - o["normal"]).max()
        dp = np.abs(g["points"]["position"][:n] - o["points"]["position"][:n]).max() if n else 0.0
        dd = np.abs(g["points"]["penetration"][:n] - o["points"]["penetration"][:n]).max() if n else 0.0
        ds = np.abs(g["points"]["separation"][:n] - o["points"]["separation"][:n]).max() if n else 0.0
        assert max(dn, dp, dd, ds) <= TOL_GEOM, "geometry differs at %s: %g %g %g %g" % (key, dn, dp, dd, ds)
        for f in ("normal_impulse",):
            gv, ov = g["points"][f][:n].astype(np.float64), o["points"][f][:n].astype(np.float64)
            assert np.all(np.abs(gv - ov) <= TOL_IMP_REL * np.abs(ov) + TOL_IMP_ABS), "%s differs at %s: %s vs %s" % (f, key, gv, ov)
        gv = g["points"]["tangent_impulse"][:n].astype(np.float64)
        ov = o["points"]["tangent_impulse"][:n].astype(np.float64)
        assert np.all(np.abs(gv - ov) <= TOL_IMP_REL * np.abs(ov) + TOL_IMP_ABS), "tangent impulse differs at %s" % (key,)
        report["manifolds"] += 1
        report["bitexact_manifolds"] += int(g.tobytes() == o.tobytes())
    order, offs = gw.solve_order()
    check_colouring(gm, order, offs, pre_bodies)
    gb, ob = gw.bodies(), ow.bodies()
    for f in ("position", "orientation", "linear_velocity", "angular_velocity"):
        d = np.abs(gb[f].astype(np.float64) - ob[f]).max() if len(gb) else 0.0
        assert d <= TOL_BODY, "body %s differs by %g" % (f, d)
        report["max_body_diff"] = max(report["max_body_diff"], float(d))
    assert np.array_equal(gb["awake"], ob["awake"]), "sleep flags differ"
    report["bitexact_bodies"] += int(np.array_equal(gb.view(np.uint8), ob.view(np.uint8)))
    if getattr(gw, "n_joints", 0):
        gj, oj = gw.joints(), ow.joints()
        for f in ("point_impulse", "axis_impulse", "limit_impulse", "swing_impulse"):
            gv, ov = gj[f].astype(np.float64), oj[f].astype(np.float64)
            assert np.all(np.abs(gv - ov) <= TOL_IMP_REL * np.abs(ov) + TOL_IMP_ABS * 10), "joint %s differs" % f
        report["bitexact_joints"] += int(np.array_equal(gj.view(np.uint8), oj.view(np.uint8)))
    if check_events and hasattr(gw, "events"):
        ge, oe = gw.events(), ow.events()
        ge = np.stack([ge["a"], ge["b"], ge["type"], ge["is_trigger"]], axis=1) if len(ge) else np.zeros((0, 4), np.uint32)
        assert np.array_equal(ge, oe), "contact events differ: gpu %d oracle %d" % (len(ge), len(oe))
        report["events"] = report.get("events", 0) + len(ge)
    report["steps"] += 1
    report["colours"] = max(report["colours"], len(offs) - 1)


def lockstep(World, OracleWorld, cfg, bodies, shapes, steps, resync=True, feed_order=True, joints=None, ignored=None,
             rounded_atan2=True, hulls=None, meshes=None):
    """feed_order: give the oracle the GPU's solve order each step; without it the oracle solves in
    its own (the reference's) manifold order. joints / ignored: a joint table and ignored body
    pairs for both sides. rounded_atan2: the oracle evaluates atan2 in double and rounds once, as
    the GPU does (False: the reference's float std::atan2, one ulp apart at most)."""
    extra = {}
    if hulls:
        extra["hulls"] = hulls
    if meshes:
        extra["meshes"] = meshes
    gw = World(cfg, bodies, shapes, **extra)
    ow = OracleWorld(cfg, bodies, shapes, **extra)
    ow.set_atan2_mode(rounded_atan2)
    if joints is not None:
        gw.set_joints(joints)
        ow.set_joints(joints)
    if ignored is not None:
        gw.set_ignored_pairs(ignored)
        ow.set_ignored_pairs(ignored)
    report = dict(steps=0, manifolds=0, bitexact_manifolds=0, bitexact_bodies=0, bitexact_joints=0, max_body_diff=0.0,
                  colours=0)
    for _ in range(steps):
        if resync:
            ow.set_bodies(gw.bodies())
            ow.set_fat_aabbs(gw.fat_aabbs())
            k, i = gw.warm_cache()
            ow.set_warm_cache(k, i)
            if joints is not None:
                ow.set_joints(gw.joints())
        pre = gw.bodies()
        if hasattr(gw, "clear_events"):
            gw.clear_events()
            ow.clear_events()
        gw.step_fixed()
        gm = gw.manifolds()
        if feed_order:
            order, _ = gw.solve_order()
            ow.set_order(np.stack([gm["shape_a"][order], gm["shape_b"][order]], axis=1) if len(order) else np.zeros((0, 2)))
        ow.step_fixed()
        compare_step(gw, ow, pre, report)
    return gw, ow, report
