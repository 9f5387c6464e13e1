"""Batched stepping (pxg_batch_*) against the CPU oracle, scene by scene.

Every checked scene is also run as its own oracle world; each step the oracle is given the batch's
solve order for that scene and the two must agree (same bar as tests/parity.py)."""
import numpy as np
import pytest

from physxcoopa_amd import scenes, _abi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_batch():
    b, s, bo, so, cfg = scenes.config3_scenes(seed=5, n_scenes=24, n_min=8, n_max=96)
    return b, s, bo, so, cfg


def compare_scene(batch, k, ow, step_report):
    gm = batch.manifolds(k)
    om = ow.manifolds()
    assert len(gm) == len(om), "scene %d: manifold count %d vs %d" % (k, len(gm), len(om))
    gk = {(int(m["shape_a"]), int(m["shape_b"])): m for m in gm}
    for o in om:
        g = gk[(int(o["shape_a"]), int(o["shape_b"]))]
        n = int(o["count"])
        assert g["count"] == o["count"]
        assert np.array_equal(g["points"]["feature_id"][:n], o["points"]["feature_id"][:n])
        assert np.abs(g["points"]["position"][:n] - o["points"]["position"][:n]).max(initial=0) <= 1e-5
        assert np.abs(g["normal"] - o["normal"]).max() <= 1e-5
        if not o["is_trigger"]:
            for f in ("normal_impulse", "tangent_impulse"):
                gv, ov = g["points"][f][:n].astype(np.float64), o["points"][f][:n].astype(np.float64)
                assert np.all(np.abs(gv - ov) <= 1e-4 * np.abs(ov) + 1e-6), (k, f)
    gb = batch.scene_bodies(k)
    ob = ow.bodies()
    for f in ("position", "orientation", "linear_velocity", "angular_velocity"):
        d = np.abs(gb[f].astype(np.float64) - ob[f]).max()
        assert d <= 1e-5, "scene %d body %s differs by %g" % (k, f, d)
    assert np.array_equal(gb["awake"], ob["awake"])
    step_report["bitexact"] += int(np.array_equal(gb[["position", "orientation", "linear_velocity",
                                                      "angular_velocity"]].tobytes(),
                                                  ob[["position", "orientation", "linear_velocity",
                                                      "angular_velocity"]].tobytes()))
    step_report["checks"] += 1


def test_batch_matches_oracle_per_scene(small_batch, oracle):
    from physxcoopa_amd.batch import Batch
    b, s, bo, so, settings = small_batch
    cfg = _abi.default_config(**settings)
    batch = Batch(cfg, b, s, bo, so)
    checked = [0, 1, 2, 3, 7, 12, 23]
    ows = {k: oracle.OracleWorld(cfg, b[bo[k]:bo[k + 1]], s[so[k]:so[k + 1]]) for k in checked}
    rep = dict(checks=0, bitexact=0)
    for step in range(120):
        batch.step(1)
        for k in checked:
            gm = batch.manifolds(k)
            order, _ = batch.solve_order(k)
            pairs = np.stack([gm["shape_a"][order], gm["shape_b"][order]], axis=1) if len(order) else np.zeros((0, 2))
            ows[k].set_order(pairs)
            ows[k].step_fixed()
            compare_scene(batch, k, ows[k], rep)
    print("batch parity:", rep, flush=True)
    st = batch.stats()
    assert np.all(st["error"] == 0)
    assert st["manifolds"].sum() > 0


def test_batch_multi_step_launch_equals_single_steps(small_batch):
    """K steps in one persistent launch give the same state as K launches of one step."""
    from physxcoopa_amd.batch import Batch
    b, s, bo, so, settings = small_batch
    cfg = _abi.default_config(**settings)
    a = Batch(cfg, b, s, bo, so)
    c = Batch(cfg, b, s, bo, so)
    for _ in range(30):
        a.step(1)
    c.step(30)
    ba, bc = a.bodies(), c.bodies()
    assert ba.tobytes() == bc.tobytes()


@pytest.mark.parametrize("block,lpt", [(64, True), (128, True), (512, True), (256, False)])
def test_batch_tunings_give_identical_results(small_batch, block, lpt):
    from physxcoopa_amd.batch import Batch
    b, s, bo, so, settings = small_batch
    cfg = _abi.default_config(**settings)
    ref = Batch(cfg, b, s, bo, so)
    ref.step(40)
    t = Batch(cfg, b, s, bo, so)
    t.set_tuning(block, lpt)
    t.step(40)
    assert ref.bodies().tobytes() == t.bodies().tobytes()


def test_batch_residual_mode_matches_oracle(small_batch, oracle):
    from physxcoopa_amd.batch import Batch
    b, s, bo, so, settings = small_batch
    cfg = _abi.default_config(residual_mode=1, residual_tolerance=1e-4, residual_max_iterations=48, **settings)
    batch = Batch(cfg, b, s, bo, so)
    k = 2
    ow = oracle.OracleWorld(cfg, b[bo[k]:bo[k + 1]], s[so[k]:so[k + 1]])
    rep = dict(checks=0, bitexact=0)
    for step in range(60):
        batch.step(1)
        gm = batch.manifolds(k)
        order, _ = batch.solve_order(k)
        ow.set_order(np.stack([gm["shape_a"][order], gm["shape_b"][order]], axis=1) if len(order) else np.zeros((0, 2)))
        ow.step_fixed()
        compare_scene(batch, k, ow, rep)
        assert batch.stats()[k]["velocity_passes"] == ow.stats()["velocity_passes"]


def test_batch_ragdolls_match_oracle_per_scene(oracle):
    """Batched ragdoll scenes (joints, ignored pairs): each checked scene equals its own oracle world
    fed the batch's contact order, joints included, bit for bit."""
    from physxcoopa_amd.batch import Batch
    b, s, bo, so, jt, jo, ig, io, settings = scenes.config5_scenes(seed=7, n_scenes=12)
    cfg = _abi.default_config(**settings)
    batch = Batch(cfg, b, s, bo, so)
    batch.set_joints(jo, jt)
    batch.set_ignored_pairs(io, ig)
    checked = [0, 3, 5, 11]
    ows = {}
    for k in checked:
        ow = oracle.OracleWorld(cfg, b[bo[k]:bo[k + 1]], s[so[k]:so[k + 1]])
        ow.set_atan2_mode(True)
        ow.set_joints(jt[jo[k]:jo[k + 1]])
        ow.set_ignored_pairs(ig[io[k]:io[k + 1]])
        ows[k] = ow
    rep = dict(checks=0, bitexact=0)
    for step in range(150):
        batch.step(1)
        for k in checked:
            gm = batch.manifolds(k)
            order, _ = batch.solve_order(k)
            pairs = np.stack([gm["shape_a"][order], gm["shape_b"][order]], axis=1) if len(order) else np.zeros((0, 2))
            ows[k].set_order(pairs)
            ows[k].step_fixed()
            compare_scene(batch, k, ows[k], rep)
    print("ragdoll batch parity:", rep, flush=True)
    assert rep["bitexact"] == rep["checks"]
    gj = batch.joints()
    for k in checked:
        assert np.array_equal(gj[jo[k]:jo[k + 1]].view(np.uint8), ows[k].joints().view(np.uint8))
    assert np.all(batch.stats()["error"] == 0)


def test_batch_learned_warm_start_matches_oracle(oracle):
    """Batched ragdolls with the learned warm start on: the network runs inside the batch kernel on
    the matrix cores; each checked scene equals its oracle world fed the batch's contact order and
    the batch's (clamped) predictions for its warm-start misses."""
    from physxcoopa_amd.batch import Batch
    from physxcoopa_amd import learned
    rng = np.random.default_rng(3)
    net = learned.init_params(rng.normal(0, 1, (512, 24)).astype(np.float32),
                              np.abs(rng.normal(0, 1, (512, 3))).astype(np.float32), seed=5)
    b, s, bo, so, jt, jo, ig, io, settings = scenes.config5_scenes(seed=9, n_scenes=8)
    cfg = _abi.default_config(**settings)
    batch = Batch(cfg, b, s, bo, so)
    batch.set_joints(jo, jt)
    batch.set_ignored_pairs(io, ig)
    batch.set_mlp(net)
    checked = [0, 4, 7]
    ows = {}
    for k in checked:
        ow = oracle.OracleWorld(cfg, b[bo[k]:bo[k + 1]], s[so[k]:so[k + 1]])
        ow.set_atan2_mode(True)
        ow.set_joints(jt[jo[k]:jo[k + 1]])
        ow.set_ignored_pairs(ig[io[k]:io[k + 1]])
        ows[k] = ow
    rep = dict(checks=0, bitexact=0)
    predicted = 0
    for step in range(100):
        batch.step(1)
        for k in checked:
            gm = batch.manifolds(k)
            order, _ = batch.solve_order(k)
            pairs = np.stack([gm["shape_a"][order], gm["shape_b"][order]], axis=1) if len(order) else np.zeros((0, 2))
            ows[k].set_order(pairs)
            keys, imp = batch.predictions(k)
            predicted += len(keys)
            ows[k].set_miss_table(keys, imp)
            ows[k].step_fixed()
            compare_scene(batch, k, ows[k], rep)
    print("learned batch parity:", rep, "predicted", predicted, flush=True)
    assert predicted > 0
    assert rep["bitexact"] == rep["checks"]


def _batch_from(scene_arrays):
    """(bodies, shapes, body offsets, shape offsets) of a list of per-scene (bodies, shapes)."""
    bo = np.zeros(len(scene_arrays) + 1, np.uint32)
    so = np.zeros(len(scene_arrays) + 1, np.uint32)
    bo[1:] = np.cumsum([len(b) for b, _ in scene_arrays])
    so[1:] = np.cumsum([len(s) for _, s in scene_arrays])
    b = np.ascontiguousarray(np.concatenate([b for b, _ in scene_arrays]))
    s = np.ascontiguousarray(np.concatenate([s for _, s in scene_arrays]))
    return b, s, bo, so


def test_batch_config3_sizes_match_oracle_per_scene(oracle):
    """Config-3 scenes at their real sizes: 64 scenes, the first eight of 200-256 dynamic bodies
    (four box pyramids, four random drops: the scenes that fill the workgroup's LDS), the rest
    N ~ U{8..256} alternating pyramids and drops. Every scene is stepped 150 times in the batch and
    in its own oracle world fed the batch's solve order, and compared after every step."""
    from physxcoopa_amd.batch import Batch
    rng = np.random.default_rng(31)
    sizes = [200, 205, 220, 230, 240, 248, 256, 256] + [int(x) for x in rng.integers(8, 257, 56)]
    arrays = []
    for k, n in enumerate(sizes):
        arrays.append(scenes.pyramid_scene(n) if k % 2 == 0 else scenes.drop_scene(n, rng))
    b, s, bo, so = _batch_from(arrays)
    cfg = _abi.default_config(velocity_iterations=8, position_iterations=3)
    batch = Batch(cfg, b, s, bo, so)
    ows = [oracle.OracleWorld(cfg, b[bo[k]:bo[k + 1]], s[so[k]:so[k + 1]]) for k in range(len(sizes))]
    rep = dict(checks=0, bitexact=0)
    max_manifolds = 0
    for step in range(150):
        batch.step(1)
        for k in range(len(sizes)):
            gm = batch.manifolds(k)
            max_manifolds = max(max_manifolds, len(gm))
            order, _ = batch.solve_order(k)
            pairs = np.stack([gm["shape_a"][order], gm["shape_b"][order]], axis=1) if len(order) else np.zeros((0, 2))
            ows[k].set_order(pairs)
            ows[k].step_fixed()
            compare_scene(batch, k, ows[k], rep)
    st = batch.stats()
    print("config-3 sizes parity:", rep, "largest manifold count", max_manifolds, "colours max", st["colours"].max(),
          flush=True)
    assert np.all(st["error"] == 0)
    assert rep["bitexact"] == rep["checks"] == 150 * len(sizes)
    assert max_manifolds > 200


def test_batch_config3_full_launch_properties():
    """The full config-3 batch (4096 scenes of 8-256 bodies) in one persistent launch: no scene
    reports an error, and the result does not depend on the tuning (block size, queue order)."""
    from physxcoopa_amd.batch import Batch
    b, s, bo, so, settings = scenes.config3_scenes(seed=3, n_scenes=4096, n_min=8, n_max=256)
    cfg = _abi.default_config(**settings)
    out = []
    for block, lpt in ((256, True), (128, False)):
        bt = Batch(cfg, b, s, bo, so)
        bt.set_tuning(block, lpt)
        bt.step(30)
        bt.step(30)
        st = bt.stats()
        assert np.all(st["error"] == 0), "scenes with errors: %s" % np.nonzero(st["error"])[0][:10]
        assert st["manifolds"].sum() > 0
        out.append(bt.bodies().tobytes())
        print("tuning", block, lpt, "manifolds", int(st["manifolds"].sum()), "max colours", int(st["colours"].max()),
              flush=True)
        bt.close()
    assert out[0] == out[1]


def test_autotuner_search_and_result_invariance(small_batch, tmp_path, monkeypatch):
    """The autotuner times the batch over its search space (threads per scene, workgroups per CU,
    queue order, steps per launch), caches its choice, and the chosen tuning (and one workgroup per
    CU) steps the scenes to exactly the default tuning's state."""
    from physxcoopa_amd import autotune
    from physxcoopa_amd.batch import Batch
    b, s, bo, so, settings = small_batch
    cfg = _abi.default_config(**settings)
    monkeypatch.setattr(autotune, "CACHE", str(tmp_path / "tuning.json"))
    choice, timings = autotune.tune(cfg, b, s, bo, so, steps=10, rounds=1)
    print("autotuner:", autotune.label(choice), len(timings), "candidates timed", flush=True)
    assert set(choice) == set(autotune.DEFAULT)
    for k, v in choice.items():
        assert v in autotune.SPACE[k]
    assert len(timings) >= sum(len(v) for v in autotune.SPACE.values()) - len(autotune.SPACE) + 1
    again, _ = autotune.tune(cfg, b, s, bo, so, steps=10, rounds=1)
    assert again == choice
    ref = Batch(cfg, b, s, bo, so)
    ref.step(30)
    for tuning in (choice, dict(autotune.DEFAULT, workgroups_per_cu=1, steps_per_launch=10)):
        t = Batch(cfg, b, s, bo, so)
        autotune.apply(t, tuning)
        for _ in range(30 // tuning["steps_per_launch"]):
            t.step(tuning["steps_per_launch"])
        assert ref.bodies().tobytes() == t.bodies().tobytes(), tuning
        t.close()
    ref.close()
