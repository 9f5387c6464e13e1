"""Benchmark of the rigid-body contact step on one MI355X.

Default workload (N=1): the 16,384-body mixed pile (40% boxes, 40% spheres, 20% capsules over a
200x200x1 m ground), dt 1/60. --settle steps (default 540; 8 velocity, 2 relax, 3 position
iterations) let the pile fall and settle whatever --warmup says; then each solver configuration
runs W untimed and K timed steps in residual mode (velocity passes until the residual <= 1e-4, at
most --cap), inputs resident in HBM. A "step" is one full step_fixed: broadphase, narrowphase,
colouring, warm start, velocity, relax and position passes, cache rebuild, islands/sleep and
integration.

Two configurations of the GPU solver are measured from the same scene: plain (the reference's
solve order, plain warm start) and batched + learned (colour batches within 32 spatial tiles, each
tile solved in one workgroup's LDS, learned warm start; --batched-order coloured for greedy colour
batches over the whole pile); the headline value is the latter. Prints ONE JSON line: value =
dynamic body-steps per second, the plain solver's figures beside it, the roofline of the batched
solver's kernel (timed with HIP events on the world's stream; equivalent-HBM algorithmic bytes 192
per manifold + 40 per contact point per velocity pass), and the CPU oracle (the reference algorithm
restated in C++; bounds, pairs and narrowphase on the host's threads, the solver serial) timed on
the first timed steps of the plain world, continued from its state with the same residual policy
(the same steps with the same pass counts: the reference order is bit-identical on both sides).

Other workloads: --workload batch (4096 independent scenes), ragdolls (2048 jointed ragdoll
scenes, learned warm start on), learned (config-1 piles, passes to the residual with and without
the learned warm start; and the 16k pile's drop phase from identical states), queries, aggregates
(FP32 against FP64).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def calibrate_copy():
    """A known-size copy kernel: 10 copies of 256 MiB (read 256 MiB, write 256 MiB each)."""
    import torch
    n = 256 << 20
    a = torch.empty(n, dtype=torch.uint8, device="cuda")
    a.fill_(1)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(10):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read_per_launch": n, "copy_bytes_written_per_launch": n, "launches": 10}), flush=True)


def batch_step_bytes(st, cfg):
    """Equivalent-HBM algorithmic bytes of one step of every scene from their counts (the canonical
    per-body, per-pair, per-manifold and per-point figures: 392 B per dynamic body; 128 B per
    candidate pair; per manifold with n points 32 + 24n written, (192 + 40n) per velocity or relax
    pass, 92 + 4n per position pass and 248 + 44n for warm start, cache rebuild and restitution)."""
    M = float(st["manifolds"].sum())
    C = float(st["points"].sum())
    P = float(st["pairs"].sum())
    vel = float(st["velocity_passes"].sum()) / max(1, len(st))  # passes per step (fixed count or residual mode)
    per_m = (32 + 192 * (vel + cfg.relax_iterations) + 92 * cfg.position_iterations + 248) * M
    per_p = (24 + 40 * (vel + cfg.relax_iterations) + 4 * cfg.position_iterations + 44) * C
    return 128.0 * P + per_m + per_p


def bench_batch(args):
    """Batched scenes stepped K at a time in one graph-captured persistent launch. batch: 4096
    independent scenes (half box pyramids, half random drops of the pile's shape mix) of 8-256
    bodies; ragdolls: 2048 scenes of one 15-bone ragdoll over 32 boxes (joints). The launch
    parameters come from the autotuner (physxcoopa_amd/autotune.py) run on the batch itself."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    from physxcoopa_amd import scenes, _abi, autotune, learned
    from physxcoopa_amd.batch import Batch
    t0 = time.time()
    jt = jo = ig = io = net = None
    if args.workload == "ragdolls":
        b, s, bo, so, jt, jo, ig, io, settings = scenes.config5_scenes(seed=5, n_scenes=2048)
        name = "config5_ragdolls_2048"
        net = learned.MlpParams.load(learned.DEFAULT_NET)  # the learned warm start is on in this workload
    else:
        b, s, bo, so, settings = scenes.config3_scenes(seed=3, n_scenes=4096, n_min=8, n_max=256)
        name = "config3_batch_4096"
    cfg = _abi.default_config(**settings)
    n_dyn = int(np.sum(b["type"] == 2))
    log("batch built in %.1fs: %d scenes, %d dynamic bodies" % (time.time() - t0, len(bo) - 1, n_dyn))

    def setup(batch):
        if jt is not None:
            batch.set_joints(jo, jt)
            batch.set_ignored_pairs(io, ig)
        if net is not None:
            batch.set_mlp(net)
    t_tune = time.time()
    choice, timings = autotune.tune(cfg, b, s, bo, so, steps=30, use_cache=False, log=log, setup=setup)
    t_tune = time.time() - t_tune
    log("autotuner chose %s in %.1f s" % (autotune.label(choice), t_tune))
    bt = Batch(cfg, b, s, bo, so)
    setup(bt)
    autotune.apply(bt, choice)
    chunk = choice["steps_per_launch"]
    done = 0
    while done < args.warmup:
        k = min(chunk, args.warmup - done)
        bt.step(k)
        done += k
    bt.synchronize()
    log("warm-up %d steps in %.1fs" % (args.warmup, time.time() - t0))
    # timed: whole launches of `chunk` steps; the statistics after each launch give the counts
    # the algorithmic bytes are taken from (one download per launch, outside the timed region)
    launches = max(1, (args.steps + chunk - 1) // chunk)
    elapsed, step_bytes = 0.0, 0.0
    for _ in range(launches):
        t1 = time.perf_counter()
        bt.step(chunk)
        bt.synchronize()
        elapsed += time.perf_counter() - t1
        st = bt.stats()
        step_bytes += chunk * batch_step_bytes(st, cfg)
    steps = launches * chunk
    assert np.all(st["error"] == 0), "batch reported errors"
    value = n_dyn * steps / elapsed
    cpu = None
    if not args.no_cpu:
        # scenes stepped concurrently on host threads (the oracle's ctypes calls release the
        # interpreter lock), each scene an oracle world of its own from the GPU state
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pxo
        bodies = bt.bodies()
        rng = np.random.default_rng(0)
        gj = bt.joints() if jt is not None else None
        threads = cpu_threads()
        order = rng.permutation(len(bo) - 1)

        def run_scene(k):
            sb, ss = bodies[bo[k]:bo[k + 1]], s[so[k]:so[k + 1]]
            ow = pxo.OracleWorld(cfg, sb, ss)
            if gj is not None:
                ow.set_joints(gj[jo[k]:jo[k + 1]])
                ow.set_ignored_pairs(ig[io[k]:io[k + 1]])
            ow.run(10)
            return int(np.sum(sb["type"] == 2)) * 10
        bs_cpu, n_sc = 0, 0
        t_cpu0 = time.perf_counter()
        with ThreadPoolExecutor(max_workers=threads) as ex:
            for start in range(0, len(order), 4 * threads):
                bs_cpu += sum(ex.map(run_scene, order[start:start + 4 * threads]))
                n_sc += len(order[start:start + 4 * threads])
                if time.perf_counter() - t_cpu0 > args.cpu_seconds:
                    break
        t_cpu = time.perf_counter() - t_cpu0
        cpu = {"value": round(bs_cpu / t_cpu, 1), "unit": "body·steps/s", "cores": threads, "kind": "port",
               "sample": "%d randomly chosen scenes of the batch, 10 steps each from the GPU state, %d body-steps "
                         "(C++ oracle, -O2, one scene per host thread, %d threads)" % (n_sc, bs_cpu, threads)}
    st = bt.stats()
    ach = step_bytes / elapsed / 1e9
    out = {
        "metric": "body·steps/s at solver residual <=1e-4 (plain vs batched+learned); HBM GB/s",
        "value": round(value, 1), "unit": "body·steps/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": round(1000.0 * elapsed / steps, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": name, "scenes": len(bo) - 1, "dynamic_bodies": n_dyn, "joints": 0 if jt is None else len(jt),
                   "velocity_iterations": int(cfg.velocity_iterations), "position_iterations": int(cfg.position_iterations),
                   "tuning": choice},
        "learned_warm_start": net is not None,
        "learned_predictions": int(st["learned_predictions"].sum()) if net is not None else 0,
        "autotune": {"seconds": round(t_tune, 1), "ms_per_step": {k: round(1e3 * v, 4) for k, v in timings.items()}},
        "manifolds": int(st["manifolds"].sum()), "points": int(st["points"].sum()),
        "roofline": {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 5), "traffic": None, "kernel": "k_batch",
                     "bytes_per_launch": round(step_bytes / launches, 1),
                     "avg_launch_us": round(1e6 * elapsed / launches, 3),
                     "note": "equivalent-HBM algorithmic bytes (the scenes live in LDS during a launch); launch time "
                             "from the host clock around each synchronised graph launch"},
        "cpu_baseline": cpu,
        "vs_cpu": round(value / cpu["value"], 2) if cpu else None,
    }
    bt.close()
    print(json.dumps(out), flush=True)


def bench_learned(args):
    """The learned warm start on config-1 piles it was not trained on, 64 of them stepped as one batch
    for 300 steps in residual mode (velocity passes until the residual <= 1e-4, at most 256): total
    velocity passes and time, plain warm start against learned."""
    import numpy as np
    from physxcoopa_amd import scenes, _abi, learned
    from physxcoopa_amd.batch import Batch
    net = learned.MlpParams.load(learned.DEFAULT_NET)
    seeds = list(range(500, 564))
    sc_list = [scenes.config1_scene(seed=k) for k in seeds]
    b = np.ascontiguousarray(np.concatenate([sc.bodies for sc in sc_list]))
    s = np.ascontiguousarray(np.concatenate([sc.shapes for sc in sc_list]))
    bo = np.concatenate([[0], np.cumsum([len(sc.bodies) for sc in sc_list])]).astype(np.uint32)
    so = np.concatenate([[0], np.cumsum([len(sc.shapes) for sc in sc_list])]).astype(np.uint32)
    n_dyn = int(np.sum(b["type"] == 2))
    cfg = _abi.default_config(residual_mode=1, residual_tolerance=1e-4, residual_max_iterations=256, **sc_list[0].config)
    steps = 300
    res = {}
    for name, mlp in (("plain", None), ("learned", net)):
        bt = Batch(cfg, b, s, bo, so)
        bt.set_mlp(mlp)
        bt.set_tuning(256, True)
        bt.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps // 30):
            bt.step(30)
        bt.synchronize()
        el = time.perf_counter() - t0
        st = bt.stats()
        assert np.all(st["error"] == 0), "batch reported errors"
        res[name] = {"seconds": el, "passes": int(st["velocity_passes_total"].sum()),
                     "residual_mean_last": float(st["residual"].mean())}
        bt.close()
        log("%s: %.2f s, %d velocity passes" % (name, el, res[name]["passes"]))
    drop = learned_drop_phase(net)
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pxo
        sc = sc_list[0]
        ow = pxo.OracleWorld(cfg, sc.bodies, sc.shapes)
        t_cpu = ow.run(steps)
        cpu = {"value": round(sc.n_dynamic * steps / t_cpu, 1), "unit": "body·steps/s", "cores": 1, "kind": "port",
               "sample": "one of the piles (seed %d), 300 residual-mode steps, plain warm start (C++ oracle, -O2, "
                         "one thread)" % seeds[0]}
    value = n_dyn * steps / res["learned"]["seconds"]
    out = {
        "metric": "body·steps/s at solver residual <=1e-4 (plain vs batched+learned); HBM GB/s",
        "value": round(value, 1), "unit": "body·steps/s", "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": round(1000.0 * res["learned"]["seconds"] / steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 solver, bf16 network", "data": "synthetic",
        "config": {"workload": "config4_learned_config1_piles_x64", "piles": len(seeds), "residual_tolerance": 1e-4,
                   "residual_max_passes": 256, "batched": True,
                   "network": "24-64-64-3, trained on 200k records of 32 other config-1 piles at 64 iterations"},
        "velocity_passes": {k: v["passes"] for k, v in res.items()},
        "passes_saved": round(1.0 - res["learned"]["passes"] / max(1, res["plain"]["passes"]), 4),
        "seconds": {k: round(v["seconds"], 3) for k, v in res.items()},
        "residual_mean_last_step": {k: round(v["residual_mean_last"], 7) for k, v in res.items()},
        "plain_gpu": {"value": round(n_dyn * steps / res["plain"]["seconds"], 1)},
        "vs_plain_gpu": round(res["plain"]["seconds"] / res["learned"]["seconds"], 3),
        "vs_cpu": round(value / cpu["value"], 2) if cpu else None,
        "roofline": None,
        "cpu_baseline": cpu,
        "drop_phase_16k": drop,
    }
    print(json.dumps(out), flush=True)


def learned_drop_phase(net, every=20, first=40, last=300):
    """The learned warm start while the 16,384-body pile falls and lands, where warm-start misses are
    common: from the same states (snapshots of one plain run in the tiled order, every `every`-th
    step), one residual-mode step (tolerance 1e-4, residual after every pass, at most 4096 passes)
    with the plain warm start, with the learned one, and with the misses seeded with the impulses
    the plain step converged to (what a perfect predictor of misses would give)."""
    import numpy as np
    from physxcoopa_amd import scenes, _abi
    from physxcoopa_amd.world import World
    sc = scenes.config2_scene(seed=2, n=16384, lattice=(64, 64, 40))
    cfg = _abi.default_config(**sc.config)
    cfg.solve_order = _abi.PXG_ORDER_TILED
    w0 = World(cfg, sc.bodies, sc.shapes)
    rc = _abi.default_config(residual_mode=1, residual_tolerance=1e-4, residual_max_iterations=4096,
                             residual_check_interval=1, **sc.config)
    rc.solve_order = _abi.PXG_ORDER_TILED
    w = World(rc, sc.bodies, sc.shapes)
    rows = []
    for step in range(1, last + 1):
        w0.step_fixed()
        if step < first or step % every:
            continue
        b, f, (k0, i0) = w0.bodies(), w0.fat_aabbs(), w0.warm_cache()

        def run(keys, imp, mlp):
            w.set_mlp(mlp)
            w.set_bodies(b)
            w.set_fat_aabbs(f)
            w.set_warm_cache(keys, imp)
            w.synchronize()
            t0 = time.perf_counter()
            w.step_fixed()
            st = w.stats()
            return st, time.perf_counter() - t0
        st, tp = run(k0, i0, None)
        kc, ic = w.warm_cache()
        known = set(k0.tolist())
        miss = np.array([x not in known for x in kc.tolist()], bool)
        stl, tl = run(k0, i0, net)
        ste, _ = run(np.concatenate([k0, kc[miss]]), np.concatenate([i0, ic[miss]]), None)
        rows.append({"step": step, "contacts": len(kc), "misses": int(miss.sum()),
                     "passes": {"plain": st["velocity_passes"], "learned": stl["velocity_passes"],
                                "exact_misses": ste["velocity_passes"]},
                     "ms": {"plain": round(1000 * tp, 2), "learned": round(1000 * tl, 2)}})
        log("drop phase step %d: %d misses of %d contacts; passes plain %d, learned %d, exact misses %d" %
            (step, rows[-1]["misses"], rows[-1]["contacts"], st["velocity_passes"], stl["velocity_passes"],
             ste["velocity_passes"]))
    w.close()
    w0.close()
    tot = {k: sum(r["passes"][k] for r in rows) for k in ("plain", "learned", "exact_misses")}
    return {"states": len(rows), "residual_tolerance": 1e-4, "residual_max_passes": 4096,
            "velocity_passes": tot, "ms": {k: round(sum(r["ms"][k] for r in rows), 1) for k in ("plain", "learned")},
            "per_state": rows}


def bench_queries(args):
    """Scene queries on the 16,384-body pile after the warm-up steps: K batches of 65,536 rays
    (from 45 m above the pile, downward with a normal jitter of 0.3, max distance 200 m), closest
    hit per ray. The timed batches run on rays resident in device memory (pxg_raycast_device: the
    per-shape world geometry pass and the raycast kernel); the host-buffer rate (pxg_raycast, PCIe
    copies and one wait included) is reported beside it. No CPU baseline: the oracle tests every
    shape per ray (the result the reference's tree walk gives), which would understate the
    reference's own tree-culled CPU cost. Beside it: K batches of 16,384 mixed sphere / box /
    capsule casts (from 5-45 m, downward-biased, max distance 3-36 m) and of 16,384 mixed overlap
    queries over the pile (host buffers both)."""
    import numpy as np
    import torch
    from physxcoopa_amd import scenes, _abi
    from physxcoopa_amd.world import World
    from physxcoopa_amd._lib import lib, check
    sc = scenes.config2_scene(seed=2, n=16384, lattice=(64, 64, 40))
    cfg = _abi.default_config(**sc.config)
    w = World(cfg, sc.bodies, sc.shapes)
    for _ in range(args.settle + args.warmup):
        w.step_fixed()
    w.synchronize()
    n = 65536
    rng = np.random.default_rng(9)
    o = np.stack([rng.uniform(-34, 34, n), rng.uniform(-34, 34, n), np.full(n, 45.0)], axis=1)
    dvec = np.tile([0.0, 0.0, -1.0], (n, 1)) + rng.normal(scale=0.3, size=(n, 3))
    dvec /= np.linalg.norm(dvec, axis=1, keepdims=True)
    rays = _abi.rays(o, dvec, 200.0)
    d_rays = torch.from_numpy(rays.view(np.uint8).copy()).to("cuda")
    d_hits = torch.empty(n * _abi.HIT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    h = w.h
    check(lib().pxg_raycast_device(h, d_rays.data_ptr(), n, None, d_hits.data_ptr()), "pxg_raycast_device")
    w.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        check(lib().pxg_raycast_device(h, d_rays.data_ptr(), n, None, d_hits.data_ptr()), "pxg_raycast_device")
    w.synchronize()
    el = time.perf_counter() - t0
    dev_hits = d_hits.cpu().numpy().view(_abi.HIT_DTYPE)
    t1 = time.perf_counter()
    for _ in range(args.steps):
        host_hits = w.raycast(rays)
    el_host = time.perf_counter() - t1
    assert dev_hits.tobytes() == host_hits.tobytes(), "device and host entry points differ"
    # batched shape casts and overlap queries over the same pile (host buffers, one pass each)
    nq = 16384
    casts = np.zeros(nq, _abi.CAST_DTYPE)
    ovq = np.zeros(nq, _abi.OVERLAP_DTYPE)
    for k in range(nq):
        c = rng.uniform([-34, -34, 5], [34, 34, 45])
        dv = rng.normal(size=3)
        dv[2] -= 1.0
        maxd = rng.uniform(3, 36)
        q = scenes.random_quaternion(rng)
        if k % 3 == 0:
            casts[k] = _abi.sphere_cast(c, rng.uniform(0.1, 0.6), dv, maxd)[0]
            ovq[k] = _abi.overlap_sphere(c * (1, 1, 0.25), rng.uniform(0.2, 1.5))[0]
        elif k % 3 == 1:
            casts[k] = _abi.box_cast(c, rng.uniform(0.1, 0.5, 3), q, dv, maxd)[0]
            ovq[k] = _abi.overlap_box(c * (1, 1, 0.25), rng.uniform(0.2, 1.2, 3), q)[0]
        else:
            casts[k] = _abi.capsule_cast(c, rng.uniform(0.1, 0.4), rng.uniform(0.1, 0.6), k % 3, q, dv, maxd)[0]
            cc = c * (1, 1, 0.25)
            ovq[k] = _abi.overlap_capsule(cc, cc + rng.uniform(-1, 1, 3), rng.uniform(0.1, 0.6))[0]
    cast_hits = w.shape_cast_batch(casts)
    ov_counts, _ = w.overlap_batch(ovq, cap=64)
    t2 = time.perf_counter()
    for _ in range(args.steps):
        w.shape_cast_batch(casts)
    el_cast = time.perf_counter() - t2
    t3 = time.perf_counter()
    for _ in range(args.steps):
        w.overlap_batch(ovq, cap=64)
    el_ov = time.perf_counter() - t3
    pairs = float(n) * len(sc.shapes)
    flop = 26.0 * pairs  # slab test against the fat AABB per (ray, shape) pair
    out = {
        "metric": "rays/s (closest hit, scene raycasts on the settled 16k-body pile)",
        "value": round(n * args.steps / el, 1), "unit": "rays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1000.0 * el / args.steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "raycasts_config2_pile_16384", "rays_per_batch": n, "shapes": int(len(sc.shapes)),
                   "hit_fraction": round(float(dev_hits["hit"].mean()), 4)},
        "host_buffers": {"value": round(n * args.steps / el_host, 1), "ms_per_batch": round(1000.0 * el_host / args.steps, 4)},
        "batched_casts": {"casts_per_s": round(nq * args.steps / el_cast, 1), "ms_per_batch": round(1000.0 * el_cast / args.steps, 4),
                          "casts_per_batch": nq, "hit_fraction": round(float(cast_hits["hit"].mean()), 4),
                          "kernel": "k_shape_cast_batch (host buffers)"},
        "batched_overlaps": {"queries_per_s": round(nq * args.steps / el_ov, 1), "ms_per_batch": round(1000.0 * el_ov / args.steps, 4),
                             "queries_per_batch": nq, "bodies_found": int(ov_counts.sum()),
                             "kernel": "k_overlap_batch (host buffers)"},
        "roofline": {"bound": "fp32", "achieved": round(flop * args.steps / el / 1e12, 3), "peak": 157.3,
                     "unit": "TFLOP/s", "frac": round(flop * args.steps / el / 1e12 / 157.3, 5), "traffic": None,
                     "kernel": "k_raycast (+ k_query_prep)", "flop_per_batch": flop},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def state_aggregates(bodies, manifolds):
    """Kinetic energy (J; linear plus rotational, the body-frame inertia from inv_inertia_local),
    the largest contact penetration (m) and the number of sleeping dynamic bodies of one state."""
    import numpy as np
    b = bodies[(bodies["type"] == 2) & (bodies["alive"] != 0)]
    v = b["linear_velocity"].astype(np.float64)
    w = b["angular_velocity"].astype(np.float64)
    q = b["orientation"].astype(np.float64)
    # the angular velocity in the body frame: rotate by the conjugate quaternion
    u, s = -q[:, :3], q[:, 3:4]
    t = 2.0 * np.cross(u, w)
    wb = w + s * t + np.cross(u, t)
    iil = b["inv_inertia_local"].astype(np.float64)
    rot = np.where(iil > 0, wb ** 2 / np.where(iil > 0, iil, 1.0), 0.0).sum(1)
    ke = float(0.5 * (b["mass"].astype(np.float64) * (v ** 2).sum(1)).sum() + 0.5 * rot.sum())
    pen = 0.0
    for m in manifolds:
        if m["valid"] and not m["is_trigger"] and m["count"]:
            pen = max(pen, float(m["points"]["penetration"][:m["count"]].max()))
    return {"kinetic_energy": ke, "max_penetration": pen, "asleep": int((b["awake"] == 0).sum())}


def bench_aggregates(args):
    """The FP32 GPU step against the FP32 and FP64 oracles over 300 steps of config 1 (64 boxes) and
    config 2 (the 16,384-body pile) at the fixed iteration counts: the GPU's state hash is compared
    with the FP32 oracle's every step, and the 300-step aggregates (kinetic energy, maximum
    penetration, sleeping bodies) and body positions of FP32 runs are compared with the FP64 run."""
    import numpy as np
    from physxcoopa_amd import scenes, _abi
    from physxcoopa_amd.world import World
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pxo
    threads = cpu_threads()
    out = {}
    t_all = time.time()
    for name, sc in (("config1", scenes.config1_scene(seed=1)),
                     ("config2", scenes.config2_scene(seed=2, n=16384, lattice=(64, 64, 40)))):
        cfg = _abi.default_config(**sc.config)
        gw = World(cfg, sc.bodies, sc.shapes)
        of = pxo.OracleWorld(cfg, sc.bodies, sc.shapes)
        od = pxo.OracleWorld(cfg, sc.bodies, sc.shapes, dbl=True)
        of.set_threads(threads)
        od.set_threads(threads)
        equal = 0
        trace = {"gpu": [], "fp32": [], "fp64": []}
        pos_err = {}
        steps = 300
        for i in range(steps):
            gw.step_fixed()
            of.step_fixed()
            od.step_fixed()
            equal += int(gw.state_hash() == of.state_hash())
            gb, fb, db = gw.bodies(), of.bodies(), od.bodies()
            trace["gpu"].append(state_aggregates(gb, gw.manifolds()))
            trace["fp32"].append(state_aggregates(fb, of.manifolds()))
            trace["fp64"].append(state_aggregates(db, od.manifolds()))
            if i + 1 in (60, 150, 300):
                d = np.abs(gb["position"].astype(np.float64) - db["position"].astype(np.float64)).max(axis=1)
                pos_err[str(i + 1)] = {"max": round(float(d.max()), 7), "median": round(float(np.median(d)), 9)}
            if (i + 1) % 50 == 0:
                log("%s: %d/%d steps, %.0f s" % (name, i + 1, steps, time.time() - t_all))
        def dev(key):
            a = np.array([t[key] for t in trace["gpu"]], np.float64)
            b = np.array([t[key] for t in trace["fp64"]], np.float64)
            return {"max_abs": round(float(np.abs(a - b).max()), 7),
                    "max_rel": round(float((np.abs(a - b) / np.maximum(np.abs(b), 1e-12)).max()), 5)}
        out[name] = {
            "dynamic_bodies": sc.n_dynamic, "steps": steps,
            "gpu_hash_equals_oracle_fp32_steps": equal,
            "final": {k: {kk: (round(vv, 7) if isinstance(vv, float) else vv) for kk, vv in trace[k][-1].items()}
                      for k in trace},
            "gpu_vs_fp64_over_steps": {"kinetic_energy": dev("kinetic_energy"), "max_penetration": dev("max_penetration"),
                                       "asleep_max_abs": int(max(abs(a["asleep"] - b["asleep"])
                                                                 for a, b in zip(trace["gpu"], trace["fp64"])))},
            "position_gpu_fp32_vs_fp64_m": pos_err,
        }
        gw.close()
    print(json.dumps({"metric": "FP32 GPU step against the FP32 and FP64 CPU oracles (300-step aggregates)",
                      "value": None, "unit": None, "n_gpus": 1, "data": "synthetic", "cpu_threads": threads,
                      "results": out}), flush=True)


def cpu_threads():
    """Host threads for the CPU baseline: the cores this process may use, at most 16 (one GPU's
    share of the host on the benchmark machines)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def pile_world(sc, order, net):
    from physxcoopa_amd import _abi
    from physxcoopa_amd.world import World
    cfg = _abi.default_config(**sc.config)
    cfg.solve_order = order
    w = World(cfg, sc.bodies, sc.shapes)
    w.set_mlp(net)
    return w, cfg


def timed_steps(w, steps, per_step_stats):
    """Times `steps` step_fixed calls (device synchronised on both sides); returns (seconds,
    per-step records or None, statistics of the last step, profile of the solve kernel). With
    per_step_stats every step is followed by a statistics read (one wait), which gives each
    step's wall time, velocity passes, contact points and residual."""
    w.set_profiling(True)
    w.profile()  # reset
    w.synchronize()
    per = [] if per_step_stats else None
    t0 = time.perf_counter()
    t_prev = t0
    for _ in range(steps):
        w.step_fixed()
        if per_step_stats:
            st = w.stats()
            t = time.perf_counter()
            per.append({"seconds": t - t_prev, "passes": int(st["velocity_passes"]), "points": int(st["points"]),
                        "residual": float(st["residual"])})
            t_prev = t
    w.synchronize()
    elapsed = time.perf_counter() - t0
    prof = w.profile()
    w.set_profiling(False)
    return elapsed, per, w.stats(), prof


def roofline_of(prof, kernel="k_solve_persistent"):
    vel_s = prof["velocity_ms"] / 1000.0
    achieved = prof["velocity_bytes"] / vel_s / 1e9 if vel_s > 0 else 0.0
    n = max(1, prof["velocity_launches"])
    # traffic: HBM bytes per launch need rocprofv3 PMC counters, which this run does not collect
    return {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None, "kernel": kernel,
            "launches": int(prof["velocity_launches"]), "avg_launch_us": round(1e6 * vel_s / n, 3),
            "bytes_per_launch": round(prof["velocity_bytes"] / n, 1)}


def profiled_traffic(key):
    """The solve kernel's PMC traffic (FETCH_SIZE and WRITE_SIZE scaled by the copy-kernel
    calibration) from the committed rocprofv3 summary of an earlier profiling run, with the state
    of the pile it was measured on, or None. Not a measurement of this run."""
    path = os.path.join(ROOT, "profiles", "solve_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f).get(key)
    except (OSError, ValueError):
        return None
    if not t:
        return None
    out = {"source": "profiles/solve_traffic.json", "key": key}
    out.update(t)
    return out


def pass_summary(per):
    """Velocity passes and residuals of per-step records: whether each step reached the tolerance."""
    import numpy as np
    passes = np.array([r["passes"] for r in per])
    return {"min": int(passes.min()), "median": float(np.median(passes)), "mean": round(float(passes.mean()), 1),
            "max": int(passes.max()), "per_step": [int(x) for x in passes]}


def bench_pile(args):
    """The 16,384-body pile, two configurations from the same scene: the plain GPU solver (the
    reference's solve order, plain warm start) and the batched + learned one (greedy colour
    batches, learned warm start). Each world first settles the pile for --settle steps at the fixed
    iteration counts (8 velocity, 2 relax, 3 position), is timed for --fixed-steps steps at those
    counts, then switches to residual mode (velocity passes until the residual <= 1e-4, at most
    --cap) for W untimed and K timed steps. The headline value is the batched + learned solver in
    residual mode; every step's pass count and final residual are reported, and the line says
    whether each timed step reached the tolerance."""
    import numpy as np
    from physxcoopa_amd import scenes, _abi, learned
    t0 = time.time()
    sc = scenes.config2_scene(seed=2, n=16384, lattice=(64, 64, 40))
    n_dyn = sc.n_dynamic
    log("scene %s built in %.1fs" % (sc.name, time.time() - t0))
    net = learned.MlpParams.load(learned.DEFAULT_NET)
    runs = {}
    worlds = {}
    snap = snap_fixed = None
    batched_order = _abi.PXG_ORDER_TILED if args.batched_order == "tiled" else _abi.PXG_ORDER_COLOURED
    plan = [("plain", _abi.PXG_ORDER_REFERENCE, None), ("batched_learned", batched_order, net)]
    if args.only:
        plan = [p for p in plan if p[0] == args.only]
    tol = 1e-4
    for name, order, mlp in plan:
        w, cfg = pile_world(sc, order, mlp)
        worlds[name] = w
        t1 = time.time()
        for i in range(args.settle):
            w.step_fixed()
            if (i + 1) % 180 == 0:
                w.synchronize()
                log("%s settle %d/%d %.1fs" % (name, i + 1, args.settle, time.time() - t1))
        w.synchronize()
        if name == "plain":
            k, imp = w.warm_cache()
            snap_fixed = (w.bodies(), w.fat_aabbs(), k, imp, cfg)
        el_f, _, st_f, prof_f = timed_steps(w, args.fixed_steps, False)
        res_cfg = _abi.default_config(**sc.config)
        res_cfg.solve_order = order
        res_cfg.residual_mode, res_cfg.residual_tolerance, res_cfg.residual_max_iterations = 1, tol, args.cap
        res_cfg.residual_check_interval = args.check_every
        w.set_config(res_cfg)
        for i in range(args.warmup):
            w.step_fixed()
            st = w.stats()
            log("%s residual-mode warm-up %d/%d: %d passes, residual %.3g" %
                (name, i + 1, args.warmup, st["velocity_passes"], st["residual"]))
        if name == "plain":
            w.synchronize()
            k, imp = w.warm_cache()
            snap = (w.bodies(), w.fat_aabbs(), k, imp, res_cfg)
        el, per, st, prof = timed_steps(w, args.steps, True)
        reached = [r["residual"] <= tol for r in per]
        runs[name] = {"value": n_dyn * args.steps / el, "ms_per_step": 1000.0 * el / args.steps,
                      "contacts_per_s": sum(r["points"] for r in per) / el, "per": per,
                      "passes": pass_summary(per), "residual_max": max(r["residual"] for r in per),
                      "reached": sum(reached), "fixed_iterations_value": n_dyn * args.fixed_steps / el_f,
                      "fixed_iterations_ms_per_step": 1000.0 * el_f / args.fixed_steps,
                      "manifolds": st["manifolds"], "points": st["points"], "colours": st["colours"],
                      "awake_bodies": st["awake_bodies"], "learned_predictions": st["learned_predictions"],
                      "fixed_manifolds": st_f["manifolds"], "prof": prof}
        log("%s: residual mode %.2f ms/step (passes %s, %d/%d steps reached %g), fixed %.2f ms/step, colours %d, "
            "tiled workgroups %d (in LDS %d, largest tile %d KiB)" %
            (name, runs[name]["ms_per_step"], runs[name]["passes"]["per_step"], runs[name]["reached"], args.steps,
             tol, runs[name]["fixed_iterations_ms_per_step"], st["colours"], st["tiled_workgroups"] & 0xFF,
             (st["tiled_workgroups"] >> 8) & 0xFF, st["tiled_workgroups"] >> 16))
        tt = prof.get("tiled_ticks") or {}
        n_pass = sum(r["passes"] for r in per)
        if tt and tt.get("colours") and n_pass:
            runs[name]["tiled_us_per_pass"] = {k: round(v * 0.01 / n_pass, 3) for k, v in tt.items()}
            log("%s: tiled solve, first tile's workgroup, us per velocity pass: %s" % (name, runs[name]["tiled_us_per_pass"]))
    head = runs.get("batched_learned") or runs["plain"]
    cpu = cpu_fixed = None
    if not args.no_cpu and snap is not None:
        # The oracle continues from the plain GPU world's state: in the reference order the GPU and
        # the oracle are bit-identical, so the oracle's steps are the same steps with the same pass
        # counts, and their times compare one to one. Bounds, pair discovery and narrowphase run on
        # a host thread pool (the stages the reference runs on its job system); the solver is serial.
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pxo
        threads = cpu_threads()

        def oracle_from(sn):
            bodies, fat, k, imp, c = sn
            ow = pxo.OracleWorld(c, bodies, sc.shapes)
            ow.set_fat_aabbs(fat)
            ow.set_warm_cache(k, imp)
            ow.set_threads(threads)
            return ow
        ow = oracle_from(snap_fixed)
        n_fix = min(args.fixed_steps, 20)
        t_fix = ow.run(n_fix)
        cpu_fixed = {"value": round(n_dyn * n_fix / t_fix, 1), "unit": "body·steps/s", "cores": threads, "kind": "port",
                     "sample": "%d steps at 8 velocity / 2 relax / 3 position iterations of the %d-body pile after %d "
                               "settling steps, from the plain GPU world's state (C++ oracle, -O2; bounds, pairs and "
                               "narrowphase on %d host threads, solver serial)" % (n_fix, n_dyn, args.settle, threads),
                     "ms_per_step": round(1000 * t_fix / n_fix, 2)}
        log("CPU oracle, fixed iterations: %.1f ms/step on %d threads" % (cpu_fixed["ms_per_step"], threads))
        ow = oracle_from(snap)
        n_cpu, t_cpu, p_cpu = 0, 0.0, []
        log("CPU oracle: residual-mode steps from the plain world's state (budget %.0f s, at least one step)" %
            args.cpu_seconds)
        while n_cpu < args.steps and (n_cpu == 0 or t_cpu < args.cpu_seconds):
            t_cpu += ow.run(1)
            p_cpu.append(int(ow.stats()["velocity_passes"]))
            n_cpu += 1
            log("CPU oracle step %d: %d passes, %.1f s total" % (n_cpu, p_cpu[-1], t_cpu))
        gpu_same = sum(r["seconds"] for r in runs["plain"]["per"][:n_cpu])
        gpu_passes = [r["passes"] for r in runs["plain"]["per"][:n_cpu]]
        cpu = {"value": round(n_dyn * n_cpu / t_cpu, 1), "unit": "body·steps/s", "cores": threads, "kind": "port",
               "sample": "the first %d of the plain GPU world's timed residual-mode steps (tolerance 1e-4, cap %d, "
                         "residual every %d passes) of the %d-body pile after %d settling steps, continued from its "
                         "state (C++ oracle, -O2; bounds, pairs and narrowphase on %d host threads, solver serial); "
                         "velocity passes per step %s (GPU: %s)" %
                         (n_cpu, args.cap, args.check_every, n_dyn, args.settle, threads, p_cpu, gpu_passes),
               "ms_per_step": round(1000 * t_cpu / n_cpu, 2),
               "plain_gpu_same_steps_ms_per_step": round(1000 * gpu_same / n_cpu, 3),
               "plain_gpu_vs_cpu_same_steps": round(t_cpu / gpu_same, 2) if gpu_same > 0 else None,
               "passes_equal_gpu": p_cpu == gpu_passes}
        # The sample is one or a few steps whose pass counts vary (4k-16k); for the whole timed
        # workload the oracle would run the plain GPU's passes (bit-identical), at the sample's
        # time per pass: the CPU rate on the same steps, scaled from the measured sample.
        all_passes = sum(r["passes"] for r in runs["plain"]["per"])
        t_cpu_all = t_cpu / max(1, sum(p_cpu)) * all_passes
        cpu["same_steps_scaled"] = {
            "value": round(n_dyn * args.steps / t_cpu_all, 1), "unit": "body·steps/s",
            "method": "the sample's %.3f ms per velocity pass x the %d velocity passes the plain GPU world (and so the "
                      "oracle, bit-identical) runs over the %d timed steps" %
                      (1000 * t_cpu / max(1, sum(p_cpu)), all_passes, args.steps)}
    if head is runs.get("batched_learned"):
        key = "config2_pile_16384/%s_learned_residual" % ("tiled" if args.batched_order == "tiled" else "coloured")
    else:
        key = "config2_pile_16384/reference_residual"
    out = {
        "metric": "body·steps/s at solver residual <=1e-4 (plain vs batched+learned); HBM GB/s",
        "value": round(head["value"], 1),
        "unit": "body·steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": sc.name, "dynamic_bodies": n_dyn,
                   "solver": ("batched+learned GPU (%s, learned warm start)" % ("colour batches within spatial tiles, each tile in one workgroup's LDS" if args.batched_order == "tiled" else "greedy colour batches")) if "batched_learned" in runs
                   else "plain GPU (reference order)",
                   "residual_tolerance": tol, "residual_max_passes": args.cap,
                   "residual_check_interval": args.check_every,
                   "settle_steps": args.settle, "settle_iterations": "8 velocity, 2 relax, 3 position",
                   "residual_mode_warmup_steps": args.warmup},
        "residual_reached": head["reached"] == args.steps,
        "steps_reaching_tolerance": head["reached"],
        "residual_max": round(float(head["residual_max"]), 7),
        "velocity_passes": head["passes"],
        "contacts_per_s": round(head["contacts_per_s"], 1),
        "manifolds": head["manifolds"], "points": head["points"], "colours": head["colours"],
        "awake_bodies": head["awake_bodies"], "learned_predictions_last_step": head["learned_predictions"],
        "fixed_iterations": {k: {"value": round(r["fixed_iterations_value"], 1),
                                 "ms_per_step": round(r["fixed_iterations_ms_per_step"], 4),
                                 "manifolds": r["fixed_manifolds"]} for k, r in runs.items()},
        "plain_gpu": None if "plain" not in runs else {
            "value": round(runs["plain"]["value"], 1), "ms_per_step": round(runs["plain"]["ms_per_step"], 4),
            "residual_reached": runs["plain"]["reached"] == args.steps,
            "steps_reaching_tolerance": runs["plain"]["reached"],
            "velocity_passes": runs["plain"]["passes"],
            "residual_max": round(float(runs["plain"]["residual_max"]), 7)},
        "vs_plain_gpu": round(head["value"] / runs["plain"]["value"], 3) if "plain" in runs and "batched_learned" in runs else None,
        "vs_cpu": round(head["value"] / cpu["value"], 2) if cpu else None,
        "vs_cpu_same_steps_scaled": round(head["value"] / cpu["same_steps_scaled"]["value"], 2) if cpu else None,
        "roofline": roofline_of(head["prof"], "k_solve_tiled" if head is runs.get("batched_learned")
                                and args.batched_order == "tiled" else "k_solve_persistent"),
        "traffic_profiled": profiled_traffic(key),
        "cpu_baseline": cpu,
        "cpu_baseline_fixed_iterations": cpu_fixed,
        "vs_cpu_fixed_iterations": round(head["fixed_iterations_value"] / cpu_fixed["value"], 2) if cpu_fixed else None,
    }
    for w in worlds.values():
        w.close()
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed steps (residual mode for the pile)")
    ap.add_argument("--warmup", type=int, default=2,
                    help="untimed steps before the timed ones (pile: residual-mode steps after settling; "
                         "batch and queries: ordinary steps)")
    ap.add_argument("--settle", type=int, default=540,
                    help="pile and queries: steps at the fixed iteration counts that let the pile fall and settle "
                         "before anything is timed")
    ap.add_argument("--cap", type=int, default=16384, help="pile: most velocity passes per step in residual mode")
    ap.add_argument("--check-every", type=int, default=8,
                    help="pile: residual mode evaluates the residual after every k-th velocity pass (and the last)")
    ap.add_argument("--fixed-steps", type=int, default=60, help="pile: steps timed at the fixed iteration counts")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline sample budget")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batched-order", choices=["tiled", "coloured"], default="tiled",
                    help="pile: the batched solver's order (colour batches within spatial tiles, each tile solved "
                         "in one workgroup's LDS; or greedy colour batches over the whole pile)")
    ap.add_argument("--only", choices=["plain", "batched_learned"], default=None,
                    help="pile: run one of the two configurations only (for profiling)")
    ap.add_argument("--workload", choices=["pile", "batch", "ragdolls", "learned", "queries", "aggregates"], default="pile",
                    help="pile: the 16k-body pile (one world); batch: 4096 independent scenes of 8-256 bodies; "
                         "ragdolls: 2048 scenes of a jointed ragdoll over 32 boxes; learned: config-1 piles, "
                         "passes and time to the residual with and without the learned warm start; queries: "
                         "batches of 65,536 raycasts against the settled pile")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain device-to-device copy of 256 MiB ten times (the reference kernel "
                         "for scaling the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only (the contact step has no multi-GPU path); "
              "got --gpus %d" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.workload in ("batch", "ragdolls"):
        bench_batch(args)
    elif args.workload == "learned":
        bench_learned(args)
    elif args.workload == "queries":
        bench_queries(args)
    elif args.workload == "aggregates":
        bench_aggregates(args)
    else:
        bench_pile(args)


if __name__ == "__main__":
    main()
