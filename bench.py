"""Benchmark of the rigid-body contact step on one MI355X.

Workload (N=1): the 16,384-body mixed pile (40% boxes, 40% spheres, 20% capsules over a
200x200x1 m ground), 8 velocity / 2 relax / 3 position iterations, dt 1/60. W untimed warm-up
steps let the pile fall and settle, then K steps are timed with the world's inputs resident in
HBM. A "step" is one full step_fixed: broadphase, narrowphase, colouring, warm start, velocity,
relax and position passes, cache rebuild, islands/sleep and integration.

Prints ONE JSON line: value = dynamic body-steps per second (whole job), the roofline of the
dominant kernel (the velocity pass, timed with HIP events on the world's stream, algorithmic
bytes 192 per manifold + 40 per contact point per pass), and the CPU oracle (the reference
algorithm restated in C++, one core) timed on a bounded sample of the same settled pile.
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


def bench_batch(args):
    """Batched scenes stepped K at a time in one graph-captured persistent launch. batch: 4096
    independent scenes (half box pyramids, half random drops of the pile's shape mix) of 8-256
    bodies; ragdolls: 2048 scenes of one 15-bone ragdoll over 32 boxes (joints)."""
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only; got --gpus %d" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import numpy as np
    from physxcoopa_amd import scenes, _abi, autotune
    from physxcoopa_amd.batch import Batch
    t0 = time.time()
    jt = jo = ig = io = None
    if args.workload == "ragdolls":
        b, s, bo, so, jt, jo, ig, io, settings = scenes.config5_scenes(seed=5, n_scenes=2048)
        name = "config5_ragdolls_2048"
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
    (block, lpt), timings = autotune.tune(cfg, b, s, bo, so, steps=8, use_cache=False, log=log, setup=setup)
    bt = Batch(cfg, b, s, bo, so)
    setup(bt)
    bt.set_tuning(block, lpt)
    chunk = 30
    done = 0
    while done < args.warmup:
        k = min(chunk, args.warmup - done)
        bt.step(k)
        done += k
    bt.synchronize()
    log("warm-up %d steps in %.1fs" % (args.warmup, time.time() - t0))
    t_start = time.perf_counter()
    done = 0
    while done < args.steps:
        k = min(chunk, args.steps - done)
        bt.step(k)
        done += k
    bt.synchronize()
    elapsed = time.perf_counter() - t_start
    st = bt.stats()
    assert np.all(st["error"] == 0), "batch reported errors"
    value = n_dyn * args.steps / elapsed
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pxo
        bodies = bt.bodies()
        rng = np.random.default_rng(0)
        t_cpu, bs_cpu = 0.0, 0
        gj = bt.joints() if jt is not None else None
        for k in rng.permutation(len(bo) - 1):
            sb, ss = bodies[bo[k]:bo[k + 1]], s[so[k]:so[k + 1]]
            ow = pxo.OracleWorld(cfg, sb, ss)
            if gj is not None:
                ow.set_joints(gj[jo[k]:jo[k + 1]])
                ow.set_ignored_pairs(ig[io[k]:io[k + 1]])
            t_cpu += ow.run(10)
            bs_cpu += int(np.sum(sb["type"] == 2)) * 10
            if t_cpu > args.cpu_seconds:
                break
        cpu = {"value": round(bs_cpu / t_cpu, 1), "unit": "body·steps/s", "cores": 1, "kind": "port",
               "sample": "randomly chosen scenes of the batch, 10 steps each from the GPU state, %d body-steps "
                         "(C++ oracle, -O2, one thread)" % bs_cpu}
    out = {
        "metric": "body·steps/s at solver residual <=1e-4 (plain vs batched+learned); HBM GB/s",
        "value": round(value, 1), "unit": "body·steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1000.0 * elapsed / args.steps, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": name, "scenes": len(bo) - 1, "dynamic_bodies": n_dyn, "joints": 0 if jt is None else len(jt),
                   "velocity_iterations": int(cfg.velocity_iterations), "position_iterations": int(cfg.position_iterations),
                   "block": block, "longest_first": lpt, "steps_per_launch": chunk},
        "autotune_seconds": timings,
        "manifolds": int(st["manifolds"].sum()), "points": int(st["points"].sum()),
        "roofline": None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=540)
    ap.add_argument("--bodies", type=int, default=16384)
    ap.add_argument("--residual", action="store_true", help="velocity passes until residual <= 1e-4 (cap 256)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline sample budget")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--order", choices=["reference", "coloured"], default="reference",
                    help="solve order: the reference's manifold order (default) or the greedy colouring")
    ap.add_argument("--workload", choices=["pile", "batch", "ragdolls"], default="pile",
                    help="pile: the 16k-body pile (one world); batch: 4096 independent scenes of 8-256 bodies; "
                         "ragdolls: 2048 scenes of a jointed ragdoll over 32 boxes")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain device-to-device copy of 256 MiB ten times (the reference kernel "
                         "for scaling the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.workload in ("batch", "ragdolls"):
        bench_batch(args)
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only (the contact step has no multi-GPU path); "
              "got --gpus %d" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from physxcoopa_amd import scenes, _abi
    from physxcoopa_amd.world import World

    t0 = time.time()
    lattice = (64, 64, 40) if args.bodies == 16384 else None
    sc = scenes.config2_scene(seed=2, n=args.bodies, lattice=lattice or (64, 64, max(1, (args.bodies + 4095) // 4096 * 10)))
    log("scene %s built in %.1fs" % (sc.name, time.time() - t0))
    cfg = _abi.default_config(**sc.config)
    cfg.solve_order = _abi.PXG_ORDER_REFERENCE if args.order == "reference" else _abi.PXG_ORDER_COLOURED
    if args.residual:
        cfg.residual_mode = 1
        cfg.residual_tolerance = 1e-4
        cfg.residual_max_iterations = 256
    w = World(cfg, sc.bodies, sc.shapes)
    n_dyn = sc.n_dynamic

    t0 = time.time()
    for i in range(args.warmup):
        w.step_fixed()
        if (i + 1) % 100 == 0:
            w.synchronize()
            st = w.stats()
            log("warmup %d/%d  %.1fs  manifolds %d points %d colours %d awake %d" %
                (i + 1, args.warmup, time.time() - t0, st["manifolds"], st["points"], st["colours"], st["awake_bodies"]))
    w.synchronize()

    # timed region (the velocity-pass launches are bracketed by HIP events on the world's stream)
    w.set_profiling(True)
    w.profile()  # reset
    pts = 0
    passes = 0
    resid = []
    w.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        w.step_fixed()
        if args.residual:
            st = w.stats()
            pts += st["points"]
            passes += st["velocity_passes"]
            resid.append(st["residual"])
    w.synchronize()
    t_end = time.perf_counter()
    prof = w.profile()
    w.set_profiling(False)
    st = w.stats()
    if not args.residual:
        pts = st["points"] * args.steps
        passes = st["velocity_passes"] * args.steps
    elapsed = t_end - t_start
    ms_per_step = 1000.0 * elapsed / args.steps
    value = n_dyn * args.steps / elapsed
    contacts_per_s = pts / elapsed

    vel_s = prof["velocity_ms"] / 1000.0
    achieved = prof["velocity_bytes"] / vel_s / 1e9 if vel_s > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                "kernel": "k_velocity" if os.environ.get("PXG_EAGER") == "1" else "k_solve_persistent",
                "launches": int(prof["velocity_launches"]),
                "avg_launch_us": round(1e6 * vel_s / max(1, prof["velocity_launches"]), 3),
                "bytes_per_launch": round(prof["velocity_bytes"] / max(1, prof["velocity_launches"]), 1)}

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pxo
        ow = pxo.OracleWorld(cfg, w.bodies(), sc.shapes)
        ow.set_fat_aabbs(w.fat_aabbs())
        k, imp = w.warm_cache()
        ow.set_warm_cache(k, imp)
        n_cpu, t_cpu = 0, 0.0
        while t_cpu < args.cpu_seconds and n_cpu < args.steps:
            t_cpu += ow.run(1)
            n_cpu += 1
        cpu = {"value": round(n_dyn * n_cpu / t_cpu, 1), "unit": "body·steps/s", "cores": 1, "kind": "port",
               "sample": "%d steps of the settled %d-body pile from the GPU state after %d steps "
                         "(C++ oracle, -O2, one thread)" % (n_cpu, n_dyn, args.warmup + args.steps),
               "ms_per_step": round(1000 * t_cpu / n_cpu, 2)}

    out = {
        "metric": "body·steps/s at solver residual <=1e-4 (plain vs batched+learned); HBM GB/s",
        "value": round(value, 1),
        "unit": "body·steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": sc.name, "dynamic_bodies": n_dyn, "solver": "plain GPU, %s order" % args.order,
                   "velocity_iterations": int(cfg.velocity_iterations), "relax_iterations": int(cfg.relax_iterations),
                   "position_iterations": int(cfg.position_iterations), "residual_mode": bool(args.residual),
                   "residual_tolerance": 1e-4 if args.residual else None},
        "contacts_per_s": round(contacts_per_s, 1),
        "velocity_passes_per_step": round(passes / args.steps, 2),
        "residual_last": (round(float(np.max(resid)), 7) if resid else round(float(st["residual"]), 7)),
        "manifolds": st["manifolds"], "points": st["points"], "colours": st["colours"],
        "awake_bodies": st["awake_bodies"],
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
