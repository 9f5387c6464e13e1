"""Autotuner for the batched step: picks the workgroup size (threads per scene) and the scene
queue order by timing a short run of each candidate on a copy of the batch. The choice is cached
per batch shape (scene count, body count bucket) in a JSON file."""
import json
import os
import time

from .batch import Batch

CANDIDATES = [(64, True), (128, True), (256, True), (512, True), (256, False)]
CACHE = os.environ.get("PXG_TUNING_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "physxcoopa_amd",
                                                        "batch_tuning.json"))


def shape_key(n_scenes, n_bodies, config):
    return "%d:%d:%d:%d:%d" % (n_scenes, n_bodies // 1024, config.velocity_iterations, config.position_iterations,
                               config.residual_mode)


def _load():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save(d):
    try:
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        with open(CACHE, "w") as f:
            json.dump(d, f, indent=1)
    except OSError:
        pass


def tune(config, bodies, shapes, body_offsets, shape_offsets, steps=8, use_cache=True, log=None, setup=None):
    """Returns ((block_size, longest_first), {candidate: seconds}) for this batch shape. `setup`,
    if given, is called on each candidate batch before timing (joints, ignored pairs, network)."""
    key = shape_key(len(body_offsets) - 1, len(bodies), config) + (":setup" if setup else "")
    cache = _load() if use_cache else {}
    if key in cache:
        c = cache[key]
        return (c["block"], c["longest_first"]), c.get("timings", {})
    timings = {}
    for block, lpt in CANDIDATES:
        b = Batch(config, bodies, shapes, body_offsets, shape_offsets)
        if setup:
            setup(b)
        b.set_tuning(block, lpt)
        b.step(1)          # graph capture + first touch
        b.synchronize()
        t0 = time.perf_counter()
        b.step(steps)
        b.synchronize()
        dt = time.perf_counter() - t0
        timings["%d/%s" % (block, "lpt" if lpt else "natural")] = dt
        if log:
            log("autotune block %d %s: %.3f s for %d steps" % (block, "lpt" if lpt else "natural", dt, steps))
        b.close()
    best = min(CANDIDATES, key=lambda c: timings["%d/%s" % (c[0], "lpt" if c[1] else "natural")])
    cache[key] = {"block": best[0], "longest_first": best[1], "timings": timings}
    _save(cache)
    return best, timings
