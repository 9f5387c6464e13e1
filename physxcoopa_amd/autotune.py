"""Autotuner of the batched step (pxg_batch_*): picks, by timing the batch itself, the launch
parameters that change only the speed, never the results:

* colour-batch width: threads per scene workgroup (64-512), i.e. how many manifolds of one colour
  the workgroup solves between two of its barriers, and how many pairs / contacts each phase of
  the step covers per sweep;
* scene tile: workgroups per compute unit (1, or as many as the kernel's LDS and registers allow),
  each taking one scene at a time from the queue;
* queue order: scenes by decreasing size (longest processing time first) or in batch order;
* steps per launch: how many steps one replay of the captured graph advances.

The search is coordinate descent from the default (256 threads, all that fit, longest first, 30
steps per launch): each parameter in turn is timed at every value with the others at their best
so far, over two rounds. The candidates of one parameter are timed interleaved on one batch (the
scenes keep stepping), best of two timings each, so a drift of the scenes' state between timings
affects every candidate alike. The choice is cached per batch shape in a JSON file.
"""
import json
import os
import time

from .batch import Batch

SPACE = {
    "block": [64, 128, 256, 512],
    "workgroups_per_cu": [0, 1],
    "longest_first": [True, False],
    "steps_per_launch": [10, 30],
}
DEFAULT = {"block": 256, "workgroups_per_cu": 0, "longest_first": True, "steps_per_launch": 30}
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


def apply(batch, choice):
    """Sets a tuning choice on a batch (steps_per_launch is the caller's: how it calls step())."""
    batch.set_tuning(choice["block"], choice["longest_first"])
    batch.set_occupancy(choice["workgroups_per_cu"])


def label(choice):
    return "block %d, %s workgroups/CU, %s, %d steps/launch" % (
        choice["block"], "max" if choice["workgroups_per_cu"] == 0 else str(choice["workgroups_per_cu"]),
        "longest first" if choice["longest_first"] else "batch order", choice["steps_per_launch"])


def _time(batch, choice, steps):
    """Seconds per step of `choice` over at least `steps` steps (one untimed launch first: graph
    capture and instantiation)."""
    apply(batch, choice)
    k = choice["steps_per_launch"]
    batch.step(k)
    batch.synchronize()
    launches = max(1, (steps + k - 1) // k)
    t0 = time.perf_counter()
    for _ in range(launches):
        batch.step(k)
    batch.synchronize()
    return (time.perf_counter() - t0) / (launches * k)


def tune(config, bodies, shapes, body_offsets, shape_offsets, steps=30, use_cache=True, log=None, setup=None,
         rounds=2, space=None):
    """Returns (choice, timings): the tuning dict (keys of SPACE) for this batch shape and the
    seconds per step of every timed candidate. `setup`, if given, is called on the batch before
    timing (joints, ignored pairs, network)."""
    space = space or SPACE
    key = shape_key(len(body_offsets) - 1, len(bodies), config) + (":setup" if setup else "")
    cache = _load() if use_cache else {}
    if key in cache and set(cache[key].get("choice", {})) == set(DEFAULT):
        return dict(cache[key]["choice"]), cache[key].get("timings", {})
    b = Batch(config, bodies, shapes, body_offsets, shape_offsets)
    if setup:
        setup(b)
    best = dict(DEFAULT)
    timings = {}
    for rnd in range(rounds):
        changed = False
        for param, values in space.items():
            cands = [dict(best, **{param: v}) for v in values]
            t = {i: float("inf") for i in range(len(cands))}
            for _ in range(2):
                for i, c in enumerate(cands):
                    t[i] = min(t[i], _time(b, c, steps))
            for i, c in enumerate(cands):
                timings[label(c)] = t[i]
                if log:
                    log("autotune round %d, %s: %.3f ms/step" % (rnd + 1, label(c), 1e3 * t[i]))
            win = min(range(len(cands)), key=lambda i: t[i])
            if cands[win] != best:
                best = cands[win]
                changed = True
        if not changed:
            break
    b.close()
    cache[key] = {"choice": best, "timings": timings}
    if use_cache:
        _save(cache)
    return best, timings
