"""Batched stepping of many independent scenes (pxg_batch_* in include/pxg.h).

Each scene behaves like its own PhysicsWorld stepped with step_fixed(h); one `step(k)` call
advances every scene k times in one persistent, graph-captured launch.
"""
import ctypes
import sys

import numpy as np

from ._abi import BODY_DTYPE, JOINT_DTYPE, SHAPE_DTYPE, MANIFOLD_DTYPE, BATCH_STATS_DTYPE, ptr
from ._lib import lib, check_batch


class Batch:
    def __init__(self, config, bodies, shapes, body_offsets, shape_offsets, layer_masks=None):
        self._cfg = config
        self.body_offsets = np.ascontiguousarray(np.asarray(body_offsets, np.uint32))
        self.shape_offsets = np.ascontiguousarray(np.asarray(shape_offsets, np.uint32))
        self.n_scenes = len(self.body_offsets) - 1
        b = np.ascontiguousarray(bodies, dtype=BODY_DTYPE)
        s = np.ascontiguousarray(shapes, dtype=SHAPE_DTYPE)
        self.n_bodies = len(b)
        h = ctypes.c_void_p()
        check_batch(lib().pxg_batch_create(ctypes.byref(config), self.n_scenes, ptr(self.body_offsets),
                                           ptr(self.shape_offsets), ptr(b), ptr(s), ctypes.byref(h)), "pxg_batch_create")
        self.h = h
        if layer_masks is not None:
            m = np.ascontiguousarray(np.asarray(layer_masks, np.uint32))
            check_batch(lib().pxg_batch_set_layer_masks(self.h, ptr(m)), "pxg_batch_set_layer_masks")
        self.tuning = (256, True)

    def close(self):
        if getattr(self, "h", None):
            lib().pxg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        # At interpreter exit the HIP runtime may already be torn down (and a profiler's exit
        # hooks may be running): destroying the handle then would call into a dead runtime.
        # The process is ending and the driver reclaims the device memory, so skip it.
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def set_joints(self, joint_offsets, joints):
        """Scene k owns joints[joint_offsets[k]:joint_offsets[k+1]] (scene-local body indices)."""
        self.joint_offsets = np.ascontiguousarray(np.asarray(joint_offsets, np.uint32))
        j = np.ascontiguousarray(joints, dtype=JOINT_DTYPE)
        self.n_joints = len(j)
        check_batch(lib().pxg_batch_set_joints(self.h, ptr(self.joint_offsets), ptr(j)), "pxg_batch_set_joints")

    def joints(self):
        out = np.zeros(getattr(self, "n_joints", 0), JOINT_DTYPE)
        check_batch(lib().pxg_batch_get_joints(self.h, ptr(out), len(out)), "pxg_batch_get_joints")
        return out

    def set_ignored_pairs(self, pair_offsets, ab):
        o = np.ascontiguousarray(np.asarray(pair_offsets, np.uint32))
        a = np.ascontiguousarray(np.asarray(ab, np.uint32).reshape(-1, 2))
        check_batch(lib().pxg_batch_set_ignored_pairs(self.h, ptr(o), ptr(a)), "pxg_batch_set_ignored_pairs")

    def set_mlp(self, net):
        """The learned warm start in every scene with `net` (a learned.MlpParams), or off with None."""
        check_batch(lib().pxg_batch_set_mlp(self.h, ctypes.byref(net) if net is not None else None), "pxg_batch_set_mlp")

    def predictions(self, scene):
        """(keys, clamped impulses [n, 3]) the learned warm start applied in `scene` in the last step."""
        n = ctypes.c_uint32()
        check_batch(lib().pxg_batch_get_predictions(self.h, scene, None, None, 0, ctypes.byref(n)), "predictions")
        k = np.zeros(max(n.value, 1), np.uint64)
        i = np.zeros((max(n.value, 1), 3), np.float32)
        check_batch(lib().pxg_batch_get_predictions(self.h, scene, ptr(k), ptr(i), n.value, ctypes.byref(n)),
                    "pxg_batch_get_predictions")
        return k[:n.value], i[:n.value]

    def set_tuning(self, block_size, longest_first=True):
        check_batch(lib().pxg_batch_set_tuning(self.h, block_size, 1 if longest_first else 0), "pxg_batch_set_tuning")
        self.tuning = (block_size, bool(longest_first))

    def set_occupancy(self, workgroups_per_cu):
        """Workgroups launched per compute unit (0: as many as fit)."""
        check_batch(lib().pxg_batch_set_occupancy(self.h, int(workgroups_per_cu)), "pxg_batch_set_occupancy")
        self.occupancy = int(workgroups_per_cu)

    def step(self, steps=1, h=None):
        check_batch(lib().pxg_batch_step(self.h, np.float32(self._cfg.fixed_dt if h is None else h), steps),
                    "pxg_batch_step")

    def synchronize(self):
        check_batch(lib().pxg_batch_synchronize(self.h), "pxg_batch_synchronize")

    def bodies(self):
        out = np.zeros(self.n_bodies, BODY_DTYPE)
        check_batch(lib().pxg_batch_get_bodies(self.h, ptr(out), self.n_bodies), "pxg_batch_get_bodies")
        return out

    def scene_bodies(self, k, all_bodies=None):
        b = self.bodies() if all_bodies is None else all_bodies
        return b[self.body_offsets[k]:self.body_offsets[k + 1]]

    def stats(self):
        out = np.zeros(self.n_scenes, BATCH_STATS_DTYPE)
        check_batch(lib().pxg_batch_get_stats(self.h, ptr(out), self.n_scenes), "pxg_batch_get_stats")
        return out

    def manifolds(self, scene):
        n = ctypes.c_uint32()
        check_batch(lib().pxg_batch_get_manifolds(self.h, scene, None, 0, ctypes.byref(n)), "pxg_batch_get_manifolds")
        out = np.zeros(max(n.value, 1), MANIFOLD_DTYPE)
        check_batch(lib().pxg_batch_get_manifolds(self.h, scene, ptr(out), n.value, ctypes.byref(n)),
                    "pxg_batch_get_manifolds")
        return out[:n.value]

    def solve_order(self, scene):
        n, nc = ctypes.c_uint32(), ctypes.c_uint32()
        check_batch(lib().pxg_batch_get_solve_order(self.h, scene, None, 0, ctypes.byref(n), None, 0,
                                                    ctypes.byref(nc)), "order")
        order = np.zeros(max(n.value, 1), np.uint32)
        offs = np.zeros(nc.value + 1, np.uint32)
        check_batch(lib().pxg_batch_get_solve_order(self.h, scene, ptr(order), n.value, ctypes.byref(n), ptr(offs),
                                                    nc.value + 1, ctypes.byref(nc)), "pxg_batch_get_solve_order")
        return order[:n.value], offs
