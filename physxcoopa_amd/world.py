"""Python mirror of physxcoopa's PhysicsWorld over the pxg C ABI.

`World` is the thin handle wrapper (scene arrays in, arrays out). `PhysicsWorld` mirrors the
reference's interface (physxcoopa/world.h): add_body/remove_body/get_body with generational
BodyIds, step(dt)/step_fixed(h), manifolds(), config(), layers(), set_gravity(), on_substep
callbacks between the collision and solve phases, and world_state_hash(). All compute runs in
libpxg.so on the GPU; there is no CPU fallback.
"""
import ctypes
import sys

import numpy as np

from . import _abi
from ._abi import BODY_DTYPE, JOINT_DTYPE, SHAPE_DTYPE, MANIFOLD_DTYPE, EVENT_DTYPE, Config, Stats, Profile, ptr
from ._lib import lib, check


class World:
    """One GPU world handle built from ABI arrays (bodies, shapes)."""

    def __init__(self, config, bodies, shapes, layer_masks=None, hulls=None, meshes=None, cloths=None,
                 strands=None):
        self._cfg = config
        h = ctypes.c_void_p()
        check(lib().pxg_create(ctypes.byref(config), ctypes.byref(h)), "pxg_create")
        self.h = h
        self.n_bodies = len(bodies)
        self.n_shapes = len(shapes)
        b = np.ascontiguousarray(bodies, dtype=BODY_DTYPE)
        s = np.ascontiguousarray(shapes, dtype=SHAPE_DTYPE)
        check(lib().pxg_set_bodies(self.h, ptr(b), self.n_bodies), "pxg_set_bodies")
        if layer_masks is not None:
            m = np.ascontiguousarray(np.asarray(layer_masks, np.uint32))
            check(lib().pxg_set_layer_masks(self.h, ptr(m)), "pxg_set_layer_masks")
        if hulls:
            self.set_hulls(hulls)
        if meshes:
            self.set_meshes(meshes)
        check(lib().pxg_set_shapes(self.h, ptr(s), self.n_shapes), "pxg_set_shapes")
        if cloths:
            self.set_cloths(cloths)
        if strands:
            self.set_strand_sets(strands)

    def close(self):
        if getattr(self, "h", None):
            lib().pxg_destroy(self.h)
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

    @property
    def config(self):
        return self._cfg

    def set_config(self, config):
        self._cfg = config
        check(lib().pxg_set_config(self.h, ctypes.byref(config)), "pxg_set_config")

    def set_layer_masks(self, masks):
        m = np.ascontiguousarray(np.asarray(masks, np.uint32))
        check(lib().pxg_set_layer_masks(self.h, ptr(m)), "pxg_set_layer_masks")

    def set_hulls(self, hulls):
        """Uploads the convex hulls (a list of _abi.Hull) that PXG_CONVEX_HULL shapes index."""
        tab = _abi.hull_table(hulls)
        self._hulls = list(hulls)  # the arrays the records point at stay alive through the call
        check(lib().pxg_set_hulls(self.h, ctypes.byref(tab), len(hulls)), "pxg_set_hulls")

    def set_meshes(self, meshes):
        """Cooks and uploads the triangle meshes (a list of _abi.TriangleMesh) that
        PXG_TRIANGLE_MESH shapes index."""
        tab = _abi.mesh_table(meshes)
        self._meshes = list(meshes)
        check(lib().pxg_set_meshes(self.h, ctypes.byref(tab), len(meshes)), "pxg_set_meshes")

    def set_cloths(self, cloths):
        """Replaces the cloth table (a list of _abi.Cloth; slot i is cloths[i])."""
        tab = _abi.cloth_table(cloths)
        self._cloths = [c.copy() for c in cloths]  # the shapes read back into
        check(lib().pxg_set_cloths(self.h, ctypes.byref(tab), len(cloths)), "pxg_set_cloths")

    def cloth(self, index):
        """Cloth slot `index` as uploaded, with its particles, multipliers, anchors and sleep state
        read back."""
        c = self._cloths[index].copy()
        d = c.desc()
        check(lib().pxg_get_cloth(self.h, index, ctypes.byref(d)), "pxg_get_cloth")
        c.take_state(d)
        return c

    def set_strand_sets(self, sets):
        """Replaces the hair table (a list of _abi.StrandSet)."""
        tab = _abi.strand_table(sets)
        self._strands = [x.copy() for x in sets]
        check(lib().pxg_set_strand_sets(self.h, ctypes.byref(tab), len(sets)), "pxg_set_strand_sets")

    def strand_set(self, index):
        s = self._strands[index].copy()
        d = s.desc()
        check(lib().pxg_get_strand_set(self.h, index, ctypes.byref(d)), "pxg_get_strand_set")
        s.take_state(d)
        return s

    def set_strand_roots(self, index, positions, orientations=None):
        """hair::set_strand_roots on set `index` (positions n x 3, orientations n x 4 xyzw)."""
        p = np.ascontiguousarray(np.asarray(positions, np.float32).reshape(-1, 3))
        q = None if orientations is None else np.ascontiguousarray(np.asarray(orientations, np.float32).reshape(-1, 4))
        check(lib().pxg_set_strand_roots(self.h, index, ptr(p), None if q is None else ptr(q), len(p)),
              "pxg_set_strand_roots")

    def set_joints(self, joints):
        """Replaces the joint table (pxg_joint records, see physxcoopa_amd.joints)."""
        j = np.ascontiguousarray(joints, dtype=JOINT_DTYPE)
        self.n_joints = len(j)
        check(lib().pxg_set_joints(self.h, ptr(j), len(j)), "pxg_set_joints")

    def joints(self):
        out = np.zeros(getattr(self, "n_joints", 0), JOINT_DTYPE)
        check(lib().pxg_get_joints(self.h, ptr(out), len(out)), "pxg_get_joints")
        return out

    def set_ignored_pairs(self, ab):
        a = np.ascontiguousarray(np.asarray(ab, np.uint32).reshape(-1, 2))
        check(lib().pxg_set_ignored_pairs(self.h, ptr(a), len(a)), "pxg_set_ignored_pairs")

    def step_fixed(self, h=None):
        check(lib().pxg_step_fixed(self.h, np.float32(self._cfg.fixed_dt if h is None else h)), "pxg_step_fixed")

    def step_phase(self, phase, h=None):
        check(lib().pxg_step_phase(self.h, phase, np.float32(self._cfg.fixed_dt if h is None else h)), "pxg_step_phase")

    def step(self, dt):
        n = ctypes.c_uint32()
        check(lib().pxg_step(self.h, np.float32(dt), ctypes.byref(n)), "pxg_step")
        return n.value

    def synchronize(self):
        check(lib().pxg_synchronize(self.h), "pxg_synchronize")

    def stats(self):
        s = Stats()
        check(lib().pxg_get_stats(self.h, ctypes.byref(s)), "pxg_get_stats")
        return {f: getattr(s, f) for f, _ in Stats._fields_ if not f.startswith("reserved")}

    def bodies(self):
        out = np.zeros(self.n_bodies, BODY_DTYPE)
        check(lib().pxg_get_bodies(self.h, ptr(out), self.n_bodies), "pxg_get_bodies")
        return out

    def set_bodies(self, bodies):
        b = np.ascontiguousarray(bodies, dtype=BODY_DTYPE)
        check(lib().pxg_set_bodies(self.h, ptr(b), len(b)), "pxg_set_bodies")
        self.n_bodies = len(b)

    def set_shapes(self, shapes):
        """Replaces the shape table; proxies are created afresh (world.h:1369-1382)."""
        s = np.ascontiguousarray(shapes, dtype=SHAPE_DTYPE)
        check(lib().pxg_set_shapes(self.h, ptr(s), len(s)), "pxg_set_shapes")
        self.n_shapes = len(s)

    def pairs(self):
        n = ctypes.c_uint32()
        check(lib().pxg_get_pairs(self.h, None, 0, ctypes.byref(n)), "pxg_get_pairs")
        out = np.zeros((max(n.value, 1), 2), np.uint32)
        check(lib().pxg_get_pairs(self.h, ptr(out), n.value, ctypes.byref(n)), "pxg_get_pairs")
        return out[:n.value]

    def manifolds(self):
        n = ctypes.c_uint32()
        check(lib().pxg_get_manifolds(self.h, None, 0, ctypes.byref(n)), "pxg_get_manifolds")
        out = np.zeros(max(n.value, 1), MANIFOLD_DTYPE)
        check(lib().pxg_get_manifolds(self.h, ptr(out), n.value, ctypes.byref(n)), "pxg_get_manifolds")
        return out[:n.value]

    def events(self):
        """Contact events queued since the last clear (world.h:744): a structured array of
        EVENT_DTYPE (a, b owner bodies; type 0 enter, 1 stay, 2 exit; is_trigger). Raises when
        the queue overflowed since it was last cleared."""
        n = ctypes.c_uint32()
        check(lib().pxg_get_events(self.h, None, 0, ctypes.byref(n)), "pxg_get_events")
        out = np.zeros(max(n.value, 1), EVENT_DTYPE)
        check(lib().pxg_get_events(self.h, ptr(out), n.value, ctypes.byref(n)), "pxg_get_events")
        return out[:n.value]

    def clear_events(self):
        check(lib().pxg_clear_events(self.h), "pxg_clear_events")

    def raycast(self, rays, query_filter=None):
        """Closest hit of each ray (RAY_DTYPE in, HIT_DTYPE out; world.h:766-791)."""
        r = np.ascontiguousarray(rays, dtype=_abi.RAY_DTYPE).reshape(-1)
        out = np.zeros(max(len(r), 1), _abi.HIT_DTYPE)
        f = None if query_filter is None else ptr(np.ascontiguousarray(query_filter, dtype=_abi.FILTER_DTYPE))
        check(lib().pxg_raycast(self.h, ptr(r), len(r), f, ptr(out)), "pxg_raycast")
        return out[:len(r)]

    def raycast_any(self, rays, query_filter=None):
        """Whether each ray hits anything (world.h:803-817)."""
        r = np.ascontiguousarray(rays, dtype=_abi.RAY_DTYPE).reshape(-1)
        out = np.zeros(max(len(r), 1), np.uint32)
        f = None if query_filter is None else ptr(np.ascontiguousarray(query_filter, dtype=_abi.FILTER_DTYPE))
        check(lib().pxg_raycast_any(self.h, ptr(r), len(r), f, ptr(out)), "pxg_raycast_any")
        return out[:len(r)].astype(bool)

    def shape_cast(self, cast, query_filter=None):
        """The first body a swept sphere, box or capsule touches (CAST_DTYPE in; world.h:877-950): a
        HIT_DTYPE record, hit = 0 when none."""
        c = np.ascontiguousarray(cast, dtype=_abi.CAST_DTYPE).reshape(1)
        f = None if query_filter is None else ptr(np.ascontiguousarray(query_filter, dtype=_abi.FILTER_DTYPE))
        out = np.zeros(1, _abi.HIT_DTYPE)
        check(lib().pxg_shape_cast(self.h, ptr(c), f, ptr(out)), "pxg_shape_cast")
        return out[0]

    def overlap_batch(self, queries, query_filter=None, cap=64):
        """Many overlap queries in one GPU pass (an OVERLAP_DTYPE array in): (counts, bodies), where
        bodies[i, :min(cap, counts[i])] lists what overlap(queries[i]) returns."""
        q = np.ascontiguousarray(queries, dtype=_abi.OVERLAP_DTYPE).reshape(-1)
        f = None if query_filter is None else ptr(np.ascontiguousarray(query_filter, dtype=_abi.FILTER_DTYPE))
        counts = np.zeros(len(q), np.uint32)
        bodies = np.zeros((len(q), cap), np.uint32)
        if len(q):
            check(lib().pxg_overlap_batch(self.h, ptr(q), len(q), f, ptr(bodies), cap, ptr(counts)), "pxg_overlap_batch")
        return counts, bodies

    def shape_cast_batch(self, casts, query_filter=None):
        """Many shape casts in one GPU pass (a CAST_DTYPE array in): a HIT_DTYPE array, entry i the
        same record shape_cast(casts[i]) returns."""
        c = np.ascontiguousarray(casts, dtype=_abi.CAST_DTYPE).reshape(-1)
        f = None if query_filter is None else ptr(np.ascontiguousarray(query_filter, dtype=_abi.FILTER_DTYPE))
        out = np.zeros(len(c), _abi.HIT_DTYPE)
        if len(c):
            check(lib().pxg_shape_cast_batch(self.h, ptr(c), len(c), f, ptr(out)), "pxg_shape_cast_batch")
        return out

    def compute_penetration_batch(self, probes, query_filter=None, cap=16):
        """Many compute_penetration queries in one GPU pass (a CAST_DTYPE array of posed probes in):
        (counts, records), where records[i, :min(cap, counts[i])] equals compute_penetration(probes[i])."""
        c = np.ascontiguousarray(probes, dtype=_abi.CAST_DTYPE).reshape(-1)
        f = None if query_filter is None else ptr(np.ascontiguousarray(query_filter, dtype=_abi.FILTER_DTYPE))
        counts = np.zeros(len(c), np.uint32)
        out = np.zeros((len(c), cap), _abi.PEN_DTYPE)
        if len(c):
            check(lib().pxg_compute_penetration_batch(self.h, ptr(c), len(c), f, ptr(out), cap, ptr(counts)),
                  "pxg_compute_penetration_batch")
        return counts, out

    def compute_penetration(self, probe, query_filter=None):
        """Every shape the posed query shape penetrates, with its push-out (world.h:963-995): the
        shape, position and orientation of a CAST_DTYPE record (direction and max_distance unused)
        in, a PEN_DTYPE array in shape-slot order out."""
        sh, pos, rot = _abi.probe_parts(np.asarray(probe, dtype=_abi.CAST_DTYPE))
        f = None if query_filter is None else ptr(np.ascontiguousarray(query_filter, dtype=_abi.FILTER_DTYPE))
        n = ctypes.c_uint32()
        check(lib().pxg_compute_penetration(self.h, ptr(sh), ptr(pos), ptr(rot), f, None, 0, ctypes.byref(n)),
              "pxg_compute_penetration")
        out = np.zeros(max(n.value, 1), _abi.PEN_DTYPE)
        check(lib().pxg_compute_penetration(self.h, ptr(sh), ptr(pos), ptr(rot), f, ptr(out), n.value, ctypes.byref(n)),
              "pxg_compute_penetration")
        return out[:n.value]

    def overlap(self, query, query_filter=None):
        """Bodies with a shape overlapping the query shape (OVERLAP_DTYPE; world.h:1002-1060), once
        per overlapping shape, in shape-slot order."""
        q = np.ascontiguousarray(query, dtype=_abi.OVERLAP_DTYPE).reshape(1)
        f = None if query_filter is None else ptr(np.ascontiguousarray(query_filter, dtype=_abi.FILTER_DTYPE))
        n = ctypes.c_uint32()
        check(lib().pxg_overlap(self.h, ptr(q), f, None, 0, ctypes.byref(n)), "pxg_overlap")
        out = np.zeros(max(n.value, 1), np.uint32)
        check(lib().pxg_overlap(self.h, ptr(q), f, ptr(out), n.value, ctypes.byref(n)), "pxg_overlap")
        return out[:n.value]

    def raycast_all(self, ray, query_filter=None):
        """Every hit of one ray, nearest first (world.h:825-847)."""
        r = np.ascontiguousarray(ray, dtype=_abi.RAY_DTYPE).reshape(1)
        f = None if query_filter is None else ptr(np.ascontiguousarray(query_filter, dtype=_abi.FILTER_DTYPE))
        n = ctypes.c_uint32()
        check(lib().pxg_raycast_all(self.h, ptr(r), f, None, 0, ctypes.byref(n)), "pxg_raycast_all")
        out = np.zeros(max(n.value, 1), _abi.HIT_DTYPE)
        check(lib().pxg_raycast_all(self.h, ptr(r), f, ptr(out), n.value, ctypes.byref(n)), "pxg_raycast_all")
        return out[:n.value]

    def set_events(self, enabled, capacity=0):
        check(lib().pxg_set_events(self.h, 1 if enabled else 0, capacity), "pxg_set_events")

    def solve_order(self):
        """(order, colour_offsets): manifold indices in solve order and the colour boundaries."""
        n, nc = ctypes.c_uint32(), ctypes.c_uint32()
        check(lib().pxg_get_solve_order(self.h, None, 0, ctypes.byref(n), None, 0, ctypes.byref(nc)), "order")
        order = np.zeros(max(n.value, 1), np.uint32)
        offs = np.zeros(nc.value + 1, np.uint32)
        check(lib().pxg_get_solve_order(self.h, ptr(order), n.value, ctypes.byref(n), ptr(offs), nc.value + 1,
                                        ctypes.byref(nc)), "pxg_get_solve_order")
        return order[:n.value], offs

    def fat_aabbs(self):
        out = np.zeros((self.n_shapes, 6), np.float32)
        check(lib().pxg_get_fat_aabbs(self.h, ptr(out), self.n_shapes), "pxg_get_fat_aabbs")
        return out

    def set_fat_aabbs(self, boxes):
        b = np.ascontiguousarray(np.asarray(boxes, np.float32).reshape(-1, 6))
        check(lib().pxg_set_fat_aabbs(self.h, ptr(b), len(b)), "pxg_set_fat_aabbs")

    def warm_cache(self):
        n = ctypes.c_uint32()
        check(lib().pxg_get_warm_cache(self.h, None, None, 0, ctypes.byref(n)), "pxg_get_warm_cache")
        keys = np.zeros(max(n.value, 1), np.uint64)
        imp = np.zeros((max(n.value, 1), 3), np.float32)
        check(lib().pxg_get_warm_cache(self.h, ptr(keys), ptr(imp), n.value, ctypes.byref(n)), "pxg_get_warm_cache")
        return keys[:n.value], imp[:n.value]

    def set_warm_cache(self, keys, imp):
        k = np.ascontiguousarray(np.asarray(keys, np.uint64))
        i = np.ascontiguousarray(np.asarray(imp, np.float32).reshape(-1, 3))
        check(lib().pxg_set_warm_cache(self.h, ptr(k), ptr(i), len(k)), "pxg_set_warm_cache")

    def state_hash(self):
        out = ctypes.c_uint64()
        check(lib().pxg_state_hash(self.h, ctypes.byref(out)), "pxg_state_hash")
        return out.value

    def set_mlp(self, net):
        """Turns the learned warm start on with `net` (a learned.MlpParams) or off with None."""
        check(lib().pxg_set_mlp(self.h, ctypes.byref(net) if net is not None else None), "pxg_set_mlp")

    def set_recording(self, on):
        check(lib().pxg_set_recording(self.h, 1 if on else 0), "pxg_set_recording")

    def records(self):
        """(features [n, 24], final impulses [n, 3], missed [n]) of the last step's active points."""
        n = ctypes.c_uint32()
        st = self.stats()
        cap = 4 * max(1, st["active_manifolds"])
        f = np.zeros((cap, 24), np.float32)
        y = np.zeros((cap, 3), np.float32)
        m = np.zeros(cap, np.uint32)
        check(lib().pxg_get_records(self.h, ptr(f), ptr(y), ptr(m), cap, ctypes.byref(n)), "pxg_get_records")
        return f[:n.value], y[:n.value], m[:n.value]

    def predictions(self):
        """(keys, clamped impulses [n, 3]) the learned warm start applied in the last step."""
        n = ctypes.c_uint32()
        check(lib().pxg_get_predictions(self.h, None, None, 0, ctypes.byref(n)), "pxg_get_predictions")
        k = np.zeros(max(n.value, 1), np.uint64)
        i = np.zeros((max(n.value, 1), 3), np.float32)
        check(lib().pxg_get_predictions(self.h, ptr(k), ptr(i), n.value, ctypes.byref(n)), "pxg_get_predictions")
        return k[:n.value], i[:n.value]

    def set_profiling(self, on):
        check(lib().pxg_set_profiling(self.h, 1 if on else 0), "pxg_set_profiling")

    def profile(self):
        p = Profile()
        check(lib().pxg_get_profile(self.h, ctypes.byref(p)), "pxg_get_profile")
        return dict(velocity_ms=p.velocity_ms, velocity_bytes=p.velocity_bytes, velocity_launches=p.velocity_launches,
                    tiled_ticks=dict(colours=p.tiled_colour_ticks, staging=p.tiled_stage_ticks,
                                     barriers=p.tiled_barrier_ticks, residuals=p.tiled_residual_ticks))


class BodyId:
    """Generational handle (physxcoopa/dynamics/body.h:61-77)."""
    INVALID = 0xFFFFFFFF

    def __init__(self, index=INVALID, generation=0):
        self.index, self.generation = index, generation

    def is_valid(self):
        return self.index != BodyId.INVALID

    def __eq__(self, o):
        return isinstance(o, BodyId) and (self.index, self.generation) == (o.index, o.generation)

    def __hash__(self):
        return hash((self.index, self.generation))

    def __repr__(self):
        return "BodyId(%d, %d)" % (self.index, self.generation)


class LayerMatrix:
    """32x32 symmetric collision matrix (physxcoopa/broadphase/layer_matrix.h:21-47)."""

    def __init__(self):
        self.mask = np.full(32, 0xFFFFFFFF, np.uint32)

    def should_collide(self, a, b):
        return bool(self.mask[a & 31] & np.uint32(1 << (b & 31)))

    def set_layer_collision(self, a, b, collide):
        a, b = a & 31, b & 31
        if collide:
            self.mask[a] |= np.uint32(1 << b)
            self.mask[b] |= np.uint32(1 << a)
        else:
            self.mask[a] &= np.uint32(~(1 << b) & 0xFFFFFFFF)
            self.mask[b] &= np.uint32(~(1 << a) & 0xFFFFFFFF)


class PhysicsWorld:
    """The reference's PhysicsWorld interface (physxcoopa/world.h) over the GPU step.

    Bodies are ABI records (numpy scalars of BODY_DTYPE); a shape is a SHAPE_DTYPE record or
    None ("Rigidbody with no Collider"). Structural edits are staged on the host and uploaded
    before the next step; body state lives on the GPU between steps.
    """

    def __init__(self, config=None):
        self._cfg = config if config is not None else _abi.default_config()
        self._bodies = np.zeros(0, BODY_DTYPE)
        self._shapes = np.zeros(0, SHAPE_DTYPE)
        self._generations = []
        self._free = []
        self._layers = LayerMatrix()
        self._world = None
        self._dirty = True
        self._state_on_device = False
        self.on_substep = []
        self._accumulator = 0.0

    # -- lifecycle (world.h:95-143) --
    def add_body(self, body, shape=None):
        self._pull()
        b = np.array(body, dtype=BODY_DTYPE).reshape(())
        b["alive"] = 1
        b["prev_position"] = b["position"]
        b["prev_orientation"] = b["orientation"]
        if self._free:
            idx = self._free.pop()
            self._bodies[idx] = b
        else:
            idx = len(self._bodies)
            self._bodies = np.concatenate([self._bodies, b.reshape(1)])
            self._generations.append(0)
        s = np.zeros((), SHAPE_DTYPE) if shape is None else np.array(shape, dtype=SHAPE_DTYPE).reshape(())
        s["owner"] = idx
        if idx < len(self._shapes):
            self._shapes[idx] = s
        else:
            self._shapes = np.concatenate([self._shapes, s.reshape(1)])
        self._dirty = True
        return BodyId(idx, self._generations[idx])

    def remove_body(self, bid):
        if not self.is_valid(bid):
            return
        self._pull()
        self._bodies[bid.index]["alive"] = 0
        self._shapes[bid.index]["enabled"] = 0
        self._generations[bid.index] += 1
        self._free.append(bid.index)
        self._dirty = True

    def is_valid(self, bid):
        return (bid.is_valid() and bid.index < len(self._bodies) and bool(self._bodies[bid.index]["alive"])
                and self._generations[bid.index] == bid.generation)

    def get_body(self, bid):
        """A copy of the body's current state, or None for a stale handle (world.h:347)."""
        if not self.is_valid(bid):
            return None
        self._pull()
        return self._bodies[bid.index].copy()

    def set_body(self, bid, body):
        if not self.is_valid(bid):
            return
        self._pull()
        self._bodies[bid.index] = body
        self._dirty = True

    def config(self):
        return self._cfg

    def layers(self):
        self._dirty = True
        return self._layers

    def set_gravity(self, g):
        for i in range(3):
            self._cfg.gravity[i] = g[i]
        if self._world is not None:
            self._world.set_config(self._cfg)

    # -- stepping (world.h:1229-1317) --
    def _sync(self):
        if self._world is None or self._dirty:
            self._pull()
            if self._world is not None:
                fat = self._world.fat_aabbs() if self._world.n_shapes == len(self._shapes) else None
                keys, imp = self._world.warm_cache()
                self._world.close()
            else:
                fat, keys, imp = None, None, None
            self._world = World(self._cfg, self._bodies, self._shapes, self._layers.mask)
            if fat is not None:
                self._world.set_fat_aabbs(fat)
            if keys is not None and len(keys):
                self._world.set_warm_cache(keys, imp)
            self._dirty = False
        self._state_on_device = True

    def _pull(self):
        if self._world is not None and self._state_on_device:
            self._bodies = self._world.bodies()
            self._state_on_device = False

    def step_fixed(self, h=None):
        h = self._cfg.fixed_dt if h is None else h
        self._sync()
        if self.on_substep:
            self._world.step_phase(_abi.PXG_PHASE_FORCES_COLLIDE, h)
            self._bodies = self._world.bodies()
            before = self._bodies.copy()
            for cb in self.on_substep:
                cb(self, h)
            if not np.array_equal(before.view(np.uint8), self._bodies.view(np.uint8)):
                self._world.set_bodies(self._bodies)
            self._world.step_phase(_abi.PXG_PHASE_SOLVE_INTEGRATE, h)
        else:
            self._world.step_fixed(h)
        self._state_on_device = True

    def step(self, dt):
        frame = min(dt, 0.25)
        self._accumulator += frame
        h = self._cfg.fixed_dt
        n = 0
        while self._accumulator >= h and n < self._cfg.max_substeps:
            self.step_fixed(h)
            self._accumulator -= h
            n += 1
        return n

    def manifolds(self):
        return self._world.manifolds() if self._world is not None else np.zeros(0, MANIFOLD_DTYPE)

    def world_state_hash(self):
        self._sync()
        return self._world.state_hash()

    def stats(self):
        return self._world.stats() if self._world is not None else {}


def cook_mesh(mesh):
    """pxg_cook_mesh of an _abi.TriangleMesh on the host: a dict of the welded vertices (n, 3),
    indices (t, 3), unit face normals (t, 3), adjacency (t, 3; 0xFFFFFFFF for none), bounds (6)
    and the non-manifold edge count (TriangleMesh::from_triangles, triangle_mesh.h:75-170)."""
    d = mesh.desc()
    ni = len(mesh.indices)
    info = _abi.MeshInfo()
    v = np.zeros((max(len(mesh.positions), 1), 3), np.float32)
    idx = np.zeros(max(ni, 3), np.uint32)
    nrm = np.zeros(max(ni, 3), np.float32)
    adj = np.zeros(max(ni, 3), np.uint32)
    check(lib().pxg_cook_mesh(ctypes.byref(d), ctypes.byref(info), ptr(v), ptr(idx), ptr(nrm), ptr(adj)), "pxg_cook_mesh")
    nv, nt = info.n_vertices, info.n_triangles
    return {"vertices": v[:nv].copy(), "indices": idx[:3 * nt].reshape(-1, 3).copy(),
            "normals": nrm[:3 * nt].reshape(-1, 3).copy(), "adjacency": adj[:3 * nt].reshape(-1, 3).copy(),
            "bounds": np.array(list(info.local_min) + list(info.local_max), np.float32),
            "non_manifold_edges": info.non_manifold_edges, "n_nodes": info.n_nodes}
