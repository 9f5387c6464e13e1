"""ORACLE -- test infrastructure only (tests/, __graft_entry__.smoke() and bench.py's cpu_baseline).

ctypes wrapper over oracle/_build/libpxo.so, the CPU restatement of physxcoopa's rigid-body
step (see oracle/pxo_step.hpp). It takes the same ABI arrays as the GPU library.
"""
import ctypes
import os
import subprocess

import numpy as np

from physxcoopa_amd._abi import BODY_DTYPE, JOINT_DTYPE, MANIFOLD_DTYPE, Hull, MeshInfo, hull_table, mesh_table, ptr
from physxcoopa_amd import _abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libpxo.so")
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, u32, f32, i32, u64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_float, ctypes.c_int, ctypes.c_uint64
        sig = {
            "pxo_create": ([vp, i32], vp), "pxo_destroy": ([vp], None), "pxo_set_config": ([vp, vp], None),
            "pxo_set_bodies": ([vp, vp, u32], None), "pxo_get_bodies": ([vp, vp, u32], None),
            "pxo_set_shapes": ([vp, vp, u32], None), "pxo_set_layer_masks": ([vp, vp], None),
            "pxo_set_brute_force": ([vp, u32], None), "pxo_set_threads": ([vp, u32], None),
            "pxo_set_hulls": ([vp, vp, u32], None), "pxo_set_kat_hulls": ([vp, u32], None),
            "pxo_shape_bounds": ([vp, vp, vp, vp, vp], None),
            "pxo_build_hull": ([vp, u32, f32, u32, vp, vp, vp, vp, vp, vp, vp, vp], i32), "pxo_set_order": ([vp, vp, u32], None),
            "pxo_set_miss_table": ([vp, vp, vp, u32], None), "pxo_step_fixed": ([vp, f32], None),
            "pxo_step_phase": ([vp, u32, f32], None), "pxo_run": ([vp, u32, f32], ctypes.c_double),
            "pxo_get_pairs": ([vp, vp, u32], u32), "pxo_get_manifolds": ([vp, vp, u32], u32),
            "pxo_get_fat": ([vp, vp, u32], None), "pxo_set_fat": ([vp, vp, u32], None),
            "pxo_get_cache": ([vp, vp, vp, u32], u32), "pxo_set_cache": ([vp, vp, vp, u32], None),
            "pxo_state_hash": ([vp], u64), "pxo_stats": ([vp, vp], None),
            "pxo_residual_trace": ([vp, vp, u32], u32),
            "pxo_want_residual_detail": ([vp, u32], None), "pxo_residual_detail": ([vp, vp, u32], u32),
            "pxo_eval_residual": ([vp, vp, u32, u32, f32], ctypes.c_double),
            "pxo_generate_contacts": ([vp, vp, vp, vp, vp, vp, i32, vp], i32),
            "pxo_box_box": ([vp, vp, vp, vp, vp, vp, i32, vp], i32),
            "pxo_sat": ([vp, vp, vp, vp, vp, vp, i32, vp, vp, vp, vp], i32),
            "pxo_seg_seg": ([vp] * 6, None), "pxo_world_bounds": ([vp, vp, vp, vp], None),
            "pxo_combine": ([f32, f32, i32, i32], f32), "pxo_orthonormal_basis": ([vp, vp, vp], None),
            "pxo_quat_rotate": ([vp, vp, vp], None), "pxo_mat3_cast": ([vp, vp], None),
            "pxo_pack_key": ([u32, u32, u32], u64),
            "pxo_set_joints": ([vp, vp, u32], None), "pxo_get_joints": ([vp, vp, u32], u32),
            "pxo_set_ignored_pairs": ([vp, vp, u32], None), "pxo_set_atan2_mode": ([vp, u32], None),
            "pxo_m3_inverse": ([vp, vp], None), "pxo_m3_det": ([vp], f32),
            "pxo_point_block_mass": ([vp, vp, f32, f32, vp, vp, vp], i32),
            "pxo_hinge_angle": ([vp, vp, vp, vp, u32], f32), "pxo_cone_swing": ([vp, vp, vp, u32], f32),
            "pxo_atan2": ([f32, f32, u32], f32),
            "pxo_get_events": ([vp, vp, u32], u32), "pxo_clear_events": ([vp], None),
            "pxo_raycast": ([vp, vp, u32, vp, vp], None), "pxo_raycast_all": ([vp, vp, vp, vp, u32], u32),
            "pxo_ray_shape": ([vp, vp, vp, vp, vp, vp], i32), "pxo_overlap": ([vp, vp, vp, vp, u32], u32),
            "pxo_shape_cast": ([vp, vp, vp, vp], i32), "pxo_ray_aabb": ([vp, vp, vp], i32),
            "pxo_compute_penetration": ([vp, vp, vp, vp, vp, vp, i32], i32),
            "pxo_set_meshes": ([vp, vp, u32], i32), "pxo_set_kat_meshes": ([vp, u32], i32),
            "pxo_set_cloths": ([vp, vp, u32], None), "pxo_get_cloth": ([vp, u32, vp], None),
            "pxo_set_strand_sets": ([vp, vp, u32], None), "pxo_get_strand_set": ([vp, u32, vp], None),
            "pxo_set_strand_roots": ([vp, u32, vp, vp, u32], None), "pxo_step_frame": ([vp, f32], u32),
            "pxo_interpolation_alpha": ([vp], f32), "pxo_set_collider_order": ([vp, u32], None),
            "pxo_cloth_grid": ([vp, vp, vp], vp), "pxo_cloth_pin": ([vp, vp, f32, vp, vp], u32),
            "pxo_cloth_tethers": ([vp], None), "pxo_cloth_sizes": ([vp, vp], None),
            "pxo_cloth_export": ([vp, vp, vp], None), "pxo_cloth_free": ([vp], None),
            "pxo_strands_make": ([vp, vp, u32, vp, vp, vp], vp), "pxo_strands_sizes": ([vp, vp], None),
            "pxo_strands_export": ([vp, vp], None), "pxo_strands_free": ([vp], None),
            "pxo_cook_mesh": ([vp, vp, vp, vp, vp, vp, vp, vp, vp], i32),
            "pxo_mesh_contacts": ([vp, vp, vp, vp, vp, vp, i32, vp], i32),
            "pxo_mesh_candidates": ([vp, vp, vp, vp, vp, vp, vp, u32], u32),
            "pxo_mat4_inverse": ([vp, vp], None), "pxo_closest_on_triangle": ([vp, vp, vp, vp, vp, vp], None),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def f32(x, n=None):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    if n is not None:
        assert a.size == n
    return a


class OracleWorld:
    """The CPU reference step over pxg ABI arrays. dbl=True runs the same algorithm in float64."""

    def __init__(self, config, bodies, shapes, dbl=False, brute_force=False, layer_masks=None, hulls=None, meshes=None,
                 cloths=None, strands=None):
        L = lib()
        self._cfg = config
        self.h = L.pxo_create(ctypes.byref(config), 1 if dbl else 0)
        self.n_bodies = len(bodies)
        self.n_shapes = len(shapes)
        b = np.ascontiguousarray(bodies)
        s = np.ascontiguousarray(shapes)
        L.pxo_set_bodies(self.h, ptr(b), self.n_bodies)
        if layer_masks is not None:
            m = np.ascontiguousarray(np.asarray(layer_masks, np.uint32))
            L.pxo_set_layer_masks(self.h, ptr(m))
        if hulls:
            tab = hull_table(hulls)
            L.pxo_set_hulls(self.h, ctypes.byref(tab), len(hulls))
        if meshes:
            tab = mesh_table(meshes)
            assert L.pxo_set_meshes(self.h, ctypes.byref(tab), len(meshes)) == 1, "mesh index out of range"
        L.pxo_set_shapes(self.h, ptr(s), self.n_shapes)
        if cloths:
            self.set_cloths(cloths)
        if strands:
            self.set_strand_sets(strands)
        if brute_force:
            L.pxo_set_brute_force(self.h, 1)

    def __del__(self):
        try:
            if self.h:
                lib().pxo_destroy(self.h)
        except Exception:
            pass

    def set_cloths(self, cloths):
        """Replaces the cloth table (a list of _abi.Cloth)."""
        self._cloths = [c.copy() for c in cloths]
        tab = _abi.cloth_table(self._cloths)
        lib().pxo_set_cloths(self.h, ctypes.byref(tab), len(cloths))

    def cloth(self, index):
        """Cloth slot `index` as an _abi.Cloth (particles, multipliers, anchors, sleep state)."""
        c = self._cloths[index].copy()
        d = c.desc()
        lib().pxo_get_cloth(self.h, index, ctypes.byref(d))
        c.take_state(d)
        return c

    def set_strand_sets(self, sets):
        self._strands = [x.copy() for x in sets]
        tab = _abi.strand_table(self._strands)
        lib().pxo_set_strand_sets(self.h, ctypes.byref(tab), len(sets))

    def strand_set(self, index):
        x = self._strands[index].copy()
        d = x.desc()
        lib().pxo_get_strand_set(self.h, index, ctypes.byref(d))
        x.take_state(d)
        return x

    def set_strand_roots(self, index, positions, orientations=None):
        p = f32(np.asarray(positions, np.float32).reshape(-1))
        q = f32(np.asarray(orientations, np.float32).reshape(-1)) if orientations is not None else None
        lib().pxo_set_strand_roots(self.h, index, ptr(p), ptr(q) if q is not None else None, len(p) // 3)

    def step(self, dt):
        """step(dt) (world.h:1229-1257): fixed substeps, then cloth and hair once. Returns the substeps."""
        return lib().pxo_step_frame(self.h, np.float32(dt))

    def interpolation_alpha(self):
        return lib().pxo_interpolation_alpha(self.h)

    def set_collider_order(self, tree):
        """True: cloth and hair meet colliders in the reference's tree order; False: in slot order."""
        lib().pxo_set_collider_order(self.h, 1 if tree else 0)

    def set_threads(self, n):
        """Host threads for the reference's job-parallel stages (bounds, pairs, narrowphase); 1 = serial."""
        lib().pxo_set_threads(self.h, int(n))

    def set_config(self, config):
        self._cfg = config
        lib().pxo_set_config(self.h, ctypes.byref(config))

    def step_fixed(self, h=None):
        lib().pxo_step_fixed(self.h, np.float32(self._cfg.fixed_dt if h is None else h))

    def step_phase(self, phase, h=None):
        lib().pxo_step_phase(self.h, phase, np.float32(self._cfg.fixed_dt if h is None else h))

    def run(self, steps, h=None):
        return lib().pxo_run(self.h, steps, np.float32(self._cfg.fixed_dt if h is None else h))

    def set_joints(self, joints):
        j = np.ascontiguousarray(joints, dtype=JOINT_DTYPE)
        self.n_joints = len(j)
        lib().pxo_set_joints(self.h, ptr(j), len(j))

    def joints(self):
        out = np.zeros(getattr(self, "n_joints", 0), JOINT_DTYPE)
        lib().pxo_get_joints(self.h, ptr(out), len(out))
        return out

    def set_ignored_pairs(self, ab):
        a = np.ascontiguousarray(np.asarray(ab, np.uint32).reshape(-1, 2))
        lib().pxo_set_ignored_pairs(self.h, ptr(a), len(a))

    def set_atan2_mode(self, rounded):
        """False: the reference's std::atan2 on floats; True: atan2 in double rounded once."""
        lib().pxo_set_atan2_mode(self.h, 1 if rounded else 0)

    def set_order(self, pairs_ab):
        a = np.ascontiguousarray(np.asarray(pairs_ab, np.uint32).reshape(-1, 2))
        lib().pxo_set_order(self.h, ptr(a), len(a))

    def set_miss_table(self, keys, impulses):
        k = np.ascontiguousarray(np.asarray(keys, np.uint64))
        i = np.ascontiguousarray(np.asarray(impulses, np.float32).reshape(-1, 3))
        lib().pxo_set_miss_table(self.h, ptr(k), ptr(i), len(k))

    def bodies(self):
        out = np.zeros(self.n_bodies, BODY_DTYPE)
        lib().pxo_get_bodies(self.h, ptr(out), self.n_bodies)
        return out

    def set_bodies(self, bodies):
        b = np.ascontiguousarray(bodies)
        lib().pxo_set_bodies(self.h, ptr(b), len(b))

    def set_shapes(self, shapes):
        """Replaces the shape table (proxies are created afresh, world.h:1369-1382)."""
        a = np.ascontiguousarray(shapes)
        self.n_shapes = len(a)
        lib().pxo_set_shapes(self.h, ptr(a), len(a))

    def pairs(self):
        n = lib().pxo_get_pairs(self.h, None, 0)
        out = np.zeros((max(n, 1), 2), np.uint32)
        lib().pxo_get_pairs(self.h, ptr(out), n)
        return out[:n]

    def events(self):
        """Contact events queued since the last clear_events(): rows (a, b, type, is_trigger), type
        0 enter, 1 stay, 2 exit (world.h:744, collision/contact_event.h:29-36)."""
        n = lib().pxo_get_events(self.h, None, 0)
        out = np.zeros((max(n, 1), 4), np.uint32)
        lib().pxo_get_events(self.h, ptr(out), n)
        return out[:n]

    def clear_events(self):
        lib().pxo_clear_events(self.h)

    def raycast(self, rays, query_filter=None):
        """Closest hit per ray (RAY_DTYPE in, HIT_DTYPE out; world.h:766-791)."""
        from physxcoopa_amd._abi import HIT_DTYPE
        r = np.ascontiguousarray(rays)
        out = np.zeros(max(len(r), 1), HIT_DTYPE)
        f = None if query_filter is None else ptr(np.ascontiguousarray(query_filter))
        lib().pxo_raycast(self.h, ptr(r), len(r), f, ptr(out))
        return out[:len(r)]

    def shape_cast(self, cast, query_filter=None):
        """shape_cast() (CAST_DTYPE in): the hit record (hit = 0: none)."""
        from physxcoopa_amd._abi import HIT_DTYPE
        c = np.ascontiguousarray(cast).reshape(1)
        f = None if query_filter is None else ptr(np.ascontiguousarray(query_filter))
        out = np.zeros(1, HIT_DTYPE)
        lib().pxo_shape_cast(self.h, ptr(c), f, ptr(out))
        return out[0]

    def compute_penetration(self, probe, query_filter=None):
        """compute_penetration() (world.h:963-995) for the shape posed in a CAST_DTYPE record: a
        PEN_DTYPE array in shape-slot order."""
        from physxcoopa_amd._abi import PEN_DTYPE, probe_parts
        sh, pos, rot = probe_parts(probe)
        f = None if query_filter is None else ptr(np.ascontiguousarray(query_filter))
        n = lib().pxo_compute_penetration(self.h, ptr(sh), ptr(pos), ptr(rot), f, None, 0)
        out = np.zeros(max(n, 1), PEN_DTYPE)
        lib().pxo_compute_penetration(self.h, ptr(sh), ptr(pos), ptr(rot), f, ptr(out), n)
        return out[:n]

    def overlap(self, query, query_filter=None):
        """Bodies with a shape overlapping the query shape, in shape-slot order (world.h:1002-1060)."""
        q = np.ascontiguousarray(query).reshape(1)
        f = None if query_filter is None else ptr(np.ascontiguousarray(query_filter))
        n = lib().pxo_overlap(self.h, ptr(q), f, None, 0)
        out = np.zeros(max(n, 1), np.uint32)
        lib().pxo_overlap(self.h, ptr(q), f, ptr(out), n)
        return out[:n]

    def raycast_all(self, ray, query_filter=None):
        """Every hit of one ray, nearest first (world.h:825-847)."""
        from physxcoopa_amd._abi import HIT_DTYPE
        r = np.ascontiguousarray(ray).reshape(1)
        f = None if query_filter is None else ptr(np.ascontiguousarray(query_filter))
        n = lib().pxo_raycast_all(self.h, ptr(r), f, None, 0)
        out = np.zeros(max(n, 1), HIT_DTYPE)
        lib().pxo_raycast_all(self.h, ptr(r), f, ptr(out), n)
        return out[:n]

    def manifolds(self):
        n = lib().pxo_get_manifolds(self.h, None, 0)
        out = np.zeros(max(n, 1), MANIFOLD_DTYPE)
        lib().pxo_get_manifolds(self.h, ptr(out), n)
        return out[:n]

    def fat_aabbs(self):
        out = np.zeros((self.n_shapes, 6), np.float32)
        lib().pxo_get_fat(self.h, ptr(out), self.n_shapes)
        return out

    def set_fat_aabbs(self, boxes):
        b = np.ascontiguousarray(np.asarray(boxes, np.float32).reshape(-1, 6))
        lib().pxo_set_fat(self.h, ptr(b), len(b))

    def warm_cache(self):
        n = lib().pxo_get_cache(self.h, None, None, 0)
        keys = np.zeros(max(n, 1), np.uint64)
        imp = np.zeros((max(n, 1), 3), np.float32)
        lib().pxo_get_cache(self.h, ptr(keys), ptr(imp), n)
        return keys[:n], imp[:n]

    def set_warm_cache(self, keys, imp):
        k = np.ascontiguousarray(np.asarray(keys, np.uint64))
        i = np.ascontiguousarray(np.asarray(imp, np.float32).reshape(-1, 3))
        lib().pxo_set_cache(self.h, ptr(k), ptr(i), len(k))

    def state_hash(self):
        return int(lib().pxo_state_hash(self.h))

    def stats(self):
        out = np.zeros(4, np.float64)
        lib().pxo_stats(self.h, ptr(out))
        return dict(velocity_passes=int(out[0]), residual=float(out[1]), predicted=int(out[2]), steps=int(out[3]))

    def residual_trace(self):
        n = lib().pxo_residual_trace(self.h, None, 0)
        out = np.zeros(max(n, 1), np.float64)
        lib().pxo_residual_trace(self.h, ptr(out), n)
        return out[:n]

    RESIDUAL_DETAIL_FIELDS = ("manifold", "point", "rn", "rt1", "rt2", "lambda_n", "lambda_t1", "lambda_t2", "vn",
                              "vt1", "vt2", "target", "separation", "a_active", "b_active", "kn")

    def want_residual_detail(self, on=True):
        """Record every contact point's residual terms after the last velocity pass of each step."""
        lib().pxo_want_residual_detail(self.h, 1 if on else 0)

    def eval_residual(self, manifolds, use_restitution=False, h=None):
        """The velocity residual of `manifolds` (MANIFOLD_DTYPE, impulses included) over the current
        bodies, without solving; the per-point terms are then in residual_detail()."""
        m = np.ascontiguousarray(manifolds, dtype=MANIFOLD_DTYPE).reshape(-1)
        return float(lib().pxo_eval_residual(self.h, ptr(m), len(m), 1 if use_restitution else 0,
                                             np.float32(self._cfg.fixed_dt if h is None else h)))

    def residual_detail(self):
        """The last step's per-point residual terms: a dict of float64 arrays (RESIDUAL_DETAIL_FIELDS)."""
        n = lib().pxo_residual_detail(self.h, None, 0)
        out = np.zeros((max(n, 1), 16), np.float64)
        lib().pxo_residual_detail(self.h, ptr(out), n)
        return {k: out[:n, i] for i, k in enumerate(self.RESIDUAL_DETAIL_FIELDS)}


# ---- single-function helpers for known-answer tests ----

def generate_contacts(shape_a, pos_a, rot_a, shape_b, pos_b, rot_b, speculative=False):
    out = np.zeros(1, MANIFOLD_DTYPE)
    sa, sb = np.ascontiguousarray(shape_a).reshape(1), np.ascontiguousarray(shape_b).reshape(1)
    keep = [f32(pos_a, 3), f32(rot_a, 4), f32(pos_b, 3), f32(rot_b, 4)]
    hit = lib().pxo_generate_contacts(ptr(sa), ptr(keep[0]), ptr(keep[1]), ptr(sb), ptr(keep[2]), ptr(keep[3]),
                                      1 if speculative else 0, ptr(out))
    return bool(hit), out[0]


def box_box(ca, hea, qa, cb, heb, qb, speculative=False):
    out = np.zeros(1, MANIFOLD_DTYPE)
    k = [f32(ca, 3), f32(hea, 3), f32(qa, 4), f32(cb, 3), f32(heb, 3), f32(qb, 4)]
    hit = lib().pxo_box_box(*[ptr(a) for a in k], 1 if speculative else 0, ptr(out))
    return bool(hit), out[0]


def sat(ca, hea, qa, cb, heb, qb, speculative=False):
    axis, sep = ctypes.c_int(), ctypes.c_int()
    depth = np.zeros(1, np.float32)
    normal = np.zeros(3, np.float32)
    k = [f32(ca, 3), f32(hea, 3), f32(qa, 4), f32(cb, 3), f32(heb, 3), f32(qb, 4)]
    hit = lib().pxo_sat(*[ptr(a) for a in k], 1 if speculative else 0, ctypes.byref(axis), ptr(depth), ptr(normal),
                        ctypes.byref(sep))
    return bool(hit), axis.value, float(depth[0]), normal, bool(sep.value)


def seg_seg(p1, q1, p2, q2):
    c1, c2 = np.zeros(3, np.float32), np.zeros(3, np.float32)
    k = [f32(p1, 3), f32(q1, 3), f32(p2, 3), f32(q2, 3)]
    lib().pxo_seg_seg(*[ptr(a) for a in k], ptr(c1), ptr(c2))
    return c1, c2


def world_bounds(shape, pos, rot):
    out = np.zeros(6, np.float32)
    s = np.ascontiguousarray(shape).reshape(1)
    k = [f32(pos, 3), f32(rot, 4)]
    lib().pxo_world_bounds(ptr(s), ptr(k[0]), ptr(k[1]), ptr(out))
    return out


def combine(a, b, ma, mb):
    return float(lib().pxo_combine(np.float32(a), np.float32(b), ma, mb))


def orthonormal_basis(n):
    t1, t2 = np.zeros(3, np.float32), np.zeros(3, np.float32)
    k = f32(n, 3)
    lib().pxo_orthonormal_basis(ptr(k), ptr(t1), ptr(t2))
    return t1, t2


def quat_rotate(q, v):
    out = np.zeros(3, np.float32)
    k = [f32(q, 4), f32(v, 3)]
    lib().pxo_quat_rotate(ptr(k[0]), ptr(k[1]), ptr(out))
    return out


def mat3_cast(q):
    out = np.zeros(9, np.float32)
    k = f32(q, 4)
    lib().pxo_mat3_cast(ptr(k), ptr(out))
    return out.reshape(3, 3)  # [col][row]


def ray_shape(shape, pos, rot, ray):
    """raycast_shape for one shape at a pose: (hit, t, normal)."""
    s = np.ascontiguousarray(np.array(shape).reshape(()))
    r = np.ascontiguousarray(np.array(ray).reshape(()))
    t = np.zeros(1, np.float32)
    n = np.zeros(3, np.float32)
    hit = lib().pxo_ray_shape(ptr(s), ptr(f32(pos, 3)), ptr(f32(rot, 4)), ptr(r), ptr(t), ptr(n))
    return bool(hit), float(t[0]), n


def ray_aabb(ray, mn, mx):
    r = np.ascontiguousarray(np.array(ray).reshape(()))
    return bool(lib().pxo_ray_aabb(ptr(r), ptr(f32(mn, 3)), ptr(f32(mx, 3))))


def pack_key(a, b, fid):
    return int(lib().pxo_pack_key(a, b, fid))


def build_hull(points, tolerance=1e-5, max_vertices=64):
    """build_convex_hull (geometry/convex_hull.h:188-449) restated: a Hull, or None for a
    degenerate cloud."""
    p = np.ascontiguousarray(np.asarray(points, np.float32).reshape(-1, 3))
    sizes = np.array([256, 512, 2048, 1024], np.uint32)
    v = np.zeros((256, 3), np.float32)
    pl = np.zeros((512, 4), np.float32)
    fs = np.zeros(513, np.uint32)
    fv = np.zeros(2048, np.uint32)
    ed = np.zeros((1024, 2), np.uint32)
    b6 = np.zeros(6, np.float32)
    vc = np.zeros(4, np.float32)
    ok = lib().pxo_build_hull(ptr(p), len(p), np.float32(tolerance), max_vertices, ptr(sizes), ptr(v), ptr(pl), ptr(fs),
                              ptr(fv), ptr(ed), ptr(b6), ptr(vc))
    if not ok:
        return None
    nv, nf, nfv, ne = (int(x) for x in sizes)
    return Hull(v[:nv], pl[:nf], fs[:nf + 1], fv[:nfv], ed[:ne], b6, volume=float(vc[0]), centroid=vc[1:4].copy())


def set_kat_hulls(hulls):
    """The hull table the single-pair helpers (generate_contacts, shape_bounds) resolve
    PXG_CONVEX_HULL shapes against (the shape's capsule_axis is the index)."""
    tab = hull_table(hulls)
    lib().pxo_set_kat_hulls(ctypes.byref(tab), len(hulls))


def shape_bounds(shape, pos, rot):
    """(world AABB as 6 floats, CCD min extent) of a posed shape (shape.h:198-228, world.h:1411-1424)."""
    s = np.ascontiguousarray(shape).reshape(1)
    out = np.zeros(6, np.float32)
    ext = np.zeros(1, np.float32)
    keep = [f32(pos, 3), f32(rot, 4)]
    lib().pxo_shape_bounds(ptr(s), ptr(keep[0]), ptr(keep[1]), ptr(out), ptr(ext))
    return out, float(ext[0])



# ---- triangle meshes (geometry/triangle_mesh.h, geometry/mesh_bvh.h, collision/mesh_contact.h) ----

class CookedMesh:
    """from_triangles' output: welded vertices (n, 3), indices (t, 3), unit face normals (t, 3),
    adjacency (t, 3; 0xFFFFFFFF for none), bounds (6), non-manifold edge count, and the triangle
    tree: node bounds (k, 6), node links (k, 4: left, right, first, count) and triangle order."""


def cook_mesh(mesh):
    """TriangleMesh::from_triangles (triangle_mesh.h:75-170) and MeshBVH::build (mesh_bvh.h:31-51)
    of a physxcoopa_amd._abi.TriangleMesh; None for an index out of range."""
    d = mesh.desc()
    np_, ni = len(mesh.positions), len(mesh.indices)
    info = MeshInfo()
    v = np.zeros((max(np_, 1), 3), np.float32)
    idx = np.zeros(max(ni, 3), np.uint32)
    nrm = np.zeros(max(ni, 3), np.float32)
    adj = np.zeros(max(ni, 3), np.uint32)
    nmax = max(1, 2 * (ni // 3) + 1)
    nb = np.zeros((nmax, 6), np.float32)
    nl = np.zeros((nmax, 4), np.int32)
    order = np.zeros(max(ni // 3, 1), np.uint32)
    if not lib().pxo_cook_mesh(ctypes.byref(d), ctypes.byref(info), ptr(v), ptr(idx), ptr(nrm), ptr(adj), ptr(nb), ptr(nl),
                               ptr(order)):
        return None
    c = CookedMesh()
    nv, nt, nn = info.n_vertices, info.n_triangles, info.n_nodes
    c.vertices = v[:nv].copy()
    c.indices = idx[:3 * nt].reshape(-1, 3).copy()
    c.normals = nrm[:3 * nt].reshape(-1, 3).copy()
    c.adjacency = adj[:3 * nt].reshape(-1, 3).copy()
    c.bounds = np.array(list(info.local_min) + list(info.local_max), np.float32)
    c.non_manifold_edges = info.non_manifold_edges
    c.node_bounds = nb[:nn].copy()
    c.node_links = nl[:nn].copy()
    c.order = order[:nt].copy()
    return c


def set_kat_meshes(meshes):
    """The mesh table the single-pair helpers (generate_contacts, shape_bounds, mesh_contacts,
    mesh_candidates) resolve PXG_TRIANGLE_MESH shapes against (capsule_axis is the index)."""
    tab = mesh_table(meshes)
    assert lib().pxo_set_kat_meshes(ctypes.byref(tab), len(meshes)) == 1, "mesh index out of range"


def mesh_contacts(shape, pos, rot, mesh_shape, mesh_pos, mesh_rot, use_bvh=True):
    """generate_mesh_contacts (mesh_contact.h:838-884): (hit, manifold), normal from the mesh
    toward the shape; use_bvh=False scans every triangle."""
    out = np.zeros(1, MANIFOLD_DTYPE)
    s, m = np.ascontiguousarray(shape).reshape(1), np.ascontiguousarray(mesh_shape).reshape(1)
    keep = [f32(pos, 3), f32(rot, 4), f32(mesh_pos, 3), f32(mesh_rot, 4)]
    hit = lib().pxo_mesh_contacts(ptr(s), ptr(keep[0]), ptr(keep[1]), ptr(m), ptr(keep[2]), ptr(keep[3]),
                                  1 if use_bvh else 0, ptr(out))
    return bool(hit), out[0]


def mesh_candidates(shape, pos, rot, mesh_shape, mesh_pos, mesh_rot, cap=1 << 16):
    """mesh_contact_candidates (mesh_contact.h:330-340): the ascending triangle indices."""
    out = np.zeros(cap, np.uint32)
    s, m = np.ascontiguousarray(shape).reshape(1), np.ascontiguousarray(mesh_shape).reshape(1)
    keep = [f32(pos, 3), f32(rot, 4), f32(mesh_pos, 3), f32(mesh_rot, 4)]
    n = lib().pxo_mesh_candidates(ptr(s), ptr(keep[0]), ptr(keep[1]), ptr(m), ptr(keep[2]), ptr(keep[3]), ptr(out), cap)
    return out[:min(n, cap)].copy()


def mat4_inverse(m):
    """glm::inverse of a 4x4 given as (4, 4) columns (m[col][row])."""
    a = f32(np.asarray(m, np.float32).reshape(16), 16)
    out = np.zeros(16, np.float32)
    lib().pxo_mat4_inverse(ptr(a), ptr(out))
    return out.reshape(4, 4)


def closest_on_triangle(p, a, b, c):
    """(closest_point_on_triangle, barycentric_ of p) for triangle abc (mesh_contact.h:61-98, 247-258)."""
    k = [f32(p, 3), f32(a, 3), f32(b, 3), f32(c, 3)]
    out, bary = np.zeros(3, np.float32), np.zeros(3, np.float32)
    lib().pxo_closest_on_triangle(ptr(k[0]), ptr(k[1]), ptr(k[2]), ptr(k[3]), ptr(out), ptr(bary))
    return out, bary


# ---- cloth and hair builders (cloth/cloth_builder.h, hair/strand_builder.h) ----

def grid_cloth(columns=21, rows=21, width=2.0, height=2.0, center=(0, 0, 0), orientation=(0, 0, 0, 1),
               total_mass=1.0, shear=True, params=None, pins=(), tethers=False):
    """make_grid_cloth (cloth_builder.h:103-239), then each pin in order -- (point, radius) for
    pin_static, (point, radius, (body, generation), (position, orientation)) for pin_to_body --
    then build_tethers when asked. An _abi.Cloth."""
    params = params if params is not None else _abi.ClothParams()
    fd = f32([width, height] + list(center) + list(orientation) + [total_mass])
    ui = np.array([columns, rows, 1 if shear else 0], np.uint32)
    b = lib().pxo_cloth_grid(ptr(fd), ptr(ui), ctypes.byref(params))
    try:
        for pin in pins:
            pt = f32(pin[0], 3)
            if len(pin) == 2:
                lib().pxo_cloth_pin(b, ptr(pt), np.float32(pin[1]), None, None)
            else:
                body = np.array(pin[2], np.uint32)
                pose = f32(list(pin[3][0]) + list(pin[3][1]), 7)
                lib().pxo_cloth_pin(b, ptr(pt), np.float32(pin[1]), ptr(body), ptr(pose))
        if tethers:
            lib().pxo_cloth_tethers(b)
        sz = np.zeros(8, np.uint32)
        lib().pxo_cloth_sizes(b, ptr(sz))
        c = _abi.Cloth(np.zeros(sz[0], _abi.PARTICLE_DTYPE), np.zeros(sz[1], _abi.CONSTRAINT_DTYPE),
                       np.zeros(sz[2], _abi.CONSTRAINT_DTYPE), np.zeros((sz[3], 2), np.uint32),
                       np.zeros((sz[4], 2), np.uint32), np.zeros(sz[5], _abi.TETHER_DTYPE),
                       np.zeros(sz[6], _abi.ANCHOR_DTYPE), np.zeros(sz[7], np.uint32))
        batches = np.zeros(2 * (int(sz[3]) + int(sz[4])) + 2, np.uint32)
        d = c.desc()
        lib().pxo_cloth_export(b, ctypes.byref(d), ptr(batches))
        c.stretch_batches[:] = batches[:2 * sz[3]].reshape(-1, 2)
        c.bend_batches[:] = batches[2 * sz[3]:2 * (sz[3] + sz[4])].reshape(-1, 2)
        c.columns, c.rows, c.params = d.columns, d.rows, d.params
        c.bounds = np.array(list(d.bounds_min) + list(d.bounds_max), np.float32)
        return c
    finally:
        lib().pxo_cloth_free(b)


def make_strands(strands, params=None, masses=None, root_orientations=None):
    """make_strands (strand_builder.h:66-130) of a list of point lists. An _abi.StrandSet."""
    params = params if params is not None else _abi.StrandParams()
    counts = np.array([len(s) for s in strands], np.uint32)
    pts = f32(np.concatenate([np.asarray(s, np.float32).reshape(-1, 3) for s in strands]).reshape(-1))
    rq = f32(np.asarray(root_orientations if root_orientations is not None else [(0, 0, 0, 1)] * len(strands),
                        np.float32).reshape(-1))
    ms = f32(np.asarray(masses if masses is not None else [0.001] * len(strands), np.float32))
    b = lib().pxo_strands_make(ptr(pts), ptr(counts), len(strands), ptr(rq), ptr(ms), ctypes.byref(params))
    try:
        sz = np.zeros(4, np.uint32)
        lib().pxo_strands_sizes(b, ptr(sz))
        x = _abi.StrandSet(np.zeros(sz[0], _abi.PARTICLE_DTYPE), np.zeros(sz[1], _abi.CONSTRAINT_DTYPE),
                           np.zeros((sz[2], 2), np.uint32), np.zeros((sz[0], 3), np.float32),
                           np.zeros((sz[0], 3), np.float32), np.zeros(sz[3], _abi.STRAND_DTYPE), params=params)
        d = x.desc()
        lib().pxo_strands_export(b, ctypes.byref(d))
        x.bounds = np.array(list(d.bounds_min) + list(d.bounds_max), np.float32)
        return x
    finally:
        lib().pxo_strands_free(b)
