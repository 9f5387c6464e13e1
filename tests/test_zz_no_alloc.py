"""pxg_overlap_batch_noalloc (device buffers in and out, scratch in a fixed device array of the
library) against pxg_overlap_batch, the entry point it is a variant of: the same counts and body
lists bit for bit, on two calls in a row (the second with fewer queries and a smaller capacity)."""
import ctypes

import numpy as np
import pytest

from physxcoopa_amd import scenes, _abi
from physxcoopa_amd._lib import lib

pytestmark = pytest.mark.gpu


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


class DeviceBuffers:
    """Device memory for the test's inputs and outputs, from the HIP runtime libpxg.so itself uses
    (torch ships another HIP runtime, which would be a second one in the process)."""

    def __init__(self):
        lib()  # libpxg.so loads /opt/rocm's libamdhip64.so.7; CDLL by soname returns that one
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipFree.argtypes = [ctypes.c_void_p]
        self.h = h
        self.ptrs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), max(nbytes, 4)) == 0
        self.ptrs.append(p)
        return p

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        p = self.alloc(a.nbytes)
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def download(self, p, dtype, n):
        out = np.zeros(max(n, 1), dtype)
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out[:n]

    def free(self):
        for p in self.ptrs:
            self.h.hipFree(p)
        self.ptrs = []


def noalloc(w, q, cap, query_filter=None):
    """Queries copied to device memory and outputs allocated before the call (filled with
    0xFF bytes); returns the status, counts and bodies (n x cap)."""
    dev = DeviceBuffers()
    n = len(q)
    d_q = dev.upload(q)
    d_bodies = dev.upload(np.full(max(n * cap, 1), 0xFFFFFFFF, np.uint32))
    d_counts = dev.upload(np.full(max(n, 1), 0xFFFFFFFF, np.uint32))
    f = None if query_filter is None else np.ascontiguousarray(query_filter, dtype=_abi.FILTER_DTYPE)
    rc = lib().pxg_overlap_batch_noalloc(w.h, d_q, n, None if f is None else ctypes.c_void_p(f.ctypes.data),
                                         d_bodies, cap, d_counts)
    w.synchronize()
    counts = dev.download(d_counts, np.uint32, n)
    bodies = dev.download(d_bodies, np.uint32, n * cap).reshape(n, cap)
    dev.free()
    return rc, counts, bodies


def same_lists(counts, bodies, ref_counts, ref_bodies, cap):
    assert np.array_equal(counts, ref_counts)
    for k in range(len(counts)):
        m = min(cap, int(counts[k]))
        assert np.array_equal(bodies[k, :m], ref_bodies[k, :m]), k


def test_noalloc_overlap_batch_equals_overlap_batch():
    from physxcoopa_amd.world import World
    b, s = scenes.mixed_pile_scene(n=300, seed=11)
    s = s.copy()
    s["is_trigger"][3::13] = 1
    w = World(_abi.default_config(), b, s)
    for _ in range(80):
        w.step_fixed()
    rng = np.random.default_rng(41)
    q = random_overlaps(rng, 700, -5, 5, 0, 3)
    flt = _abi.query_filter(include_triggers=False, ignore_body=int(s["owner"][40]))
    # first call: 700 queries, capacity 64, no filter
    rc, counts, bodies = noalloc(w, q, 64)
    assert rc == 0
    ref_counts, ref_bodies = w.overlap_batch(q, cap=64)
    same_lists(counts, bodies, ref_counts, ref_bodies, 64)
    assert counts.sum() > 700
    # second call right after: fewer queries, a smaller capacity, a filter
    q2 = q[100:350]
    rc, counts2, bodies2 = noalloc(w, q2, 3, flt)
    assert rc == 0
    ref_counts2, ref_bodies2 = w.overlap_batch(q2, flt, cap=3)
    same_lists(counts2, bodies2, ref_counts2, ref_bodies2, 3)
    assert counts2.max() > 3  # the capacity truncated some lists
    # after more steps (new poses): the scratch's geometry is rebuilt by the call
    for _ in range(5):
        w.step_fixed()
    rc, counts3, bodies3 = noalloc(w, q[:50], 16)
    ref_counts3, ref_bodies3 = w.overlap_batch(q[:50], cap=16)
    same_lists(counts3, bodies3, ref_counts3, ref_bodies3, 16)
    w.close()


def test_noalloc_overlap_batch_flags_unknown_types_and_limits():
    """A query record of another type gets the count PXG_NO_BODY (the host cannot check records in
    device memory); a world beyond the scratch's 65536 shape slots is refused with PXG_ERR_CAPACITY
    before anything is enqueued."""
    from physxcoopa_amd.world import World
    b, s = scenes.mixed_pile_scene(n=200, seed=12)
    w = World(_abi.default_config(), b, s)
    w.step_fixed()
    q = random_overlaps(np.random.default_rng(5), 40, -5, 5, 0, 3)
    q["type"][7] = 9
    rc, counts, _ = noalloc(w, q, 8)
    assert rc == 0 and counts[7] == 0xFFFFFFFF
    ok = np.ones(40, bool)
    ok[7] = False
    ref_counts, _ = w.overlap_batch(q[ok], cap=8)
    assert np.array_equal(counts[ok], ref_counts)
    w.close()
    n = 65537
    sb = scenes.SceneBuilder()
    for k in range(n):
        sb.dynamic_sphere((float(k % 256), float(k // 256), 0.5), 0.4, mass=1.0)
    bb, ss = sb.arrays()
    big = World(_abi.default_config(), bb, ss)
    rc, _, _ = noalloc(big, q[:4], 4)
    assert rc == _abi.PXG_ERR_CAPACITY
    big.close()
