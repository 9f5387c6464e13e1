"""Loader of libpxg.so, the HIP contact step. There is no CPU fallback: every entry point of
this package fails loudly when the library is missing or no GPU is present."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpxg.so")
_lib = None

_SIGNATURES = {
    "pxg_config_default": ([ctypes.c_void_p], None),
    "pxg_create": ([ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "pxg_destroy": ([ctypes.c_void_p], ctypes.c_int),
    "pxg_get_config": ([ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "pxg_set_config": ([ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "pxg_set_gravity": ([ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "pxg_set_bodies": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32], ctypes.c_int),
    "pxg_get_bodies": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32], ctypes.c_int),
    "pxg_set_hulls": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32], ctypes.c_int),
    "pxg_set_meshes": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32], ctypes.c_int),
    "pxg_cook_mesh": ([ctypes.c_void_p] * 6, ctypes.c_int),
    "pxg_set_cloths": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32], ctypes.c_int),
    "pxg_get_cloth": ([ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p], ctypes.c_int),
    "pxg_set_strand_sets": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32], ctypes.c_int),
    "pxg_get_strand_set": ([ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p], ctypes.c_int),
    "pxg_set_strand_roots": ([ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32],
                             ctypes.c_int),
    "pxg_set_bodies_subset": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32], ctypes.c_int),
    "pxg_get_poses": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32], ctypes.c_int),
    "pxg_set_shapes": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32], ctypes.c_int),
    "pxg_set_layer_masks": ([ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "pxg_set_joints": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32], ctypes.c_int),
    "pxg_get_joints": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32], ctypes.c_int),
    "pxg_set_ignored_pairs": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32], ctypes.c_int),
    "pxg_batch_set_joints": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "pxg_batch_set_mlp": ([ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "pxg_batch_get_predictions": ([ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                   ctypes.c_void_p], ctypes.c_int),
    "pxg_batch_get_joints": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32], ctypes.c_int),
    "pxg_batch_set_ignored_pairs": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "pxg_step_fixed": ([ctypes.c_void_p, ctypes.c_float], ctypes.c_int),
    "pxg_step_phase": ([ctypes.c_void_p, ctypes.c_uint32, ctypes.c_float], ctypes.c_int),
    "pxg_step": ([ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p], ctypes.c_int),
    "pxg_step_soft": ([ctypes.c_void_p, ctypes.c_float, ctypes.c_uint32], ctypes.c_int),
    "pxg_synchronize": ([ctypes.c_void_p], ctypes.c_int),
    "pxg_get_stats": ([ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "pxg_get_pairs": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p], ctypes.c_int),
    "pxg_get_manifolds": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p], ctypes.c_int),
    "pxg_get_solve_order": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p,
                             ctypes.c_uint32, ctypes.c_void_p], ctypes.c_int),
    "pxg_get_fat_aabbs": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32], ctypes.c_int),
    "pxg_set_fat_aabbs": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32], ctypes.c_int),
    "pxg_get_warm_cache": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p],
                           ctypes.c_int),
    "pxg_set_warm_cache": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32], ctypes.c_int),
    "pxg_state_hash": ([ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "pxg_get_events": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p], ctypes.c_int),
    "pxg_clear_events": ([ctypes.c_void_p], ctypes.c_int),
    "pxg_raycast": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "pxg_shape_cast": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "pxg_overlap_batch": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p,
                           ctypes.c_uint32, ctypes.c_void_p], ctypes.c_int),
    "pxg_overlap_batch_noalloc": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_uint32, ctypes.c_void_p], ctypes.c_int),
    "pxg_shape_cast_batch": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p],
                             ctypes.c_int),
    "pxg_compute_penetration_batch": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p], ctypes.c_int),
    "pxg_compute_penetration": ([ctypes.c_void_p] * 6 + [ctypes.c_uint32, ctypes.c_void_p], ctypes.c_int),
    "pxg_overlap": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p],
                    ctypes.c_int),
    "pxg_raycast_device": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p],
                           ctypes.c_int),
    "pxg_raycast_any": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "pxg_raycast_all": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                         ctypes.c_void_p], ctypes.c_int),
    "pxg_set_events": ([ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32], ctypes.c_int),
    "pxg_set_mlp": ([ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "pxg_set_recording": ([ctypes.c_void_p, ctypes.c_uint32], ctypes.c_int),
    "pxg_get_records": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                         ctypes.c_void_p], ctypes.c_int),
    "pxg_get_predictions": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p],
                            ctypes.c_int),
    "pxg_mlp_infer": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p], ctypes.c_int),
    "pxg_mlp_gradients": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
                           ctypes.c_void_p], ctypes.c_int),
    "pxg_mlp_train": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
                       ctypes.c_void_p], ctypes.c_int),
    "pxg_mlp_last_error": ([], ctypes.c_char_p),
    "pxg_set_profiling": ([ctypes.c_void_p, ctypes.c_uint32], ctypes.c_int),
    "pxg_get_profile": ([ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "pxg_batch_create": ([ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                          ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "pxg_batch_destroy": ([ctypes.c_void_p], ctypes.c_int),
    "pxg_batch_set_tuning": ([ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32], ctypes.c_int),
    "pxg_batch_set_occupancy": ([ctypes.c_void_p, ctypes.c_uint32], ctypes.c_int),
    "pxg_batch_set_layer_masks": ([ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
    "pxg_batch_step": ([ctypes.c_void_p, ctypes.c_float, ctypes.c_uint32], ctypes.c_int),
    "pxg_batch_synchronize": ([ctypes.c_void_p], ctypes.c_int),
    "pxg_batch_get_bodies": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32], ctypes.c_int),
    "pxg_batch_get_stats": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32], ctypes.c_int),
    "pxg_batch_get_manifolds": ([ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p],
                                ctypes.c_int),
    "pxg_batch_get_solve_order": ([ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p], ctypes.c_int),
    "pxg_batch_last_error": ([], ctypes.c_char_p),
    "pxg_last_error": ([], ctypes.c_char_p),
    "pxg_version": ([], ctypes.c_char_p),
}


class PxgError(RuntimeError):
    pass


def lib():
    """The loaded libpxg.so with argument types set. Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PxgError("libpxg.so is not built (run `python -m physxcoopa_amd.build`); "
                           "there is no CPU fallback for the contact step")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != 0:
        msg = lib().pxg_last_error().decode()
        raise PxgError("%s failed (%d): %s" % (what, rc, msg))


def check_mlp(rc, what=""):
    if rc != 0:
        msg = lib().pxg_mlp_last_error().decode()
        raise PxgError("%s failed (%d): %s" % (what, rc, msg))


def check_batch(rc, what=""):
    if rc != 0:
        msg = lib().pxg_batch_last_error().decode()
        raise PxgError("%s failed (%d): %s" % (what, rc, msg))
