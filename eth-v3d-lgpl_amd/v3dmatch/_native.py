"""ctypes binding of libv3dmatch.so (declarations: include/v3dm.h).

The library is built in-tree by eth-v3d-lgpl_amd/build.py; a missing library is an
error -- there is no fallback implementation.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_DIR = os.path.join(_HERE, "_lib")

V3DM_OK = 0
ERROR_NAMES = {1: "INVALID", 2: "STATE", 3: "CAPACITY", 4: "HIP", 5: "UNSUPPORTED", 6: "NOMEM", 7: "IO"}


class V3DMError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"v3dm error {code} ({ERROR_NAMES.get(code, '?')}): {message}")
        self.code = code
        self.message = message


class Stats(C.Structure):
    _fields_ = [
        ("nLeft", C.c_int), ("nRight", C.c_int), ("nMatches", C.c_int), ("nRechecked", C.c_int),
        ("precision", C.c_int), ("nFallback", C.c_int), ("kernelMs", C.c_double),
        ("kernelLaunches", C.c_int), ("nRescored", C.c_int), ("nSampled", C.c_int), ("nSampleMissed", C.c_int),
    ]


def lib_path(name="libv3dmatch.so"):
    return os.path.join(_LIB_DIR, name)


_lib = None

_P = C.c_void_p
_I = C.c_int
_F = C.c_float
_IP = C.POINTER(C.c_int)
_FP = C.c_void_p  # raw buffer pointers are passed as addresses

SIGNATURES = {
    "v3dm_create": (_I, [C.POINTER(_P), _I]),
    "v3dm_destroy": (None, [_P]),
    "v3dm_last_error": (C.c_char_p, [_P]),
    "v3dm_allocate": (_I, [_P, _I, _I, _I]),
    "v3dm_deallocate": (_I, [_P]),
    "v3dm_set_flags": (_I, [_P, _I, _I]),
    "v3dm_set_left": (_I, [_P, _I, _FP]),
    "v3dm_set_right": (_I, [_P, _I, _FP]),
    "v3dm_find_putative_matches": (_I, [_P, _F, _F, _FP, _I, _IP]),
    "v3dm_run_guided_F": (_I, [_P, _FP, _F, _FP, _FP, _F, _F, _FP, _I, _IP]),
    "v3dm_run_guided_H": (_I, [_P, _FP, _F, _FP, _FP, _F, _F, _FP, _I, _IP]),
    "v3dm_set_precision": (_I, [_P, _I]),
    "v3dm_set_certification": (_I, [_P, _F]),
    "v3dm_set_stream": (_I, [_P, _P]),
    "v3dm_set_learned_topk": (_I, [_P, _I]),
    "v3dm_set_learned_margin": (_I, [_P, _F]),
    "v3dm_sampson_errors": (_I, [_FP, _FP, _I, _FP]),
    "v3dm_essentials_from_five": (_I, [_FP, _FP, _FP, _FP, _FP, _IP]),
    "v3dm_relative_pose_from_essential": (_I, [_FP, _I, _FP, _FP, _FP, _FP, _FP, _FP, _IP]),
    "v3dm_refine_relative_pose": (_I, [_FP, _I, _FP, _FP, _FP, _FP, C.c_double]),
    "v3dm_robust_orientation": (_I, [_FP, _I, _FP, _FP, C.c_double, _I, _I, _I, C.c_uint64, _P, _FP]),
    "v3dm_set_derived_cache": (_I, [_P, _I]),
    "v3dm_set_fallback": (_I, [_P, _I, _F]),
    "v3dm_fallback_allowed_misses": (_I, [_I, _F]),
    "v3dm_fallback_allowed_misses_binomial": (_I, [_I, _F]),
    "v3dm_set_fallback_rule": (_I, [_P, _I]),
    "v3dm_batch_set_tile_shape": (_I, [_P, _I, _I]),
    "v3dm_batch_get_tile_shape": (_I, [_P, _I, _IP, _IP]),
    "v3dm_debug_mutual_wait": (_I, [_P, _I, _I]),
    "v3dm_set_left_device": (_I, [_P, _I, _FP]),
    "v3dm_set_right_device": (_I, [_P, _I, _FP]),
    "v3dm_set_left_u8": (_I, [_P, _I, _FP]),
    "v3dm_set_right_u8": (_I, [_P, _I, _FP]),
    "v3dm_find_putative_matches_async": (_I, [_P, _F, _F]),
    "v3dm_get_matches": (_I, [_P, _FP, _I, _IP]),
    "v3dm_device_results": (_I, [_P, C.POINTER(_P), C.POINTER(_P)]),
    "v3dm_get_maxima": (_I, [_P, _FP, _FP]),
    "v3dm_synchronize": (_I, [_P]),
    "v3dm_set_profiling": (_I, [_P, _I]),
    "v3dm_get_stats": (_I, [_P, C.POINTER(Stats)]),
    "v3dm_get_second_best": (_I, [_P, _FP, _FP]),
    "v3dm_time_dominant_kernel": (_I, [_P, _I, C.POINTER(C.c_float)]),
    "v3dm_batch_set_images_u8": (_I, [_P, _I, _FP, _FP]),
    "v3dm_batch_set_images": (_I, [_P, _I, _FP, _FP]),
    "v3dm_batch_set_pairs": (_I, [_P, _I, _FP, _FP]),
    "v3dm_batch_match_async": (_I, [_P, _F, _F]),
    "v3dm_batch_get_matches": (_I, [_P, _FP, _FP, _I, _IP]),
    "v3dm_batch_match": (_I, [_P, _F, _F, _FP, _FP, _I, _IP]),
    "v3dm_batch_set_tuning": (_I, [_P, _I, _I, _I, _I]),
    "v3dm_batch_get_tuning": (_I, [_P, _I, _IP, _IP, _IP, C.POINTER(C.c_float)]),
    "v3dm_batch_use_graph": (_I, [_P, _I]),
    "v3dm_batch_last_ms": (_I, [_P, C.POINTER(C.c_float)]),
    "v3dm_batch_get_fallback": (_I, [_P, _FP]),
    "v3dm_batch_get_check": (_I, [_P, _FP]),
    "v3dm_batch_time_dominant_kernel": (_I, [_P, _I, C.POINTER(C.c_float)]),
    "v3dm_score_block": (_I, [_P, _I, _I, _I, _FP]),
    "v3dm_match_device_noalloc": (_I, [_FP, _I, _FP, _I, _F, _F, _FP, _FP, _P]),
    "v3dm_noalloc_last_error": (C.c_char_p, []),
    "v3dm_learned_set_state": (_I, [_P, _FP, _FP, _FP, _I]),
    "v3dm_learned_get_state": (_I, [_P, _FP, _FP, _FP, _IP]),
    "v3dm_learned_set_code_scale": (_I, [_P, _F]),
    "v3dm_learned_set_training": (_I, [_P, _I, _FP, _I, _FP]),
    "v3dm_learned_train": (_I, [_P, _I, _I, _F, _F, C.c_uint64, _FP]),
    "v3dm_learned_gradients": (_I, [_P, _I, _FP, _F, C.POINTER(C.c_float), _FP, _FP]),
    "v3dm_learned_embed": (_I, [_P, _I, _FP, _FP]),
    "v3dm_features_count": (_I, [C.c_char_p, _IP, _IP]),
    "v3dm_features_read": (_I, [C.c_char_p, _I, _FP, _FP, _FP, _FP, _FP, _FP, _IP]),
    "v3dm_features_write": (_I, [C.c_char_p, _I, _FP, _FP, _FP, _FP, _FP, _FP, _I]),
    "v3dm_voctree_sizes": (_I, [_I, _I, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "v3dm_voctree_create": (_I, [C.POINTER(_P), _I, _I, _I]),
    "v3dm_voctree_destroy": (None, [_P]),
    "v3dm_voctree_set_nodes": (_I, [_P, _FP]),
    "v3dm_voctree_load": (_I, [_P, C.c_char_p]),
    "v3dm_voctree_traverse": (_I, [_P, _I, _FP, _FP, _I]),
    "v3dm_voctree_traverse_device": (_I, [_P, _I, _FP, _FP, _I]),
    "v3dm_voctree_synchronize": (_I, [_P, _IP]),
}


def load_library(path=None):
    """Load libv3dmatch.so once and declare the C signatures."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or lib_path()
    if not os.path.exists(p):
        raise V3DMError(5, f"{p} is missing: build it with eth-v3d-lgpl_amd/build.py "
                           "(there is no non-HIP implementation)")
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib
