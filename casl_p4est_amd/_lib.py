"""ctypes bindings of libcasl_reinit.so (the C ABI of include/casl_reinit.h).

The library is loaded from this package directory.  There is no fallback:
if the shared library is missing, importing the bindings raises.
"""
import ctypes
import weakref
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcasl_reinit.so")

CASL_OK = 0
CASL_ERR_INVALID_ARGUMENT = -1
CASL_ERR_NO_DEVICE = -2
CASL_ERR_HIP = -3
CASL_ERR_OUT_OF_MEMORY = -4
CASL_ERR_INEXACT_DISTANCES = -5
CASL_ERR_INVALID_STENCIL = -6
CASL_ERR_NOT_BUILT = -7

CASL_PATH_TILED = 0
CASL_PATH_TWO_PASS = 1

CASL_PRECISION_FP64 = 0
CASL_PRECISION_MIXED = 1

CASL_LS_CALLBACK = 0
CASL_LS_STAR = 1
CASL_LS_SPHERES = 2
CASL_LS_GYROID = 3
CASL_LS_SPHERE = 4
CASL_LS_PLANE = 5
CASL_LS_MODULATED = 6

c_int32_p = ctypes.POINTER(ctypes.c_int32)
c_int64_p = ctypes.POINTER(ctypes.c_int64)
c_double_p = ctypes.POINTER(ctypes.c_double)
c_uint8_p = ctypes.POINTER(ctypes.c_uint8)
c_int8_p = ctypes.POINTER(ctypes.c_int8)

LEVELSET_FN = ctypes.CFUNCTYPE(ctypes.c_double, c_double_p, ctypes.c_void_p)


class FlatStencil(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32),
        ("n_nodes", ctypes.c_int64),
        ("nbr", c_int32_p),
        ("d", c_double_p),
        ("d_tr", c_double_p),
        ("wall", c_uint8_p),
        ("zero_distance_threshold", ctypes.c_double),
    ]


class LevelSetSpec(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("n_items", ctypes.c_int32),
        ("params", c_double_p),
        ("fn", LEVELSET_FN),
        ("ctx", ctypes.c_void_p),
    ]


class CaslError(RuntimeError):
    """Non-zero status from the native library."""

    def __init__(self, code, func, message):
        super().__init__(f"{func} failed with status {code}: {message}")
        self.code = code


# The signatures exported by include/casl_reinit.h: name -> (restype, argtypes)
SIGNATURES = {
    "casl_last_error": (ctypes.c_char_p, []),
    "casl_stencil_create": (ctypes.c_int, [ctypes.POINTER(FlatStencil), ctypes.POINTER(ctypes.c_void_p)]),
    "casl_stencil_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "casl_stencil_get_path": (ctypes.c_int, [ctypes.c_void_p, c_int32_p]),
    "casl_stencil_info": (ctypes.c_int, [ctypes.c_void_p, c_int32_p, c_int64_p]),
    "casl_stencil_set_path": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "casl_stencil_tile_info": (ctypes.c_int, [ctypes.c_void_p, c_int32_p, c_int64_p, c_int64_p, c_int64_p,
                                              c_int64_p]),
    "casl_tile_plan": (ctypes.c_int, [ctypes.POINTER(FlatStencil), c_int64_p, c_int64_p, c_int64_p, c_int64_p,
                                      c_int64_p]),
    "casl_stencil_tile_check": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(FlatStencil), c_int32_p]),
    "casl_second_derivatives": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "casl_reinit": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_void_p]),
    "casl_reinit_host": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                        ctypes.c_void_p]),
    "casl_reinit_graph": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_void_p]),
    "casl_reinit_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                           ctypes.c_void_p]),
    "casl_reinit_masked": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_void_p]),
    "casl_band_linf": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_double, c_double_p, ctypes.c_void_p]),
    "casl_profile_enable": (ctypes.c_int, [ctypes.c_int]),
    "casl_profile_reset": (ctypes.c_int, []),
    "casl_profile_query": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), c_int64_p]),
    "casl_device_cache_set_limit": (ctypes.c_int, [ctypes.c_size_t]),
    "casl_device_cache_trim": (ctypes.c_int, []),
    "casl_device_cache_held": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_size_t)]),
    "casl_batch_create": (ctypes.c_int, [ctypes.POINTER(FlatStencil), ctypes.c_int32,
                                         ctypes.POINTER(ctypes.c_void_p)]),
    "casl_batch_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "casl_batch_info": (ctypes.c_int, [ctypes.c_void_p, c_int64_p, c_int32_p, c_int32_p]),
    "casl_batch_set_schedule": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    "casl_batch_get_schedule": (ctypes.c_int, [ctypes.c_void_p, c_int32_p, c_int32_p]),
    "casl_batch_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_void_p, c_int32_p, c_int32_p, ctypes.POINTER(ctypes.c_float)]),
    "casl_batch_tile_info": (ctypes.c_int, [ctypes.c_void_p, c_int32_p, c_int64_p]),
    "casl_batch_reinit": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_void_p]),
    "casl_batch_reinit_graph": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_void_p]),
    "casl_advect_in_normal_direction": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                       ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p]),
    "casl_advect_cfl_dt": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                                          ctypes.c_void_p]),
    "casl_first_derivatives": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "casl_normals": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "casl_mean_curvature": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "casl_curvature_samples": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_int,
                                              ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                              c_int64_p, ctypes.c_void_p]),
    "casl_curvature_normalize": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]),
    "casl_curvature_hybrid": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_float,
                                             ctypes.c_float, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    "casl_curvature_grid_samples": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                   ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                   ctypes.c_int64, c_int64_p, c_int64_p, ctypes.c_void_p]),
    "casl_curvature_normalize_3d": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                                   ctypes.c_void_p]),
    "casl_curvature_hybrid_3d": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_float,
                                                ctypes.c_float, ctypes.c_int, ctypes.c_float, ctypes.c_void_p,
                                                ctypes.c_void_p]),
    "casl_curvature_learned": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_float,
                                              ctypes.c_float, ctypes.c_int, ctypes.c_float, ctypes.c_void_p,
                                              ctypes.c_void_p]),
    "casl_mlp_create": (ctypes.c_int, [c_int32_p, ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)]),
    "casl_mlp_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "casl_mlp_num_params": (ctypes.c_int, [ctypes.c_void_p, c_int64_p, c_int64_p]),
    "casl_mlp_set_scaler": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float),
                                           ctypes.POINTER(ctypes.c_float)]),
    "casl_mlp_get_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float),
                                           ctypes.POINTER(ctypes.c_float)]),
    "casl_mlp_set_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float),
                                           ctypes.POINTER(ctypes.c_float)]),
    "casl_mlp_get_grads": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float),
                                          ctypes.POINTER(ctypes.c_float)]),
    "casl_mlp_forward": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                        ctypes.c_void_p]),
    "casl_mlp_backward": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_float), ctypes.c_void_p]),
    "casl_mlp_adam_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                          ctypes.c_float, ctypes.c_void_p]),
    "casl_bf16_gemm": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                      ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                      ctypes.c_void_p]),
    "casl_levelset_eval": (ctypes.c_int, [ctypes.POINTER(LevelSetSpec), ctypes.c_int, ctypes.c_int64,
                                          c_double_p, c_double_p]),
    "casl_forest_new": (ctypes.c_int, [ctypes.c_int, c_double_p, c_double_p, ctypes.POINTER(ctypes.c_void_p)]),
    "casl_forest_new_brick": (ctypes.c_int, [ctypes.c_int, c_double_p, c_double_p, c_int32_p, c_int32_p,
                                             ctypes.POINTER(ctypes.c_void_p)]),
    "casl_forest_trees": (ctypes.c_int, [ctypes.c_void_p, c_int32_p, c_int32_p, c_int32_p, c_int32_p]),
    "casl_forest_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "casl_forest_refine_levelset": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                                   ctypes.c_double, ctypes.POINTER(LevelSetSpec)]),
    "casl_forest_build_nodes": (ctypes.c_int, [ctypes.c_void_p]),
    "casl_forest_sizes": (ctypes.c_int, [ctypes.c_void_p, c_int64_p, c_int64_p, c_int32_p]),
    "casl_forest_get_leaves": (ctypes.c_int, [ctypes.c_void_p, c_int32_p, c_int8_p, c_int32_p]),
    "casl_forest_get_nodes": (ctypes.c_int, [ctypes.c_void_p, c_int32_p, c_double_p]),
    "casl_forest_sample": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(LevelSetSpec), c_double_p]),
    "casl_forest_init_neighbors": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(FlatStencil)]),
    "casl_device_grid_build": (ctypes.c_int, [ctypes.c_int, c_double_p, c_double_p, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_double, ctypes.c_double, ctypes.POINTER(LevelSetSpec),
                                              ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    "casl_device_grid_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "casl_device_grid_sizes": (ctypes.c_int, [ctypes.c_void_p, c_int64_p, c_int64_p, c_int32_p]),
    "casl_device_grid_tables": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(FlatStencil)]),
    "casl_device_grid_download": (ctypes.c_int, [ctypes.c_void_p, c_int32_p, c_double_p, c_double_p, c_uint8_p,
                                                  c_int32_p]),
    "casl_device_grid_sample": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(LevelSetSpec), ctypes.c_void_p,
                                               ctypes.c_void_p]),
    "casl_device_grid_stencil": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
}

_lib = None


def own(obj, handle, destroy):
    """Ties a native handle to obj: the C function `destroy` frees it when obj
    is collected or, at the latest, when the interpreter exits (before the
    modules and the HIP runtime are torn down)."""
    obj._h = handle
    obj._finalizer = weakref.finalize(obj, _destroy, destroy, handle.value)


def _destroy(name, value):
    if value:
        getattr(lib(), name)(ctypes.c_void_p(value))


def lib():
    """Loads the native library once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} not found: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(code, func):
    if code != CASL_OK:
        msg = lib().casl_last_error()
        raise CaslError(code, func, msg.decode() if msg else "")
    return code


def call(name, *args):
    """Calls a C-ABI function and raises CaslError on a non-zero status."""
    return check(getattr(lib(), name)(*args), name)
