"""ctypes binding of libhdog.so (include/hdog.h).

torch is imported first so that libhdog.so resolves its HIP runtime dependency to the
runtime torch already loaded: device pointers and streams are then shared.  There is
no fallback: if the library is missing this module raises at import.
"""
import ctypes
import os

import torch  # noqa: F401  (loads the HIP runtime libhdog.so binds to)

_HERE = os.path.dirname(os.path.abspath(__file__))
# HDOG_LIBRARY selects another build of the same library (e.g. a tuning variant)
LIB_PATH = os.environ.get("HDOG_LIBRARY") or os.path.join(_HERE, "libhdog.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build it with `make -C cubic-informatics_amd` "
        "(or __graft_entry__.build()); the HDoG path has no CPU fallback")

lib = ctypes.CDLL(LIB_PATH)

HDOG_OK, HDOG_ERR_ARG, HDOG_ERR_HIP, HDOG_ERR_STATE, HDOG_ERR_CAPACITY = range(5)
HDOG_FP32, HDOG_FP16, HDOG_BF16, HDOG_FP16_DOG32, HDOG_BF16_DOG32 = range(5)

_vp = ctypes.c_void_p
_i = ctypes.c_int
_f = ctypes.c_float
_d = ctypes.c_double
_i64 = ctypes.c_int64


class Regions(ctypes.Structure):
    _fields_ = [("capacity", _i64), ("count", _i64), ("label", _vp), ("size", _vp),
                ("voxels", _vp), ("centroid_x", _vp), ("centroid_y", _vp), ("centroid_z", _vp),
                ("structureness", _vp), ("blobness", _vp), ("max_normalized", _vp),
                ("hessian", _vp), ("intensity_sum", _vp), ("intensity_max", _vp)]


class Params(ctypes.Structure):
    _fields_ = [("width", _i), ("height", _i), ("depth", _i), ("radius_xy", _i),
                ("radius_z", _i), ("min_intensity", _f), ("sigma_xy1", _f), ("sigma_z1", _f),
                ("sigma_xy2", _f), ("sigma_z2", _f), ("truncate", _d), ("gamma", _f),
                ("is_ave_mode", _i), ("precision", _i), ("minor_band", _f)]


class Tile(ctypes.Structure):
    _fields_ = [("image", _vp), ("width", _i), ("height", _i), ("depth", _i)]


def _sig(name, args):
    fn = getattr(lib, name)
    fn.argtypes = args
    fn.restype = _i
    return fn


hdog_create = _sig("hdog_create", [ctypes.POINTER(_vp), _i])
lib.hdog_destroy.argtypes = [_vp]
lib.hdog_destroy.restype = None
hdog_destroy = lib.hdog_destroy
lib.hdog_last_error.argtypes = [_vp]
lib.hdog_last_error.restype = ctypes.c_char_p
hdog_last_error = lib.hdog_last_error
hdog_release_scratch = _sig("hdog_release_scratch", [_vp])
hdog_normalize3d = _sig("hdog_normalize3d", [_vp, _vp, _vp, _vp, _f, _i, _i, _i, _i, _i])
hdog_normalize3d_ed = _sig("hdog_normalize3d_ed", [_vp, _vp, _vp, _vp, _vp, _vp, _f, _i, _i, _i,
                                                   _i, _i])
hdog_init_gaussian3d = _sig("hdog_init_gaussian3d", [_vp, _f, _f, _f, _f, _d, ctypes.POINTER(_i)])
hdog_dog3d = _sig("hdog_dog3d", [_vp, _vp, _vp, _vp, _i, _i, _i, _f])
hdog_hessian_pd = _sig("hdog_hessian_pd", [_vp, _vp, _vp, _vp, _vp, _i, _i, _i])
hdog_ccl3d = _sig("hdog_ccl3d", [_vp, _vp, _vp, _vp, _i, _i, _i])
hdog_ccl3d_noalloc = _sig("hdog_ccl3d_noalloc", [_vp, _vp, _vp, _vp, _i, _i, _i])
hdog_regional_features = _sig("hdog_regional_features", [_vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i,
                                                         _i, ctypes.POINTER(_i64)])
hdog_regions_download = _sig("hdog_regions_download", [_vp, _vp, ctypes.POINTER(Regions)])
hdog_process_substack = _sig("hdog_process_substack", [_vp, _vp, ctypes.POINTER(Params), _vp, _i,
                                                       ctypes.POINTER(Regions)])
hdog_run_device = _sig("hdog_run_device", [_vp, _vp, ctypes.POINTER(Params), _vp, _vp, _vp, _vp,
                                           _vp, _vp, ctypes.POINTER(_i64)])
hdog_host_alloc = _sig("hdog_host_alloc", [_vp, ctypes.c_size_t, ctypes.POINTER(_vp)])
hdog_host_free = _sig("hdog_host_free", [_vp, _vp])
hdog_substack_upload = _sig("hdog_substack_upload", [_vp, ctypes.POINTER(Params), _vp, _i, _i])
hdog_substack_run = _sig("hdog_substack_run", [_vp, _vp, ctypes.POINTER(Params), _i,
                                               ctypes.POINTER(Regions)])
hdog_substack_upload_overlap = _sig("hdog_substack_upload_overlap",
                                    [_vp, ctypes.POINTER(Params), _i, _i, _vp, _i, _i])
hdog_autotune = _sig("hdog_autotune", [_vp, _vp, ctypes.POINTER(Params), _vp, ctypes.c_char_p,
                                       ctypes.POINTER(_i), ctypes.POINTER(_i)])
hdog_batch_plan = _sig("hdog_batch_plan", [_vp, ctypes.POINTER(Params), ctypes.POINTER(Tile), _i,
                                           _i64, _i64])
hdog_batch_run = _sig("hdog_batch_run", [_vp, _vp])
hdog_batch_tile_info = _sig("hdog_batch_tile_info", [_vp, _i, ctypes.POINTER(_i64),
                                                     ctypes.POINTER(_i)])
hdog_batch_download = _sig("hdog_batch_download", [_vp, _vp, _i, ctypes.POINTER(Regions)])
hdog_batch_release = _sig("hdog_batch_release", [_vp])
hdog_synth_volume = _sig("hdog_synth_volume", [_vp, _vp, _vp, _vp, _i, _i, _i, ctypes.c_uint64,
                                               _d, _d, _d, _vp, _i64])

hdog_soft_normalize_forward = _sig("hdog_soft_normalize_forward", [_vp, _vp, _vp, _vp, _vp, _vp,
                                                                   _f, _i, _i, _i, _i, _i])
hdog_soft_normalize_backward = _sig("hdog_soft_normalize_backward", [_vp, _vp, _vp, _vp, _vp, _vp,
                                                                     _vp, _f, _i, _i, _i, _i, _i])
hdog_set_profiling = _sig("hdog_set_profiling", [_vp, _i])
hdog_stage_times = _sig("hdog_stage_times", [_vp, ctypes.POINTER(_d), ctypes.POINTER(ctypes.c_char_p), _i])

_d3 = ctypes.c_double * 3
_i3 = ctypes.c_int * 3
hdog_soft_forward = _sig("hdog_soft_forward", [_vp, _vp, _vp, _vp, _i, _i, _i, _d3, _d3, _i3, _i3,
                                               _d, _d, _vp, _vp, ctypes.POINTER(_d),
                                               ctypes.POINTER(_i64)])
hdog_soft_backward = _sig("hdog_soft_backward", [_vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _d3, _d3,
                                                 _i3, _i3, _d, _d, _i64, _d, _vp, _d3, _d3,
                                                 ctypes.POINTER(_d)])

EXPORTED = ["hdog_create", "hdog_destroy", "hdog_last_error", "hdog_release_scratch",
            "hdog_normalize3d", "hdog_normalize3d_ed", "hdog_init_gaussian3d", "hdog_dog3d", "hdog_hessian_pd",
            "hdog_ccl3d", "hdog_ccl3d_noalloc", "hdog_regional_features", "hdog_regions_download",
            "hdog_process_substack", "hdog_run_device", "hdog_synth_volume",
            "hdog_set_profiling", "hdog_stage_times", "hdog_soft_forward", "hdog_soft_backward",
            "hdog_soft_normalize_forward", "hdog_soft_normalize_backward",
            "hdog_autotune", "hdog_batch_plan", "hdog_batch_run", "hdog_batch_tile_info", "hdog_batch_download",
            "hdog_batch_release", "hdog_host_alloc", "hdog_host_free", "hdog_substack_upload",
            "hdog_substack_run", "hdog_substack_upload_overlap"]
