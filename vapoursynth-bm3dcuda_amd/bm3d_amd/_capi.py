"""ctypes binding of libbm3d_mi355x.so (include/bm3d_mi355x.h).

The library is the product: there is no Python or CPU fallback. Loading
fails loudly when the shared object is missing.
"""
from __future__ import annotations

import ctypes
import os

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_ROOT, "lib", "libbm3d_mi355x.so")

BM3D_OK = 0
BM3D_EINVAL = 1
BM3D_EHIP = 2
BM3D_EUNSUPPORTED = 3

PRECISION_FP32 = 0
PRECISION_MIXED = 1

c_ll = ctypes.c_longlong
c_int = ctypes.c_int
c_float = ctypes.c_float
c_void_p = ctypes.c_void_p


class Desc(ctypes.Structure):
    _fields_ = [
        ("width", c_int), ("height", c_int), ("stride", c_int),
        ("sigma", c_float * 3),
        ("block_step", c_int), ("bm_range", c_int), ("radius", c_int),
        ("ps_num", c_int), ("ps_range", c_int),
        ("chroma", c_int), ("final_", c_int), ("extractor_exp", c_int),
        ("precision", c_int),
        ("bm_error", c_int), ("transform_2d", c_int), ("transform_1d", c_int),
        ("mixed_temporal_keys", c_int),
    ]


class Clip(ctypes.Structure):
    _fields_ = [("data", c_void_p), ("frame_stride", c_ll), ("plane_stride", c_ll),
                ("num_frames", c_int)]


class VideoInfo(ctypes.Structure):
    _fields_ = [("color_family", c_int), ("sample_type", c_int), ("bits_per_sample", c_int),
                ("subsampling_w", c_int), ("subsampling_h", c_int), ("num_planes", c_int),
                ("width", c_int), ("height", c_int), ("num_frames", c_int)]


class Args(ctypes.Structure):
    _fields_ = [
        ("n_sigma", c_int), ("sigma", ctypes.c_double * 3),
        ("n_block_step", c_int), ("block_step", c_ll * 3),
        ("n_bm_range", c_int), ("bm_range", c_ll * 3),
        ("n_radius", c_int), ("radius", c_ll),
        ("n_ps_num", c_int), ("ps_num", c_ll * 3),
        ("n_ps_range", c_int), ("ps_range", c_ll * 3),
        ("n_chroma", c_int), ("chroma", c_ll),
        ("n_device_id", c_int), ("device_id", c_ll),
        ("n_fast", c_int), ("fast", c_ll),
        ("n_extractor_exp", c_int), ("extractor_exp", c_ll),
        ("n_zero_init", c_int), ("zero_init", c_ll),
        ("n_precision", c_int), ("precision", c_ll),
        ("n_bm_error_s", c_int), ("bm_error_s", ctypes.c_char_p * 3),
        ("n_transform_2d_s", c_int), ("transform_2d_s", ctypes.c_char_p * 3),
        ("n_transform_1d_s", c_int), ("transform_1d_s", ctypes.c_char_p * 3),
    ]


class Frames(ctypes.Structure):
    _fields_ = [("data", c_void_p * 3), ("frame_stride", c_ll * 3), ("stride", c_int * 3)]


class HostFrame(ctypes.Structure):
    _fields_ = [("data", c_void_p * 3), ("stride", c_int * 3)]


class Tuning(ctypes.Structure):
    _fields_ = [("match_rows", c_int), ("filter_tile_x", c_int), ("filter_tile_y", c_int),
                ("filter_waves", c_int), ("batch_frames", c_int), ("chunks", c_int), ("match_width", c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchJob(ctypes.Structure):
    _fields_ = [("desc", Desc), ("passes", c_int), ("src", Clip), ("basic", c_void_p), ("out", c_void_p)]


class BM3DError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"libbm3d_mi355x.so not built ({LIB_PATH}); run "
            "vapoursynth-bm3dcuda_amd/build_lib.py")
    L = ctypes.CDLL(LIB_PATH)
    P = c_void_p
    L.bm3d_plan_create.argtypes = [ctypes.POINTER(Desc), ctypes.POINTER(P)]
    L.bm3d_plan_create.restype = c_int
    L.bm3d_plan_destroy.argtypes = [P]
    L.bm3d_plan_destroy.restype = None
    L.bm3d_stack_size.argtypes = [P]
    L.bm3d_stack_size.restype = c_ll
    L.bm3d_num_blocks.argtypes = [P]
    L.bm3d_num_blocks.restype = c_ll
    L.bm3d_plan_matching_kernel.argtypes = [P]
    L.bm3d_plan_matching_kernel.restype = c_int
    L.bm3d_plan_filter_kernel.argtypes = [P]
    L.bm3d_plan_filter_kernel.restype = c_int
    L.bm3d_plan_select_kernels.argtypes = [P, c_int, c_int]
    L.bm3d_plan_select_kernels.restype = c_int
    L.bm3d_plan_get_tuning.argtypes = [P, ctypes.POINTER(Tuning)]
    L.bm3d_plan_get_tuning.restype = c_int
    L.bm3d_plan_set_tuning.argtypes = [P, ctypes.POINTER(Tuning)]
    L.bm3d_plan_set_tuning.restype = c_int
    L.bm3d_plan_autotune.argtypes = [P, ctypes.POINTER(Clip), ctypes.POINTER(Clip), c_int, c_int, P]
    L.bm3d_plan_autotune.restype = c_int
    L.bm3d_batch_create.argtypes = [ctypes.POINTER(BatchJob), c_int, c_ll, ctypes.POINTER(P)]
    L.bm3d_batch_create.restype = c_int
    L.bm3d_batch_execute.argtypes = [P, P]
    L.bm3d_batch_execute.restype = c_int
    L.bm3d_batch_get_tuning.argtypes = [P, ctypes.POINTER(Tuning)]
    L.bm3d_batch_get_tuning.restype = c_int
    L.bm3d_batch_set_tuning.argtypes = [P, ctypes.POINTER(Tuning)]
    L.bm3d_batch_set_tuning.restype = c_int
    L.bm3d_batch_autotune.argtypes = [P, P]
    L.bm3d_batch_autotune.restype = c_int
    L.bm3d_batch_info.argtypes = [P, ctypes.POINTER(c_int), ctypes.POINTER(c_int), ctypes.POINTER(c_ll)]
    L.bm3d_batch_info.restype = c_int
    L.bm3d_batch_destroy.argtypes = [P]
    L.bm3d_batch_destroy.restype = None
    L.bm3d_execute.argtypes = [P, P, c_ll, c_ll, ctypes.POINTER(Clip), ctypes.POINTER(Clip),
                               c_int, c_int, P]
    L.bm3d_execute.restype = c_int
    L.bm3d_execute_vagg.argtypes = [P, P, c_ll, c_ll, c_int, ctypes.POINTER(Clip), ctypes.POINTER(Clip),
                                    c_int, c_int, P]
    L.bm3d_execute_vagg.restype = c_int
    L.bm3d_execute_window.argtypes = [P, P, P, P]
    L.bm3d_execute_window.restype = c_int
    L.bm3d_execute_noalloc.argtypes = [P, P, c_ll, c_ll, ctypes.POINTER(Clip), ctypes.POINTER(Clip),
                                       c_int, c_int, P]
    L.bm3d_execute_noalloc.restype = c_int
    L.bm3d_match.argtypes = [P, P, ctypes.POINTER(Clip), ctypes.POINTER(Clip), c_int, c_int, P]
    L.bm3d_match.restype = c_int
    L.bm3d_aggregate.argtypes = [P, c_ll, c_ll, c_int, P, c_ll, c_ll, c_int, c_int, c_int, c_int,
                                 c_int, P]
    L.bm3d_aggregate.restype = c_int
    L.bm3d_vaggregate.argtypes = [P, c_ll, c_ll, c_int, P, c_ll, c_ll, c_int, c_int, c_int, c_int,
                                  c_int, c_int, c_int, c_int, c_int, P]
    L.bm3d_vaggregate.restype = c_int
    L.bm3d_profile_enable.argtypes = [c_int]
    L.bm3d_profile_enable.restype = c_int
    L.bm3d_profile_reset.argtypes = []
    L.bm3d_profile_reset.restype = c_int
    L.bm3d_profile_read.argtypes = [c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(c_ll)]
    L.bm3d_profile_read.restype = c_int
    L.bm3d_count_candidates.argtypes = [P, ctypes.POINTER(Clip), ctypes.POINTER(Clip), c_int, c_int,
                                        ctypes.POINTER(ctypes.c_ulonglong), P]
    L.bm3d_count_candidates.restype = c_int
    PV = ctypes.POINTER(VideoInfo)
    PA = ctypes.POINTER(Args)
    PF = ctypes.POINTER(Frames)
    for name in ("bm3d_filter_create", "bm3dv2_filter_create"):
        fn = getattr(L, name)
        fn.argtypes = [PA, PV, PV, ctypes.POINTER(P)]
        fn.restype = c_int
    L.bm3d_vaggregate_create.argtypes = [ctypes.POINTER(c_ll), c_int, PV, PV, ctypes.POINTER(P)]
    L.bm3d_vaggregate_create.restype = c_int
    L.bm3d_filter_info.argtypes = [P, PV, ctypes.POINTER(c_int), ctypes.POINTER(c_int)]
    L.bm3d_filter_info.restype = c_int
    L.bm3d_filter_get_frames.argtypes = [P, PF, PF, PF, c_int, c_int, P]
    L.bm3d_filter_get_frames.restype = c_int
    L.bm3d_filter_destroy.argtypes = [P]
    L.bm3d_filter_destroy.restype = None
    L.bm3d_filter_host_range.argtypes = [P, c_int, c_int, ctypes.POINTER(c_int), ctypes.POINTER(c_int)]
    L.bm3d_filter_host_range.restype = c_int
    PH = ctypes.POINTER(HostFrame)
    L.bm3d_filter_get_frames_host.argtypes = [P, PH, PH, c_int, PH, c_int, c_int]
    L.bm3d_filter_get_frames_host.restype = c_int
    L.bm3d_debug_inject_failure.argtypes = [c_int, c_ll]
    L.bm3d_debug_inject_failure.restype = c_int
    L.bm3d_last_error.argtypes = []
    L.bm3d_last_error.restype = ctypes.c_char_p
    L.bm3d_version.argtypes = []
    L.bm3d_version.restype = ctypes.c_char_p
    _lib = L
    return L


def check(rc):
    if rc != BM3D_OK:
        msg = lib().bm3d_last_error().decode()
        raise BM3DError(rc, msg or f"bm3d error {rc}")
    return rc


# Exported names of include/bm3d_mi355x.h (checked by the CPU test suite).
EXPORTS = (
    "bm3d_plan_create", "bm3d_plan_destroy", "bm3d_stack_size", "bm3d_execute", "bm3d_execute_vagg",
    "bm3d_execute_window", "bm3d_aggregate", "bm3d_vaggregate", "bm3d_match",
    "bm3d_num_blocks", "bm3d_plan_matching_kernel", "bm3d_plan_filter_kernel", "bm3d_plan_select_kernels",
    "bm3d_plan_get_tuning", "bm3d_plan_set_tuning", "bm3d_plan_autotune", "bm3d_last_error", "bm3d_version",
    "bm3d_profile_enable",
    "bm3d_profile_reset", "bm3d_profile_read", "bm3d_count_candidates",
    "bm3d_batch_create", "bm3d_batch_execute", "bm3d_batch_get_tuning", "bm3d_batch_set_tuning",
    "bm3d_batch_autotune", "bm3d_batch_info", "bm3d_batch_destroy",
    "bm3d_filter_create", "bm3d_vaggregate_create", "bm3dv2_filter_create", "bm3d_filter_info",
    "bm3d_filter_get_frames", "bm3d_filter_destroy", "bm3d_filter_host_range", "bm3d_filter_get_frames_host",
    "bm3d_execute_noalloc", "bm3d_debug_inject_failure",
)

INJECT_DEVICE_ALLOC = 1
INJECT_HOST_ALLOC = 2
INJECT_LAUNCH = 3

KERNEL_MATCH = 0
KERNEL_FILTER = 1


def profile_enable(on=True):
    check(lib().bm3d_profile_enable(int(bool(on))))


def profile_reset():
    check(lib().bm3d_profile_reset())


def profile_read(kernel):
    """(total milliseconds, launches) of one kernel since the last reset."""
    ms = ctypes.c_double()
    n = c_ll()
    check(lib().bm3d_profile_read(int(kernel), ctypes.byref(ms), ctypes.byref(n)))
    return ms.value, n.value
