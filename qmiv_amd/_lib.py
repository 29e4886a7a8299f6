"""ctypes bindings of libivqm_hip.so (include/ivqm_hip.h).

Loading fails loudly when the library is missing: there is no CPU fallback on
the product path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libivqm_hip.so")

IVQM_M_PSNR = 1 << 0
IVQM_M_WSPSNR = 1 << 1
IVQM_M_IVPSNR = 1 << 2
IVQM_M_SSIM = 1 << 3
IVQM_M_MSSSIM = 1 << 4
IVQM_M_IVSSIM = 1 << 5
IVQM_M_IVMSSSIM = 1 << 6
IVQM_M_MSE = 1 << 7
IVQM_M_ALL = 0xFF

# Every entry point include/*.h declares (checked by tests/test_abi.py).
EXPORTED = [
    "ivqm_config_init", "ivqm_create", "ivqm_destroy", "ivqm_set_stream", "ivqm_strerror", "ivqm_last_error",
    "ivqm_upload", "ivqm_bind", "ivqm_gcd", "ivqm_ssd", "ivqm_mse", "ivqm_psnr", "ivqm_wspsnr", "ivqm_ivpsnr",
    "ivqm_gen_scp", "ivqm_ssim", "ivqm_msssim", "ivqm_ivssim", "ivqm_ivmsssim", "ivqm_frame",
    "ivqm_frame_enqueue", "ivqm_frame_collect", "ivqm_get_rows", "ivqm_get_scp", "ivqm_enable_best_map",
    "ivqm_get_best_map", "ivqm_upsample420", "ivqm_downsample", "ivqm_profile", "ivqm_upload_scp",
    "ivqm_bind_mask", "ivqm_upload_mask", "ivqm_batch_enqueue_masked", "ivqm_convert_colour",
    "ivqm_batch_create", "ivqm_batch_destroy", "ivqm_batch_set_stream", "ivqm_batch_set_tuning",
    "ivqm_batch_get_tuning", "ivqm_batch_autotune", "ivqm_batch_run", "ivqm_batch_enqueue", "ivqm_batch_collect",
    "ivqm_batch_profile", "ivqm_batch_run_noalloc",
    "ivqm_server_config_init", "ivqm_server_create", "ivqm_server_destroy", "ivqm_server_add_sequence",
    "ivqm_server_remove_sequence", "ivqm_server_submit", "ivqm_server_poll", "ivqm_server_drain",
    "ivqm_server_get_stats",
]
MAX_JOB_FRAMES = 8
NOALLOC_MAX_PAIRS = 16
NOALLOC_MAX_ROWS = 32768


class Config(C.Structure):
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32), ("bit_depth", C.c_int32), ("search_range", C.c_int32),
        ("cmp_weights_search", C.c_int32 * 4), ("cmp_weights_average", C.c_int32 * 4),
        ("unnoticeable_coef", C.c_float * 4), ("erp", C.c_int32), ("lon_range_deg", C.c_int32),
        ("lat_range_deg", C.c_int32), ("ssim_window", C.c_int32), ("ssim_stride", C.c_int32),
        ("fake_vals_for_exact", C.c_int32), ("gcd_sd_wrap32", C.c_int32), ("search_tile_rows", C.c_int32),
        ("ssim_wide_sums", C.c_int32), ("search_mode_cap", C.c_int32), ("ssim_mode", C.c_int32),
        ("ssim_border", C.c_int32), ("reserved", C.c_int32 * 3),
    ]


class Picture(C.Structure):
    _fields_ = [("plane", C.c_void_p * 3), ("stride", C.c_int32 * 3), ("chroma_420", C.c_int32),
                ("bytes_per_sample", C.c_int32)]


class Mask(C.Structure):
    _fields_ = [("plane", C.c_void_p), ("stride", C.c_int32), ("bytes_per_sample", C.c_int32),
                ("bit_depth", C.c_int32), ("reserved", C.c_int32)]


class Pair(C.Structure):
    _fields_ = [("p0", Picture), ("p1", Picture), ("width", C.c_int32), ("height", C.c_int32),
                ("bit_depth", C.c_int32)]


class FrameResult(C.Structure):
    _fields_ = [
        ("metrics", C.c_uint32), ("gcd_r2t", C.c_int32 * 4), ("ssd", C.c_uint64 * 4), ("exact", C.c_int32 * 4),
        ("mse", C.c_double * 4), ("psnr", C.c_double * 4), ("wspsnr", C.c_double * 4),
        ("psnr_pic", C.c_double), ("wspsnr_pic", C.c_double),
        ("ivpsnr", C.c_double), ("ivpsnr_r2t", C.c_double), ("ivpsnr_t2r", C.c_double),
        ("ssim", C.c_double * 4), ("msssim", C.c_double * 4), ("ssim_pic", C.c_double), ("msssim_pic", C.c_double),
        ("ivssim", C.c_double), ("ivssim_r2t", C.c_double), ("ivssim_t2r", C.c_double),
        ("ivmsssim", C.c_double), ("ivmsssim_r2t", C.c_double), ("ivmsssim_t2r", C.c_double),
    ]

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if hasattr(v, "__len__") else v
        return out


class ServerConfig(C.Structure):
    _fields_ = [("params", Config), ("memory_budget", C.c_uint64), ("max_batch_pairs", C.c_int32),
                ("max_inflight", C.c_int32), ("reserved", C.c_int32 * 8)]


class SequenceDesc(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("bit_depth", C.c_int32), ("chroma_420", C.c_int32),
                ("bytes_per_sample", C.c_int32), ("num_frames", C.c_int32), ("planes", C.POINTER(C.c_void_p)),
                ("strides", C.POINTER(C.c_int32))]


class Job(C.Structure):
    _fields_ = [("tag", C.c_uint64), ("seq_id", C.c_int32), ("first_frame", C.c_int32), ("num_frames", C.c_int32),
                ("metrics", C.c_uint32), ("planes_on_device", C.c_int32), ("reserved", C.c_int32),
                ("planes", C.POINTER(C.c_void_p)), ("strides", C.POINTER(C.c_int32)),
                ("mask_planes", C.POINTER(C.c_void_p)), ("mask_strides", C.POINTER(C.c_int32)),
                ("mask_bit_depth", C.c_int32), ("mask_bytes_per_sample", C.c_int32)]


class JobResult(C.Structure):
    _fields_ = [("tag", C.c_uint64), ("status", C.c_int32), ("num_frames", C.c_int32), ("latency_us", C.c_double),
                ("queue_us", C.c_double), ("frames", FrameResult * MAX_JOB_FRAMES), ("stage_us", C.c_double),
                ("wait_us", C.c_double), ("device_us", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("jobs_submitted", C.c_uint64), ("jobs_completed", C.c_uint64), ("jobs_failed", C.c_uint64),
                ("pairs_completed", C.c_uint64), ("batches", C.c_uint64), ("budget_bytes", C.c_uint64),
                ("budget_in_use", C.c_uint64), ("budget_peak", C.c_uint64), ("resident_bytes", C.c_uint64),
                ("max_batch_fill", C.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


_lib = None


def load():
    """Loads libivqm_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m qmiv_amd.build`")
    # One HIP runtime per process: torch ships its own libamdhip64.so.7. Loading
    # torch first lets the library's libamdhip64.so.7 dependency resolve to that
    # copy; loaded the other way round, torch finds no HIP device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    L.ivqm_strerror.restype = C.c_char_p
    L.ivqm_last_error.restype = C.c_char_p
    L.ivqm_create.argtypes = [C.POINTER(Config), C.POINTER(C.c_void_p)]
    L.ivqm_destroy.argtypes = [C.c_void_p]
    L.ivqm_set_stream.argtypes = [C.c_void_p, C.c_void_p]
    for fn in ("ivqm_upload", "ivqm_bind", "ivqm_upload_scp"):
        getattr(L, fn).argtypes = [C.c_void_p, C.c_int32, C.POINTER(Picture)]
    L.ivqm_gcd.argtypes = [C.c_void_p, C.POINTER(C.c_int32)]
    L.ivqm_ssd.argtypes = [C.c_void_p, C.POINTER(C.c_uint64)]
    for fn in ("ivqm_mse", "ivqm_psnr", "ivqm_wspsnr", "ivqm_ssim", "ivqm_msssim"):
        getattr(L, fn).argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    L.ivqm_ivpsnr.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_double), C.POINTER(C.c_double),
                              C.POINTER(C.c_double)]
    L.ivqm_gen_scp.argtypes = [C.c_void_p, C.POINTER(C.c_int32)]
    for fn in ("ivqm_ivssim", "ivqm_ivmsssim"):
        getattr(L, fn).argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.ivqm_frame.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(FrameResult)]
    L.ivqm_frame_enqueue.argtypes = [C.c_void_p, C.c_uint32, C.c_int32]
    L.ivqm_frame_collect.argtypes = [C.c_void_p, C.c_int32, C.POINTER(FrameResult)]
    L.ivqm_get_rows.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_uint64)]
    L.ivqm_get_scp.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.POINTER(C.c_uint16))]
    L.ivqm_enable_best_map.argtypes = [C.c_void_p, C.c_int32]
    L.ivqm_get_best_map.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_uint8)]
    L.ivqm_upsample420.argtypes = [C.POINTER(Picture), C.c_int32, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p]
    L.ivqm_downsample.argtypes = [C.POINTER(Picture), C.c_int32, C.c_int32, C.c_void_p, C.c_void_p]
    L.ivqm_convert_colour.argtypes = [C.POINTER(Picture), C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                      C.c_void_p, C.c_int32, C.c_void_p]
    L.ivqm_profile.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_double)]
    L.ivqm_batch_create.argtypes = [C.POINTER(Config), C.POINTER(C.c_void_p)]
    L.ivqm_batch_destroy.argtypes = [C.c_void_p]
    L.ivqm_batch_set_stream.argtypes = [C.c_void_p, C.c_void_p]
    L.ivqm_batch_set_tuning.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]
    L.ivqm_batch_get_tuning.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    L.ivqm_batch_autotune.argtypes = [C.c_void_p, C.POINTER(Pair), C.c_int32, C.c_uint32, C.POINTER(C.c_int32),
                                      C.POINTER(C.c_int32)]
    L.ivqm_batch_run.argtypes = [C.c_void_p, C.POINTER(Pair), C.c_int32, C.c_uint32, C.POINTER(FrameResult)]
    L.ivqm_batch_enqueue.argtypes = [C.c_void_p, C.POINTER(Pair), C.c_int32, C.c_uint32, C.c_int32]
    L.ivqm_batch_collect.argtypes = [C.c_void_p, C.c_int32, C.POINTER(FrameResult)]
    L.ivqm_bind_mask.argtypes = [C.c_void_p, C.POINTER(Mask)]
    L.ivqm_upload_mask.argtypes = [C.c_void_p, C.POINTER(Mask)]
    L.ivqm_batch_enqueue_masked.argtypes = [C.c_void_p, C.POINTER(Pair), C.POINTER(Mask), C.c_int32, C.c_uint32,
                                            C.c_int32]
    L.ivqm_batch_profile.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_double)]
    L.ivqm_batch_run_noalloc.argtypes = [C.POINTER(Config), C.POINTER(Pair), C.c_int32, C.c_uint32,
                                         C.POINTER(FrameResult), C.c_void_p]
    L.ivqm_server_config_init.argtypes = [C.POINTER(ServerConfig)]
    L.ivqm_server_create.argtypes = [C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]
    L.ivqm_server_destroy.argtypes = [C.c_void_p]
    L.ivqm_server_add_sequence.argtypes = [C.c_void_p, C.POINTER(SequenceDesc), C.POINTER(C.c_int32)]
    L.ivqm_server_remove_sequence.argtypes = [C.c_void_p, C.c_int32]
    L.ivqm_server_submit.argtypes = [C.c_void_p, C.POINTER(Job)]
    L.ivqm_server_poll.argtypes = [C.c_void_p, C.POINTER(JobResult), C.c_int32, C.POINTER(C.c_int32), C.c_int32]
    L.ivqm_server_drain.argtypes = [C.c_void_p]
    L.ivqm_server_get_stats.argtypes = [C.c_void_p, C.POINTER(ServerStats)]
    _lib = L
    return L


class IvqmError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"ivqm status {status}: {msg}")
        self.status = status


def check(status: int) -> None:
    if status != 0:
        L = load()
        raise IvqmError(status, L.ivqm_last_error().decode(errors="replace"))
