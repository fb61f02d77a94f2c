"""ctypes binding of the C ABI in ``include/occastft_hip.h``.

The library ``occastft_amd/libocca_stft_hip.so`` is the product: there is no
CPU fallback.  If it is missing, every entry point raises ``NativeMissing``.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libocca_stft_hip.so")

# status codes (occastft_hip.h)
OK = 0
ERR_INVALID_ARG = 1
ERR_UNSUPPORTED = 2
ERR_OPTION = 3
ERR_CAPACITY = 4
ERR_HIP = 5
ERR_NO_DEVICE = 6
ERR_TOO_LARGE = 7
ERR_TIMEOUT = 8
ERR_SHUTDOWN = 9
ERR_NOT_PREPARED = 10

# OSTFT_ABI_VERSION of the header these bindings follow
ABI_VERSION = 2

# windows
WIN_NONE, WIN_HANNING, WIN_HAMMING, WIN_BLACKMAN, WIN_NUTTALL = 0, 1, 2, 3, 4
WIN_BLACKMAN_NUTTALL, WIN_BLACKMAN_HARRIS, WIN_FLATTOP, WIN_GAUSSIAN = 5, 6, 7, 8

# outputs
OUT_COMPAT_POWER, OUT_HALF_COMPLEX, OUT_MAGNITUDE, OUT_POWER, OUT_HALF_SPECTRUM = 0, 1, 2, 3, 4

# flags
FLAG_REMOVE_DC = 0x1
FLAG_EXACT_ODD = 0x2
FLAG_EXACT_WINDOW = 0x4
FLAG_SIGMA_EXACT = 0x8
FLAG_MFMA_FP16 = 0x10
FLAG_MFMA_BF16 = 0x20

# every symbol the header declares
EXPORTS = (
    "ostft_parse_options",
    "ostft_get_geometry",
    "ostft_create",
    "ostft_destroy",
    "ostft_run_device",
    "ostft_run_device_noalloc",
    "ostft_activate",
    "ostft_strerror",
    "ostft_last_error",
    "ostft_abi_version",
    "ostft_window_coefficients",
    "ostft_run_batch_device",
    "ostft_run_strided_device",
    "ostft_activate_geometry",
    "ostft_batch_config",
    "ostft_autotune_batch",
    "ostft_engine_variant",
    "ostft_scratch_budget",
    "ostft_server_create",
    "ostft_server_destroy",
    "ostft_server_submit",
    "ostft_server_wait",
    "ostft_server_stats_get",
    "ostft_server_last_error",
)


class NativeMissing(RuntimeError):
    """The HIP engine library is not built (run occastft_amd.build)."""


class Params(ctypes.Structure):
    _fields_ = [
        ("exp", ctypes.c_int32),
        ("overlap", ctypes.c_float),
        ("window", ctypes.c_int32),
        ("sigma", ctypes.c_float),
        ("output", ctypes.c_int32),
        ("flags", ctypes.c_uint32),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class Geometry(ctypes.Structure):
    _fields_ = [
        ("full_size", ctypes.c_uint64),
        ("window_size", ctypes.c_uint64),
        ("frames", ctypes.c_uint64),
        ("hop", ctypes.c_uint64),
        ("out_full", ctypes.c_uint64),
        ("out_half", ctypes.c_uint64),
        ("out_floats", ctypes.c_uint64),
        ("compute_frames", ctypes.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


class Job(ctypes.Structure):
    _fields_ = [
        ("in_ptr", ctypes.c_void_p),
        ("full_size", ctypes.c_uint64),
        ("out_ptr", ctypes.c_void_p),
        ("out_cap", ctypes.c_uint64),
        ("params", Params),
    ]


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("arena_bytes", ctypes.c_uint64),
        ("max_batch", ctypes.c_uint32),
        ("max_inflight", ctypes.c_uint32),
        ("use_graphs", ctypes.c_int32),
        ("autotune", ctypes.c_int32),
        ("autotune_like", ctypes.c_void_p),
        ("autotune_samples", ctypes.c_uint64),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_uint64),
        ("completed", ctypes.c_uint64),
        ("failed", ctypes.c_uint64),
        ("batches", ctypes.c_uint64),
        ("frames", ctypes.c_uint64),
        ("bytes_in", ctypes.c_uint64),
        ("bytes_out", ctypes.c_uint64),
        ("latency_p50_ms", ctypes.c_double),
        ("latency_p99_ms", ctypes.c_double),
        ("latency_max_ms", ctypes.c_double),
        ("latency_mean_ms", ctypes.c_double),
        ("mean_batch", ctypes.c_double),
        ("span_s", ctypes.c_double),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib = None


def lib() -> ctypes.CDLL:
    """Load the engine; raise NativeMissing if it was never built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NativeMissing(f"{LIB_PATH} not found; build it with `python -m occastft_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    fp = ctypes.POINTER(ctypes.c_float)
    L.ostft_parse_options.restype = ctypes.c_int
    L.ostft_parse_options.argtypes = [ctypes.c_char_p, ctypes.c_int32, ctypes.c_float, ctypes.POINTER(Params)]
    L.ostft_get_geometry.restype = ctypes.c_int
    L.ostft_get_geometry.argtypes = [ctypes.POINTER(Params), ctypes.c_uint64, ctypes.POINTER(Geometry)]
    L.ostft_create.restype = ctypes.c_int
    L.ostft_create.argtypes = [ctypes.c_int32, ctypes.POINTER(vp)]
    L.ostft_destroy.restype = None
    L.ostft_destroy.argtypes = [vp]
    L.ostft_run_device.restype = ctypes.c_int
    L.ostft_run_device.argtypes = [vp, ctypes.POINTER(Params), vp, ctypes.c_uint64, vp, ctypes.c_uint64, vp]
    L.ostft_run_device_noalloc.restype = ctypes.c_int
    L.ostft_run_device_noalloc.argtypes = [vp, ctypes.POINTER(Params), vp, ctypes.c_uint64, vp, ctypes.c_uint64, vp]
    L.ostft_activate.restype = ctypes.c_int
    L.ostft_activate.argtypes = [vp, fp, ctypes.c_uint64, ctypes.c_int32, ctypes.c_float, ctypes.c_char_p, fp,
                                 ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
    L.ostft_strerror.restype = ctypes.c_char_p
    L.ostft_strerror.argtypes = [ctypes.c_int]
    L.ostft_last_error.restype = ctypes.c_char_p
    L.ostft_last_error.argtypes = [vp]
    L.ostft_abi_version.restype = ctypes.c_int
    L.ostft_abi_version.argtypes = []
    if L.ostft_abi_version() != ABI_VERSION:
        raise NativeMissing(f"{LIB_PATH} has ABI version {L.ostft_abi_version()}, these bindings {ABI_VERSION}: "
                            "rebuild it with `python -m occastft_amd.build`")
    L.ostft_window_coefficients.restype = ctypes.c_int
    L.ostft_window_coefficients.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_float, fp]
    i32p = ctypes.POINTER(ctypes.c_int32)
    u64p = ctypes.POINTER(ctypes.c_uint64)
    L.ostft_run_batch_device.restype = ctypes.c_int
    L.ostft_run_batch_device.argtypes = [vp, ctypes.POINTER(Job), ctypes.c_uint32, vp]
    L.ostft_run_strided_device.restype = ctypes.c_int
    L.ostft_run_strided_device.argtypes = [vp, ctypes.POINTER(Params), vp, ctypes.c_uint64, ctypes.c_uint64,
                                           ctypes.c_uint32, vp, ctypes.c_uint64, vp]
    L.ostft_activate_geometry.restype = ctypes.c_int
    L.ostft_activate_geometry.argtypes = [vp, ctypes.POINTER(Params), fp, ctypes.c_uint64, ctypes.c_uint64,
                                          ctypes.c_uint64, fp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
    L.ostft_batch_config.restype = ctypes.c_int
    L.ostft_batch_config.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, i32p]
    L.ostft_autotune_batch.restype = ctypes.c_int
    L.ostft_autotune_batch.argtypes = [vp, ctypes.c_int32, ctypes.c_uint64, ctypes.POINTER(Params), i32p, i32p,
                                       i32p, ctypes.POINTER(ctypes.c_double)]
    L.ostft_engine_variant.restype = ctypes.c_int
    L.ostft_engine_variant.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, i32p, i32p]
    L.ostft_scratch_budget.restype = ctypes.c_int
    L.ostft_scratch_budget.argtypes = [vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
    L.ostft_server_create.restype = ctypes.c_int
    L.ostft_server_create.argtypes = [ctypes.c_int32, ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)]
    L.ostft_server_destroy.restype = None
    L.ostft_server_destroy.argtypes = [vp]
    L.ostft_server_submit.restype = ctypes.c_int
    L.ostft_server_submit.argtypes = [vp, vp, ctypes.c_uint64, ctypes.POINTER(Params), vp, ctypes.c_uint64, u64p]
    L.ostft_server_wait.restype = ctypes.c_int
    L.ostft_server_wait.argtypes = [vp, ctypes.c_uint64, ctypes.c_int32, u64p]
    L.ostft_server_stats_get.restype = ctypes.c_int
    L.ostft_server_stats_get.argtypes = [vp, ctypes.POINTER(ServerStats), ctypes.c_int32]
    L.ostft_server_last_error.restype = ctypes.c_char_p
    L.ostft_server_last_error.argtypes = [vp]
    _lib = L
    return L


def strerror(status: int) -> str:
    return lib().ostft_strerror(status).decode()
