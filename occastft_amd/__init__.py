"""occastft_amd -- MI355X (gfx950) batched STFT engine.

A drop-in for the device path of Rliop913/OCCAstft's runners: the fused HIP
kernels live in ``libocca_stft_hip.so`` (C ABI in ``include/occastft_hip.h``);
this package is the thin host binding used by tests, the benchmark and Python
callers.
"""
from ._native import (  # noqa: F401
    FLAG_EXACT_ODD,
    FLAG_EXACT_WINDOW,
    FLAG_MFMA_BF16,
    FLAG_MFMA_FP16,
    FLAG_REMOVE_DC,
    FLAG_SIGMA_EXACT,
    OUT_COMPAT_POWER,
    OUT_HALF_COMPLEX,
    OUT_HALF_SPECTRUM,
    OUT_MAGNITUDE,
    OUT_POWER,
    WIN_BLACKMAN,
    WIN_BLACKMAN_HARRIS,
    WIN_BLACKMAN_NUTTALL,
    WIN_FLATTOP,
    WIN_GAUSSIAN,
    WIN_HAMMING,
    WIN_HANNING,
    WIN_NONE,
    WIN_NUTTALL,
    NativeMissing,
    Params,
)
from .runner import Runner, Server, STFTError, geometry, make_params, parse_options  # noqa: F401

__all__ = [
    "Runner",
    "Server",
    "STFTError",
    "NativeMissing",
    "Params",
    "geometry",
    "make_params",
    "parse_options",
]
