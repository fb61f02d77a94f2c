"""Host-side mirror of the reference runner interface for the STFT path.

``Runner.activate_stft`` has the meaning of ``Runner::ActivateSTFT``
(StandAlone/cross_gpgpu/GLOBAL/RunnerInterface.hpp:122-127): signal, window
exponent, overlap ratio and the option string in; the OFullSize-float result
out, or ``None`` where the reference returns ``std::nullopt`` (the CUDA runner's
behaviour on a pipeline error, StandAlone/cross_gpgpu/CUDA/src/CudaImpl.cpp:497-501).
``Runner.stft`` is the same call raising ``STFTError`` instead.
``Runner.run_device`` is the device-resident form over torch tensors: torch is
used only for device memory and the stream; the work is the HIP engine.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from . import _native as N


class STFTError(RuntimeError):
    def __init__(self, status: int, detail: str = ""):
        self.status = status
        msg = N.strerror(status)
        if detail:
            msg = f"{msg}: {detail}"
        super().__init__(msg)


def parse_options(options: str, exp: int, overlap: float) -> N.Params:
    """Default_Pipeline's option dispatch (RunnerCommon.cpp:216-248, 309-314)."""
    p = N.Params()
    rc = N.lib().ostft_parse_options(options.encode(), int(exp), float(overlap), ctypes.byref(p))
    if rc != N.OK:
        raise STFTError(rc, options)
    return p


def make_params(exp: int, overlap: float, window: int = N.WIN_NONE, sigma: float = 0.0,
                output: int = N.OUT_COMPAT_POWER, flags: int = 0) -> N.Params:
    return N.Params(int(exp), float(overlap), int(window), float(sigma), int(output), int(flags))


def geometry(params: N.Params, full_size: int) -> dict:
    """toQuot / OMove / OFullSize (RunnerInterface.hpp:26-35, SerialImpl.cpp:36-43)."""
    g = N.Geometry()
    rc = N.lib().ostft_get_geometry(ctypes.byref(params), int(full_size), ctypes.byref(g))
    if rc != N.OK:
        raise STFTError(rc)
    return g.as_dict()


class Runner:
    """One engine context on one HIP device (not reentrant: one per thread)."""

    def __init__(self, device: int = 0):
        L = N.lib()
        h = ctypes.c_void_p()
        rc = L.ostft_create(int(device), ctypes.byref(h))
        if rc != N.OK:
            raise STFTError(rc, f"device {device}")
        self._h = h
        self.device = device

    # -- lifetime ---------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None):
            N.lib().ostft_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def last_error(self) -> str:
        return N.lib().ostft_last_error(self._h).decode()

    # -- ActivateSTFT ------------------------------------------------------
    def stft(self, in_data: np.ndarray, window_size_exp: int, overlap_ratio: float, options: str) -> np.ndarray:
        x = np.ascontiguousarray(in_data, dtype=np.float32)
        p = parse_options(options, window_size_exp, overlap_ratio)
        g = geometry(p, len(x))
        out = np.empty(max(g["out_floats"], 1), dtype=np.float32)
        n = ctypes.c_uint64()
        fp = ctypes.POINTER(ctypes.c_float)
        xin = x if len(x) else np.zeros(1, np.float32)
        rc = N.lib().ostft_activate(self._h, xin.ctypes.data_as(fp), len(x), int(window_size_exp),
                                    float(overlap_ratio), options.encode(), out.ctypes.data_as(fp),
                                    out.size, ctypes.byref(n))
        if rc != N.OK:
            raise STFTError(rc, self.last_error)
        return out[: n.value]

    def activate_stft(self, in_data: np.ndarray, window_size_exp: int, overlap_ratio: float,
                      options: str) -> Optional[np.ndarray]:
        try:
            return self.stft(in_data, window_size_exp, overlap_ratio, options)
        except STFTError:
            return None

    # -- device-resident ---------------------------------------------------
    def run_device(self, x, params: N.Params, out=None, stream=None):
        """STFT of a float32 CUDA tensor into `out` (allocated if None), queued on
        `stream` (a torch.cuda.Stream; default: torch's current stream)."""
        import torch

        if x.dtype != torch.float32 or not x.is_cuda or not x.is_contiguous():
            raise ValueError("x must be a contiguous float32 tensor on the GPU")
        g = geometry(params, x.numel())
        if out is None:
            out = torch.empty(max(g["out_floats"], 1), dtype=torch.float32, device=x.device)
        if out.dtype != torch.float32 or not out.is_cuda or not out.is_contiguous():
            raise ValueError("out must be a contiguous float32 tensor on the GPU")
        s = stream if stream is not None else torch.cuda.current_stream(x.device)
        rc = N.lib().ostft_run_device(self._h, ctypes.byref(params), ctypes.c_void_p(x.data_ptr()), x.numel(),
                                      ctypes.c_void_p(out.data_ptr()), out.numel(), ctypes.c_void_p(s.cuda_stream))
        if rc != N.OK:
            raise STFTError(rc, self.last_error)
        return out[: g["out_floats"]]
