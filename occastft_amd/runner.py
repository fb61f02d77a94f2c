"""Host-side mirror of the reference runner interface for the STFT path.

``Runner.activate_stft`` has the meaning of ``Runner::ActivateSTFT``
(StandAlone/cross_gpgpu/GLOBAL/RunnerInterface.hpp:122-127): signal, window
exponent, overlap ratio and the option string in; the OFullSize-float result
out, or ``None`` where the reference returns ``std::nullopt`` (the CUDA runner's
behaviour on a pipeline error, StandAlone/cross_gpgpu/CUDA/src/CudaImpl.cpp:497-501).
``Runner.stft`` is the same call raising ``STFTError`` instead.
``Runner.run_device`` is the device-resident form over torch tensors: torch is
used only for device memory and the stream; the work is the HIP engine.
``Runner.run_batch_device`` runs many such requests in one launch sequence, and
``Server`` is the continuous-batching engine for host-buffer requests (neither
has a counterpart in the reference, which serves one request per
ActivateSTFT call).
"""
from __future__ import annotations

import ctypes
from typing import NamedTuple, Optional

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
    """One engine context on one HIP device.  It may be shared by several host
    threads: host-buffer calls run concurrently on the context's staging
    lanes, and `last_error` is the calling thread's."""

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

    def run_strided_device(self, x, params: N.Params, out=None, stream=None):
        """STFT of every row of a (count, L) float32 CUDA tensor in one launch
        (ostft_run_strided_device); returns a (count, out_floats) view of
        `out` (allocated with rows padded to 4 floats if None)."""
        import torch

        if x.dtype != torch.float32 or not x.is_cuda or x.dim() != 2 or x.stride(1) != 1:
            raise ValueError("x must be a (count, L) float32 CUDA tensor with contiguous rows")
        count, L = x.shape
        g = geometry(params, L)
        row = (g["out_floats"] + 3) // 4 * 4
        if out is None:
            out = torch.empty(count, row, dtype=torch.float32, device=x.device)
        if out.dim() != 2 or out.shape[0] != count or out.stride(1) != 1:
            raise ValueError("out must be (count, >= out_floats) with contiguous rows")
        s = stream if stream is not None else torch.cuda.current_stream(x.device)
        rc = N.lib().ostft_run_strided_device(self._h, ctypes.byref(params), ctypes.c_void_p(x.data_ptr()), L,
                                              x.stride(0), count, ctypes.c_void_p(out.data_ptr()), out.stride(0),
                                              ctypes.c_void_p(s.cuda_stream))
        if rc != N.OK:
            raise STFTError(rc, self.last_error)
        return out[:, : g["out_floats"]]

    # -- batches -----------------------------------------------------------
    def run_batch_device(self, jobs, stream=None):
        """Run [(x, params, out), ...] of float32 CUDA tensors in one batch
        (ostft_run_batch_device); `out` may be None (allocated).  Returns the
        output tensors, each trimmed to its layout's length."""
        import torch

        arr = (N.Job * max(len(jobs), 1))()
        outs = []
        dev = None
        for i, (x, params, out) in enumerate(jobs):
            if x.dtype != torch.float32 or not x.is_cuda or not x.is_contiguous():
                raise ValueError("inputs must be contiguous float32 tensors on the GPU")
            dev = x.device
            try:
                n_out = geometry(params, x.numel())["out_floats"]
            except STFTError:
                n_out = 0  # the engine rejects the batch and names this job
            if out is None:
                out = torch.empty(max(n_out, 4), dtype=torch.float32, device=x.device)
            arr[i].in_ptr = x.data_ptr()
            arr[i].full_size = x.numel()
            arr[i].out_ptr = out.data_ptr()
            arr[i].out_cap = out.numel()
            arr[i].params = params
            outs.append(out[:n_out])
        if not jobs:
            return []
        s = stream if stream is not None else torch.cuda.current_stream(dev)
        rc = N.lib().ostft_run_batch_device(self._h, arr, len(jobs), ctypes.c_void_p(s.cuda_stream))
        if rc != N.OK:
            raise STFTError(rc, self.last_error)
        return outs

    def batch_config(self, exp: int = 0, blocks_per_cu: int = 0, use_graphs: int = -1) -> int:
        """blocks_per_cu > 0: persistent form with that many workgroups per CU;
        < 0: one workgroup per tile; 0: unchanged.  Returns the setting in
        force (0 = one workgroup per tile)."""
        cur = ctypes.c_int32()
        rc = N.lib().ostft_batch_config(self._h, int(exp), int(blocks_per_cu), int(use_graphs), ctypes.byref(cur))
        if rc != N.OK:
            raise STFTError(rc, self.last_error)
        return cur.value

    def scratch_budget(self, nbytes: int = 0) -> int:
        """Scratch budget (bytes) of the two-pass path for windows 2^16..2^20;
        0 leaves it.  Returns the budget in force."""
        cur = ctypes.c_uint64()
        rc = N.lib().ostft_scratch_budget(self._h, int(nbytes), ctypes.byref(cur))
        if rc != N.OK:
            raise STFTError(rc, self.last_error)
        return cur.value

    def engine_variant(self, exp: int, shape: int = -1, plan: int = -1):
        """-> (shape, plan) in force for window exponent `exp` (6..15), after
        setting either (-1 leaves it).  plan: radix plan of the FP32 engine
        for every request of that window on this context (0: radix-16 passes,
        remainder radix last; 1: remainder first; 2, 3: the same with radix-8
        passes, exponents 8..14; plans 1 and 3 only where the size has a
        remainder radix).  shape: frames per workgroup of the batch kernels
        (0: 256 threads, 1: 512, 2: 128; plans 2 and 3 shape 0 only; shapes 1
        and 2 run one workgroup per tile)."""
        cs = ctypes.c_int32()
        cp = ctypes.c_int32()
        rc = N.lib().ostft_engine_variant(self._h, int(exp), int(shape), int(plan), ctypes.byref(cs),
                                          ctypes.byref(cp))
        if rc != N.OK:
            raise STFTError(rc, self.last_error)
        return cs.value, cp.value

    def autotune_batch(self, exp: int, samples: int = 1 << 24, like: N.Params | None = None) -> "AutotuneChoice":
        """Times every built variant and launch form of window exponent `exp`
        on a synthetic ragged batch, keeps the fastest for this context and
        returns it: blocks_per_cu (0 = one workgroup per tile, k > 0 =
        persistent with k workgroups per CU), shape, plan, ms.  The synthetic
        requests take the overlap, window, output layout and flags of `like`
        (a Params, e.g. from parse_options; its exp is ignored), or without
        it overlap 0.5, Hann, DC removal and the reference power layout."""
        best = ctypes.c_int32()
        shape = ctypes.c_int32()
        plan = ctypes.c_int32()
        ms = ctypes.c_double()
        rc = N.lib().ostft_autotune_batch(self._h, int(exp), int(samples),
                                          ctypes.byref(like) if like is not None else None, ctypes.byref(best),
                                          ctypes.byref(shape), ctypes.byref(plan), ctypes.byref(ms))
        if rc != N.OK:
            raise STFTError(rc, self.last_error)
        return AutotuneChoice(best.value, shape.value, plan.value, ms.value)


class AutotuneChoice(NamedTuple):
    """What Runner.autotune_batch kept for one window size."""
    blocks_per_cu: int  # 0: one workgroup per tile; k > 0: persistent, k per CU
    shape: int          # workgroup shape of the batch kernels
    plan: int           # radix plan of the FP32 engine
    ms: float           # GPU time of one synthetic batch


class Server:
    """Continuous-batching engine for host-buffer requests (ostft_server_*).

    ``submit`` queues a request and returns a ticket; ``wait`` returns the
    result array.  Input and output arrays are kept referenced here until the
    request is waited for."""

    def __init__(self, device: int = 0, arena_bytes: int = 0, max_batch: int = 0, max_inflight: int = 0,
                 use_graphs: bool = True, autotune: bool = False, autotune_like: N.Params | None = None,
                 autotune_samples: int = 0):
        like = ctypes.cast(ctypes.pointer(autotune_like), ctypes.c_void_p) if autotune_like is not None else None
        cfg = N.ServerConfig(int(arena_bytes), int(max_batch), int(max_inflight), 1 if use_graphs else 0,
                             1 if autotune else 0, like, int(autotune_samples))
        h = ctypes.c_void_p()
        rc = N.lib().ostft_server_create(int(device), ctypes.byref(cfg), ctypes.byref(h))
        if rc != N.OK:
            raise STFTError(rc, f"server on device {device}")
        self._h = h
        self._live = {}

    def close(self) -> None:
        if getattr(self, "_h", None):
            N.lib().ostft_server_destroy(self._h)
            self._h = None
            self._live.clear()

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
        return N.lib().ostft_server_last_error(self._h).decode()

    def submit(self, x: np.ndarray, params: N.Params, out: Optional[np.ndarray] = None) -> int:
        x = np.ascontiguousarray(x, dtype=np.float32)
        g = geometry(params, x.size)
        if out is None:
            out = np.empty(max(g["out_floats"], 1), dtype=np.float32)
        t = ctypes.c_uint64()
        rc = N.lib().ostft_server_submit(self._h, ctypes.c_void_p(x.ctypes.data), x.size, ctypes.byref(params),
                                         ctypes.c_void_p(out.ctypes.data), out.size, ctypes.byref(t))
        if rc != N.OK:
            raise STFTError(rc, self.last_error)
        self._live[t.value] = (x, out)
        return t.value

    def wait(self, ticket: int, timeout_ms: int = -1) -> np.ndarray:
        n = ctypes.c_uint64()
        rc = N.lib().ostft_server_wait(self._h, int(ticket), int(timeout_ms), ctypes.byref(n))
        if rc == N.ERR_TIMEOUT:
            raise TimeoutError(f"request {ticket} still running")
        x, out = self._live.pop(ticket, (None, None))
        if rc != N.OK:
            raise STFTError(rc, self.last_error)
        return out[: n.value]

    def stats(self, reset: bool = False) -> dict:
        s = N.ServerStats()
        rc = N.lib().ostft_server_stats_get(self._h, ctypes.byref(s), 1 if reset else 0)
        if rc != N.OK:
            raise STFTError(rc)
        return s.as_dict()
