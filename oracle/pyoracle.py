"""ctypes bindings for the oracle -- TEST INFRASTRUCTURE ONLY.

Loads ``oracle/liboracle.so`` (the C restatement of the reference Serial runner
plus the FP64 truth pipeline) and, when it was built, ``oracle/_ref/*.so`` (the
reference's own translated CPU kernels behind ``ref_harness.cpp``).  Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg may
import this module; the product never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_SERIAL = os.path.join(HERE, "_ref", "libref_serial.so")
REF_OPENMP = os.path.join(HERE, "_ref", "libref_openmp.so")

WINDOWS = {
    "none": 0,
    "hanning": 1,
    "hamming": 2,
    "blackman": 3,
    "nuttall": 4,
    "blackman_nuttall": 5,
    "blackman_harris": 6,
    "flattop": 7,
    "gaussian": 8,
}
WINDOW_OPTION = {
    "hanning": "--hanning_window",
    "hamming": "--hamming_window",
    "blackman": "--blackman_window",
    "nuttall": "--nuttall_window",
    "blackman_nuttall": "--blackman_nuttall_window",
    "blackman_harris": "--blackman_harris_window",
    "flattop": "--flattop_window",
}
# oracle_stft64 output modes
OUT_COMPAT_POWER, OUT_HALF_COMPLEX, OUT_MAGNITUDE, OUT_POWER, OUT_HALF_SPECTRUM = range(5)

_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_i32p = ctypes.POINTER(ctypes.c_int32)


def build() -> None:
    """Compile liboracle.so (and oracle/_ref when the reference is present)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


class _Opts(ctypes.Structure):
    _fields_ = [
        ("window", ctypes.c_int32),
        ("sigma", ctypes.c_float),
        ("remove_dc", ctypes.c_int32),
        ("half_complex", ctypes.c_int32),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_toquot.restype = ctypes.c_int32
        L.oracle_toquot.argtypes = [ctypes.c_float] * 3
        L.oracle_geometry.argtypes = [ctypes.c_uint32, ctypes.c_int32, ctypes.c_float, _i32p, _u32p, _u32p, _u32p]
        L.oracle_window_coef.restype = ctypes.c_float
        L.oracle_window_coef.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_float]
        L.oracle_window_coef64.restype = ctypes.c_double
        L.oracle_window_coef64.argtypes = [ctypes.c_int32, ctypes.c_int64, ctypes.c_int64, ctypes.c_double]
        L.oracle_overlap.argtypes = [_f32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _f32p]
        L.oracle_apply_window.argtypes = [_f32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32, ctypes.c_float]
        L.oracle_remove_dc.argtypes = [_f32p, ctypes.c_uint32, ctypes.c_uint32]
        L.oracle_stockham_stage.argtypes = [_f32p, _f32p, _f32p, _f32p] + [ctypes.c_uint32] * 4
        L.oracle_to_power.argtypes = [_f32p, _f32p, _f32p, ctypes.c_uint32]
        L.oracle_half_complex.argtypes = [_f32p, _f32p, _f32p, ctypes.c_uint32, ctypes.c_int32]
        L.oracle_parse_options.restype = ctypes.c_int32
        L.oracle_parse_options.argtypes = [ctypes.c_char_p, ctypes.POINTER(_Opts)]
        L.oracle_pipeline.restype = ctypes.c_int32
        L.oracle_pipeline.argtypes = [_f32p, ctypes.c_uint32, ctypes.c_int32, ctypes.c_float, ctypes.c_char_p, _f32p]
        L.oracle_dft64.argtypes = [_f64p, _f64p, ctypes.c_int64, _f64p, _f64p]
        L.oracle_fft64.argtypes = [_f64p, _f64p, ctypes.c_int64]
        L.oracle_stft64.restype = ctypes.c_int64
        L.oracle_stft64.argtypes = [
            _f32p, ctypes.c_uint64, ctypes.c_int32, ctypes.c_uint64, ctypes.c_uint64,
            ctypes.c_int32, ctypes.c_double, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _f64p,
        ]
        _lib = L
    return _lib


def _p32(a):
    return a.ctypes.data_as(_f32p)


def _p64(a):
    return a.ctypes.data_as(_f64p)


def toquot(L: float, ov: float, N: float) -> int:
    return int(lib().oracle_toquot(L, ov, N))


def geometry(L: int, exp: int, ov: float):
    qt, of, oh, om = ctypes.c_int32(), ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
    lib().oracle_geometry(L, exp, ov, ctypes.byref(qt), ctypes.byref(of), ctypes.byref(oh), ctypes.byref(om))
    return dict(frames=qt.value, ofull=of.value, ohalf=oh.value, hop=om.value)


def parse_options(options: str):
    o = _Opts()
    rc = lib().oracle_parse_options(options.encode(), ctypes.byref(o))
    return rc, dict(window=o.window, sigma=o.sigma, remove_dc=o.remove_dc, half_complex=o.half_complex)


def window_coef(kind: int, index: int, N: int, sigma: float = 0.0) -> float:
    return float(lib().oracle_window_coef(kind, index, N, sigma))


def window64(kind: int, N: int, sigma: float = 0.0) -> np.ndarray:
    f = lib().oracle_window_coef64
    return np.array([f(kind, n, N, sigma) for n in range(N)])


def pipeline(x: np.ndarray, exp: int, ov: float, options: str):
    """Restated Serial ActivateSTFT.  Returns (status, out[OFullSize])."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    g = geometry(len(x), exp, ov)
    out = np.zeros(max(g["ofull"], 1), dtype=np.float32)
    xin = x if len(x) else np.zeros(1, np.float32)
    rc = lib().oracle_pipeline(_p32(xin), len(x), exp, ov, options.encode(), _p32(out))
    return rc, out[: g["ofull"]]


def stft64(x: np.ndarray, exp: int, hop: int, frames: int, window: int = 0, sigma: float = 0.0,
           remove_dc: bool = False, out_mode: int = OUT_COMPAT_POWER, quirk_odd: bool = True) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32)
    N = 1 << exp
    per = {0: N, 1: N, 2: N, 3: N // 2 + 1, 4: N + 2}[out_mode]
    out = np.zeros(frames * per, dtype=np.float64)
    xin = x if len(x) else np.zeros(1, np.float32)
    n = lib().oracle_stft64(_p32(xin), len(x), exp, hop, frames, window, sigma, int(remove_dc), out_mode,
                            int(quirk_odd), _p64(out))
    return out[: max(n, 0)] if out_mode != 1 else out


def dft64(re: np.ndarray, im: np.ndarray):
    re = np.ascontiguousarray(re, np.float64)
    im = np.ascontiguousarray(im, np.float64)
    n = len(re)
    ore, oim = np.zeros(n), np.zeros(n)
    lib().oracle_dft64(_p64(re), _p64(im), n, _p64(ore), _p64(oim))
    return ore + 1j * oim


def fft64(z: np.ndarray) -> np.ndarray:
    re = np.ascontiguousarray(z.real, np.float64).copy()
    im = np.ascontiguousarray(z.imag, np.float64).copy()
    lib().oracle_fft64(_p64(re), _p64(im), len(re))
    return re + 1j * im


# -- stage kernels (used by the unit tests of the restatement) ------------------

def overlap(x, ofull, exp, hop):
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros(ofull, np.float32)
    lib().oracle_overlap(_p32(x), ofull, len(x), exp, hop, _p32(out))
    return out


def apply_window(data, N, kind, sigma=0.0):
    d = np.ascontiguousarray(data, np.float32).copy()
    lib().oracle_apply_window(_p32(d), len(d), N, kind, sigma)
    return d


def remove_dc(data, N):
    d = np.ascontiguousarray(data, np.float32).copy()
    lib().oracle_remove_dc(_p32(d), len(d), N)
    return d


def stockham(re, im, exp, stages=None):
    """Run `stages` radix-2 Stockham stages (default: all) over frame-major re/im.
    Returns the buffer written by the last stage."""
    N = 1 << exp
    a_re = np.ascontiguousarray(re, np.float32).copy()
    a_im = np.ascontiguousarray(im, np.float32).copy()
    b_re, b_im = np.zeros_like(a_re), np.zeros_like(a_im)
    stages = exp if stages is None else stages
    for s in range(stages):
        lib().oracle_stockham_stage(_p32(a_re), _p32(a_im), _p32(b_re), _p32(b_im), N // 2, s, len(a_re) // 2, exp)
        a_re, b_re = b_re, a_re
        a_im, b_im = b_im, a_im
    return a_re, a_im


# -- the reference's own kernels (oracle/_ref) -------------------------------------

_ref = {}


def have_ref() -> bool:
    return os.path.exists(REF_SERIAL)


def ref_lib(kind: str = "serial"):
    if kind not in _ref:
        path = REF_SERIAL if kind == "serial" else REF_OPENMP
        L = ctypes.CDLL(path)
        L.ref_activate.restype = ctypes.c_int
        L.ref_activate.argtypes = [_f32p, ctypes.c_uint32, ctypes.c_int, ctypes.c_float, ctypes.c_char_p, _f32p]
        L.ref_geometry.argtypes = [ctypes.c_uint32, ctypes.c_int, ctypes.c_float, _i32p, _u32p, _u32p]
        _ref[kind] = L
    return _ref[kind]


def ref_geometry(L: int, exp: int, ov: float):
    qt, of, om = ctypes.c_int32(), ctypes.c_uint32(), ctypes.c_uint32()
    ref_lib().ref_geometry(L, exp, ov, ctypes.byref(qt), ctypes.byref(of), ctypes.byref(om))
    return dict(frames=qt.value, ofull=of.value, hop=om.value)


def ref_activate(x: np.ndarray, exp: int, ov: float, options: str, kind: str = "serial"):
    x = np.ascontiguousarray(x, dtype=np.float32)
    g = ref_geometry(len(x), exp, ov)
    out = np.zeros(max(g["ofull"], 1), dtype=np.float32)
    xin = x if len(x) else np.zeros(1, np.float32)
    rc = ref_lib(kind).ref_activate(_p32(xin), len(x), exp, ov, options.encode(), _p32(out))
    return rc, out[: g["ofull"]]
