"""The C ABI without a GPU: the library loads, exports every symbol the header
declares, and its host logic (option parsing, frame geometry) matches the
oracle's restatement of Default_Pipeline / toQuot."""
import os
import re

import numpy as np
import pytest

import occastft_amd as ostft
from occastft_amd import _native as N

import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    syms = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            text = open(os.path.join(inc, fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            syms |= set(re.findall(r"\b(ostft_[a-z_0-9]+)\s*\(", text))
    return syms


def test_library_exports_every_declared_symbol():
    L = N.lib()
    syms = header_symbols()
    assert syms == set(N.EXPORTS)
    for s in syms:
        assert hasattr(L, s), s
    assert L.ostft_abi_version() == N.ABI_VERSION == 2


def test_native_missing_is_loud(monkeypatch, tmp_path):
    monkeypatch.setattr(N, "_lib", None)
    monkeypatch.setattr(N, "LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(N.NativeMissing):
        N.lib()


OPTION_STRINGS = [
    "",
    "--hanning_window",
    "--hamming_window --remove_dc",
    "--blackman_window --half_complex_return",
    "--nuttall_window",
    "--blackman_nuttall_window --remove_dc --half_complex_return",
    "--blackman_harris_window",
    "--flattop_window --hanning_window",
    "--gaussian_window=0.35<<sigma",
    "--gaussian_window=2.5<<sigma --remove_dc",
    "--gaussian_window=10<<sigma",
    "--gaussian_window=0.5",
    "<<sigma --gaussian_window=0.1",
    "--gaussian_window=1e-50<<sigma",
    "--gaussian_window=x<<sigma",
    "--gaussian_window=3<<sigma",
    "--gaussian_window= 0.7<<sigma",
]


@pytest.mark.parametrize("opts", OPTION_STRINGS)
def test_parse_options_matches_oracle(opts):
    rc_o, o = po.parse_options(opts)
    if rc_o != 0:
        with pytest.raises(ostft.STFTError) as ei:
            ostft.parse_options(opts, 10, 0.5)
        assert ei.value.status == N.ERR_OPTION
        return
    p = ostft.parse_options(opts, 10, 0.5)
    assert p.window == o["window"]
    if o["window"] == po.WINDOWS["gaussian"]:
        assert p.sigma == pytest.approx(o["sigma"], rel=0, abs=0)
    assert bool(p.flags & N.FLAG_REMOVE_DC) == bool(o["remove_dc"])
    assert (p.output == N.OUT_HALF_COMPLEX) == bool(o["half_complex"])


def test_new_options():
    p = ostft.parse_options("--hanning_window --power_return --exact_odd_exp --exact_window", 12, 0.5)
    assert p.output == N.OUT_POWER and p.flags & N.FLAG_EXACT_ODD and p.flags & N.FLAG_EXACT_WINDOW
    assert ostft.parse_options("--magnitude_return", 12, 0).output == N.OUT_MAGNITUDE
    assert ostft.parse_options("--half_spectrum_return", 12, 0).output == N.OUT_HALF_SPECTRUM
    # the reference's own flag wins over the new layouts
    assert ostft.parse_options("--half_spectrum_return --half_complex_return", 12, 0).output == N.OUT_HALF_COMPLEX
    p = ostft.parse_options("--gaussian_window=2.5<<sigma --exact_sigma_parse", 12, 0)
    assert p.sigma == pytest.approx(2.5)


def test_mfma_options_and_limits():
    p = ostft.parse_options("--hanning_window --mfma_fp16", 10, 0.5)
    assert p.flags & N.FLAG_MFMA_FP16 and not p.flags & N.FLAG_MFMA_BF16
    assert ostft.parse_options("--mfma_bf16", 8, 0.5).flags & N.FLAG_MFMA_BF16
    for exp in (6, 8, 10):
        ostft.geometry(ostft.make_params(exp, 0.5, flags=N.FLAG_MFMA_FP16), 5000)
    for exp in (7, 9, 11, 12, 16):
        with pytest.raises(ostft.STFTError) as ei:
            ostft.geometry(ostft.make_params(exp, 0.5, flags=N.FLAG_MFMA_BF16), 5000)
        assert ei.value.status == N.ERR_UNSUPPORTED
    with pytest.raises(ostft.STFTError) as ei:
        ostft.geometry(ostft.make_params(10, 0.5, flags=N.FLAG_MFMA_FP16 | N.FLAG_MFMA_BF16), 5000)
    assert ei.value.status == N.ERR_INVALID_ARG


def test_geometry_matches_oracle(rng):
    for _ in range(500):
        L = int(rng.integers(0, 1 << 30))
        exp = int(rng.integers(6, 21))
        ov = float(np.float32(rng.choice([0.0, 0.25, 0.5, 0.75, 0.6, 0.33, rng.uniform(0, 0.9)])))
        p = ostft.make_params(exp, ov)
        g = ostft.geometry(p, L)
        q = po.toquot(float(np.float32(L)), ov, float(1 << exp))
        assert g["frames"] == q
        assert g["hop"] == int(np.float32(1 << exp) * (np.float32(1) - np.float32(ov)))
        assert g["out_full"] == q * (1 << exp)
        if L < (1 << 32) and q * (1 << exp) < (1 << 32):
            o = po.geometry(L, exp, ov)
            assert (o["frames"], o["ofull"], o["hop"]) == (g["frames"], g["out_full"], g["hop"])


def test_geometry_output_sizes():
    g = ostft.geometry(ostft.make_params(10, 0.5, output=N.OUT_POWER), 1 << 20)
    assert g["out_floats"] == 2049 * 513
    g = ostft.geometry(ostft.make_params(10, 0.5, output=N.OUT_HALF_SPECTRUM), 1 << 20)
    assert g["out_floats"] == 2049 * 1026
    g = ostft.geometry(ostft.make_params(10, 0.5, output=N.OUT_HALF_COMPLEX), 1 << 20)
    assert g["out_floats"] == 2049 * 1024 and g["compute_frames"] == 1025


def test_parameter_validation():
    for bad in (ostft.make_params(5, 0.0), ostft.make_params(21, 0.0)):
        with pytest.raises(ostft.STFTError) as ei:
            ostft.geometry(bad, 1000)
        assert ei.value.status == N.ERR_UNSUPPORTED
    for ov in (1.0, -0.1, float("nan")):
        with pytest.raises(ostft.STFTError) as ei:
            ostft.geometry(ostft.make_params(10, ov), 1000)
        assert ei.value.status == N.ERR_INVALID_ARG
    with pytest.raises(ostft.STFTError):
        ostft.geometry(ostft.make_params(10, 0.0, window=9), 1000)


def test_context_calls_reject_a_null_context():
    import ctypes

    cur = ctypes.c_uint64(7)
    assert N.lib().ostft_scratch_budget(None, 1 << 20, ctypes.byref(cur)) == N.ERR_INVALID_ARG
    assert cur.value == 7
    best = ctypes.c_int32()
    assert N.lib().ostft_batch_config(None, 10, 1, -1, ctypes.byref(best)) == N.ERR_INVALID_ARG
    shape, plan = ctypes.c_int32(-5), ctypes.c_int32(-5)
    assert N.lib().ostft_engine_variant(None, 10, 0, 0, ctypes.byref(shape), ctypes.byref(plan)) == N.ERR_INVALID_ARG
    assert shape.value == -5 and plan.value == -5
    ms = ctypes.c_double(-1.0)
    assert N.lib().ostft_autotune_batch(None, 10, 1 << 20, None, ctypes.byref(best), ctypes.byref(shape),
                                        ctypes.byref(plan), ctypes.byref(ms)) == N.ERR_INVALID_ARG
    assert ms.value == -1.0


def test_create_without_gpu_is_a_clean_error():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(ostft.STFTError) as ei:
        ostft.Runner(0)
    assert ei.value.status == N.ERR_NO_DEVICE


@pytest.mark.parametrize("kind", range(1, 9))
def test_default_window_table_is_the_reference_window(kind):
    """The engine's default coefficients equal the restated reference window
    bit for bit (the restatement itself is pinned to the reference kernels)."""
    import ctypes

    for n in (64, 1000, 4096, 1 << 16):
        out = np.zeros(n, np.float32)
        rc = N.lib().ostft_window_coefficients(kind, n, 0.4, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        assert rc == N.OK
        ref = np.array([po.window_coef(kind, i, n, 0.4) for i in range(n)], np.float32)
        assert np.array_equal(out.view(np.uint32), ref.view(np.uint32)), (kind, n)
