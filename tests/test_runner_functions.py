"""The record-and-fuse runner functions (runnerFunction::*, the reference's
per-stage plugin API, StandAlone/cross_gpgpu/GLOBAL/RunnerInterface.hpp:36-83).

CPU: the library defines exactly the reference's stage functions, with its
signatures (so a runner built from the reference's RunnerCommon.cpp links);
without a GPU the pipeline driver reports the missing device and returns
zeros.  GPU: driving the functions in Default_Pipeline's order
(occastft_amd/runnerfn_pipeline) gives bit for bit the engine's own result
(ostft_activate) for every path: windows, the Gaussian sigma parse, DC
removal, the fixed-size and general FFT stages, both output stages, and the
early exits with the reference's status strings.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import occastft_amd as ostft

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(os.path.dirname(HERE), "occastft_amd")
LIB = os.path.join(PKG, "libocca_stft_runnerfn.so")
DRIVER = os.path.join(PKG, "runnerfn_pipeline")

pytestmark = pytest.mark.skipif(not (os.path.exists(LIB) and os.path.exists(DRIVER)),
                                reason="runner-function library not built")

STAGES = ["Overlap", "Hanning", "Hamming", "Blackman", "Nuttall", "Blackman_Nuttall", "Blackman_Harris",
          "FlatTop", "Gaussian", "RemoveDC", "EXP6", "EXP7", "EXP8", "EXP9", "EXP10", "EXP11", "EXPC",
          "HalfComplex", "ToPower"]
# RunnerInterface.hpp:38-83, parameters as the C++ ABI sees them (CUI = const unsigned int by value)
U = "unsigned int"
SIGNATURES = {
    "Overlap": f"void*, void*, {U}, {U}, {U}, {U}, void*",
    "Gaussian": f"void*, void*, {U}, {U}, float",
    "RemoveDC": f"void*, void*, {U}, {U}, {U}",
    "EXP6": f"void*, void*, void*, {U}",
    "EXPC": f"void*, void*, void*, void*, void*, void*, {U}, {U}, {U}, void*, void*",
    "HalfComplex": f"void*, void*, void*, void*, {U}, {U}",
    "ToPower": f"void*, void*, void*, void*, {U}",
    "Hanning": f"void*, void*, {U}, {U}",
}


@pytest.mark.skipif(shutil.which("nm") is None, reason="nm not available")
def test_library_defines_the_reference_stage_functions():
    out = subprocess.run(["nm", "-DC", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    defined = {}
    for line in out.splitlines():
        if "runnerFunction::" in line:
            sig = line.split("runnerFunction::", 1)[1]
            name, args = sig.split("(", 1)
            defined[name] = args.rstrip(")")
    assert set(defined) == set(STAGES)
    for name, args in SIGNATURES.items():
        assert defined[name] == args, name


def run_driver(tmp_path, x, exp, ov, opts):
    f_in, f_out = tmp_path / "in.f32", tmp_path / "out.f32"
    np.asarray(x, np.float32).tofile(f_in)
    r = subprocess.run([DRIVER, str(f_in), str(exp), str(ov), opts, str(f_out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    return lines[0], lines[1:], np.fromfile(f_out, np.float32)


def gpu_available():
    import torch

    return torch.cuda.is_available()


def test_driver_without_a_gpu_returns_zeros(tmp_path):
    if gpu_available():
        pytest.skip("a GPU is present")
    x = np.ones(5000, np.float32)
    status, notes, y = run_driver(tmp_path, x, 10, 0.5, "--hanning_window")
    assert status == "--hanning_window"  # the first stage whose result Default_Pipeline checks
    assert any("no HIP device" in n for n in notes)
    assert y.size == ostft.geometry(ostft.parse_options("", 10, 0.5), 5000)["out_full"] and not y.any()


CASES = [
    (10, 0.5, "--hanning_window --remove_dc"),
    (6, 0.0, "--hamming_window"),
    (8, 0.75, "--gaussian_window=x2.5<<sigma --remove_dc"),
    (11, 0.5, "--blackman_harris_window --half_complex_return"),
    (12, 0.25, "--flattop_window --remove_dc"),
    (13, 0.5, "--nuttall_window --remove_dc"),          # odd exponent: the reference's penultimate-stage result
    (13, 0.5, "--blackman_nuttall_window --half_complex_return"),
    (16, 0.5, "--blackman_window --remove_dc"),         # the two-pass path
    (9, 0.6, ""),                                       # no window, non-dyadic overlap
]


@pytest.mark.gpu
@pytest.mark.parametrize("exp,ov,opts", CASES)
def test_pipeline_through_the_stage_functions_equals_the_engine(tmp_path, exp, ov, opts):
    x = np.random.default_rng(exp * 7 + int(ov * 100)).uniform(-1, 1, 70_001).astype(np.float32)
    status, notes, y = run_driver(tmp_path, x, exp, ov, opts)
    assert status == "OK", notes
    with ostft.Runner(0) as r:
        ref = r.stft(x, exp, ov, opts)
    assert y.shape == ref.shape
    assert np.array_equal(y.view(np.uint32), ref.view(np.uint32))


@pytest.mark.gpu
@pytest.mark.parametrize("exp,opts,status", [(5, "--hanning_window", "Operation not supported"),
                                             (10, "--gaussian_window=1<<sigma", "--gaussian_window")])
def test_early_exits_keep_the_reference_status_and_zeros(tmp_path, exp, opts, status):
    x = np.random.default_rng(1).uniform(-1, 1, 9000).astype(np.float32)
    got, notes, y = run_driver(tmp_path, x, exp, 0.5, opts)
    assert got == status
    assert not y.any()
