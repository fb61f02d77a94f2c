"""Pins for the oracle (CPU only).

The float32 restatement (oracle/refstft.c) is checked against
  * the golden vectors written from the reference's own translated kernels
    (tests/golden/ref_serial_vectors.npz), bit for bit;
  * the reference kernels themselves when oracle/_ref was built here;
  * closed forms and an FP64 DFT, so a dropped term, a sign, an index or a
    transposed operand fails somewhere even without the reference.
"""
import os
import sys

import numpy as np
import pytest

from conftest import frame_errors

import pyoracle as po

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
import make_golden  # noqa: E402


@pytest.fixture(scope="module")
def golden():
    return make_golden.load()


# ---------------------------------------------------------------- golden vectors

def test_restatement_matches_reference_golden_bitwise(golden):
    assert len(golden) >= 16
    for c in golden:
        rc, y = po.pipeline(c["x"], c["exp"], c["overlap"], c["options"])
        assert rc == c["status"], c["name"]
        assert y.size == c["y"].size, c["name"]
        assert np.array_equal(y.view(np.uint32), c["y"].view(np.uint32)), c["name"]


def test_golden_covers_the_quirk_paths(golden):
    names = {c["name"] for c in golden}
    for must in ("exp13_odd_hamming_dc", "exp15_odd_bh_dc", "blackman_hc_odd_frames", "gauss_sigma_parse",
                 "gauss_sigma_fail", "flattop_dc_4096"):
        assert must in names


def test_golden_matches_fp64_closed_forms(golden):
    """The reference outputs themselves agree with the FP64 statement of each
    quirk: sqrt(2)|Re X|, the buffer-half layout, and the penultimate stage."""
    for c in golden:
        if c["status"] != 0:
            assert not c["y"].any()
            continue
        rc, o = po.parse_options(c["options"])
        g = po.geometry(len(c["x"]), c["exp"], c["overlap"])
        mode = po.OUT_HALF_COMPLEX if o["half_complex"] else po.OUT_COMPAT_POWER
        t = po.stft64(c["x"], c["exp"], g["hop"], g["frames"], o["window"], o["sigma"], bool(o["remove_dc"]),
                      mode, quirk_odd=True)
        N = 1 << c["exp"]
        per = N
        y = c["y"].astype(np.float64)
        # skip frames dominated by the window edge (the reference's float
        # window is inaccurate there; see test_window_edge_precision)
        yy, tt = y.reshape(-1, per), t.reshape(-1, per)
        energy = np.linalg.norm(tt, axis=1)
        keep = energy > 1e-3 * energy.max()
        mx, _, _ = frame_errors(yy[keep], tt[keep], per)
        assert mx < 2e-5, (c["name"], mx)


# ---------------------------------------------------------------- live reference

@pytest.mark.skipif(not po.have_ref(), reason="oracle/_ref not built (no reference checkout here)")
@pytest.mark.parametrize("exp", [6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16])
def test_restatement_matches_reference_kernels_live(exp, rng):
    for ov in (0.0, 0.5, 0.6):
        for opts in ("", "--hanning_window --remove_dc", "--gaussian_window=0.45<<sigma --half_complex_return",
                     "--flattop_window --remove_dc --half_complex_return"):
            L = int(rng.integers(1, 4)) * (1 << exp) + int(rng.integers(0, 300))
            x = rng.uniform(-1, 1, L).astype(np.float32)
            rc1, a = po.pipeline(x, exp, ov, opts)
            rc2, b = po.ref_activate(x, exp, ov, opts)
            assert rc1 == rc2
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (exp, ov, opts)


# ---------------------------------------------------------------- geometry

def test_toquot_known_frame_counts():
    # frame counts in the float32 arithmetic of toQuot
    assert po.geometry(1 << 20, 10, 0.5)["frames"] == 2049
    assert po.geometry(100_000_000, 10, 0.5)["frames"] == 195_313
    assert po.geometry(1 << 27, 20, 0.75)["frames"] == 513
    assert po.geometry(1 << 27, 16, 0.0)["frames"] == 2049
    assert po.geometry(100_000_000, 6, 0.0)["frames"] == 1_562_501
    assert po.geometry(1000, 10, 0.6)["hop"] == 409


def test_toquot_is_float32_arithmetic(rng):
    f32 = np.float32
    for _ in range(2000):
        L = int(rng.integers(0, 1 << 28))
        exp = int(rng.integers(6, 21))
        ov = float(f32(rng.choice([0.0, 0.25, 0.5, 0.75, 0.6, 0.33, rng.uniform(0, 0.95)])))
        N = f32(1 << exp)
        if ov == 0.0:
            q = f32(L) / N + f32(1)
        else:
            q = f32(L) / (N * (f32(1) - f32(ov))) + f32(1)
        assert po.toquot(float(f32(L)), ov, float(N)) == int(q)


# ---------------------------------------------------------------- stage kernels

def test_overlap_closed_form(rng):
    x = rng.standard_normal(1000).astype(np.float32)
    exp, hop = 6, 24
    g = po.toquot(1000.0, 1 - hop / 64.0, 64.0)
    ofull = g * 64
    out = po.overlap(x, ofull, exp, hop).reshape(-1, 64)
    for q in range(g):
        for n in range(64):
            s = q * hop + n
            assert out[q, n] == (x[s] if s < 1000 else 0.0)


def test_windows_against_fp64_formulas():
    for kind in range(1, 9):
        for N in (64, 1024, 8192):
            w32 = np.array([po.window_coef(kind, n, N, 0.4) for n in range(N)], np.float64)
            w64 = po.window64(kind, N, 0.4)
            # float angle rounding (<= 2^-24 * 2*pi*j for cos(j theta)) plus float
            # cos rounding, weighted by the coefficients: < 6e-7 for every window
            assert np.max(np.abs(w32 - w64)) < 6e-7, (kind, N)


def test_window_closed_form_values():
    N = 1024
    han = po.window64(po.WINDOWS["hanning"], N)
    assert han[0] == 0.0 and abs(han[N - 1]) < 1e-15  # symmetric (N-1) form
    ham = po.window64(po.WINDOWS["hamming"], N)
    assert abs(ham[0] - 0.08) < 1e-15
    blk = po.window64(po.WINDOWS["blackman"], N)
    assert abs(blk[0]) < 1e-15 and abs(blk[N // 2] - 1.0) < 1e-15  # periodic form peaks at N/2
    assert np.allclose(blk[1:], blk[1:][::-1], atol=1e-15)
    ft = po.window64(po.WINDOWS["flattop"], N)
    assert abs(ft[N // 2] - (0.21557895 + 0.41663158 + 0.277263158 + 0.083578947 + 0.006947368)) < 1e-12
    ga = po.window64(po.WINDOWS["gaussian"], N, 0.5)
    assert ga[N // 2] == 1.0 and abs(ga[0] - np.exp(-0.5 * (1 / 0.5) ** 2)) < 1e-15


def test_window_edge_precision():
    """The reference's float Hanning (1 - cosf) is coarse where w ~ 0: that is why
    the engine can reproduce the reference's roundings (compat window)."""
    N = 16384
    w32 = np.array([po.window_coef(1, n, N) for n in range(1, 40)], np.float64)
    w64 = po.window64(1, N)[1:40]
    rel = np.abs(w32 - w64) / w64
    assert rel.max() > 1e-3


def test_remove_dc_is_mean_subtraction(rng):
    N = 512
    x = (rng.standard_normal(N * 5) + 3.0).astype(np.float32)
    y = po.remove_dc(x, N).reshape(-1, N).astype(np.float64)
    xx = x.reshape(-1, N).astype(np.float64)
    expect = xx - xx.mean(axis=1, keepdims=True)
    assert np.max(np.abs(y - expect)) < 1e-5
    # planted mistake: subtracting the mean of the un-windowed or wrong frame is caught
    wrong = xx - xx.mean()
    assert np.max(np.abs(y - wrong)) > 1e-3


@pytest.mark.parametrize("exp", [6, 9, 12, 13])
def test_stockham_equals_dft(exp, rng):
    N = 1 << exp
    frames = 3
    re = rng.standard_normal(N * frames).astype(np.float32)
    im = rng.standard_normal(N * frames).astype(np.float32)
    o_re, o_im = po.stockham(re, im, exp)
    got = (o_re.astype(np.float64) + 1j * o_im).reshape(frames, N)
    z = (re.astype(np.float64) + 1j * im).reshape(frames, N)
    ref = np.fft.fft(z, axis=1)
    err = np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert err.max() < 1e-5
    # the penultimate stage is [DFT_{N/2}(even), DFT_{N/2}(odd)]
    p_re, p_im = po.stockham(re, im, exp, stages=exp - 1)
    pen = (p_re.astype(np.float64) + 1j * p_im).reshape(frames, N)
    half = np.concatenate([np.fft.fft(z[:, 0::2], axis=1), np.fft.fft(z[:, 1::2], axis=1)], axis=1)
    err = np.linalg.norm(pen - half, axis=1) / np.linalg.norm(half, axis=1)
    assert err.max() < 1e-5


def test_fp64_fft_against_dft_and_numpy(rng):
    for n in (1, 2, 4, 8, 16, 32, 64, 256, 1024):
        z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        d = po.dft64(z.real, z.imag)
        f = po.fft64(z)
        assert np.max(np.abs(f - d)) < 1e-9 * max(1, n)
        assert np.max(np.abs(f - np.fft.fft(z))) < 1e-9 * max(1, n)


def test_fp64_dft_closed_forms():
    n = 64
    imp = np.zeros(n)
    imp[0] = 1
    assert np.allclose(po.dft64(imp, np.zeros(n)), np.ones(n), atol=1e-15)
    assert np.allclose(po.dft64(np.ones(n), np.zeros(n)), np.r_[n, np.zeros(n - 1)], atol=1e-12)
    k0 = 5
    c = np.cos(2 * np.pi * k0 * np.arange(n) / n)
    X = po.dft64(c, np.zeros(n))
    expect = np.zeros(n)
    expect[k0] = expect[n - k0] = n / 2
    assert np.allclose(X, expect, atol=1e-11)


def test_compat_power_and_half_complex_closed_forms(rng):
    re = rng.standard_normal(256).astype(np.float32)
    im = rng.standard_normal(256).astype(np.float32)
    out = np.zeros(256, np.float32)
    L = po.lib()
    L.oracle_to_power(po._p32(out), po._p32(re), po._p32(re), 256)
    assert np.allclose(out, np.sqrt(2.0) * np.abs(re), rtol=2e-7)
    hc = np.zeros(256, np.float32)
    L.oracle_half_complex(po._p32(hc), po._p32(re), po._p32(im), 128, 6)
    assert np.array_equal(hc[0::2], re[:128]) and np.array_equal(hc[1::2], im[:128])


def test_pipeline_order_window_then_dc(golden):
    """Planted mistake: DC removal before the window (the dead fused kernels'
    order) does not reproduce the reference."""
    c = next(c for c in golden if c["name"] == "hann_dc_1024")
    x, exp, ov = c["x"], c["exp"], c["overlap"]
    N = 1 << exp
    g = po.geometry(len(x), exp, ov)
    fr = po.overlap(x, g["ofull"], exp, g["hop"])
    wrong = po.apply_window(po.remove_dc(fr, N), N, po.WINDOWS["hanning"])
    re, im = po.stockham(wrong, np.zeros_like(wrong), exp)
    y = np.sqrt(re.astype(np.float64) ** 2 * 2)
    mx, _, _ = frame_errors(y, c["y"], N)
    assert mx > 1e-3


# ---------------------------------------------------------------- option parser

def test_option_parser_quirks():
    assert po.parse_options("--gaussian_window=0.35<<sigma")[1]["sigma"] == pytest.approx(0.35)
    # the first sigma character is skipped: "2.5" parses as ".5"
    assert po.parse_options("--gaussian_window=2.5<<sigma")[1]["sigma"] == pytest.approx(0.5)
    # a single digit leaves nothing to parse: error
    assert po.parse_options("--gaussian_window=3<<sigma")[0] == 1
    # "10" parses as "0": no window, no error
    rc, o = po.parse_options("--gaussian_window=10<<sigma")
    assert rc == 0 and o["window"] == po.WINDOWS["none"]
    # precedence follows the if/else-if chain, not the string order
    assert po.parse_options("--flattop_window --hanning_window")[1]["window"] == po.WINDOWS["hanning"]
    assert po.parse_options("--blackman_nuttall_window")[1]["window"] == po.WINDOWS["blackman_nuttall"]
    assert po.parse_options("--remove_dc --half_complex_return")[1] == dict(
        window=0, sigma=0.0, remove_dc=1, half_complex=1)


def test_pipeline_rejects_small_exp():
    rc, y = po.pipeline(np.ones(100, np.float32), 5, 0.0, "")
    assert rc == 2 and not y.any()
