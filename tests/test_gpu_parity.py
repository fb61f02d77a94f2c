"""GPU parity: the HIP engine (through the C ABI) against the oracle.

Bars: per-frame relative L2 error <= 1e-5 (frames with ||r|| < 1e-20 judged by
absolute max <= 1e-6) against the float32 restatement of the reference Serial
runner (oracle/refstft.c, itself pinned bit for bit to the reference kernels),
and against the FP64 pipeline (oracle/fp64.c) where the reference has no
defined answer (int32 twiddle overflow for 2^17..2^20) or the layout is new.
"""
import numpy as np
import pytest

from conftest import frame_errors

import occastft_amd as ostft
from occastft_amd import _native as N

import pyoracle as po

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def runner():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    r = ostft.Runner(0)
    yield r
    r.close()


def check(y, r, per, tol=TOL, what=""):
    mx, p50, zabs = frame_errors(y, r, per)
    assert mx <= tol, f"{what}: max per-frame rel err {mx:.3e} (p50 {p50:.3e})"
    assert zabs <= 1e-6, f"{what}: zero-frame abs {zabs:.3e}"
    return mx


def gpu_vs_oracle(runner, x, exp, ov, opts, tol=TOL):
    y = runner.stft(x, exp, ov, opts)
    rc, r = po.pipeline(x, exp, ov, opts)
    assert rc == 0
    assert y.shape == r.shape
    return check(y, r, 1 << exp, tol, f"exp={exp} ov={ov} opts='{opts}'")


# --------------------------------------------------------------------- configs

def test_config1_hann_dc_1024(runner):
    """2^20 uniform samples, N=1024, 50%, Hanning, DC removal, default output."""
    x = np.random.default_rng(0x5EED0001).uniform(-1, 1, 1 << 20).astype(np.float32)
    err = gpu_vs_oracle(runner, x, 10, 0.5, "--hanning_window --remove_dc")
    assert err < TOL
    y = runner.stft(x, 10, 0.5, "--hanning_window --remove_dc")
    assert y.size == 2049 * 1024


def test_golden_vectors(runner):
    import sys, os
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    import make_golden

    for c in make_golden.load():
        if c["status"] != 0:
            # the Serial runner ignores the pipeline error and returns zeros;
            # the engine reports it (as the CUDA runner does, with nullopt)
            assert runner.activate_stft(c["x"], c["exp"], c["overlap"], c["options"]) is None
            continue
        y = runner.stft(c["x"], c["exp"], c["overlap"], c["options"])
        check(y, c["y"], 1 << c["exp"], TOL, c["name"])


@pytest.mark.parametrize("exp", list(range(6, 17)))
def test_sweep_against_serial(runner, exp):
    rng = np.random.default_rng(1000 + exp)
    N_ = 1 << exp
    opts_list = [
        "--hanning_window --remove_dc",
        "--hamming_window",
        "--blackman_window --remove_dc --half_complex_return",
        "--nuttall_window --half_complex_return",
        "--blackman_nuttall_window --remove_dc",
        "--blackman_harris_window",
        "--flattop_window --remove_dc",
        "--gaussian_window=0.3<<sigma --remove_dc",
        "",
    ]
    for ov in (0.0, 0.5, 0.75, 0.6):
        for opts in opts_list[(exp + int(ov * 4)) % 3::3]:
            frames = 3 if exp >= 14 else 9
            L = int(N_ * frames * (1 - ov)) + int(rng.integers(0, N_))
            x = rng.uniform(-1, 1, L).astype(np.float32)
            gpu_vs_oracle(runner, x, exp, ov, opts)


@pytest.mark.parametrize("exp", [17, 18, 19, 20])
def test_large_windows_against_fp64(runner, exp):
    """2^17..2^20: the reference's int32 twiddle index overflows (undefined
    behaviour), so the FP64 pipeline is the oracle; the float restatement with
    a 64-bit index agrees too."""
    rng = np.random.default_rng(2000 + exp)
    N_ = 1 << exp
    L = 2 * N_ + 1000
    x = (rng.standard_normal(L) + 0.5 * np.cos(np.pi * np.arange(L) ** 2 / (2 * L))).astype(np.float32)
    # every frame, the window-edge ones included: with --exact_window both
    # sides use the FP64 window formula (the reference's float coefficients
    # differ from it by up to 4e-4 relative where the window is near zero)
    for ov, opts, quirk in ((0.0, "--blackman_window --remove_dc --exact_window", True),
                            (0.75, "--blackman_window --remove_dc --exact_odd_exp --exact_window", False)):
        y = runner.stft(x, exp, ov, opts)
        g = ostft.geometry(ostft.parse_options(opts, exp, ov), L)
        t = po.stft64(x, exp, g["hop"], g["frames"], po.WINDOWS["blackman"], 0.0, True, po.OUT_COMPAT_POWER,
                      quirk_odd=quirk)
        check(y, t, N_, TOL, f"exp={exp} ov={ov}")
    rc, r = po.pipeline(x[: N_ + 77], exp, 0.0, "--blackman_window --remove_dc")
    y = runner.stft(x[: N_ + 77], exp, 0.0, "--blackman_window --remove_dc")
    check(y, r, N_, TOL, f"restatement exp={exp}")


def test_multipass_long_signal_every_frame(runner):
    """A 2^16 window at 75% overlap over 700 frames (180 MB of spectra): every
    frame against the FP64 pipeline; then with the two-pass path's scratch
    budget cut to 128 frames (32 MiB, six chunks) and to one frame per chunk,
    every frame bitwise equal to the one-chunk result."""
    exp, ov, opts = 16, 0.75, "--blackman_window --remove_dc --power_return --exact_window"
    N_ = 1 << exp
    hop = N_ // 4
    frames = 700
    L = (frames - 1) * hop + N_
    rng = np.random.default_rng(4242)
    x = (rng.standard_normal(L) + 0.5 * np.cos(np.pi * np.arange(L) ** 2 / (2 * L))).astype(np.float32)
    g = ostft.geometry(ostft.parse_options(opts, exp, ov), L)
    assert g["frames"] >= frames
    per = N_ // 2 + 1
    default = runner.scratch_budget()
    assert default >= g["frames"] * N_ * 4  # the whole request in one chunk
    y = runner.stft(x, exp, ov, opts).reshape(-1, per)
    t = po.stft64(x, exp, hop, g["frames"], po.WINDOWS["blackman"], 0.0, True, po.OUT_POWER).reshape(-1, per)
    check(y, t, per, TOL, "700 frames")
    try:
        for budget in (128 * N_ * 4, 1):
            assert runner.scratch_budget(budget) == budget
            yc = runner.stft(x, exp, ov, opts).reshape(-1, per)
            assert np.array_equal(yc.view(np.uint32), y.view(np.uint32)), budget
    finally:
        runner.scratch_budget(default)


@pytest.mark.parametrize("exp", [13, 15])
def test_odd_exponent_result_buffer(runner, exp):
    """Odd exponents >= 13 return the penultimate radix-2 stage (the reference's
    by-value result pointers); --exact_odd_exp returns the FFT."""
    rng = np.random.default_rng(3000 + exp)
    N_ = 1 << exp
    x = rng.uniform(-1, 1, 3 * N_ + 5).astype(np.float32)
    for opts in ("--hamming_window --remove_dc", "--hanning_window --half_complex_return", "--remove_dc"):
        gpu_vs_oracle(runner, x, exp, 0.5, opts)
    y = runner.stft(x, exp, 0.5, "--hamming_window --remove_dc --exact_odd_exp")
    g = ostft.geometry(ostft.make_params(exp, 0.5), len(x))
    t = po.stft64(x, exp, g["hop"], g["frames"], po.WINDOWS["hamming"], 0.0, True, po.OUT_COMPAT_POWER,
                  quirk_odd=False)
    check(y, t, N_, TOL, "exact odd")


@pytest.mark.parametrize("exp", [6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16])
def test_new_output_layouts_against_fp64(runner, exp):
    rng = np.random.default_rng(4000 + exp)
    N_ = 1 << exp
    x = rng.standard_normal(5 * N_ + 123).astype(np.float32)
    for flag, mode, per in (("--magnitude_return", po.OUT_MAGNITUDE, N_),
                            ("--power_return", po.OUT_POWER, N_ // 2 + 1),
                            ("--half_spectrum_return", po.OUT_HALF_SPECTRUM, N_ + 2)):
        opts = f"--hanning_window --remove_dc --exact_window --exact_odd_exp {flag}"
        y = runner.stft(x, exp, 0.5, opts)
        g = ostft.geometry(ostft.parse_options(opts, exp, 0.5), len(x))
        t = po.stft64(x, exp, g["hop"], g["frames"], po.WINDOWS["hanning"], 0.0, True, mode, quirk_odd=False)
        assert y.size == t.size
        check(y, t, per, TOL, f"{flag} exp={exp}")


@pytest.mark.parametrize("exp", [13, 15])
def test_magnitude_keeps_odd_quirk(runner, exp):
    """--magnitude_return is the CUDA runner's |X| over the same result buffer."""
    N_ = 1 << exp
    x = np.random.default_rng(5).uniform(-1, 1, 2 * N_).astype(np.float32)
    y = runner.stft(x, exp, 0.0, "--magnitude_return")
    g = ostft.geometry(ostft.make_params(exp, 0.0), len(x))
    t = po.stft64(x, exp, g["hop"], g["frames"], 0, 0.0, False, po.OUT_MAGNITUDE, quirk_odd=True)
    check(y, t, N_, TOL, "magnitude quirk")


@pytest.mark.parametrize("exp", [10, 15, 16, 17])
def test_dc_offset_signals_against_fp64(runner, exp):
    """A signal riding on a large offset: windowed and DC-removed, the single
    kernel (time-domain mean) and the multi-pass path (frequency-domain k = 0
    correction) both stay at the float32 FFT's error.  (Without a window the DC
    bin carries the float32 rounding of the mean in any float32 pipeline, the
    reference's included.)"""
    N_ = 1 << exp
    x = (1000.0 + np.random.default_rng(6000 + exp).standard_normal(5 * N_)).astype(np.float32)
    opts = "--hanning_window --remove_dc"
    y = runner.stft(x, exp, 0.5, opts)
    g = ostft.geometry(ostft.parse_options(opts, exp, 0.5), len(x))
    t = po.stft64(x, exp, g["hop"], g["frames"], po.WINDOWS["hanning"], 0.0, True, po.OUT_COMPAT_POWER,
                  quirk_odd=True)
    check(y, t, N_, 1e-6, f"offset 1000, exp={exp}")


def test_exact_window_matches_fp64_on_edge_frames(runner):
    """A tail frame holding a few samples at the window's rising edge: the
    default (reference) window follows the Serial runner, --exact_window
    follows the FP64 window."""
    exp = 14
    N_ = 1 << exp
    x = np.random.default_rng(6).uniform(-1, 1, 6 * N_ + 37).astype(np.float32)
    gpu_vs_oracle(runner, x, exp, 0.0, "--hanning_window")
    y = runner.stft(x, exp, 0.0, "--hanning_window --exact_window")
    g = ostft.geometry(ostft.make_params(exp, 0.0), len(x))
    t = po.stft64(x, exp, g["hop"], g["frames"], po.WINDOWS["hanning"], 0.0, False, po.OUT_COMPAT_POWER)
    check(y, t, N_, TOL, "exact window")


# --------------------------------------------------------------------- edges

@pytest.mark.parametrize("L", [0, 1, 63, 64, 65, 1000, 1024, 4096])
def test_short_and_ragged_signals(runner, L):
    x = np.random.default_rng(L).uniform(-1, 1, L).astype(np.float32)
    for exp, ov in ((6, 0.0), (6, 0.5), (10, 0.75), (12, 0.0)):
        y = runner.stft(x, exp, ov, "--hanning_window --remove_dc")
        rc, r = po.pipeline(x, exp, ov, "--hanning_window --remove_dc")
        assert y.shape == r.shape
        check(y, r, 1 << exp, TOL, f"L={L} exp={exp} ov={ov}")


def test_non_dyadic_overlaps(runner):
    rng = np.random.default_rng(7)
    for ov in (0.1, 0.33, 0.6, 0.9, 0.99):
        x = rng.uniform(-1, 1, 20000).astype(np.float32)
        gpu_vs_oracle(runner, x, 9, ov, "--blackman_harris_window --remove_dc")


def test_errors_are_reported(runner):
    x = np.ones(1000, np.float32)
    assert runner.activate_stft(x, 5, 0.0, "") is None
    with pytest.raises(ostft.STFTError) as ei:
        runner.stft(x, 5, 0.0, "")
    assert ei.value.status == N.ERR_UNSUPPORTED
    with pytest.raises(ostft.STFTError) as ei:
        runner.stft(x, 10, 0.0, "--gaussian_window=3<<sigma")
    assert ei.value.status == N.ERR_OPTION
    with pytest.raises(ostft.STFTError) as ei:
        runner.stft(x, 10, 1.0, "")
    assert ei.value.status == N.ERR_INVALID_ARG


# --------------------------------------------------------------------- device path

def test_run_device_matches_host_path(runner):
    import torch

    x = np.random.default_rng(8).uniform(-1, 1, 300_000).astype(np.float32)
    host = runner.stft(x, 11, 0.5, "--hanning_window --remove_dc")
    p = ostft.parse_options("--hanning_window --remove_dc", 11, 0.5)
    xd = torch.from_numpy(x).cuda()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        out = runner.run_device(xd, p, stream=s)
    s.synchronize()
    assert np.array_equal(out.cpu().numpy().view(np.uint32), host.view(np.uint32))


def test_two_pass_requests_on_two_streams(runner):
    """Windows 2^16..2^20 go through the context's one scratch buffer: four
    requests queued alternately on two streams, with no wait between them,
    give bitwise the results each gives alone."""
    import torch

    rng = np.random.default_rng(77)
    opts = "--blackman_window --remove_dc --power_return"
    xs = [rng.standard_normal(40 * 65536 + 123 * i).astype(np.float32) for i in range(4)]
    alone = [runner.stft(x, 16, 0.5, opts) for x in xs]
    p = ostft.parse_options(opts, 16, 0.5)
    xd = [torch.from_numpy(x).cuda() for x in xs]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    torch.cuda.synchronize()
    outs = [runner.run_device(x, p, stream=streams[i % 2]) for i, x in enumerate(xd)]
    torch.cuda.synchronize()
    for i, (o, a) in enumerate(zip(outs, alone)):
        assert np.array_equal(o.cpu().numpy().view(np.uint32), a.view(np.uint32)), i


@pytest.mark.parametrize("exp", [8, 10, 14, 16])
def test_frames_are_independent_of_the_batch(runner, exp):
    """A frame's result does not depend on which other frames share the launch
    (every frame is its own transform): bitwise equal alone and in a batch."""
    N_ = 1 << exp
    hop = N_ // 2
    x = np.random.default_rng(9).uniform(-1, 1, 40 * hop).astype(np.float32)
    full = runner.stft(x, exp, 0.5, "--hanning_window --remove_dc").reshape(-1, N_)
    for start in (10, 11):
        sub = runner.stft(x[start * hop: start * hop + 12 * hop], exp, 0.5,
                          "--hanning_window --remove_dc").reshape(-1, N_)
        assert np.array_equal(sub[:10].view(np.uint32), full[start:start + 10].view(np.uint32))


def test_config2_sampled_frames_full_size(runner):
    """1e8 samples of 8 sinusoids + noise, windows 2^6..2^12, overlap 0 / 50%,
    Hanning + DC, both outputs: sampled frames against the oracle run on each
    frame's own span."""
    import torch

    rng = np.random.default_rng(0x5EED0002)
    L = 100_000_000
    n = torch.arange(L, dtype=torch.float64, device="cuda")
    xd = torch.zeros(L, dtype=torch.float64, device="cuda")
    for _ in range(8):
        f, a, ph = rng.uniform(0, 0.5), rng.uniform(0.1, 1.0), rng.uniform(0, 2 * np.pi)
        xd += a * torch.cos(2 * np.pi * f * n + ph)
    xd += 0.05 * torch.randn(L, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(2))
    xd = xd.float().contiguous()
    del n
    for exp in range(6, 13):
        N_ = 1 << exp
        for ov in (0.0, 0.5):
            for opts in ("--hanning_window --remove_dc", "--hanning_window --remove_dc --half_complex_return"):
                p = ostft.parse_options(opts, exp, ov)
                g = ostft.geometry(p, L)
                out = runner.run_device(xd, p)
                torch.cuda.synchronize()
                frames = g["compute_frames"]
                pick = np.unique(np.r_[0, frames - 1, rng.integers(0, frames, 24)])
                for f in pick:
                    s0 = int(f) * g["hop"]
                    seg = xd[s0: s0 + N_].cpu().numpy()
                    rc, r = po.pipeline(seg, exp, 0.0, opts.replace(" --half_complex_return", ""))
                    if "--half_complex_return" in opts:
                        # frame f's spectrum, re/im interleaved, from the un-halved oracle
                        rr, ri = spectrum_of(seg, exp, opts)
                        ref = np.empty(2 * N_)
                        ref[0::2], ref[1::2] = rr, ri
                        lo = 2 * int(f) * N_
                        got = out[lo: lo + 2 * N_].cpu().numpy()
                        n_ok = min(2 * N_, 2 * g["out_half"] - lo)
                        check(got[:n_ok], ref[:n_ok], n_ok, TOL, f"hc exp={exp} ov={ov} f={f}")
                    else:
                        got = out[int(f) * N_: (int(f) + 1) * N_].cpu().numpy()
                        check(got, r[:N_], N_, TOL, f"exp={exp} ov={ov} f={f}")
                del out


def test_config3_sampled_frames_full_size(runner):
    """2^27 samples of N(0,1) noise + a linear chirp, windows 2^16..2^20,
    overlap 0 / 75%, Blackman + DC, power spectrum, with the default scratch
    budget (the 75% requests run in three chunks of frames): the first and
    last frames and those either side of every third of the request against
    the FP64 pipeline run on each frame's own span."""
    import torch

    L = 1 << 27
    rng = np.random.default_rng(0x5EED0003)
    xd = torch.empty(L, dtype=torch.float32, device="cuda")
    gen = torch.Generator(device="cuda").manual_seed(3)
    for s in range(0, L, 1 << 24):
        n = torch.arange(s, s + (1 << 24), dtype=torch.int64, device="cuda")
        ph = ((n * n) % (4 * L)).double() * (np.pi / (2 * L))
        xd[s: s + (1 << 24)] = (torch.randn(n.shape, dtype=torch.float64, device="cuda", generator=gen)
                                + 0.5 * torch.cos(ph)).float()
    del n, ph
    opts = "--blackman_window --remove_dc --power_return --exact_window"
    for exp in range(16, 21):
        N_ = 1 << exp
        per = N_ // 2 + 1
        for ov in (0.0, 0.75):
            p = ostft.parse_options(opts, exp, ov)
            g = ostft.geometry(p, L)
            out = runner.run_device(xd, p)
            torch.cuda.synchronize()
            frames, hop = g["frames"], g["hop"]
            thirds = [frames // 3, 2 * frames // 3]
            pick = sorted({0, 1, frames - 2, frames - 1, *[t + d for t in thirds for d in (-1, 0, 1)],
                           *rng.integers(0, frames, 4).tolist()})
            pairs = []
            for f in pick:
                s0 = f * hop
                seg = np.zeros(N_, np.float32)
                part = xd[s0: s0 + N_].cpu().numpy()
                seg[: part.size] = part
                t = po.stft64(seg, exp, hop, 1, po.WINDOWS["blackman"], 0.0, True, po.OUT_POWER)
                pairs.append((f, out[f * per: (f + 1) * per].cpu().numpy(), t))
            for f, got, t in pairs:
                # every sampled frame (FP64 window on both sides); frames past
                # the signal are all zero
                check(got, t, per, TOL, f"exp={exp} ov={ov} f={f}")
                if f * hop >= L:
                    assert not np.any(got), f
            del out


def spectrum_of(seg, exp, opts):
    """Re/Im of the first frame of `seg` from the restated Serial half-complex
    output (two frames' worth: the buffer half holds frame 0 whole)."""
    N_ = 1 << exp
    pad = np.zeros(2 * N_, np.float32)
    pad[: len(seg)] = seg
    rc, r = po.pipeline(pad, exp, 0.0, "--hanning_window --remove_dc --half_complex_return")
    return r[0: 2 * N_: 2].astype(np.float64), r[1: 2 * N_: 2].astype(np.float64)
