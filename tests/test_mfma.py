"""The split-precision MFMA form (--mfma_fp16 / --mfma_bf16) and strided
batches.

Bar for the FP16 form: per-frame relative L2 error <= 1e-5 against the FP64
pipeline (oracle/fp64.c), the same bar as the FP32 path, for every window kind
and output layout on windows 64, 256, 1024, including signals scaled by 1e-20
and 1e20 (every GEMM column is scaled by a power of two from its own maximum)
and all-zero signals.  The BF16 form carries 16 significant bits per operand
pair (about 1.5e-5 per term at worst, 5-7e-6 measured per frame): it is held
to the same 1e-5 bar on noise signals.
"""
import numpy as np
import pytest

from conftest import frame_errors

import occastft_amd as ostft
from occastft_amd import _native as N

import pyoracle as po

pytestmark = pytest.mark.gpu

EXPS = (6, 8, 10)
FP64_MODE = {N.OUT_COMPAT_POWER: po.OUT_COMPAT_POWER, N.OUT_HALF_COMPLEX: po.OUT_HALF_COMPLEX,
             N.OUT_MAGNITUDE: po.OUT_MAGNITUDE, N.OUT_POWER: po.OUT_POWER,
             N.OUT_HALF_SPECTRUM: po.OUT_HALF_SPECTRUM}
WINDOW_NAME = {N.WIN_NONE: "none", N.WIN_HANNING: "hanning", N.WIN_HAMMING: "hamming",
               N.WIN_BLACKMAN: "blackman", N.WIN_BLACKMAN_HARRIS: "blackman_harris"}
# against FP64 the window is the FP64 formula (--exact_window): the reference's
# float coefficients differ from it by up to 4e-4 near the tapers, which is not
# what these tests measure
EXACT = N.FLAG_EXACT_WINDOW


@pytest.fixture(scope="module")
def runner():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    r = ostft.Runner(0)
    yield r
    r.close()


def run(runner, x, p):
    import torch

    y = runner.run_device(torch.from_numpy(x).cuda(), p)
    torch.cuda.synchronize()
    return y.cpu().numpy()


def fp64(x, p):
    g = ostft.geometry(p, x.size)
    win = po.WINDOWS[WINDOW_NAME[p.window]]
    frames = g["frames"]
    return po.stft64(x, p.exp, g["hop"], frames, win, 0.0, bool(p.flags & N.FLAG_REMOVE_DC),
                     FP64_MODE[p.output]), g


def per_frame(p):
    N_ = 1 << p.exp
    return {N.OUT_POWER: N_ // 2 + 1, N.OUT_HALF_SPECTRUM: N_ + 2}.get(p.output, N_)


def check_vs_fp64(runner, x, p, tol):
    y = run(runner, x, p)
    r, g = fp64(x, p)
    r = np.asarray(r, np.float64)[: y.size]
    assert y.size == g["out_floats"]
    mx, p50, zabs = frame_errors(y, r, per_frame(p))
    assert mx <= tol and zabs <= 1e-6 * max(1.0, float(np.abs(x).max())), \
        (p.exp, p.window, p.output, float(np.abs(x).max()), mx)
    return mx


@pytest.mark.parametrize("exp", EXPS)
@pytest.mark.parametrize("out", sorted(FP64_MODE))
def test_fp16_form_against_fp64(runner, exp, out):
    x = np.random.default_rng(exp * 10 + out).uniform(-1, 1, 40000).astype(np.float32)
    for win in (N.WIN_NONE, N.WIN_HANNING, N.WIN_BLACKMAN_HARRIS):
        p = ostft.make_params(exp, 0.5, win, 0.0, out, N.FLAG_REMOVE_DC | N.FLAG_MFMA_FP16 | EXACT)
        check_vs_fp64(runner, x, p, 1e-5)


@pytest.mark.parametrize("exp", EXPS)
def test_fp16_form_is_scale_invariant(runner, exp):
    """Columns are scaled by powers of two: tiny and huge signals keep the
    same relative error (no FP16 underflow or overflow).  The complex layout
    is used: magnitudes square the values, which leaves the FP32 range at
    1e+-20 on any path."""
    base = np.random.default_rng(3).standard_normal(30000)
    for scale in (1e-20, 1e-6, 1.0, 1e6, 1e20):
        x = (base * scale).astype(np.float32)
        p = ostft.make_params(exp, 0.75, N.WIN_HAMMING, 0.0, N.OUT_HALF_SPECTRUM, N.FLAG_MFMA_FP16 | EXACT)
        check_vs_fp64(runner, x, p, 1e-5)


@pytest.mark.parametrize("exp", EXPS)
def test_fp16_form_pure_tones_and_zeros(runner, exp):
    n = np.arange(50000)
    N_ = 1 << exp
    x = (np.cos(2 * np.pi * 5 * n / N_) + 1e-3 * np.sin(2 * np.pi * 0.31 * n)).astype(np.float32)
    x[20000:30000] = 0.0  # all-zero frames
    p = ostft.make_params(exp, 0.5, N.WIN_HANNING, 0.0, N.OUT_POWER, N.FLAG_REMOVE_DC | N.FLAG_MFMA_FP16 | EXACT)
    check_vs_fp64(runner, x, p, 1e-5)


@pytest.mark.parametrize("exp", EXPS)
def test_bf16_form_error_is_bounded(runner, exp):
    x = np.random.default_rng(exp).uniform(-1, 1, 40000).astype(np.float32)
    p = ostft.make_params(exp, 0.5, N.WIN_HANNING, 0.0, N.OUT_MAGNITUDE,
                          N.FLAG_REMOVE_DC | N.FLAG_MFMA_BF16 | EXACT)
    mx = check_vs_fp64(runner, x, p, 1e-5)
    assert mx > 1e-7  # it is the 16-bit form, not the FP32 path


def test_mfma_form_against_the_serial_runner(runner):
    """Default layout (the Serial runner's sqrt(2)|Re X|): the FP16 form meets
    the 1e-5 bar against the restated reference too."""
    x = np.random.default_rng(9).uniform(-1, 1, 1 << 16).astype(np.float32)
    for exp in EXPS:
        y = run(runner, x, ostft.parse_options("--hanning_window --remove_dc --mfma_fp16", exp, 0.5))
        rc, r = po.pipeline(x, exp, 0.5, "--hanning_window --remove_dc")
        assert rc == 0 and y.shape == r.shape
        mx, _, zabs = frame_errors(y, r, 1 << exp)
        assert mx <= 1e-5 and zabs <= 1e-6, (exp, mx)


@pytest.mark.parametrize("flags", [0, N.FLAG_MFMA_FP16])
def test_strided_batch_matches_single_requests(runner, flags):
    import torch

    rng = np.random.default_rng(4)
    for exp in (6, 8, 10, 13, 16):
        if flags and exp >= 13:
            continue
        # window 2^16: the two-pass path, one signal after the other; the
        # 3-float row offset leaves the sample pairs unaligned
        count, L = (3, 3 * 65536 + 77) if exp == 16 else (37, 20000 + exp)
        x = torch.from_numpy(rng.uniform(-1, 1, (count, L + 5)).astype(np.float32)).cuda()[:, 3: 3 + L]
        p = ostft.make_params(exp, 0.5, N.WIN_HANNING, 0.0, N.OUT_COMPAT_POWER, N.FLAG_REMOVE_DC | flags)
        out = runner.run_strided_device(x, p)
        torch.cuda.synchronize()
        for y in sorted({0, 1, min(17, count - 1), count - 1}):
            one = runner.run_device(x[y].contiguous(), p)
            torch.cuda.synchronize()
            assert np.array_equal(out[y].cpu().numpy().view(np.uint32), one.cpu().numpy().view(np.uint32))


def test_config4_full_size(runner):
    """4096 signals x 2^18 samples of unit-variance 1/f noise, windows 64 / 256
    / 1024, 50%, Hann, half spectrum, one strided launch per form: every frame
    of four signals against the FP64 pipeline (FP32 and FP16 forms <= 1e-5,
    BF16 <= 1e-5)."""
    import torch

    S, L = 4096, 1 << 18
    x = torch.empty(S, L, dtype=torch.float32, device="cuda")
    gen = torch.Generator(device="cuda").manual_seed(4)
    f = torch.arange(L // 2 + 1, dtype=torch.float32, device="cuda")
    shape = torch.where(f > 0, 1.0 / torch.sqrt(torch.clamp(f, min=1.0)), torch.zeros_like(f))
    for s in range(0, S, 256):
        w = torch.randn(256, L, dtype=torch.float32, device="cuda", generator=gen)
        y = torch.fft.irfft(torch.fft.rfft(w, dim=1) * shape, n=L, dim=1)
        x[s: s + 256] = y / y.std(dim=1, keepdim=True)
    del w, y
    rows = (0, 1, 2047, 4095)
    xs = {r: x[r].cpu().numpy() for r in rows}
    for exp in EXPS:
        for flags, tol in ((0, 1e-5), (N.FLAG_MFMA_FP16, 1e-5), (N.FLAG_MFMA_BF16, 1e-5)):
            p = ostft.make_params(exp, 0.5, N.WIN_HANNING, 0.0, N.OUT_HALF_SPECTRUM, EXACT | flags)
            out = runner.run_strided_device(x, p)
            torch.cuda.synchronize()
            for r in rows:
                ref, g = fp64(xs[r], p)
                got = out[r, : g["out_floats"]].cpu().numpy()
                mx, _, zabs = frame_errors(got, ref, (1 << exp) + 2)
                assert mx <= tol and zabs <= 1e-6, (exp, flags, r, mx)
            del out


# ------------------------------------------------------------ edge geometry

def _activate_geometry(runner, x, p, hop, frames):
    import ctypes

    per = {N.OUT_POWER: (1 << p.exp) // 2 + 1, N.OUT_HALF_SPECTRUM: (1 << p.exp) + 2}.get(p.output, 1 << p.exp)
    out = np.empty(max(frames * per, 1), np.float32)
    n = ctypes.c_uint64(0)
    rc = N.lib().ostft_activate_geometry(runner._h, ctypes.byref(p), x.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                         x.size, hop, frames, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                         out.size, ctypes.byref(n))
    assert rc == N.OK, runner.last_error
    return out[: n.value]


@pytest.mark.parametrize("exp", EXPS)
@pytest.mark.parametrize("hop_of", ["1", "N/2+3", "N+1", "3N+5"])
@pytest.mark.parametrize("flags", [N.FLAG_MFMA_FP16, N.FLAG_MFMA_BF16])
def test_matrix_core_form_any_hop(runner, exp, hop_of, flags):
    """The matrix-core kernels read frame f at f*hop for any hop: 1 (every
    frame overlaps the next in all but one sample), odd, and beyond the
    window (gaps), with a last frame running past the end of the signal and a
    frame count that leaves the last tile partial."""
    Nw = 1 << exp
    hop = {"1": 1, "N/2+3": Nw // 2 + 3, "N+1": Nw + 1, "3N+5": 3 * Nw + 5}[hop_of]
    rng = np.random.default_rng(exp * 13 + hop)
    frames = 37  # not a multiple of any tile (16, 8 or 2 frames)
    L = (frames - 1) * hop + Nw // 2
    x = rng.standard_normal(L).astype(np.float32) + np.float32(0.3)
    for out_mode in (N.OUT_POWER, N.OUT_HALF_SPECTRUM, N.OUT_COMPAT_POWER):
        p = ostft.make_params(exp, 0.0, N.WIN_HANNING, 0.0, out_mode, N.FLAG_REMOVE_DC | EXACT | flags)
        y = _activate_geometry(runner, x, p, hop, frames)
        ref = po.stft64(x, exp, hop, frames, window=N.WIN_HANNING, remove_dc=True, out_mode=FP64_MODE[out_mode])
        mx, _, zabs = frame_errors(y, np.asarray(ref, np.float64)[: y.size], per_frame(p))
        assert mx <= 1e-5 and zabs <= 1e-6, (exp, hop, out_mode, flags, mx, zabs)


@pytest.mark.parametrize("exp", EXPS)
@pytest.mark.parametrize("case", ["shorter_than_window", "one_tile_and_a_frame", "misaligned_input"])
def test_matrix_core_form_short_and_misaligned_signals(runner, exp, case):
    """Signals shorter than one window (one zero-padded frame), a frame count
    of one tile plus one, and a device input 4 bytes past a 16-byte boundary
    (the window-64 kernel's vector loads fall back to scalar ones)."""
    import torch

    Nw = 1 << exp
    tile = {6: 16, 8: 8, 10: 2}[exp]
    rng = np.random.default_rng(exp + len(case))
    L = {"shorter_than_window": Nw // 2 + 5, "one_tile_and_a_frame": tile * (Nw // 2) + Nw // 2,
         "misaligned_input": 40 * Nw + 7}[case]
    x = rng.uniform(-1, 1, L).astype(np.float32)
    for out_mode in (N.OUT_HALF_COMPLEX, N.OUT_POWER):
        p = ostft.make_params(exp, 0.5, N.WIN_BLACKMAN, 0.0, out_mode, N.FLAG_REMOVE_DC | EXACT | N.FLAG_MFMA_FP16)
        if case == "misaligned_input":
            buf = torch.zeros(L + 4, dtype=torch.float32, device="cuda")
            buf[1: 1 + L] = torch.from_numpy(x).cuda()
            xin = buf[1: 1 + L]
            assert xin.data_ptr() % 16 == 4
            y = runner.run_device(xin, p)
            torch.cuda.synchronize()
            y = y.cpu().numpy()
        else:
            y = run(runner, x, p)
        r, g = fp64(x, p)
        r = np.asarray(r, np.float64)[: y.size]
        assert y.size == g["out_floats"] and g["frames"] >= 1
        mx, _, zabs = frame_errors(y, r, per_frame(p))
        assert mx <= 1e-5 and zabs <= 1e-6, (exp, case, out_mode, mx)


def test_matrix_core_form_empty_signal(runner):
    """An empty signal gives an empty result, as on the FP32 path."""
    import torch

    for exp in EXPS:
        p = ostft.make_params(exp, 0.5, N.WIN_HANNING, 0.0, N.OUT_POWER, N.FLAG_MFMA_FP16)
        g = ostft.geometry(p, 0)
        y = runner.run_device(torch.zeros(0, dtype=torch.float32, device="cuda"), p)
        torch.cuda.synchronize()
        assert y.numel() == g["out_floats"]


@pytest.mark.parametrize("exp", [8])
@pytest.mark.parametrize("flags", [N.FLAG_MFMA_FP16, N.FLAG_MFMA_BF16])
def test_vector_and_scalar_loads_agree_bitwise(runner, exp, flags):
    """Window 256: the 16-byte-load path (input, hop and stride multiples of
    4 floats, each quad's blocks transposed across its lanes) and the
    per-sample path (the same signal 4 bytes off a 16-byte boundary) hand
    every lane the same samples, so results are bitwise equal."""
    import torch

    Nw = 1 << exp
    rng = np.random.default_rng(900 + exp)
    L = 300 * Nw + 11
    x = torch.from_numpy(rng.uniform(-1, 1, L).astype(np.float32)).cuda()
    buf = torch.zeros(L + 4, dtype=torch.float32, device="cuda")
    buf[1: 1 + L] = x
    xm = buf[1: 1 + L]
    assert x.data_ptr() % 16 == 0 and xm.data_ptr() % 16 == 4
    for out_mode in (N.OUT_POWER, N.OUT_HALF_COMPLEX, N.OUT_COMPAT_POWER):
        p = ostft.make_params(exp, 0.5, N.WIN_HANNING, 0.0, out_mode, N.FLAG_REMOVE_DC | flags)
        a = runner.run_device(x, p)
        b = runner.run_device(xm, p)
        torch.cuda.synchronize()
        a, b = a.cpu().numpy(), b.cpu().numpy()
        assert a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32)), (exp, out_mode)
