"""One context shared by several host threads; hops beyond the window; the
matrix-core form inside batches and the server.

The reference runner calls ActivateSTFT on IXWebSocket's per-connection
threads with no locking (StandAlone/cross_gpgpu/GLOBAL/RunnerCommon.cpp:77-120),
so the drop-in context must give every concurrent caller the result it would
get alone.  runnerFunction::Overlap receives OMove from the caller
(StandAlone/cross_gpgpu/GLOBAL/RunnerInterface.hpp:38-40): a negative
overlap ratio there gives a hop larger than the window, frames with gaps
between them.
"""
import ctypes
import threading

import numpy as np
import pytest

from conftest import frame_errors

import occastft_amd as ostft
from occastft_amd import _native as N

import pyoracle as po

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def runner():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    r = ostft.Runner(0)
    yield r
    r.close()


def bitwise_equal(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def activate_geometry(runner, x, p, hop, frames):
    per = {N.OUT_POWER: (1 << p.exp) // 2 + 1, N.OUT_HALF_SPECTRUM: (1 << p.exp) + 2}.get(p.output, 1 << p.exp)
    out = np.empty(frames * per, np.float32)
    n = ctypes.c_uint64(0)
    rc = N.lib().ostft_activate_geometry(runner._h, ctypes.byref(p), x.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                         x.size, hop, frames, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                         out.size, ctypes.byref(n))
    assert rc == N.OK, runner.last_error
    return out[: n.value]


# ------------------------------------------------------------------ hop > N

@pytest.mark.parametrize("exp", [6, 8, 10, 12, 13, 16])
@pytest.mark.parametrize("hop_of", ["N+1", "2N", "3N+5"])
def test_hop_beyond_the_window_against_fp64(runner, exp, hop_of):
    """Each frame reads its own N samples at q*hop (gaps skipped), zero past
    the end; the single kernel then stages fewer frames per workgroup tile."""
    Nw = 1 << exp
    hop = {"N+1": Nw + 1, "2N": 2 * Nw, "3N+5": 3 * Nw + 5}[hop_of]
    rng = np.random.default_rng(exp * 7 + hop)
    frames = 70 if exp < 16 else 5
    L = (frames - 1) * hop + Nw // 2  # the last frame runs past the end
    x = rng.standard_normal(L).astype(np.float32) + np.float32(0.2)
    for out_mode in (N.OUT_POWER, N.OUT_HALF_SPECTRUM):
        p = ostft.make_params(exp, 0.0, N.WIN_HANNING, 0.0, out_mode, N.FLAG_REMOVE_DC | N.FLAG_EXACT_WINDOW)
        y = activate_geometry(runner, x, p, hop, frames)
        ref = po.stft64(x, exp, hop, frames, window=N.WIN_HANNING, remove_dc=True, out_mode=out_mode)
        per = Nw // 2 + 1 if out_mode == N.OUT_POWER else Nw + 2
        mx, _, zabs = frame_errors(y, ref, per)
        assert mx <= 1e-5 and zabs <= 1e-6, (exp, hop, out_mode, mx, zabs)


# ------------------------------------------------------------------ threads

def _requests(seed, n):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        exp = int(rng.choice([6, 8, 10, 11, 12, 13, 14, 16, 17]))
        ov = float(rng.choice([0.0, 0.5, 0.75]))
        opts = str(rng.choice(["--hanning_window --remove_dc", "--blackman_window", "--hamming_window "
                               "--half_complex_return", "--hanning_window --power_return"]))
        L = int(rng.integers(1 << exp, 1 << 19))
        reqs.append((rng.uniform(-1, 1, L).astype(np.float32), exp, ov, opts))
    return reqs


def test_one_context_four_threads_host_buffers(runner):
    """4 host threads x 12 requests each on ONE context through ostft_activate
    (staging lanes, shared two-pass scratch for windows 2^16..2^17): every
    result bitwise equal to the same request run alone."""
    reqs = _requests(5, 48)
    alone = [runner.stft(x, e, ov, o) for x, e, ov, o in reqs]
    got = [None] * len(reqs)
    errors = []

    def worker(k):
        try:
            for rep in range(2):
                for i in range(k, len(reqs), 4):
                    x, e, ov, o = reqs[i]
                    got[i] = runner.stft(x, e, ov, o)
        except Exception as exc:  # noqa: BLE001 -- reported below
            errors.append(exc)

    th = [threading.Thread(target=worker, args=(k,)) for k in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    for i, (g, a) in enumerate(zip(got, alone)):
        assert bitwise_equal(g, a), f"request {i} {reqs[i][1:]}"


def test_one_context_threads_mixing_device_batch_and_host_calls(runner):
    """Threads on one context using different entry points at once: device
    calls on their own streams (single kernel and two-pass), batches, and
    host-buffer calls; each result bitwise equal to its run alone."""
    import torch

    reqs = _requests(9, 24)
    alone = [runner.stft(x, e, ov, o) for x, e, ov, o in reqs]
    errors = []
    got = {}

    def dev_worker(k):
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                for i in range(k, len(reqs), 3):
                    x, e, ov, o = reqs[i]
                    xd = torch.from_numpy(x).cuda()
                    y = runner.run_device(xd, ostft.parse_options(o, e, ov), stream=s)
                    s.synchronize()
                    got[("dev", i)] = y.cpu().numpy()
        except Exception as exc:  # noqa: BLE001
            errors.append(exc)

    def batch_worker():
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                idx = list(range(0, len(reqs), 2))
                jobs = [(torch.from_numpy(reqs[i][0]).cuda(), ostft.parse_options(reqs[i][3], reqs[i][1],
                                                                                   reqs[i][2]), None)
                        for i in idx]
                outs = runner.run_batch_device(jobs, stream=s)
                s.synchronize()
                for i, o in zip(idx, outs):
                    got[("batch", i)] = o.cpu().numpy()
        except Exception as exc:  # noqa: BLE001
            errors.append(exc)

    def host_worker():
        try:
            for i in range(1, len(reqs), 2):
                x, e, ov, o = reqs[i]
                got[("host", i)] = runner.stft(x, e, ov, o)
        except Exception as exc:  # noqa: BLE001
            errors.append(exc)

    th = [threading.Thread(target=dev_worker, args=(k,)) for k in range(3)]
    th += [threading.Thread(target=batch_worker), threading.Thread(target=host_worker)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    assert len(got) == len(reqs) + len(range(0, len(reqs), 2)) + len(range(1, len(reqs), 2))
    for (kind, i), g in got.items():
        assert bitwise_equal(g, alone[i]), f"{kind} request {i} {reqs[i][1:]}"


def test_last_error_is_per_thread(runner):
    """A failing call on one thread does not overwrite another thread's error
    text on the same context."""
    bar = threading.Barrier(2)
    seen = {}

    def fail(tag, exp):
        p = ostft.make_params(exp, 0.5, N.WIN_HANNING, 0.0, N.OUT_COMPAT_POWER, 0)
        x = np.ones(4096, np.float32)
        out = np.empty(16, np.float32)  # too small
        bar.wait()
        rc = N.lib().ostft_activate(runner._h, x.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), x.size,
                                    exp if tag == "cap" else 5, 0.5, b"", out.ctypes.data_as(
                                        ctypes.POINTER(ctypes.c_float)), out.size, None)
        bar.wait()
        seen[tag] = (rc, runner.last_error)

    th = [threading.Thread(target=fail, args=("cap", 10)), threading.Thread(target=fail, args=("exp", 10))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert seen["cap"][0] == N.ERR_CAPACITY and "too small" in seen["cap"][1]
    assert seen["exp"][0] == N.ERR_UNSUPPORTED and "not supported" in seen["exp"][1]


# ------------------------------------------------------------------ MFMA in batches

def test_batch_with_a_matrix_core_request(runner):
    """A batch holding --mfma_fp16 / --mfma_bf16 requests: those run on their
    own launches; every request equals its run alone, bitwise."""
    import torch

    rng = np.random.default_rng(3)
    reqs = []
    for i, opts in enumerate(["--hanning_window", "--hanning_window --mfma_fp16 --half_spectrum_return",
                              "--blackman_window --remove_dc", "--hanning_window --mfma_bf16 --power_return",
                              "--hamming_window --half_complex_return"]):
        exp = [10, 6, 12, 8, 16][i]
        x = rng.uniform(-1, 1, 50_000 + 777 * i).astype(np.float32)
        reqs.append((x, ostft.parse_options(opts, exp, 0.5)))
    alone = []
    for x, p in reqs:
        y = runner.run_device(torch.from_numpy(x).cuda(), p)
        torch.cuda.synchronize()
        alone.append(y.cpu().numpy())
    outs = runner.run_batch_device([(torch.from_numpy(x).cuda(), p, None) for x, p in reqs])
    torch.cuda.synchronize()
    for i, (o, a) in enumerate(zip(outs, alone)):
        assert bitwise_equal(o.cpu().numpy(), a), f"request {i}"


def test_server_with_a_matrix_core_request():
    """Through the continuous-batching server: a --mfma_fp16 request among
    FP32 ones; nobody's request fails and each result equals its run alone."""
    import torch

    rng = np.random.default_rng(4)
    reqs = []
    for i in range(12):
        opts = "--hanning_window --power_return" + (" --mfma_fp16" if i == 5 else "")
        exp = 8 if i == 5 else int(rng.choice([8, 9, 10, 11]))
        reqs.append((rng.uniform(-1, 1, 30_000 + 1000 * i).astype(np.float32), ostft.parse_options(opts, exp, 0.5)))
    with ostft.Runner(0) as r:
        alone = []
        for x, p in reqs:
            y = r.run_device(torch.from_numpy(x).cuda(), p)
            torch.cuda.synchronize()
            alone.append(y.cpu().numpy())
    with ostft.Server(0, arena_bytes=64 << 20) as srv:
        tickets = [srv.submit(x, p) for x, p in reqs]
        outs = [srv.wait(t) for t in tickets]
        st = srv.stats()
    assert st["failed"] == 0 and st["completed"] == len(reqs)
    for i, (o, a) in enumerate(zip(outs, alone)):
        assert bitwise_equal(o, a), f"request {i}"
