"""Batches (ostft_run_batch_device) and continuous-batching serving
(ostft_server_*).

The bar for both is the one-request path, which the parity tests pin to the
oracle: every request's result in a batch or in the server's mix is bitwise
identical to the same request run alone with ostft_run_device (each frame is
its own transform and the batch kernels run the same per-frame code).  A few
requests are also checked against the oracle directly.
"""
import threading
import time

import numpy as np
import pytest

from conftest import frame_errors

import occastft_amd as ostft
from occastft_amd import _native as N

import pyoracle as po

pytestmark = pytest.mark.gpu

OUTS = (N.OUT_COMPAT_POWER, N.OUT_HALF_COMPLEX, N.OUT_MAGNITUDE, N.OUT_POWER, N.OUT_HALF_SPECTRUM)
WINS = (N.WIN_NONE, N.WIN_HANNING, N.WIN_HAMMING, N.WIN_BLACKMAN, N.WIN_FLATTOP, N.WIN_GAUSSIAN)


@pytest.fixture(scope="module")
def runner():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    r = ostft.Runner(0)
    yield r
    r.close()


def ragged_requests(seed, n, exps=range(6, 15), max_len=1 << 17, min_len=None):
    """n requests of mixed window exponent, overlap, window, layout, DC flag
    and length (including lengths below one window and not multiples of 4)."""
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        exp = int(rng.choice(list(exps)))
        ov = float(rng.choice([0.0, 0.25, 0.5, 0.75, 0.6]))
        win = int(rng.choice(WINS))
        out = int(rng.choice(OUTS))
        flags = N.FLAG_REMOVE_DC if rng.random() < 0.7 else 0
        if rng.random() < 0.2:
            flags |= N.FLAG_EXACT_ODD
        if min_len is None:
            L = int(rng.choice([1, 63, (1 << exp) - 1, int(np.exp(rng.uniform(np.log(64), np.log(max_len))))]))
        else:
            L = int(rng.integers(min_len, max_len))
        p = ostft.make_params(exp, ov, win, 1.5 if win == N.WIN_GAUSSIAN else 0.0, out, flags)
        x = rng.uniform(-1, 1, L).astype(np.float32) + np.float32(0.3)
        reqs.append((x, p))
    return reqs


def run_alone(runner, x, p):
    import torch

    xd = torch.from_numpy(x).cuda()
    y = runner.run_device(xd, p)
    torch.cuda.synchronize()
    return y.cpu().numpy()


def device_inputs(reqs, misalign=False):
    """Inputs as CUDA tensors; with `misalign` every other one starts one float
    past a 16-byte boundary (the kernels' scalar-load path)."""
    import torch

    xs = []
    for i, (x, _) in enumerate(reqs):
        if misalign and i % 2:
            buf = torch.empty(x.size + 1, dtype=torch.float32, device="cuda")
            buf[1:] = torch.from_numpy(x).cuda()
            xs.append(buf[1:])
        else:
            xs.append(torch.from_numpy(x).cuda())
    return xs


def bitwise_equal(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


# ------------------------------------------------------------------ batches

@pytest.mark.parametrize("misalign", [False, True])
@pytest.mark.parametrize("form", ["dispatch", "persistent"])
def test_batch_matches_requests_run_alone(runner, misalign, form):
    import torch

    reqs = ragged_requests(11 + misalign, 48, exps=list(range(6, 15)) + [15, 16])
    alone = [run_alone(runner, x, p) for x, p in reqs]
    xs = device_inputs(reqs, misalign)
    for e in range(6, 15):
        runner.batch_config(e, blocks_per_cu=2 if form == "persistent" else -1)
    try:
        outs = runner.run_batch_device([(xd, p, None) for xd, (_, p) in zip(xs, reqs)])
        torch.cuda.synchronize()
    finally:
        for e in range(6, 15):
            runner.batch_config(e, blocks_per_cu=-1)
    for i, (o, a) in enumerate(zip(outs, alone)):
        assert bitwise_equal(o.cpu().numpy(), a), f"request {i}: exp {reqs[i][1].exp} out {reqs[i][1].output}"


def test_batch_against_oracle(runner):
    import torch

    rng = np.random.default_rng(5)
    cases = [(7, 0.5, "--hanning_window --remove_dc"), (10, 0.75, "--blackman_window"),
             (12, 0.25, "--hamming_window --remove_dc --half_complex_return"),
             (13, 0.5, "--hanning_window --remove_dc"), (9, 0.0, "--nuttall_window --remove_dc")]
    xs = [rng.uniform(-1, 1, int(rng.integers(5000, 60000))).astype(np.float32) for _ in cases]
    jobs = [(torch.from_numpy(x).cuda(), ostft.parse_options(o, e, ov), None) for x, (e, ov, o) in zip(xs, cases)]
    outs = runner.run_batch_device(jobs)
    torch.cuda.synchronize()
    for x, (e, ov, o), y in zip(xs, cases, outs):
        rc, r = po.pipeline(x, e, ov, o)
        assert rc == 0 and y.numel() == r.size
        mx, _, zabs = frame_errors(y.cpu().numpy(), r, 1 << e)
        assert mx <= 1e-5 and zabs <= 1e-6, (e, ov, o, mx)


def test_batch_graph_replays_match_direct_launches(runner):
    """The captured launch sequence is replayed for every batch with the same
    set of window sizes: replays, with other work and synchronisations between
    them, give the direct launches' results every time."""
    import torch

    reqs = ragged_requests(21, 24, exps=[8, 10, 12])
    xs = device_inputs(reqs)
    runner.batch_config(use_graphs=0)
    ref = [o.cpu().numpy() for o in runner.run_batch_device([(xd, p, None) for xd, (_, p) in zip(xs, reqs)])]
    runner.batch_config(use_graphs=1)
    try:
        for rep in range(4):
            outs = runner.run_batch_device([(xd, p, None) for xd, (_, p) in zip(xs, reqs)])
            torch.cuda.synchronize()
            for i, (o, r) in enumerate(zip(outs, ref)):
                assert bitwise_equal(o.cpu().numpy(), r), f"replay {rep} request {i}"
            # eager work and a batch of another set of window sizes in between
            run_alone(runner, reqs[0][0], reqs[0][1])
            one = runner.run_batch_device([(xs[1], reqs[1][1], None)])
            torch.cuda.synchronize()
            assert bitwise_equal(one[0].cpu().numpy(), ref[1]), f"replay {rep}: single-request batch"
    finally:
        runner.batch_config(use_graphs=1)


def test_batch_rejects_a_bad_request_before_launching(runner):
    import torch

    x = torch.ones(4096, device="cuda")
    good = ostft.make_params(10, 0.5, N.WIN_HANNING)
    bad = ostft.make_params(5, 0.5, N.WIN_HANNING)
    out = torch.full((16384,), -7.0, device="cuda")
    with pytest.raises(ostft.STFTError) as ei:
        runner.run_batch_device([(x, good, out), (x, bad, None)])
    assert ei.value.status == N.ERR_UNSUPPORTED and "job 1" in str(ei.value)
    torch.cuda.synchronize()
    assert bool((out == -7.0).all())  # nothing ran
    small = torch.empty(16, device="cuda")
    with pytest.raises(ostft.STFTError) as ei:
        runner.run_batch_device([(x, good, small)])
    assert ei.value.status == N.ERR_CAPACITY


def test_empty_batch(runner):
    assert runner.run_batch_device([]) == []


def test_autotune_keeps_results(runner):
    import torch

    choice = runner.autotune_batch(10, 1 << 24)
    assert choice.blocks_per_cu >= 0 and choice.ms > 0  # 0: one workgroup per tile
    assert runner.batch_config(10) == choice.blocks_per_cu
    assert runner.engine_variant(10) == (choice.shape, choice.plan)
    reqs = ragged_requests(31, 12, exps=[10])
    alone = [run_alone(runner, x, p) for x, p in reqs]
    xs = device_inputs(reqs)
    outs = runner.run_batch_device([(xd, p, None) for xd, (_, p) in zip(xs, reqs)])
    torch.cuda.synchronize()
    for o, a in zip(outs, alone):
        assert bitwise_equal(o.cpu().numpy(), a)
    for bpc in (1, 2, -1):  # persistent at 1 and 2 per CU, then one workgroup per tile
        runner.batch_config(10, blocks_per_cu=bpc)
        outs = runner.run_batch_device([(xd, p, None) for xd, (_, p) in zip(xs, reqs)])
        torch.cuda.synchronize()
        for o, a in zip(outs, alone):
            assert bitwise_equal(o.cpu().numpy(), a)
    runner.engine_variant(10, 0, 0)


# ------------------------------------------------------------------ serving

def test_server_results_alone_and_in_the_mix(runner):
    """Continuous batching: the same requests submitted one at a time
    (each alone on the GPU) and all at once from several threads (mixed into
    shared batches, later arrivals admitted while earlier batches run) give
    bitwise the results of ostft_run_device."""
    reqs = ragged_requests(41, 60, exps=list(range(8, 14)), max_len=1 << 18)
    alone = [run_alone(runner, x, p) for x, p in reqs]
    with ostft.Server(0, arena_bytes=64 << 20) as srv:
        for i, (x, p) in enumerate(reqs[:10]):
            y = srv.wait(srv.submit(x, p))
            assert bitwise_equal(y, alone[i]), f"alone {i}"
        results = {}

        def client(idx):
            tickets = [(i, srv.submit(reqs[i][0], reqs[i][1])) for i in idx]
            for i, t in tickets:
                results[i] = srv.wait(t)

        ts = [threading.Thread(target=client, args=(list(range(k, len(reqs), 4)),)) for k in range(4)]
        for t in ts:
            t.start()
        for t in ts:
            t.join(timeout=120)
        assert len(results) == len(reqs)
        for i in range(len(reqs)):
            assert bitwise_equal(results[i], alone[i]), f"mixed {i}"
        st = srv.stats()
    assert st["completed"] == len(reqs) + 10 and st["failed"] == 0
    assert st["batches"] < st["completed"]  # requests shared batches
    assert 0 < st["latency_p50_ms"] <= st["latency_p99_ms"] <= st["latency_max_ms"]


def test_server_autotuned_for_a_template(runner):
    """A server autotuned at creation for a template request (its context
    keeps whatever variants the tuner picked, so results need not be bitwise
    those of another context) returns every request within the parity bar
    of the FP64 pipeline."""
    opts = "--hanning_window --remove_dc --exact_window --exact_odd_exp --power_return"
    like = ostft.parse_options(opts, 10, 0.75)
    rng = np.random.default_rng(77)
    with ostft.Server(0, arena_bytes=64 << 20, autotune=True, autotune_like=like, autotune_samples=1 << 20) as srv:
        for e in (8, 11, 13):
            n = 1 << e
            x = rng.uniform(-1, 1, 7 * n + 123).astype(np.float32)
            p = ostft.parse_options(opts, e, 0.75)
            y = srv.wait(srv.submit(x, p))
            g = ostft.geometry(p, len(x))
            t = po.stft64(x, e, g["hop"], g["frames"], po.WINDOWS["hanning"], 0.0, True, po.OUT_POWER,
                          quirk_odd=False)
            mx, p50, zabs = frame_errors(y, t, n // 2 + 1)
            assert mx <= 1e-5 and zabs <= 1e-6, f"exp {e}: max per-frame rel err {mx:.3e}"


def test_server_full_request_distribution(runner):
    """200 requests of the serving workload's distribution (lengths
    log-uniform 2^14..2^22, windows 2^8..2^13, overlap 0/25/50/75%, Hann /
    Hamming / Blackman, power spectrum) submitted at once from four threads
    to a server with the default arena: every result bitwise that of
    ostft_run_device alone."""
    rng = np.random.default_rng(5)
    wins = (N.WIN_HANNING, N.WIN_HAMMING, N.WIN_BLACKMAN)
    reqs = []
    for _ in range(200):
        L = int(np.exp(rng.uniform(np.log(2.0 ** 14), np.log(2.0 ** 22))))
        p = ostft.make_params(int(rng.integers(8, 14)), float(rng.choice([0.0, 0.25, 0.5, 0.75])),
                              int(rng.choice(wins)), 0.0, N.OUT_POWER, N.FLAG_REMOVE_DC)
        f, a = rng.uniform(0, 0.5), rng.uniform(0.1, 1.0)
        x = (rng.uniform(-1, 1, L) + a * np.sin(2 * np.pi * f * np.arange(L))).astype(np.float32)
        reqs.append((x, p))
    alone = [run_alone(runner, x, p) for x, p in reqs]
    results = {}
    with ostft.Server(0) as srv:
        def client(idx):
            tickets = [(i, srv.submit(reqs[i][0], reqs[i][1])) for i in idx]
            for i, t in tickets:
                results[i] = srv.wait(t)

        ts = [threading.Thread(target=client, args=(list(range(k, len(reqs), 4)),)) for k in range(4)]
        for t in ts:
            t.start()
        for t in ts:
            t.join(timeout=120)
        st = srv.stats()
    assert len(results) == len(reqs) and st["failed"] == 0
    for i in range(len(reqs)):
        assert bitwise_equal(results[i], alone[i]), i


def test_server_admission_by_footprint(runner):
    """A small arena: requests wait for memory to free up and still all
    complete correctly; a request larger than the arena is refused at submit."""
    # 0.5-1 MB of input and up to ~4 MB of output each: at most a few fit at once
    reqs = ragged_requests(51, 40, exps=[9, 10, 11], min_len=1 << 17, max_len=1 << 18)
    alone = [run_alone(runner, x, p) for x, p in reqs]
    with ostft.Server(0, arena_bytes=8 << 20, max_inflight=2) as srv:
        tickets = [srv.submit(x, p) for x, p in reqs]
        for i, t in enumerate(tickets):
            assert bitwise_equal(srv.wait(t), alone[i]), f"request {i}"
        big = np.zeros(2 << 20, np.float32)
        with pytest.raises(ostft.STFTError) as ei:
            srv.submit(big, ostft.make_params(10, 0.5))
        assert ei.value.status == N.ERR_TOO_LARGE
        st = srv.stats()
        assert st["completed"] == len(reqs) and st["batches"] >= 8


def test_server_errors_and_timeouts(runner):
    with ostft.Server(0, arena_bytes=16 << 20) as srv:
        with pytest.raises(ostft.STFTError) as ei:
            srv.submit(np.ones(100, np.float32), ostft.make_params(5, 0.0))
        assert ei.value.status == N.ERR_UNSUPPORTED
        with pytest.raises(ostft.STFTError) as ei:
            srv.submit(np.ones(100, np.float32), ostft.make_params(8, 0.0), out=np.empty(3, np.float32))
        assert ei.value.status == N.ERR_CAPACITY
        x = np.random.default_rng(3).uniform(-1, 1, 1 << 20).astype(np.float32)
        t = srv.submit(x, ostft.make_params(10, 0.5, N.WIN_HANNING))
        try:
            y = srv.wait(t, timeout_ms=0)
        except TimeoutError:
            y = srv.wait(t)
        assert y.size == ostft.geometry(ostft.make_params(10, 0.5), x.size)["out_floats"]
        with pytest.raises(ostft.STFTError):
            srv.wait(t)  # released


def test_server_shutdown_finishes_submitted_work(runner):
    reqs = ragged_requests(61, 20, exps=[10, 12])
    srv = ostft.Server(0, arena_bytes=32 << 20)
    outs = []
    for x, p in reqs:
        out = np.full(ostft.geometry(p, x.size)["out_floats"] + 1, -1.0, np.float32)
        srv.submit(x, p, out=out)
        outs.append(out)
    t0 = time.time()
    srv.close()  # destroy: drains the queue first
    assert time.time() - t0 < 60
    for (x, p), out in zip(reqs, outs):
        assert bitwise_equal(out[:-1], run_alone(runner, x, p))
