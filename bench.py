#!/usr/bin/env python3
"""Benchmark: the plain fused FP32 small-window sweep on one MI355X.

Workload (the reference-path configuration the headline metric is quoted on):
one float32 signal of 10^8 samples -- 8 sinusoids (frequency U(0, 0.5)
cycles/sample, amplitude U[0.1, 1], random phase) plus Gaussian noise of sigma
0.05 -- transformed with windows 2^6 .. 2^12, overlap 0 and 50%, Hanning window,
DC removal, in both the default (power) and --half_complex_return layouts:
28 requests.  One step = the whole sweep.  The signal and the output buffer are
resident in HBM before the timed region.

value = request frames per second over the sweep, frames counted as the
reference's qtConst per request.  roofline = the dominant kernel (largest share
of the step): algorithmic bytes per launch (4 B per input sample of the span the
launch reads + 4 B per output float) / its mean launch time from HIP events on
the launch stream, measured cold (Infinity Cache evicted before each launch),
against 8 TB/s; traffic = the HBM bytes per launch that rocprofv3 counters
measured for that kernel (profiles/traffic.json), when profiled.

cpu_baseline = the reference's own OpenMP kernels (oracle/_ref, compiled from
the reference sources) or, without them, the oracle's OpenMP port, timed on the
host cores over a bounded prefix of the same signal and the same 28 requests.

Other workloads (--workload; one JSON line each, same fields):
  large  -- 2^27-sample signal (N(0,1) noise + 0.5-amplitude linear chirp
            0 -> 0.5 cycles/sample), windows 2^16..2^20, overlap 0 and 75%,
            Blackman + DC removal, power spectrum (N/2+1 bins): the multi-pass
            path; 10 requests per step.
  batch  -- 4096 signals of 2^18 samples of unit-variance 1/f (pink) noise,
            windows 64, 256, 1024, overlap 50%, Hanning, complex half
            spectrum, in four forms (FP32 through the ragged tile queue, FP32
            strided, MFMA FP16 / BF16 split precision strided), each checked
            against a numpy FP64 DFT on sampled signals; value = the fastest
            form within 1e-5; 3 launches per step.
  serve  -- 2000 requests arriving as a Poisson process (--rate per second),
            lengths log-uniform in [2^14, 2^22] (U[-1,1] noise + one sinusoid),
            window uniform over 2^8..2^13, overlap over {0, 25, 50, 75}%, window
            type over Hanning/Hamming/Blackman, power spectrum: host buffers
            through the continuous-batching server; p50/p99 latency from submit
            to result in the caller's buffer.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
SIGNAL_LEN = 100_000_000
EXPS = list(range(6, 13))
OVERLAPS = (0.0, 0.5)
LAYOUTS = (("power", "--hanning_window --remove_dc"),
           ("half_complex", "--hanning_window --remove_dc --half_complex_return"))
SEED = 0x5EED0002


def sweep():
    return [(e, ov, name, opts) for e in EXPS for ov in OVERLAPS for name, opts in LAYOUTS]


def make_signal(L: int, seed: int):
    """C2 synthetic signal on the GPU (float64 synthesis, float32 result)."""
    import torch

    rng = np.random.default_rng(seed)
    x = torch.zeros(L, dtype=torch.float32, device="cuda")
    chunk = 1 << 24
    comps = [(rng.uniform(0, 0.5), rng.uniform(0.1, 1.0), rng.uniform(0, 2 * np.pi)) for _ in range(8)]
    g = torch.Generator(device="cuda").manual_seed(seed & 0xFFFFFFFF)
    for s in range(0, L, chunk):
        n = torch.arange(s, min(L, s + chunk), dtype=torch.float64, device="cuda")
        acc = torch.zeros_like(n)
        for f, a, ph in comps:
            acc += a * torch.cos(2 * np.pi * f * n + ph)
        acc += 0.05 * torch.randn(n.shape, dtype=torch.float64, device="cuda", generator=g)
        x[s: s + n.numel()] = acc.float()
    return x


def algorithmic_bytes(g: dict, N: int) -> int:
    """4 B per input sample of the span the launch reads + 4 B per output float."""
    span = min(g["full_size"], (g["compute_frames"] - 1) * g["hop"] + N)
    return 4 * span + 4 * g["out_floats"]


def measured_traffic(kernel_key: str):
    """HBM bytes per launch of `kernel_key` from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE, scaled by a plain copy's readings), or
    None when that kernel/config has not been profiled."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as fh:
            table = json.load(fh)
    except (OSError, ValueError):
        return None
    ent = table.get(kernel_key)
    return int(ent["hbm_bytes_per_launch"]) if ent else None


def cpu_baseline(x_dev, gpu_runner, budget_s: float):
    """Time the reference CPU path on a prefix of the signal, all 28 requests."""
    import torch

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    try:
        ctypes.CDLL("libgomp.so.1").omp_set_num_threads(threads)
    except OSError:
        pass
    if po.have_ref() and os.path.exists(po.REF_OPENMP):
        kind = "reference"
        run = lambda xs, e, ov, o: po.ref_activate(xs, e, ov, o, kind="openmp")  # noqa: E731
    else:
        kind = "port"
        run = po.pipeline
    reqs = sweep()

    def one_pass(n):
        xs = x_dev[:n].cpu().numpy()
        frames, t0, errs = 0, time.perf_counter(), []
        outs = []
        for e, ov, _, opts in reqs:
            rc, y = run(xs, e, ov, opts)
            frames += po.geometry(n, e, ov)["frames"]
            outs.append(y)
        dt = time.perf_counter() - t0
        # checker: the GPU result of the same requests against the reference
        for (e, ov, _, opts), y in zip(reqs, outs):
            gy = gpu_runner.stft(xs, e, ov, opts)
            N = 1 << e
            a, b = gy.reshape(-1, N).astype(np.float64), y.reshape(-1, N).astype(np.float64)
            den = np.linalg.norm(b, axis=1)
            ok = den > 1e-20
            errs.append(float((np.linalg.norm(a - b, axis=1)[ok] / den[ok]).max()))
        return frames, dt, max(errs)

    n = 1 << 18
    frames, dt, _ = one_pass(n)
    n = int(min(len(x_dev), max(n, n * budget_s / max(dt, 1e-3))))
    n = max(1 << 18, (n >> 12) << 12)
    frames, dt, err = one_pass(n)
    torch.cuda.synchronize()
    return {
        "value": frames / dt,
        "unit": "frames/s",
        "cores": threads,
        "kind": kind,
        "sample": f"first {n} samples of the 1e8-sample signal, same 28 requests "
                  f"({frames} frames in {dt:.2f} s, ActivateSTFT-equivalent incl. buffer allocation); "
                  f"GPU vs this CPU result: max per-frame rel err {err:.2e}",
    }


def copy_probe(args) -> int:
    """torch's device copy of 1 GiB: 2 GiB of HBM traffic per launch."""
    import torch

    n = 1 << 28
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(args.warmup):
        dst.copy_(src)
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    s.record()
    for _ in range(args.steps):
        dst.copy_(src)
    e.record()
    torch.cuda.synchronize()
    ms = s.elapsed_time(e) / args.steps
    print(json.dumps({"copy_bytes_per_launch": 8 * n, "ms": round(ms, 4),
                      "gbps": round(8 * n / (ms / 1e3) / 1e9, 1)}))
    return 0


def header(args, value, unit, ms_per_step, config, dtype="f32", higher_is_better=True):
    return {
        "metric": "frames/s & HBM GB/s (% of peak), plain vs fused/MFMA at fixed error; p99 ms",
        "value": round(value, 1),
        "unit": unit,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": higher_is_better,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": config,
    }


def roofline(achieved_gbps, kernel, bytes_per_launch, launch_ms, traffic=None, **extra):
    r = {"bound": "hbm", "achieved": round(achieved_gbps, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
         "frac": round(achieved_gbps / HBM_PEAK_GBPS, 4), "traffic": traffic, "kernel": kernel,
         "bytes_per_launch": int(bytes_per_launch), "launch_ms": round(launch_ms, 4)}
    r.update(extra)
    return r


def cold_time(fn, stream, reps):
    """Mean ms of fn() between HIP events on `stream`, Infinity Cache evicted
    before each call by reading a 1 GiB buffer: a read leaves only clean lines
    behind, whereas overwriting a buffer would leave up to 256 MB of dirty
    lines whose write-back then lands inside the timed launch."""
    import torch

    flush = torch.zeros(1 << 28, dtype=torch.float32, device="cuda")
    ts = []
    for _ in range(reps):
        flush.sum()
        s0, s1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s0.record(stream)
        fn()
        s1.record(stream)
        torch.cuda.synchronize()
        ts.append(s0.elapsed_time(s1))
    del flush
    return float(np.mean(ts))


# ---------------------------------------------------------------- large windows

LARGE_LEN = 1 << 27
LARGE_EXPS = list(range(16, 21))
LARGE_OVERLAPS = (0.0, 0.75)
LARGE_OPTS = "--blackman_window --remove_dc --power_return"


def make_chirp(L: int, seed: int):
    """N(0,1) noise + 0.5 cos(phi(n)), phi = 2 pi (0.5 n^2 / (2L)): a linear
    chirp from 0 to 0.5 cycles/sample; the phase is reduced exactly in int64
    (n^2 mod 4L)."""
    import torch

    x = torch.empty(L, dtype=torch.float32, device="cuda")
    g = torch.Generator(device="cuda").manual_seed(seed)
    chunk = 1 << 24
    for s in range(0, L, chunk):
        n = torch.arange(s, min(L, s + chunk), dtype=torch.int64, device="cuda")
        ph = ((n * n) % (4 * L)).double() * (np.pi / (2 * L))
        acc = torch.randn(n.shape, dtype=torch.float64, device="cuda", generator=g) + 0.5 * torch.cos(ph)
        x[s: s + n.numel()] = acc.float()
    return x


def bench_large(args) -> dict:
    import torch

    import occastft_amd as ostft

    runner = ostft.Runner(0)
    x = make_chirp(LARGE_LEN, SEED + 3)
    reqs = []
    max_out = 0
    for e in LARGE_EXPS:
        for ov in LARGE_OVERLAPS:
            p = ostft.parse_options(LARGE_OPTS, e, ov)
            g = ostft.geometry(p, LARGE_LEN)
            reqs.append((e, ov, p, g))
            max_out = max(max_out, g["out_floats"])
    out = torch.empty(max_out, dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        for e, ov, p, g in reqs:
            runner.run_device(x, p, out=out, stream=stream)
    torch.cuda.synchronize()
    ev = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in reqs]
          for _ in range(args.steps)]
    t0 = time.perf_counter()
    for st in range(args.steps):
        for i, (e, ov, p, g) in enumerate(reqs):
            ev[st][i][0].record(stream)
            runner.run_device(x, p, out=out, stream=stream)
            ev[st][i][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per_ms = np.array([[ev[st][i][0].elapsed_time(ev[st][i][1]) for i in range(len(reqs))]
                       for st in range(args.steps)]).mean(axis=0)
    frames = sum(g["frames"] for *_, g in reqs)
    ms_per_step = wall * 1000.0 / args.steps
    rows = []
    for (e, ov, p, g), ms in zip(reqs, per_ms):
        b = algorithmic_bytes(g, 1 << e)
        rows.append(dict(window=1 << e, overlap=ov, ms=round(float(ms), 4),
                         frames_per_s=round(g["frames"] / (ms / 1e3)), gbps=round(b / (ms / 1e3) / 1e9, 1)))
    dom = int(np.argmax(per_ms))
    e, ov, p, g = reqs[dom]
    b = algorithmic_bytes(g, 1 << e)
    cold_ms = cold_time(lambda: runner.run_device(x, p, out=out, stream=stream), stream, max(3, args.steps))
    total_bytes = sum(algorithmic_bytes(g, 1 << e) for e, ov, p, g in reqs)
    res = header(args, frames / (ms_per_step / 1e3), "frames/s", ms_per_step, {
        "workload": "plain FP32 large-window multi-pass path: 2^27-sample signal (N(0,1) + 0.5 linear chirp "
                    "0->0.5 cycles/sample), windows 2^16..2^20, overlap 0/75%, Blackman + DC removal, power "
                    "spectrum N/2+1 bins (10 requests per step)",
        "signal_samples": LARGE_LEN, "requests_per_step": len(reqs), "frames_per_step": frames})
    res["hbm_gbps_sweep"] = round(total_bytes / (ms_per_step / 1e3) / 1e9, 1)
    res["roofline"] = roofline(b / (cold_ms / 1e3) / 1e9,
                               f"multi-pass request window {1 << e} overlap {ov} (column passes + split, "
                               f"algorithmic bytes of the request over its launches)", b, cold_ms,
                               launch_ms_in_sweep=round(float(per_ms[dom]), 4))
    res["sweep"] = rows
    res["cpu_baseline"] = None
    runner.close()
    return res


# ---------------------------------------------------------------- many signals

BATCH_SIGNALS = 4096
BATCH_LEN = 1 << 18
BATCH_EXPS = (6, 8, 10)


def make_pink(S: int, L: int, seed: int):
    """S signals of unit-variance 1/f noise: white Gaussian noise shaped by
    1/sqrt(f) in the frequency domain (DC removed), per-signal normalised."""
    import torch

    x = torch.empty(S, L, dtype=torch.float32, device="cuda")
    g = torch.Generator(device="cuda").manual_seed(seed)
    f = torch.arange(L // 2 + 1, dtype=torch.float32, device="cuda")
    shape = torch.where(f > 0, 1.0 / torch.sqrt(torch.clamp(f, min=1.0)), torch.zeros_like(f))
    for s in range(0, S, 256):
        w = torch.randn(min(256, S - s), L, dtype=torch.float32, device="cuda", generator=g)
        y = torch.fft.irfft(torch.fft.rfft(w, dim=1) * shape, n=L, dim=1)
        x[s: s + w.shape[0]] = y / y.std(dim=1, keepdim=True)
    return x


def batch_jobs(x, exp, out, opts):
    """ctypes job array: every row of x with the same parameters, outputs packed
    into `out` (16-byte aligned slices)."""
    import occastft_amd as ostft
    from occastft_amd import _native as N

    S, L = x.shape
    p = ostft.parse_options(opts, exp, 0.5)
    g = ostft.geometry(p, L)
    stride = (g["out_floats"] + 3) // 4 * 4
    assert out.numel() >= S * stride
    arr = (N.Job * S)()
    for i in range(S):
        arr[i].in_ptr = x[i].data_ptr()
        arr[i].full_size = L
        arr[i].out_ptr = out.data_ptr() + 4 * i * stride
        arr[i].out_cap = g["out_floats"]
        arr[i].params = p
    return arr, p, g, stride


BATCH_OPTS = "--hanning_window --exact_window --half_spectrum_return"
BATCH_FORMS = (("fp32_tile_queue", ""), ("fp32_strided", ""), ("mfma_fp16_strided", " --mfma_fp16"),
               ("mfma_bf16_strided", " --mfma_bf16"))
BATCH_CHECK_ROWS = (0, 1, 2047, 4095)


def fp64_half_spectrum(x: np.ndarray, N: int, hop: int, frames: int) -> np.ndarray:
    """FP64 reference for the batch workload (numpy, independent of the
    engine): frames of x (zero past the end) times the symmetric Hann window
    0.5 (1 - cos(2 pi n / (N - 1))), bins 0..N/2 of the DFT, as (frames, N+2)
    interleaved re/im."""
    xs = np.zeros((frames - 1) * hop + N, np.float64)
    xs[: min(len(x), len(xs))] = x[: len(xs)]
    idx = np.arange(frames)[:, None] * hop + np.arange(N)[None, :]
    w = 0.5 * (1.0 - np.cos(2.0 * np.pi * np.arange(N) / (N - 1)))
    X = np.fft.rfft(xs[idx] * w, axis=1)
    out = np.empty((frames, N + 2))
    out[:, 0::2], out[:, 1::2] = X.real, X.imag
    return out


def bench_batch(args) -> dict:
    import torch

    import occastft_amd as ostft
    from occastft_amd import _native as N

    runner = ostft.Runner(0)
    x = make_pink(BATCH_SIGNALS, BATCH_LEN, SEED + 4)
    need = 0
    for e in BATCH_EXPS:
        g = ostft.geometry(ostft.parse_options(BATCH_OPTS, e, 0.5), BATCH_LEN)
        need = max(need, BATCH_SIGNALS * ((g["out_floats"] + 3) // 4 * 4))
    out = torch.empty(need, dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()
    lib = N.lib()
    sp = ctypes.c_void_p(stream.cuda_stream)
    xs_host = {r: x[r].cpu().numpy() for r in BATCH_CHECK_ROWS}
    for e in BATCH_EXPS:
        runner.autotune_batch(e, 1 << 27)

    def launcher(form, e):
        p = ostft.parse_options(BATCH_OPTS + dict(BATCH_FORMS)[form], e, 0.5)
        g = ostft.geometry(p, BATCH_LEN)
        stride = (g["out_floats"] + 3) // 4 * 4
        if form == "fp32_tile_queue":
            arr, _, _, _ = batch_jobs(x, e, out, BATCH_OPTS)

            def go():
                rc = lib.ostft_run_batch_device(runner._h, arr, BATCH_SIGNALS, sp)
                if rc != N.OK:
                    raise ostft.STFTError(rc, runner.last_error)
        else:
            def go():
                rc = lib.ostft_run_strided_device(runner._h, ctypes.byref(p), ctypes.c_void_p(x.data_ptr()),
                                                  BATCH_LEN, BATCH_LEN, BATCH_SIGNALS,
                                                  ctypes.c_void_p(out.data_ptr()), stride, sp)
                if rc != N.OK:
                    raise ostft.STFTError(rc, runner.last_error)
        return go, p, g, stride

    rows = []
    for form, _ in BATCH_FORMS:
        for e in BATCH_EXPS:
            go, p, g, stride = launcher(form, e)
            for _ in range(args.warmup):
                go()
            torch.cuda.synchronize()
            s0, s1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            s0.record(stream)
            for _ in range(args.steps):
                go()
            s1.record(stream)
            torch.cuda.synchronize()
            ms = s0.elapsed_time(s1) / args.steps
            # error of this form against FP64 on the sampled signals
            Nw = 1 << e
            o = out[: BATCH_SIGNALS * stride].view(BATCH_SIGNALS, stride)
            errs = []
            for r in BATCH_CHECK_ROWS:
                y = o[r, : g["out_floats"]].cpu().numpy().astype(np.float64).reshape(-1, Nw + 2)
                ref = fp64_half_spectrum(xs_host[r], Nw, g["hop"], g["frames"])
                den = np.linalg.norm(ref, axis=1)
                ok = den > 1e-20
                errs.append(float((np.linalg.norm(y - ref, axis=1)[ok] / den[ok]).max()))
            b = BATCH_SIGNALS * algorithmic_bytes(g, Nw)
            rows.append(dict(form=form, window=Nw, ms=round(ms, 4),
                             frames_per_s=round(BATCH_SIGNALS * g["frames"] / (ms / 1e3)),
                             gbps=round(b / (ms / 1e3) / 1e9, 1), max_rel_err_vs_fp64=float(f"{max(errs):.3e}")))
    frames = sum(BATCH_SIGNALS * ostft.geometry(ostft.parse_options(BATCH_OPTS, e, 0.5), BATCH_LEN)["frames"]
                 for e in BATCH_EXPS)
    # the fastest form that meets 1e-5 against FP64 on every window
    per_form = {}
    for r in rows:
        d = per_form.setdefault(r["form"], {"ms": 0.0, "err": 0.0})
        d["ms"] += r["ms"]
        d["err"] = max(d["err"], r["max_rel_err_vs_fp64"])
    ok = {f: d for f, d in per_form.items() if d["err"] <= 1e-5}
    best = min(ok, key=lambda f: ok[f]["ms"])
    ms_per_step = per_form[best]["ms"]
    dom_row = max((r for r in rows if r["form"] == best), key=lambda r: r["ms"])
    go, p, g, stride = launcher(best, int(np.log2(dom_row["window"])))
    b = BATCH_SIGNALS * algorithmic_bytes(g, dom_row["window"])
    cold_ms = cold_time(go, stream, max(3, args.steps))
    res = header(args, frames / (ms_per_step / 1e3), "frames/s", ms_per_step, {
        "workload": "many signals: 4096 signals x 2^18 samples of unit-variance pink noise, windows 64/256/1024, "
                    "overlap 50%, Hanning (FP64 coefficients), complex half spectrum; forms: FP32 through the "
                    "ragged tile queue, FP32 strided, MFMA FP16 and BF16 split precision strided; value = the "
                    "fastest form within 1e-5 of FP64 (one launch per window, 3 per step)",
        "signals": BATCH_SIGNALS, "samples_per_signal": BATCH_LEN, "frames_per_step": frames,
        "form": best, "error_check": f"signals {list(BATCH_CHECK_ROWS)} against a numpy FP64 DFT"})
    res["forms"] = {f: {"ms_per_step": round(d["ms"], 4), "frames_per_s": round(frames / (d["ms"] / 1e3)),
                        "max_rel_err_vs_fp64": d["err"]} for f, d in per_form.items()}
    res["roofline"] = roofline(b / (cold_ms / 1e3) / 1e9,
                               f"{best} window {dom_row['window']}, 4096 signals in one launch", b, cold_ms,
                               launch_ms_in_sweep=dom_row["ms"])
    res["sweep"] = rows
    res["cpu_baseline"] = None
    runner.close()
    return res


# ---------------------------------------------------------------- serving

SERVE_REQUESTS = 2000
SERVE_EXPS = tuple(range(8, 14))
SERVE_OVERLAPS = (0.0, 0.25, 0.5, 0.75)
SERVE_WINDOWS = ("--hanning_window", "--hamming_window", "--blackman_window")


def make_trace(n: int, rate: float, seed: int):
    """n requests: (arrival s, signal, params); arrivals Poisson at `rate`/s
    (all at 0 when rate <= 0).  Signals are generated on the GPU and kept in
    host memory, as a server receives them."""
    import torch

    import occastft_amd as ostft

    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, n) if rate > 0 else np.zeros(n)
    arrivals = np.cumsum(gaps) - gaps[0]
    g = torch.Generator(device="cuda").manual_seed(seed)
    reqs = []
    for i in range(n):
        L = int(np.exp(rng.uniform(np.log(2.0 ** 14), np.log(2.0 ** 22))))
        e = int(rng.choice(SERVE_EXPS))
        ov = float(rng.choice(SERVE_OVERLAPS))
        opts = str(rng.choice(SERVE_WINDOWS)) + " --power_return"
        f, a, ph = rng.uniform(0, 0.5), rng.uniform(0.1, 1.0), rng.uniform(0, 2 * np.pi)
        t = torch.arange(L, dtype=torch.float64, device="cuda")
        sig = (torch.rand(L, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
               + a * torch.sin(2 * np.pi * f * t + ph)).float()
        reqs.append((float(arrivals[i]), sig.cpu().numpy(), ostft.parse_options(opts, e, ov)))
    return reqs


def replay(srv, reqs, outs):
    """Submit on schedule from one thread, wait in order on this one; returns
    the server's statistics for the replay."""
    import queue
    import threading

    q = queue.Queue()

    def submitter():
        t0 = time.perf_counter()
        for (t, x, p), out in zip(reqs, outs):
            dt = t0 + t - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            q.put(srv.submit(x, p, out=out))
        q.put(None)

    th = threading.Thread(target=submitter)
    th.start()
    while True:
        tk = q.get()
        if tk is None:
            break
        srv.wait(tk)
    th.join()
    return srv.stats(reset=True)


def bench_serve(args) -> dict:
    import torch

    import occastft_amd as ostft
    from occastft_amd import _native as N

    reqs = make_trace(SERVE_REQUESTS, args.rate, SEED + 5)
    outs = []
    for _, x, p in reqs:
        o = np.empty(ostft.geometry(p, x.size)["out_floats"], np.float32)
        o.fill(0.0)  # touch the pages: the server's copy-out should not pay first-touch faults
        outs.append(o)
    total_frames = sum(ostft.geometry(p, x.size)["frames"] for _, x, p in reqs)
    srv = ostft.Server(0, arena_bytes=2 << 30, autotune=True)
    # warm-up: every window size once (tables, captured launches)
    warm = [(0.0, np.ones(1 << 16, np.float32), ostft.parse_options(o + " --power_return", e, 0.5))
            for e in SERVE_EXPS for o in SERVE_WINDOWS]
    wouts = [np.empty(ostft.geometry(p, x.size)["out_floats"], np.float32) for _, x, p in warm]
    for _ in range(max(1, args.warmup)):
        replay(srv, warm, wouts)
    steps = max(1, min(args.steps, 3))
    stats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        stats.append(replay(srv, reqs, outs))
    wall = time.perf_counter() - t0
    srv.close()
    st = stats[-1]
    span = st["span_s"]
    # the same requests resident in HBM, one batch: the persistent kernels'
    # bandwidth on the ragged mix, measured cold
    runner = ostft.Runner(0)
    dins = [torch.from_numpy(x).cuda() for _, x, _ in reqs]
    sizes = [(ostft.geometry(p, x.size)["out_floats"] + 3) // 4 * 4 for _, x, p in reqs]
    dout = torch.empty(sum(sizes), dtype=torch.float32, device="cuda")
    arr = (N.Job * len(reqs))()
    off = 0
    alg = 0
    for i, ((_, x, p), xd, sz) in enumerate(zip(reqs, dins, sizes)):
        g = ostft.geometry(p, x.size)
        arr[i].in_ptr = xd.data_ptr()
        arr[i].full_size = x.size
        arr[i].out_ptr = dout.data_ptr() + 4 * off
        arr[i].out_cap = g["out_floats"]
        arr[i].params = p
        off += sz
        alg += algorithmic_bytes(g, 1 << p.exp)
    for e in SERVE_EXPS:
        runner.autotune_batch(e, 1 << 27)
    stream = torch.cuda.current_stream()
    sp = ctypes.c_void_p(stream.cuda_stream)
    lib = N.lib()

    def launch():
        rc = lib.ostft_run_batch_device(runner._h, arr, len(reqs), sp)
        if rc != N.OK:
            raise ostft.STFTError(rc, runner.last_error)

    launch()
    torch.cuda.synchronize()
    cold_ms = cold_time(launch, stream, 5)
    runner.close()
    res = header(args, st["frames"] / span if span > 0 else 0.0, "frames/s", wall * 1000.0 / steps, {
        "workload": "ragged serving: 2000 requests, Poisson arrivals, lengths log-uniform [2^14, 2^22] "
                    "(U[-1,1] + one sinusoid), windows 2^8..2^13, overlap 0/25/50/75%, Hanning/Hamming/"
                    "Blackman, power spectrum N/2+1 bins; host buffers through the continuous-batching server "
                    "(one step = one replay of the trace)",
        "requests": SERVE_REQUESTS, "arrival_rate_per_s": args.rate, "frames_per_step": total_frames,
        "arena_bytes": 2 << 30})
    res["latency_ms"] = {"p50": round(st["latency_p50_ms"], 3), "p99": round(st["latency_p99_ms"], 3),
                         "max": round(st["latency_max_ms"], 3), "mean": round(st["latency_mean_ms"], 3)}
    res["serving"] = {"completed": st["completed"], "failed": st["failed"], "batches": st["batches"],
                      "mean_batch": round(st["mean_batch"], 2), "span_s": round(span, 3),
                      "requests_per_s": round(st["completed"] / span, 1) if span > 0 else None,
                      "host_gbps": round((st["bytes_in"] + st["bytes_out"]) / span / 1e9, 2) if span else None}
    res["roofline"] = roofline(alg / (cold_ms / 1e3) / 1e9,
                               "stft_batch_kernel, the 2000 requests resident in HBM as one batch (6 window "
                               "sizes: 6 persistent launches + descriptor upload)", alg, cold_ms)
    res["cpu_baseline"] = None
    return res


def launch_only(args) -> int:
    """Counter-collection runs (no timing, no JSON): three launches of each
    `exp:overlap:layout` of --launch on the 1e8-sample sweep signal
    (layout power | hc), or with --launch-strided two strided launches per
    window 64/256/1024 of 512 x 2^18 Gaussian samples, FP32 and MFMA FP16,
    complex half spectrum."""
    import torch

    import occastft_amd as ostft
    from occastft_amd import _native as N

    r = ostft.Runner(0)
    if args.launch_strided:
        x = torch.randn(512, 1 << 18, device="cuda", generator=torch.Generator(device="cuda").manual_seed(7))
        for e in (6, 8, 10):
            for flags in (0, N.FLAG_MFMA_FP16):
                p = ostft.make_params(e, 0.5, N.WIN_HANNING, 0.0, N.OUT_HALF_SPECTRUM, flags)
                for _ in range(2):
                    r.run_strided_device(x, p)
    else:
        x = make_signal(SIGNAL_LEN, SEED)
        out = torch.empty(2 * SIGNAL_LEN + 65536, dtype=torch.float32, device="cuda")
        for spec in args.launch:
            e, ov, lay = spec.split(":")
            opts = "--hanning_window --remove_dc" + (" --half_complex_return" if lay == "hc" else "")
            p = ostft.parse_options(opts, int(e), float(ov))
            for _ in range(3):
                r.run_device(x, p, out=out)
    torch.cuda.synchronize()
    r.close()
    print("launches done", flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline time budget")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only time a plain 1 GiB device copy (achievable HBM bandwidth; counter calibration)")
    ap.add_argument("--workload", choices=("sweep", "large", "batch", "serve"), default="sweep")
    ap.add_argument("--rate", type=float, default=1000.0, help="serve: Poisson arrival rate per second")
    ap.add_argument("--launch", nargs="+", metavar="EXP:OVERLAP:LAYOUT",
                    help="counter runs: only launch these sweep configs (3 times each)")
    ap.add_argument("--launch-strided", action="store_true",
                    help="counter runs: only launch the strided FP32 / MFMA batches")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: frames are independent, so N GPUs would each run this "
              "same kernel on their own share of the frames; run with --gpus 1.", file=sys.stderr)
        return 2

    import torch

    import occastft_amd as ostft

    torch.cuda.set_device(0)
    if args.copy_probe:
        return copy_probe(args)
    if args.launch or args.launch_strided:
        return launch_only(args)
    if args.workload != "sweep":
        fn = {"large": bench_large, "batch": bench_batch, "serve": bench_serve}[args.workload]
        print(json.dumps(fn(args)))
        return 0
    runner = ostft.Runner(0)
    x = make_signal(SIGNAL_LEN, SEED)
    reqs = []
    max_out = 0
    for e, ov, name, opts in sweep():
        p = ostft.parse_options(opts, e, ov)
        g = ostft.geometry(p, SIGNAL_LEN)
        reqs.append((e, ov, name, p, g))
        max_out = max(max_out, g["out_floats"])
    out = torch.empty(max_out, dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()

    def run_all(events=None):
        for i, (e, ov, name, p, g) in enumerate(reqs):
            if events is not None:
                events[i][0].record(stream)
            runner.run_device(x, p, out=out, stream=stream)
            if events is not None:
                events[i][1].record(stream)

    for _ in range(args.warmup):
        run_all()
    torch.cuda.synchronize()

    ev = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in reqs]
          for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        run_all(ev[s])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    per_ms = np.array([[ev[s][i][0].elapsed_time(ev[s][i][1]) for i in range(len(reqs))]
                       for s in range(args.steps)]).mean(axis=0)
    frames_per_step = sum(g["frames"] for *_, g in reqs)
    ms_per_step = wall * 1000.0 / args.steps
    value = frames_per_step / (ms_per_step / 1000.0)

    rows = []
    for (e, ov, name, p, g), ms in zip(reqs, per_ms):
        b = algorithmic_bytes(g, 1 << e)
        rows.append(dict(window=1 << e, overlap=ov, layout=name, ms=round(float(ms), 4),
                         frames_per_s=round(g["frames"] / (ms / 1e3)),
                         gbps=round(b / (ms / 1e3) / 1e9, 1)))
    dom = int(np.argmax(per_ms))
    e, ov, name, p, g = reqs[dom]
    b = algorithmic_bytes(g, 1 << e)
    # Roofline of the dominant kernel measured cold: no input of the previous
    # request stays in the 256 MB Infinity Cache; only the kernel is between
    # the events.
    cold_ms = cold_time(lambda: runner.run_device(x, p, out=out, stream=stream), stream, max(3, args.steps))
    achieved = b / (cold_ms / 1e3) / 1e9
    total_bytes = sum(algorithmic_bytes(g, 1 << e) for e, ov, name, p, g in reqs)
    kernel_key = f"stft_fused_kernel<{e - 1}, 0> window {1 << e} overlap {ov} {name}"
    traffic = measured_traffic(kernel_key)

    result = {
        "metric": "frames/s & HBM GB/s (% of peak), plain vs fused/MFMA at fixed error; p99 ms",
        "value": round(value, 1),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "plain fused FP32 small-window sweep: 1e8-sample signal (8 sinusoids + N(0,0.05^2)), "
                        "windows 2^6..2^12, overlap 0/50%, Hanning + DC removal, power and half-complex "
                        "layouts (28 requests per step)",
            "signal_samples": SIGNAL_LEN,
            "requests_per_step": len(reqs),
            "frames_per_step": frames_per_step,
        },
        "hbm_gbps_sweep": round(total_bytes / (ms_per_step / 1e3) / 1e9, 1),
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": traffic,
            "kernel": kernel_key,
            "bytes_per_launch": b,
            "launch_ms": round(cold_ms, 4),
            "launch_ms_in_sweep": round(float(per_ms[dom]), 4),
        },
        "sweep": rows,
    }
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(x, runner, args.cpu_seconds)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result))
    runner.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
