#!/usr/bin/env python3
"""Benchmark of the MI355X STFT engine on one GPU: the headline small-window
sweep plus the large-window, many-signal and serving workloads, each with its
roofline, its HBM traffic and a CPU baseline of the reference path.

Headline (`value`): the plain fused FP32 small-window sweep.  One float32
signal of 10^8 samples -- 8 sinusoids (frequency U(0, 0.5) cycles/sample,
amplitude U[0.1, 1], random phase) plus Gaussian noise of sigma 0.05 --
transformed with windows 2^6 .. 2^12, overlap 0 and 50%, Hanning window, DC
removal, in both the default (power) and --half_complex_return layouts: 28
requests.  One step = the whole sweep; the signal and the output buffer are
resident in HBM before the timed region.  value = transformed frames per
second: the frames each request's layout needs (all qtConst frames for the
power layout, ceil(qtConst/2) for the reference's half-complex layout, which
returns the first half of the frame-major buffer); the qtConst count is
reported beside it as `qtconst_frames_per_s`.

Sub-objects of the same JSON line (each with roofline, traffic, cpu_baseline):
  large  -- 2^27-sample signal (N(0,1) noise + 0.5-amplitude linear chirp
            0 -> 0.5 cycles/sample), windows 2^16..2^20, overlap 0 and 75%,
            Blackman + DC removal, power spectrum (N/2+1 bins): the two-pass
            path; 10 requests per step.
  batch  -- 4096 signals of 2^18 samples of unit-variance 1/f (pink) noise,
            windows 64, 256, 1024, overlap 50%, Hanning, complex half
            spectrum, in four forms (FP32 through the ragged tile queue, FP32
            strided, MFMA FP16 / BF16 split precision strided), each checked
            against a numpy FP64 DFT on sampled signals; value = the fastest
            form within 1e-5.
  serve  -- 2000 requests arriving as a Poisson process (1000/s), lengths
            log-uniform in [2^14, 2^22] (U[-1,1] noise + one sinusoid), window
            uniform over 2^8..2^13, overlap over {0, 25, 50, 75}%, window type
            over Hanning/Hamming/Blackman, power spectrum: host buffers through
            the continuous-batching server; p50/p99 latency from submit to
            result in the caller's buffer; also the same trace submitted at
            once (capacity), and resident in HBM as one batch (roofline).

roofline: the dominant kernel (or two-pass request) of each workload,
algorithmic bytes per launch (4 B per input sample of the span the launch
reads + 4 B per output float) over its mean launch time from HIP events on the
launch stream, measured cold (Infinity Cache evicted before each launch),
against 8 TB/s.  traffic: HBM bytes per launch from rocprofv3 FETCH_SIZE and
WRITE_SIZE runs of the same launches in a child process, each scaled by the
counter's reading for a plain 1 GiB device copy in the same run; if rocprofv3
is unavailable, the committed counter summary (profiles/traffic.json), or null.

cpu_baseline: the reference's own OpenMP kernels (oracle/_ref, compiled from
the reference sources; else the oracle's port) on all host threads the process
may use (OMP_NUM_THREADS when set), median of 5 timed runs after one warm-up,
ActivateSTFT-equivalent (per-call allocation included), on a bounded sample of
each workload stated in `sample`; the CPU model is recorded.
"""
from __future__ import annotations

import argparse
import ctypes
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
METRIC = "frames/s & HBM GB/s (% of peak), plain vs fused/MFMA at fixed error; p99 ms"
SIGNAL_LEN = 100_000_000
EXPS = list(range(6, 13))
OVERLAPS = (0.0, 0.5)
LAYOUTS = (("power", "--hanning_window --remove_dc"),
           ("half_complex", "--hanning_window --remove_dc --half_complex_return"))
SEED = 0x5EED0002
T_START = time.perf_counter()


def log(msg: str) -> None:
    print(f"[bench {time.perf_counter() - T_START:7.1f}s] {msg}", file=sys.stderr, flush=True)


def sweep():
    return [(e, ov, name, opts) for e in EXPS for ov in OVERLAPS for name, opts in LAYOUTS]


def make_signal(L: int, seed: int):
    """Sweep signal on the GPU (float64 synthesis, float32 result)."""
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


def header(args, value, unit, ms_per_step, config, dtype="f32", higher_is_better=True):
    return {
        "metric": METRIC,
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


def free_gpu():
    import gc

    import torch

    gc.collect()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


# ---------------------------------------------------------------- CPU baseline

def cpu_env():
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    return model, threads


def cpu_reference():
    """(kind, run(x, exp, overlap, options) -> (status, out), threads, model)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    model, threads = cpu_env()
    try:
        ctypes.CDLL("libgomp.so.1").omp_set_num_threads(threads)
    except OSError:
        pass
    if po.have_ref() and os.path.exists(po.REF_OPENMP):
        return "reference", (lambda x, e, ov, o: po.ref_activate(x, e, ov, o, kind="openmp")), threads, model
    return "port", po.pipeline, threads, model


def median_runs(fn, runs=5, warmup=1):
    for _ in range(warmup):
        fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), ts


def cpu_result(kind, threads, model, frames, med, ts, unit_extra, sample):
    return {"value": frames / med, "unit": "frames/s", "cores": threads, "kind": kind, "cpu_model": model,
            "median_s": round(med, 4), "runs_s": [round(t, 4) for t in ts], **unit_extra, "sample": sample}


def sized_prefix(one_pass, n0, n_max, target_s):
    """The prefix length whose pass takes about target_s (one calibration run)."""
    t0 = time.perf_counter()
    one_pass(n0)
    dt = max(time.perf_counter() - t0, 1e-3)
    n = int(min(n_max, max(n0, n0 * target_s / dt)))
    return max(n0, (n >> 12) << 12) if n < n_max else n_max


# ---------------------------------------------------------------- counters

def counter_run(plan: dict, counter: str, out_dir: str, timeout_s: int):
    """One rocprofv3 pass of `counter` over the plan's launches (child
    process); returns [(kernel name, value KiB)] in dispatch order."""
    exe = shutil.which("rocprofv3")
    if not exe:
        return None
    d = os.path.join(out_dir, counter.lower())
    cmd = ["timeout", "-s", "KILL", str(timeout_s), exe, "--pmc", counter, "-f", "csv", "-d", d, "-o", "k", "--",
           sys.executable, os.path.join(ROOT, "bench.py"), "--counters-run", json.dumps(plan)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if r.returncode not in (0, 139) or not files:  # rocprofv3 may segfault at exit after a complete run
        log(f"counter pass {counter} failed (rc {r.returncode}): {r.stderr[-400:]}")
        return None
    import csv

    rows = []
    for f in files:
        with open(f) as fh:
            for row in csv.DictReader(fh):
                rows.append((int(row.get("Dispatch_Id", 0) or 0), row["Kernel_Name"], row["Counter_Name"],
                             float(row["Counter_Value"])))
    rows.sort()
    return [(k, v) for _, k, c, v in rows if c == counter]


ENGINE_KERNELS = ("stft_", "first_pass_kernel", "last_pass_kernel")
COUNTER_LAUNCHES = 2  # launches of each workload in the counter child


def measure_traffic(plan: dict, budget_s: int):
    """HBM bytes per launch of each workload in `plan` from FETCH_SIZE and
    WRITE_SIZE, each scaled by a 1 GiB device copy's reading in the same run.
    The child separates the workloads with a torch fill kernel; the engine's
    dispatches between two fills are one workload's.  Returns
    {w: {"bytes", "fetch", "write", "fetch_scale", "write_scale"}} or None."""
    out_dir = tempfile.mkdtemp(prefix="ostft_pmc_")
    try:
        res = {}
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            rows = counter_run(plan, counter, out_dir, budget_s)
            if rows is None:
                return None
            marks = [i for i, (k, _) in enumerate(rows) if "fill" in k.lower()]
            if len(marks) < len(plan) + 1:
                log(f"counter pass {counter}: {len(marks)} markers for {len(plan)} workloads")
                return None
            marks = marks[-(len(plan) + 1):]  # the generation code's own fills come earlier
            copy = [v for k, v in rows[marks[-1] + 1:] if v > 1e5]  # the 1 GiB copies (KiB readings)
            if not copy:
                log(f"counter pass {counter}: no copy kernel found")
                return None
            scale = (1 << 30) / (1024.0 * float(np.median(copy)))  # bytes per KiB unit
            for j, w in enumerate(plan):
                seg = rows[marks[j] + 1: marks[j + 1]]
                v = sum(val for k, val in seg if any(p in k for p in ENGINE_KERNELS))
                res.setdefault(w, {})[counter] = v * 1024.0 * scale / COUNTER_LAUNCHES
                res[w][counter + "_scale"] = round(scale, 4)
        return {w: {"bytes": int(d["FETCH_SIZE"] + d["WRITE_SIZE"]), "fetch": int(d["FETCH_SIZE"]),
                    "write": int(d["WRITE_SIZE"]), "fetch_scale": d["FETCH_SIZE_scale"],
                    "write_scale": d["WRITE_SIZE_scale"]} for w, d in res.items()}
    finally:
        shutil.rmtree(out_dir, ignore_errors=True)


def counters_child(plan: dict) -> int:
    """Under rocprofv3: set every input up first; then, for each workload in
    plan order, a marker (a torch fill kernel) and the workload's measured
    launch twice; then a last marker and two 1 GiB device copies."""
    import torch

    import occastft_amd as ostft

    r = ostft.Runner(0)
    jobs = []
    if "c2" in plan:
        c = plan["c2"]
        x2 = make_signal(SIGNAL_LEN, SEED)
        p2 = ostft.parse_options(dict(LAYOUTS)[c["layout"]], c["exp"], c["ov"])
        o2 = torch.empty(ostft.geometry(p2, SIGNAL_LEN)["out_floats"] + 4, dtype=torch.float32, device="cuda")
        jobs.append(lambda: r.run_device(x2, p2, out=o2))
    if "c3" in plan:
        c = plan["c3"]
        x3 = make_chirp(LARGE_LEN, SEED + 3)
        p3 = ostft.parse_options(LARGE_OPTS, c["exp"], c["ov"])
        o3 = torch.empty(ostft.geometry(p3, LARGE_LEN)["out_floats"] + 4, dtype=torch.float32, device="cuda")
        jobs.append(lambda: r.run_device(x3, p3, out=o3))
    if "c4" in plan:
        c = plan["c4"]
        x4 = make_pink(BATCH_SIGNALS, BATCH_LEN, SEED + 4)
        go4 = batch_launcher(r, x4, c["form"], c["exp"])[0]
        jobs.append(go4)
    if "c5" in plan:
        reqs = make_trace(SERVE_REQUESTS, 0.0, SEED + 5, on_gpu=True)
        go5 = resident_batch(r, reqs)[0]
        jobs.append(go5)
    src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    mark = torch.empty(1024, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    for go in jobs:
        mark.fill_(1.0)
        torch.cuda.synchronize()
        for _ in range(COUNTER_LAUNCHES):
            go()
        torch.cuda.synchronize()
    mark.fill_(2.0)
    torch.cuda.synchronize()
    for _ in range(2):
        dst.copy_(src)
    torch.cuda.synchronize()
    r.close()
    print("counter launches done", flush=True)
    return 0


def committed_traffic(kernel_key: str):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as fh:
            ent = json.load(fh).get(kernel_key)
    except (OSError, ValueError):
        return None
    return int(ent["hbm_bytes_per_launch"]) if ent else None


# ---------------------------------------------------------------- C2: sweep

def bench_sweep(args, runner):
    import torch

    import occastft_amd as ostft

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
    # the timed steps carry no per-request events (each would add a command
    # between the launches); the per-request times come from as many steps
    # run again with events
    t0 = time.perf_counter()
    for s in range(args.steps):
        run_all()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in reqs]
          for _ in range(args.steps)]
    for s in range(args.steps):
        run_all(ev[s])
    torch.cuda.synchronize()
    per_ms = np.array([[ev[s][i][0].elapsed_time(ev[s][i][1]) for i in range(len(reqs))]
                       for s in range(args.steps)]).mean(axis=0)
    frames = sum(g["compute_frames"] for *_, g in reqs)
    qt_frames = sum(g["frames"] for *_, g in reqs)
    ms_per_step = wall * 1000.0 / args.steps
    rows = []
    for (e, ov, name, p, g), ms in zip(reqs, per_ms):
        b = algorithmic_bytes(g, 1 << e)
        rows.append(dict(window=1 << e, overlap=ov, layout=name, ms=round(float(ms), 4),
                         frames_per_s=round(g["compute_frames"] / (ms / 1e3)), gbps=round(b / (ms / 1e3) / 1e9, 1)))
    dom = int(np.argmax(per_ms))
    e, ov, name, p, g = reqs[dom]
    b = algorithmic_bytes(g, 1 << e)
    cold_ms = cold_time(lambda: runner.run_device(x, p, out=out, stream=stream), stream, max(3, args.steps))
    total_bytes = sum(algorithmic_bytes(g, 1 << e) for e, ov, name, p, g in reqs)
    kernel_key = f"stft_fused_kernel<{e - 1}, 0> window {1 << e} overlap {ov} {name}"
    res = header(args, frames / (ms_per_step / 1e3), "frames/s", ms_per_step, {
        "workload": "plain fused FP32 small-window sweep: 1e8-sample signal (8 sinusoids + N(0,0.05^2)), "
                    "windows 2^6..2^12, overlap 0/50%, Hanning + DC removal, power and half-complex "
                    "layouts (28 requests per step)",
        "signal_samples": SIGNAL_LEN, "requests_per_step": len(reqs), "frames_per_step": frames,
        "qtconst_frames_per_step": qt_frames})
    res["qtconst_frames_per_s"] = round(qt_frames / (ms_per_step / 1e3), 1)
    res["hbm_gbps_sweep"] = round(total_bytes / (ms_per_step / 1e3) / 1e9, 1)
    res["roofline"] = roofline(b / (cold_ms / 1e3) / 1e9, kernel_key, b, cold_ms,
                               launch_ms_in_sweep=round(float(per_ms[dom]), 4))
    res["sweep"] = rows
    plan = {"exp": e, "ov": ov, "layout": name}
    cpu_in = x[: 1 << 26].cpu().numpy()  # the CPU sample is a prefix of this signal
    del x, out
    free_gpu()
    return res, plan, kernel_key, cpu_in


def cpu_sweep(runner, xs_full, target_s):
    kind, run, threads, model = cpu_reference()
    import occastft_amd as ostft

    reqs = sweep()

    def one_pass(n):
        xs = xs_full[:n]
        return [run(xs, e, ov, o) for e, ov, _, o in reqs]

    n = sized_prefix(one_pass, 1 << 18, len(xs_full), target_s)
    xs = xs_full[:n]
    med, ts = median_runs(lambda: one_pass(n))
    outs = one_pass(n)
    frames = sum(ostft.geometry(ostft.parse_options(o, e, ov), n)["compute_frames"] for e, ov, _, o in reqs)
    errs = []
    for (e, ov, _, opts), (rc, y) in zip(reqs, outs):
        gy = runner.stft(xs, e, ov, opts)
        Nw = 1 << e
        a, b = gy.reshape(-1, Nw).astype(np.float64), y.reshape(-1, Nw).astype(np.float64)
        den = np.linalg.norm(b, axis=1)
        ok = den > 1e-20
        errs.append(float((np.linalg.norm(a - b, axis=1)[ok] / den[ok]).max()))
    return cpu_result(kind, threads, model, frames, med, ts, {}, (
        f"first {n} samples of the 1e8-sample signal, the same 28 requests ({frames} transformed frames per "
        f"run); GPU result of the same sample vs this CPU result: max per-frame rel err {max(errs):.2e}"))


# ---------------------------------------------------------------- C3: large windows

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


def bench_large(args, runner) -> dict:
    import torch

    import occastft_amd as ostft

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
    frames = sum(g["compute_frames"] for *_, g in reqs)
    ms_per_step = wall * 1000.0 / args.steps
    rows = []
    for (e, ov, p, g), ms in zip(reqs, per_ms):
        b = algorithmic_bytes(g, 1 << e)
        rows.append(dict(window=1 << e, overlap=ov, ms=round(float(ms), 4),
                         frames_per_s=round(g["compute_frames"] / (ms / 1e3)), gbps=round(b / (ms / 1e3) / 1e9, 1)))
    dom = int(np.argmax(per_ms))
    e, ov, p, g = reqs[dom]
    b = algorithmic_bytes(g, 1 << e)
    cold_ms = cold_time(lambda: runner.run_device(x, p, out=out, stream=stream), stream, max(3, args.steps))
    total_bytes = sum(algorithmic_bytes(g, 1 << e) for e, ov, p, g in reqs)
    res = header(args, frames / (ms_per_step / 1e3), "frames/s", ms_per_step, {
        "workload": "plain FP32 large-window two-pass path: 2^27-sample signal (N(0,1) + 0.5 linear chirp "
                    "0->0.5 cycles/sample), windows 2^16..2^20, overlap 0/75%, Blackman + DC removal, power "
                    "spectrum N/2+1 bins (10 requests per step)",
        "signal_samples": LARGE_LEN, "requests_per_step": len(reqs), "frames_per_step": frames})
    res["hbm_gbps_sweep"] = round(total_bytes / (ms_per_step / 1e3) / 1e9, 1)
    res["roofline"] = roofline(b / (cold_ms / 1e3) / 1e9,
                               f"two-pass request window {1 << e} overlap {ov} (first_pass_kernel + "
                               f"last_pass_kernel over its chunks; algorithmic bytes of the request)", b, cold_ms,
                               launch_ms_in_sweep=round(float(per_ms[dom]), 4))
    res["sweep"] = rows
    plan = {"exp": e, "ov": ov}
    cpu_in = x[: 1 << 26].cpu().numpy()
    del x, out
    free_gpu()
    return res, plan, cpu_in


def cpu_large(xs_full, target_s):
    """The reference path is defined up to window 2^16 (its int32 twiddle
    index overflows beyond): the two 2^16 rows of the workload."""
    kind, run, threads, model = cpu_reference()
    import occastft_amd as ostft

    opts = "--blackman_window --remove_dc"  # the reference's own (Serial-compatible) power layout
    rows = [(16, 0.0), (16, 0.75)]

    def one_pass(n):
        return [run(xs_full[:n], e, ov, opts) for e, ov in rows]

    n = sized_prefix(one_pass, 1 << 18, len(xs_full), target_s)
    med, ts = median_runs(lambda: one_pass(n))
    frames = sum(ostft.geometry(ostft.parse_options(opts, e, ov), n)["compute_frames"] for e, ov in rows)
    return cpu_result(kind, threads, model, frames, med, ts, {}, (
        f"first {n} samples of the 2^27-sample chirp signal, window 2^16 at overlap 0 and 75% (the windows "
        f"where the reference is defined), Blackman + DC removal, the reference's power layout; {frames} "
        f"frames per run"))


# ---------------------------------------------------------------- C4: many signals

BATCH_SIGNALS = 4096
BATCH_LEN = 1 << 18
BATCH_EXPS = (6, 8, 10)
BATCH_OPTS = "--hanning_window --exact_window --half_spectrum_return"
BATCH_FORMS = (("fp32_tile_queue", ""), ("fp32_strided", ""), ("mfma_fp16_strided", " --mfma_fp16"),
               ("mfma_bf16_strided", " --mfma_bf16"))
BATCH_CHECK_ROWS = (0, 1, 2047, 4095)


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


def batch_launcher(runner, x, form, e, out=None):
    """(go, params, geometry, out stride, out tensor) for one form and window
    over all signals of x."""
    import torch

    import occastft_amd as ostft
    from occastft_amd import _native as N

    S, L = x.shape
    p = ostft.parse_options(BATCH_OPTS + dict(BATCH_FORMS)[form], e, 0.5)
    g = ostft.geometry(p, L)
    stride = (g["out_floats"] + 3) // 4 * 4
    if out is None:
        out = torch.empty(S * stride, dtype=torch.float32, device="cuda")
    lib = N.lib()
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if form == "fp32_tile_queue":
        arr = (N.Job * S)()
        for i in range(S):
            arr[i].in_ptr = x[i].data_ptr()
            arr[i].full_size = L
            arr[i].out_ptr = out.data_ptr() + 4 * i * stride
            arr[i].out_cap = g["out_floats"]
            arr[i].params = p

        def go():
            rc = lib.ostft_run_batch_device(runner._h, arr, S, sp)
            if rc != N.OK:
                raise ostft.STFTError(rc, runner.last_error)
    else:
        def go():
            rc = lib.ostft_run_strided_device(runner._h, ctypes.byref(p), ctypes.c_void_p(x.data_ptr()), L, L, S,
                                              ctypes.c_void_p(out.data_ptr()), stride, sp)
            if rc != N.OK:
                raise ostft.STFTError(rc, runner.last_error)
    return go, p, g, stride, out


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


def bench_batch(args, runner):
    import torch

    import occastft_amd as ostft

    x = make_pink(BATCH_SIGNALS, BATCH_LEN, SEED + 4)
    need = 0
    for e in BATCH_EXPS:
        g = ostft.geometry(ostft.parse_options(BATCH_OPTS, e, 0.5), BATCH_LEN)
        need = max(need, BATCH_SIGNALS * ((g["out_floats"] + 3) // 4 * 4))
    out = torch.empty(need, dtype=torch.float32, device="cuda")
    stream = torch.cuda.current_stream()
    xs_host = {r: x[r].cpu().numpy() for r in BATCH_CHECK_ROWS}
    # tuned for the requests it runs (the variants rank differently per layout)
    tuned = {str(1 << e): runner.autotune_batch(e, 1 << 27, like=ostft.parse_options(BATCH_OPTS, e, 0.5))
             for e in BATCH_EXPS}
    rows = []
    for form, _ in BATCH_FORMS:
        for e in BATCH_EXPS:
            go, p, g, stride, _ = batch_launcher(runner, x, form, e, out)
            # at least two: the batch path alternates two descriptor blocks,
            # each with its own captured launches, made on first use
            for _ in range(max(2, args.warmup)):
                go()
            torch.cuda.synchronize()
            s0, s1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            s0.record(stream)
            for _ in range(args.steps):
                go()
            s1.record(stream)
            torch.cuda.synchronize()
            ms = s0.elapsed_time(s1) / args.steps
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
                             frames_per_s=round(BATCH_SIGNALS * g["compute_frames"] / (ms / 1e3)),
                             gbps=round(b / (ms / 1e3) / 1e9, 1), max_rel_err_vs_fp64=float(f"{max(errs):.3e}")))
    frames = sum(BATCH_SIGNALS * ostft.geometry(ostft.parse_options(BATCH_OPTS, e, 0.5), BATCH_LEN)["compute_frames"]
                 for e in BATCH_EXPS)
    per_form = {}
    for r in rows:
        d = per_form.setdefault(r["form"], {"ms": 0.0, "err": 0.0})
        d["ms"] += r["ms"]
        d["err"] = max(d["err"], r["max_rel_err_vs_fp64"])
    ok = {f: d for f, d in per_form.items() if d["err"] <= 1e-5}
    best = min(ok, key=lambda f: ok[f]["ms"])
    ms_per_step = per_form[best]["ms"]
    dom_row = max((r for r in rows if r["form"] == best), key=lambda r: r["ms"])
    dom_exp = int(np.log2(dom_row["window"]))
    go, p, g, stride, _ = batch_launcher(runner, x, best, dom_exp, out)
    b = BATCH_SIGNALS * algorithmic_bytes(g, dom_row["window"])
    cold_ms = cold_time(go, stream, max(3, args.steps))
    res = header(args, frames / (ms_per_step / 1e3), "frames/s", ms_per_step, {
        "workload": "many signals: 4096 signals x 2^18 samples of unit-variance pink noise, windows 64/256/1024, "
                    "overlap 50%, Hanning (FP64 coefficients), complex half spectrum; forms: FP32 through the "
                    "ragged tile queue, FP32 strided, MFMA FP16 and BF16 split precision strided; value = the "
                    "fastest form within 1e-5 of FP64 (one launch per window, 3 per step)",
        "signals": BATCH_SIGNALS, "samples_per_signal": BATCH_LEN, "frames_per_step": frames,
        "form": best, "error_check": f"signals {list(BATCH_CHECK_ROWS)} against a numpy FP64 DFT, every frame"})
    res["forms"] = {f: {"ms_per_step": round(d["ms"], 4), "frames_per_s": round(frames / (d["ms"] / 1e3)),
                        "max_rel_err_vs_fp64": d["err"]} for f, d in per_form.items()}
    by = {(r["form"], r["window"]): r for r in rows}
    res["mfma_fp16_vs_fp32_strided"] = {str(1 << e): round(by[("fp32_strided", 1 << e)]["ms"]
                                                           / by[("mfma_fp16_strided", 1 << e)]["ms"], 3)
                                        for e in BATCH_EXPS}
    res["roofline"] = roofline(b / (cold_ms / 1e3) / 1e9,
                               f"{best} window {dom_row['window']}, 4096 signals in one launch", b, cold_ms,
                               launch_ms_in_sweep=dom_row["ms"])
    res["sweep"] = rows
    res["autotune"] = {w: {"blocks_per_cu": c.blocks_per_cu, "shape": c.shape, "plan": c.plan} for w, c in tuned.items()}
    res["autotune_like"] = BATCH_OPTS + ", overlap 0.5, 2^27 samples in 32 requests"
    plan = {"form": best, "exp": dom_exp}
    cpu_in = x[:256].cpu().numpy()
    del x, out
    free_gpu()
    return res, plan, cpu_in


def cpu_batch(xs, target_s):
    kind, run, threads, model = cpu_reference()
    import occastft_amd as ostft

    opts = "--hanning_window"  # the reference's own power layout on the same frames

    def one_pass(k):
        return [run(xs[i], e, 0.5, opts) for i in range(k) for e in BATCH_EXPS]

    t0 = time.perf_counter()
    one_pass(1)
    dt = max(time.perf_counter() - t0, 1e-3)
    k = int(max(1, min(len(xs), target_s / dt)))
    med, ts = median_runs(lambda: one_pass(k))
    frames = k * sum(ostft.geometry(ostft.parse_options(opts, e, 0.5), BATCH_LEN)["compute_frames"]
                     for e in BATCH_EXPS)
    return cpu_result(kind, threads, model, frames, med, ts, {}, (
        f"{k} of the 4096 pink-noise signals (2^18 samples each), windows 64/256/1024 at 50%, Hanning, the "
        f"reference's power layout; {frames} frames per run (a rate: the full 4096 signals scale it by "
        f"{4096 / k:.0f}x in time, not in frames/s)"))


# ---------------------------------------------------------------- C5: serving

SERVE_REQUESTS = 2000
SERVE_EXPS = tuple(range(8, 14))
SERVE_OVERLAPS = (0.0, 0.25, 0.5, 0.75)
SERVE_WINDOWS = ("--hanning_window", "--hamming_window", "--blackman_window")
SERVE_LIKE = "--hanning_window --power_return"  # the request the serving path is autotuned for
SERVE_TUNE_SAMPLES = 32 << 18  # 32 requests of the trace's median length


def make_trace(n: int, rate: float, seed: int, on_gpu: bool = False):
    """n requests: (arrival s, signal, params, options); arrivals Poisson at
    `rate`/s (all at 0 when rate <= 0).  Signals are generated on the GPU and
    kept in host memory, as a server receives them (or left on the GPU)."""
    import torch

    import occastft_amd as ostft

    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, n) if rate > 0 else np.zeros(n)
    arrivals = np.cumsum(gaps) - gaps[0]
    meta = []
    for i in range(n):
        L = int(np.exp(rng.uniform(np.log(2.0 ** 14), np.log(2.0 ** 22))))
        e = int(rng.choice(SERVE_EXPS))
        ov = float(rng.choice(SERVE_OVERLAPS))
        wopt = str(rng.choice(SERVE_WINDOWS))
        meta.append((L, e, ov, wopt, rng.uniform(0, 0.5), rng.uniform(0.1, 1.0), rng.uniform(0, 2 * np.pi)))
    # every signal in one buffer, generated a chunk at a time (a few kernels
    # per chunk, not per request)
    lens = np.array([m[0] for m in meta], np.int64)
    padded = (lens + 15) // 16 * 16  # every signal starts 64-byte aligned
    starts = np.concatenate([[0], np.cumsum(padded)[:-1]])
    total = int(padded.sum())
    buf = torch.empty(total, dtype=torch.float32, device="cuda")
    g = torch.Generator(device="cuda").manual_seed(seed)
    st_d = torch.from_numpy(starts).cuda()
    fr = torch.tensor([m[4] for m in meta], dtype=torch.float64, device="cuda")
    am = torch.tensor([m[5] for m in meta], dtype=torch.float64, device="cuda")
    ph = torch.tensor([m[6] for m in meta], dtype=torch.float64, device="cuda")
    chunk = 1 << 25
    for s in range(0, total, chunk):
        idx = torch.arange(s, min(total, s + chunk), dtype=torch.int64, device="cuda")
        req = torch.searchsorted(st_d, idx, right=True) - 1
        t = (idx - st_d[req]).double()
        noise = torch.rand(idx.numel(), dtype=torch.float32, device="cuda", generator=g) * 2 - 1
        buf[s: s + idx.numel()] = noise + (am[req] * torch.sin(2 * np.pi * fr[req] * t + ph[req])).float()
    del idx, req, t, noise
    src = buf if on_gpu else buf.cpu().numpy()
    reqs = []
    for i, (L, e, ov, wopt, *_) in enumerate(meta):
        reqs.append((float(arrivals[i]), src[starts[i]: starts[i] + L],
                     ostft.parse_options(wopt + " --power_return", e, ov), wopt))
    return reqs


def replay(srv, reqs, outs):
    """Submit on schedule from one thread, wait in order on this one; returns
    the server's statistics for the replay."""
    import queue
    import threading

    q = queue.Queue()

    def submitter():
        t0 = time.perf_counter()
        for (t, x, p, _), out in zip(reqs, outs):
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


def resident_batch(runner, reqs):
    """The trace's requests resident in HBM as one batch: (go, algorithmic bytes)."""
    import torch

    import occastft_amd as ostft
    from occastft_amd import _native as N

    dins = [x if torch.is_tensor(x) else torch.from_numpy(x).cuda() for _, x, _, _ in reqs]
    geos = [ostft.geometry(p, xd.numel()) for (_, _, p, _), xd in zip(reqs, dins)]
    sizes = [(g["out_floats"] + 3) // 4 * 4 for g in geos]
    dout = torch.empty(sum(sizes), dtype=torch.float32, device="cuda")
    arr = (N.Job * len(reqs))()
    off = 0
    alg = 0
    for i, ((_, _, p, _), xd, sz, g) in enumerate(zip(reqs, dins, sizes, geos)):
        arr[i].in_ptr = xd.data_ptr()
        arr[i].full_size = xd.numel()
        arr[i].out_ptr = dout.data_ptr() + 4 * off
        arr[i].out_cap = g["out_floats"]
        arr[i].params = p
        off += sz
        alg += algorithmic_bytes(g, 1 << p.exp)
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    lib = N.lib()

    def go():
        rc = lib.ostft_run_batch_device(runner._h, arr, len(reqs), sp)
        if rc != N.OK:
            raise ostft.STFTError(rc, runner.last_error)

    go.keep = (dins, dout, arr)
    return go, alg


def bench_serve(args, runner):
    import torch

    import occastft_amd as ostft

    reqs = make_trace(SERVE_REQUESTS, args.rate, SEED + 5)
    outs = []
    for _, x, p, _ in reqs:
        o = np.empty(ostft.geometry(p, x.size)["out_floats"], np.float32)
        o.fill(0.0)  # touch the pages: the server's copy-out should not pay first-touch faults
        outs.append(o)
    total_frames = sum(ostft.geometry(p, x.size)["compute_frames"] for _, x, p, _ in reqs)
    # autotuned for the trace's kind of request: power spectrum (N/2+1 bins),
    # overlap 0.5 the middle of its overlaps, 32 requests of the trace's
    # median length (2^18 samples) per tuning batch
    like = ostft.parse_options(SERVE_LIKE, SERVE_EXPS[0], 0.5)
    srv = ostft.Server(0, arena_bytes=2 << 30, autotune=True, autotune_like=like,
                       autotune_samples=SERVE_TUNE_SAMPLES)
    warm = [(0.0, np.ones(1 << 16, np.float32), ostft.parse_options(o + " --power_return", e, 0.5), o)
            for e in SERVE_EXPS for o in SERVE_WINDOWS]
    wouts = [np.empty(ostft.geometry(p, x.size)["out_floats"], np.float32) for _, x, p, _ in warm]
    for _ in range(max(1, min(args.warmup, 2))):
        replay(srv, warm, wouts)
    steps = max(1, min(args.steps, 3))
    stats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        stats.append(replay(srv, reqs, outs))
    wall = time.perf_counter() - t0
    # capacity: the same requests submitted at once
    burst = [(0.0, x, p, o) for _, x, p, o in reqs]
    cap = replay(srv, burst, outs)
    srv.close()
    st = stats[-1]
    span = st["span_s"]
    go, alg = resident_batch(runner, reqs)
    choices = {}
    for e in SERVE_EXPS:
        c = runner.autotune_batch(e, SERVE_TUNE_SAMPLES, like=like)
        choices[str(1 << e)] = {"blocks_per_cu": c.blocks_per_cu, "shape": c.shape, "plan": c.plan}
    # both descriptor blocks' captured launches (re-made after the autotuner
    # changed a variant) before the timed launches
    go()
    go()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    cold_ms = cold_time(go, stream, 5)
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
    res["capacity_all_at_once"] = {
        "requests_per_s": round(cap["completed"] / cap["span_s"], 1) if cap["span_s"] > 0 else None,
        "frames_per_s": round(cap["frames"] / cap["span_s"], 1) if cap["span_s"] > 0 else None,
        "p50_ms": round(cap["latency_p50_ms"], 3), "p99_ms": round(cap["latency_p99_ms"], 3),
        "mean_batch": round(cap["mean_batch"], 2), "failed": cap["failed"]}
    # blocks_per_cu 0: one workgroup per tile; shape: frames per workgroup
    # (0: 256 threads, 1: 512, 2: 128); plan 1: the remainder radix first
    res["autotune"] = choices
    res["autotune_like"] = f"{SERVE_LIKE}, overlap 0.5, {SERVE_TUNE_SAMPLES} samples in 32 requests"
    res["roofline"] = roofline(alg / (cold_ms / 1e3) / 1e9,
                               "the 2000 requests resident in HBM as one batch (tile-queue launches per window "
                               "size + descriptor upload)", alg, cold_ms)
    cpu_reqs = [(x, p, wopt) for _, x, p, wopt in reqs[:400]]
    del go
    free_gpu()
    return res, {"rate": 0.0}, cpu_reqs


def cpu_serve(cpu_reqs, target_s):
    kind, run, threads, model = cpu_reference()
    import occastft_amd as ostft

    def one_pass(k):
        return [run(x, p.exp, p.overlap, wopt) for x, p, wopt in cpu_reqs[:k]]

    t0 = time.perf_counter()
    one_pass(20)
    dt = max(time.perf_counter() - t0, 1e-3) / 20
    k = int(max(10, min(len(cpu_reqs), target_s / dt)))
    med, ts = median_runs(lambda: one_pass(k))
    frames = sum(ostft.geometry(p, x.size)["compute_frames"] for x, p, _ in cpu_reqs[:k])
    return cpu_result(kind, threads, model, frames, med, ts, {"requests_per_s": round(k / med, 1)}, (
        f"the first {k} requests of the 2000-request trace run back to back (ActivateSTFT each, the "
        f"reference's power layout, same windows and overlaps); {frames} frames per run"))


# ---------------------------------------------------------------- single-purpose modes

def copy_probe(args) -> int:
    """Achievable HBM rates of plain torch kernels (ceilings for this device):
    a 1 GiB device copy (2 GiB of traffic; the counter calibration), a
    write-only fill of 800 MB, and the 1:2 read:write mix of the sweep's
    dominant request (400 MB read, each value written twice: 1.2 GB)."""
    import torch

    n = 1 << 28
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)

    def rate(fn, nbytes):
        for _ in range(args.warmup):
            fn()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        s.record()
        for _ in range(args.steps):
            fn()
        e.record()
        torch.cuda.synchronize()
        ms = s.elapsed_time(e) / args.steps
        return {"bytes": nbytes, "ms": round(ms, 4), "gbps": round(nbytes / (ms / 1e3) / 1e9, 1)}

    res = {"copy_bytes_per_launch": 8 * n}
    c = rate(lambda: dst.copy_(src), 8 * n)
    res.update({"ms": c["ms"], "gbps": c["gbps"]})
    m = 100_000_000
    x = src[:m]
    y = dst[: 2 * m]
    res["write_only_fill"] = rate(lambda: y.fill_(1.0), 8 * m)
    res["read1_write2"] = rate(lambda: y.view(-1, 2).copy_(x.view(-1, 1).expand(-1, 2)), 12 * m)
    print(json.dumps(res))
    return 0


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
        reqs = []
        for spec in args.launch:
            e, ov, lay = spec.split(":")
            opts = "--hanning_window --remove_dc" + (" --half_complex_return" if lay == "hc" else "")
            reqs.append(ostft.parse_options(opts, int(e), float(ov)))
        need = max(ostft.geometry(p, SIGNAL_LEN)["out_floats"] for p in reqs)
        out = torch.empty(need + 4, dtype=torch.float32, device="cuda")
        for p in reqs:
            for _ in range(3):
                r.run_device(x, p, out=out)
    torch.cuda.synchronize()
    r.close()
    print("launches done", flush=True)
    return 0


def traffic_entry(meas, key, committed_key=None):
    if meas and key in meas:
        m = meas[key]
        return m["bytes"], {"fetch_bytes": m["fetch"], "write_bytes": m["write"],
                            "fetch_scale": m["fetch_scale"], "write_scale": m["write_scale"],
                            "source": "rocprofv3 FETCH_SIZE/WRITE_SIZE in this run, scaled by a 1 GiB copy"}
    if committed_key:
        t = committed_traffic(committed_key)
        if t is not None:
            return t, {"source": "profiles/traffic.json (committed counter summary)"}
    return None, {"source": "not measured"}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=2.5, help="CPU baseline: seconds per timed run")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-counters", action="store_true", help="skip the rocprofv3 traffic passes")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only time a plain 1 GiB device copy (achievable HBM bandwidth; counter calibration)")
    ap.add_argument("--workload", choices=("all", "sweep", "large", "batch", "serve"), default="all",
                    help="all (default): the sweep headline with the other workloads as sub-objects")
    ap.add_argument("--rate", type=float, default=1000.0, help="serve: Poisson arrival rate per second")
    ap.add_argument("--launch", nargs="+", metavar="EXP:OVERLAP:LAYOUT",
                    help="counter runs: only launch these sweep configs (3 times each)")
    ap.add_argument("--launch-strided", action="store_true",
                    help="counter runs: only launch the strided FP32 / MFMA batches")
    ap.add_argument("--counters-run", metavar="PLAN_JSON", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: frames are independent, so N GPUs would each run this "
              "same kernel on their own share of the frames; run with --gpus 1.", file=sys.stderr)
        return 2

    import torch

    import occastft_amd as ostft

    torch.cuda.set_device(0)
    if args.counters_run:
        return counters_child(json.loads(args.counters_run))
    if args.copy_probe:
        return copy_probe(args)
    if args.launch or args.launch_strided:
        return launch_only(args)

    runner = ostft.Runner(0)
    plan, subs = {}, {}
    want = {"all": ("sweep", "large", "batch", "serve")}.get(args.workload, (args.workload,))
    cpu_inputs = {}
    c2_key = None
    if "sweep" in want:
        log("sweep (C2)")
        res, plan["c2"], c2_key, cpu_inputs["c2"] = bench_sweep(args, runner)
        subs["c2"] = res
    if "large" in want:
        log("large windows (C3)")
        subs["c3"], plan["c3"], cpu_inputs["c3"] = bench_large(args, runner)
    if "batch" in want:
        log("many signals (C4)")
        subs["c4"], plan["c4"], cpu_inputs["c4"] = bench_batch(args, runner)
    if "serve" in want:
        log("serving (C5)")
        subs["c5"], plan["c5"], cpu_inputs["c5"] = bench_serve(args, runner)

    if not args.no_cpu:
        for w, fn in (("c2", lambda: cpu_sweep(runner, cpu_inputs["c2"], args.cpu_seconds)),
                      ("c3", lambda: cpu_large(cpu_inputs["c3"], args.cpu_seconds)),
                      ("c4", lambda: cpu_batch(cpu_inputs["c4"], args.cpu_seconds)),
                      ("c5", lambda: cpu_serve(cpu_inputs["c5"], args.cpu_seconds))):
            if w in subs:
                log(f"CPU baseline {w}")
                subs[w]["cpu_baseline"] = fn()
    for w in subs:
        subs[w].setdefault("cpu_baseline", None)
    runner.close()
    free_gpu()

    meas = None
    if not args.no_counters and time.perf_counter() - T_START < 400:
        log("rocprofv3 traffic passes")
        meas = measure_traffic(plan, 180)
    for w, res in subs.items():
        t, info = traffic_entry(meas, w, c2_key if w == "c2" else None)
        res["roofline"]["traffic"] = t
        res["roofline"]["traffic_detail"] = info
        if t:
            res["roofline"]["traffic_vs_algorithmic"] = round(t / res["roofline"]["bytes_per_launch"], 4)

    if "c2" in subs:
        result = subs.pop("c2")
        names = {"c3": "large", "c4": "batch", "c5": "serve"}
        for w, res in subs.items():
            result[names[w]] = res
    else:
        result = next(iter(subs.values()))
    log("done")
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
