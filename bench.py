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


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline time budget")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only time a plain 1 GiB device copy (achievable HBM bandwidth; counter calibration)")
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
    # Roofline of the dominant kernel measured cold: before each launch a 1 GiB
    # buffer is overwritten so that no input stays in the 256 MB Infinity Cache
    # from the previous request; only the kernel is between the events.
    flush = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
    cold = []
    for _ in range(max(3, args.steps)):
        flush.zero_()
        s0, s1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s0.record(stream)
        runner.run_device(x, p, out=out, stream=stream)
        s1.record(stream)
        torch.cuda.synchronize()
        cold.append(s0.elapsed_time(s1))
    del flush
    cold_ms = float(np.mean(cold))
    achieved = b / (cold_ms / 1e3) / 1e9
    total_bytes = sum(algorithmic_bytes(g, 1 << e) for e, ov, name, p, g in reqs)
    kernel_key = f"stft_fused_kernel<{e - 1}> window {1 << e} overlap {ov} {name}"
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
