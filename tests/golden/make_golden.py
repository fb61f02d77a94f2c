"""Write tests/golden/ref_serial_vectors.npz from the reference's own kernels.

Run in a container that has the reference checkout at /root/reference:

    make -C oracle ref          # builds oracle/_ref/libref_serial.so
    python tests/golden/make_golden.py

Each case is a seeded float32 signal, the reference Serial runner's request
parameters (window exponent, overlap ratio, option string) and the output of
StandAlone/cross_gpgpu/Serial/kernel/compiled.hpp driven through the Serial
runner's pipeline (oracle/ref_harness.cpp).  The cases cover the behaviours a
drop-in must reproduce: sqrt(2)|Re X| as the default output, the buffer-half
layout of --half_complex_return, the penultimate-stage result for odd window
exponents >= 13, the Gaussian sigma parse, every window kind and DC removal.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyoracle as po  # noqa: E402

CASES = [
    # name, L, exp, overlap, options
    ("hann_dc_1024", 20000, 10, 0.5, "--hanning_window --remove_dc"),
    ("blackman_hc_odd_frames", 4990, 8, 0.5, "--blackman_window --half_complex_return"),
    ("exp13_odd_hamming_dc", 3 * 8192 + 100, 13, 0.0, "--hamming_window --remove_dc"),
    ("exp13_odd_hc", 2 * 8192 + 77, 13, 0.5, "--hanning_window --half_complex_return"),
    ("gauss_sigma_parse", 3000, 7, 0.75, "--gaussian_window=0.35<<sigma"),
    ("gauss_sigma_fail", 3000, 7, 0.75, "--gaussian_window=3<<sigma --remove_dc"),
    ("gauss_sigma_zero", 1500, 6, 0.5, "--gaussian_window=10<<sigma"),
    ("nuttall_06", 1000, 6, 0.6, "--nuttall_window"),
    ("blackman_nuttall_06", 1000, 6, 0.6, "--blackman_nuttall_window --remove_dc"),
    ("blackman_harris_033", 1000, 7, 0.33, "--blackman_harris_window"),
    ("flattop_dc_4096", 3 * 4096 + 7, 12, 0.5, "--flattop_window --remove_dc"),
    ("rect_512", 2000, 9, 0.0, ""),
    ("precedence_hamming_over_flattop", 900, 6, 0.25, "--flattop_window --hamming_window"),
    ("nuttall_exp14", 2 * 16384 + 5, 14, 0.25, "--nuttall_window"),
    ("exp15_odd_bh_dc", 32768 + 10, 15, 0.0, "--blackman_harris_window --remove_dc"),
    ("exp11_hc", 5000, 11, 0.75, "--hamming_window --remove_dc --half_complex_return"),
]


def main() -> None:
    if not po.have_ref():
        raise SystemExit("oracle/_ref/libref_serial.so is missing: run `make -C oracle ref` first")
    rng = np.random.default_rng(0x601DE7)
    arrays = {}
    index = []
    for i, (name, L, exp, ov, opts) in enumerate(CASES):
        x = rng.uniform(-1.0, 1.0, L).astype(np.float32)
        rc, y = po.ref_activate(x, exp, ov, opts)
        arrays[f"x{i}"] = x
        arrays[f"y{i}"] = y
        index.append(dict(name=name, L=L, exp=exp, overlap=ov, options=opts, status=int(rc), out_len=int(y.size)))
    arrays["index"] = np.frombuffer(json.dumps(index).encode(), dtype=np.uint8)
    path = os.path.join(HERE, "ref_serial_vectors.npz")
    np.savez_compressed(path, **arrays)
    print(f"wrote {path} ({os.path.getsize(path) / 1e6:.2f} MB, {len(CASES)} cases)")


def load() -> list[dict]:
    """Read the fixture: a list of cases with inputs `x` and outputs `y`."""
    path = os.path.join(HERE, "ref_serial_vectors.npz")
    with np.load(path, allow_pickle=False) as z:
        index = json.loads(bytes(z["index"]).decode())
        for i, c in enumerate(index):
            c["x"] = z[f"x{i}"]
            c["y"] = z[f"y{i}"]
    return index


if __name__ == "__main__":
    main()
