"""Build the HIP engine (gfx950) into ``occastft_amd/libocca_stft_hip.so``.

Every source under ``csrc/`` is compiled with hipcc for gfx950 only; objects go
to ``occastft_amd/_build/`` and the shared library is linked in-tree so that it
travels with the source tree.  ``python -m occastft_amd.build`` runs it.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libocca_stft_hip.so")
SOURCES = ["capi.hip", "fused_small.hip", "multipass.hip", "tables.hip", "batch.hip", "server.hip",
           "mfma_real_dft.hip"]
ARCH = "gfx950"


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the ROCm toolchain is required to build occastft_amd")


def common_flags(defines: tuple = ()) -> list[str]:
    return [*[f"-D{d}" for d in defines],
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        # keep the compiler from forming packed-FP32 (v_pk_*_f32) instructions
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        # a one-lane atomic stays a plain atomic: the wave-scan rewrite waits
        # for its result at once, which would stall the persistent batch
        # kernel's prefetched tile claim behind the previous tile's stores
        "-mllvm", "-amdgpu-atomic-optimizer-strategy=None",
        "-Wall",
        "-Wno-unused-function",
        "-I", os.path.join(HERE, "..", "include"),
    ]


def _compile(src: str, build_dir: str = BUILD, defines: tuple = ()) -> str:
    obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *common_flags(defines), "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False, out: str = LIB, build_dir: str = BUILD, defines: tuple = ()) -> str:
    """Compile every source and link `out`; `defines` are extra -D macros."""
    os.makedirs(build_dir, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(4, len(SOURCES))) as ex:
        objs = list(ex.map(lambda s: _compile(s, build_dir, defines), SOURCES))
    tmp = out + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    return out


RUNNER_SOURCES = ["runner/hip_run.cpp", "runner/request_codec.cpp", "runner/websocket.cpp"]
RUNNER = os.path.join(HERE, "hipRun")


def build_runner(verbose: bool = False) -> str:
    """Link the `hipRun` runner executable (host C++) against the engine."""
    cxx = shutil.which("g++") or "g++"
    tmp = RUNNER + ".tmp"
    cmd = [cxx, "-O2", "-std=c++17", "-pthread", "-Wall", "-I", os.path.join(HERE, "..", "include"),
           *[os.path.join(CSRC, s) for s in RUNNER_SOURCES], "-L", HERE, "-locca_stft_hip",
           "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipRun build failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, RUNNER)
    if verbose:
        print(f"built {RUNNER}")
    return RUNNER


PLUGIN_LIB = os.path.join(HERE, "libocca_stft_runnerfn.so")
PLUGIN_DRIVER = os.path.join(HERE, "runnerfn_pipeline")


def build_plugin(verbose: bool = False) -> str:
    """Link the record-and-fuse runner functions (libocca_stft_runnerfn.so,
    host C++) and the pipeline driver that exercises them."""
    cxx = shutil.which("g++") or "g++"
    inc = ["-I", os.path.join(HERE, "..", "include")]
    link = ["-L", HERE, "-locca_stft_hip", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib"]
    steps = [
        (PLUGIN_LIB, [cxx, "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", *inc,
                      os.path.join(CSRC, "plugin", "runner_functions.cpp"), *link]),
        (PLUGIN_DRIVER, [cxx, "-O2", "-std=c++17", "-Wall", *inc, os.path.join(CSRC, "plugin", "pipeline_driver.cpp"),
                         "-L", HERE, "-locca_stft_runnerfn", *link]),
    ]
    for out, cmd in steps:
        tmp = out + ".tmp"
        r = subprocess.run([*cmd, "-o", tmp], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"build failed for {out}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        os.replace(tmp, out)
        if verbose:
            print(f"built {out}")
    return PLUGIN_LIB


if __name__ == "__main__":
    build(verbose=True)
    build_runner(verbose=True)
    build_plugin(verbose=True)
    sys.exit(0)
