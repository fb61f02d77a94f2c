"""Seeded fuzz of the runner's wire code under AddressSanitizer and
UndefinedBehaviorSanitizer (host code only, no GPU).

tests/fuzz/fuzz_wire.cpp drives the Cap'n Proto request codec
(occastft_amd/csrc/runner/request_codec.cpp: round trips, mutated and random
messages) and the WebSocket server (websocket.cpp: valid, fragmented and
malformed handshakes and frames from many loopback connections).  Any
sanitizer report aborts the driver (-fno-sanitize-recover=all), so a clean
exit with its closing "ok" is the pass condition.  The build line is

    g++ -std=c++17 -O1 -g -pthread -fsanitize=address,undefined \
        -fno-sanitize-recover=all -fno-omit-frame-pointer \
        -I occastft_amd/csrc/runner tests/fuzz/fuzz_wire.cpp \
        occastft_amd/csrc/runner/request_codec.cpp occastft_amd/csrc/runner/websocket.cpp
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUNNER = os.path.join(ROOT, "occastft_amd", "csrc", "runner")


@pytest.fixture(scope="module")
def fuzz_binary(tmp_path_factory):
    cxx = shutil.which("g++")
    if not cxx:
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("fuzz") / "fuzz_wire")
    cmd = [cxx, "-std=c++17", "-O1", "-g", "-pthread", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-I", RUNNER,
           os.path.join(ROOT, "tests", "fuzz", "fuzz_wire.cpp"), os.path.join(RUNNER, "request_codec.cpp"),
           os.path.join(RUNNER, "websocket.cpp"), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0 and "asan" in (r.stderr or "").lower():
        pytest.skip("the sanitizer runtimes are not installed: " + r.stderr[-300:])
    assert r.returncode == 0, r.stderr
    return out


@pytest.mark.parametrize("seed", [1, 7])
def test_wire_fuzz_sanitized(fuzz_binary, seed):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([fuzz_binary, str(seed), "30000", "150"], capture_output=True, text=True, timeout=300,
                       env=env)
    assert r.returncode == 0 and r.stdout.rstrip().endswith("ok"), (r.stdout[-2000:], r.stderr[-4000:])
    assert "runtime error" not in r.stderr, r.stderr[-4000:]
