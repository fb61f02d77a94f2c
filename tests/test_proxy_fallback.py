"""Runtime fallback around `hipRun`, through a client with STFTProxy's semantics
(StandAlone/src/STFTProxy.cpp:4-120, 160-211; runtimeChecker.cpp:138-222).

The CPU cases drive the proxy logic itself: an unavailable runtime is skipped,
an exhausted fallback list sets "Run Out Fallback", KillRunner gets
CLOSE_COMPLETE.  The GPU case is the fault injection: hipRun serves a request,
is killed with SIGKILL, the next request is lost and the next one is
served by the CPU fallback runner; both answers agree with the Serial
restatement to the 1e-5 per-frame bar.
"""
import os
import signal

import numpy as np
import pytest

import proxy_harness as ph
import wire_client as wc
from conftest import frame_errors

import pyoracle as po

OPTS = "--hanning_window --remove_dc"


def _signal(n, seed):
    rng = np.random.default_rng(seed)
    t = np.arange(n, dtype=np.float64)
    return (np.sin(2e-3 * t * t / n * 400) + 0.3 * rng.standard_normal(n)).astype(np.float32)


def test_unavailable_runtime_is_skipped(tmp_path):
    p = ph.Proxy([("HIP", str(tmp_path / "no_such_dir")), ("SERIAL", None)])
    try:
        assert p.status == "OK" and p.runtime == "SERIAL"
        assert 49152 <= p.port <= 65535
        x = _signal(5000, 1)
        y = p.request(x, 10, 0.5, OPTS)
        rc, r = po.pipeline(x, 10, 0.5, OPTS)
        assert rc == 0 and y is not None and np.array_equal(y, r)
        # a second request on the same connection
        y2 = p.request(x[:3000], 9, 0.25, "--hanning_window")
        assert np.array_equal(y2, po.pipeline(x[:3000], 9, 0.25, "--hanning_window")[1])
        assert p.cpu.served == 2
    finally:
        p.close()


def test_run_out_fallback(tmp_path):
    p = ph.Proxy([("HIP", str(tmp_path)), ("HIP", None)])
    try:
        assert p.status == "Run Out Fallback" and p.runtime is None
        assert p.request(_signal(2048, 2), 10, 0.5, OPTS) is None
    finally:
        p.close()


def test_kill_runner_close_handshake():
    p = ph.Proxy([("SERIAL", None)])
    assert p.kill_runner() is True
    assert p.status == "NOFALLBACK"
    assert p.request(_signal(2048, 3), 10, 0.5, OPTS) is None
    p.close()


def test_err_answer_keeps_runtime():
    """An ERR answer is not a fallback trigger: hipRun on a host without a HIP
    device answers ERR, stays the runtime, and ends on CLOSE_REQUEST."""
    if not os.path.exists(wc.HIPRUN) or not ph.runtime_available("HIP", wc.HIPRUN):
        pytest.skip("hipRun or the HIP runtime library is absent")
    import torch
    if torch.cuda.is_available():
        pytest.skip("a HIP device is present: the GPU case covers this runner")
    p = ph.Proxy([("HIP", os.path.dirname(wc.HIPRUN)), ("SERIAL", None)], tries=40)
    try:
        assert p.status == "OK" and p.runtime == "HIP"
        assert p.request(_signal(3000, 6), 10, 0.5, OPTS) is None
        assert p.status == "OK" and p.runtime == "HIP" and p.cpu is None
        proc = p.proc
        assert p.kill_runner() is True
        assert proc.poll() == 0
    finally:
        p.close()


def test_killed_runner_process_falls_back():
    """SIGKILL of the runner process (hipRun without a device here): the next
    request is lost and the one after it is served by the CPU runner."""
    if not os.path.exists(wc.HIPRUN) or not ph.runtime_available("HIP", wc.HIPRUN):
        pytest.skip("hipRun or the HIP runtime library is absent")
    p = ph.Proxy([("HIP", wc.HIPRUN), ("SERIAL", None)], tries=240)
    try:
        assert p.runtime == "HIP"
        proc = p.proc
        os.kill(proc.pid, signal.SIGKILL)
        assert proc.wait(timeout=30) == -signal.SIGKILL
        x = _signal(3000, 7)
        assert p.request(x, 10, 0.5, OPTS) is None
        assert p.status == "OK" and p.runtime == "SERIAL" and p.proc is None
        assert np.array_equal(p.request(x, 10, 0.5, OPTS), po.pipeline(x, 10, 0.5, OPTS)[1])
    finally:
        p.close()


def test_runner_death_falls_back():
    """A runner that drops the connection moves the proxy to the next runtime."""
    p = ph.Proxy([("SERIAL", None), ("SERIAL", None)])
    try:
        x = _signal(4096, 4)
        assert p.request(x, 10, 0.5, OPTS) is not None
        p.cpu.stop()
        p.client.s.shutdown(2)  # the connection ends under the next request
        assert p.request(x, 10, 0.5, OPTS) is None
        assert p.status == "OK" and p.runtime == "SERIAL"
        y = p.request(x, 10, 0.5, OPTS)
        assert np.array_equal(y, po.pipeline(x, 10, 0.5, OPTS)[1])
        assert p.fallback == []
    finally:
        p.close()


@pytest.mark.gpu
def test_hiprun_killed_falls_back_to_cpu():
    if not os.path.exists(wc.HIPRUN):
        pytest.fail("hipRun is not built")
    # the reference's connect window (5 tries, 500 ms apart): hipRun listens
    # before it brings the device up, so a cold start does not miss it
    p = ph.Proxy([("HIP", os.path.dirname(wc.HIPRUN)), ("SERIAL", None)])
    try:
        assert p.status == "OK" and p.runtime == "HIP", p.status
        x = _signal(200_000, 5)
        rc, r = po.pipeline(x, 10, 0.5, OPTS)
        assert rc == 0
        y = p.request(x, 10, 0.5, OPTS)
        assert y is not None and y.size == r.size
        assert frame_errors(y, r, 1 << 10)[0] <= 1e-5  # per 1024-sample frame

        # fault injection: SIGKILL the runner; the next request finds its
        # connection gone, is lost, and the proxy moves to the CPU runner
        os.kill(p.proc.pid, signal.SIGKILL)
        assert p.proc.wait(timeout=30) == -signal.SIGKILL
        assert p.request(x, 10, 0.5, OPTS) is None
        assert p.status == "OK" and p.runtime == "SERIAL"
        y2 = p.request(x, 10, 0.5, OPTS)
        assert y2 is not None and np.array_equal(y2, r)
        assert p.kill_runner() is True
    finally:
        p.close()
