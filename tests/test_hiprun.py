"""`hipRun`, the runner process: wire format, WebSocket behaviour, shared
memory, errors.  CPU tests exercise the codec (`hipRun --decode`) and the
server without a GPU (every request is then answered with "ERR"); GPU tests
run requests end to end against the oracle."""
import json
import os
import signal
import subprocess
import threading

import numpy as np
import pytest

import wire_client as wc
from conftest import frame_errors

import pyoracle as po

pytestmark = pytest.mark.skipif(not os.path.exists(wc.HIPRUN), reason="hipRun not built")


def gpu_available():
    import torch

    return torch.cuda.is_available()


def decode_with_runner(tmp_path, msg: bytes, reencode: bool = False):
    f = tmp_path / "req.bin"
    f.write_bytes(msg)
    args = [wc.HIPRUN, "--decode", str(f)]
    if reencode:
        args.append(str(tmp_path / "out.bin"))
    r = subprocess.run(args, capture_output=True, text=True, timeout=60)
    return r.returncode, json.loads(r.stdout), (tmp_path / "out.bin")


def sample_request(n=1000, far="none"):
    x = np.random.default_rng(n).standard_normal(n).astype(np.float32)
    r = wc.Req(sharedMemory="", mappedID="42", options="--hanning_window --remove_dc", memPTR=0xDEADBEEF,
               posixFileDes=-1, windowsHandlePTR=7, windowRadix=9, overlapRatio=0.75, dataLength=n,
               overlapdataLength=wc.overlap_length(n, 0.75, 512), data=x)
    return r, wc.encode(r, far=far)


# ------------------------------------------------------------------ codec (CPU)

@pytest.mark.parametrize("far", ["none", "single", "double"])
def test_decode_fields_and_data(tmp_path, far):
    r, msg = sample_request(5000, far)
    rc, d, _ = decode_with_runner(tmp_path, msg)
    assert rc == 0 and d["ok"]
    for k in ("sharedMemory", "mappedID", "options", "memPTR", "posixFileDes", "windowsHandlePTR", "windowRadix",
              "dataLength", "overlapdataLength"):
        assert d[k] == getattr(r, k), k
    assert d["overlapRatio"] == pytest.approx(0.75)
    assert d["data_count"] == 5000
    assert d["data_sum"] == pytest.approx(float(np.sum(r.data.astype(np.float64))), rel=1e-12)


def test_reencode_round_trip(tmp_path):
    r, msg = sample_request(777, "double")
    rc, _, out = decode_with_runner(tmp_path, msg, reencode=True)
    assert rc == 0
    back = wc.decode(out.read_bytes())
    assert back.mappedID == r.mappedID and back.options == r.options and back.windowRadix == r.windowRadix
    assert back.dataLength == r.dataLength and np.array_equal(back.data, r.data)
    # single segment, as the reference's messageToFlatArray of a small message
    assert int.from_bytes(out.read_bytes()[:4], "little") == 0


@pytest.mark.parametrize("cut", [0, 4, 8, 17, 40, 100])
def test_truncated_messages_are_rejected(tmp_path, cut):
    _, msg = sample_request(300)
    bad = msg[:cut] if cut < len(msg) else msg
    bad = bad[: len(bad) // 8 * 8]
    rc, d, _ = decode_with_runner(tmp_path, bad)
    assert rc == 1 and not d["ok"]


def test_corrupt_pointers_are_rejected(tmp_path):
    _, msg = sample_request(300)
    b = bytearray(msg)
    b[8 + 8 * 8 + 4: 8 + 8 * 8 + 8] = b"\xff\xff\xff\x7f"  # data pointer: huge element count
    rc, d, _ = decode_with_runner(tmp_path, bytes(b))
    assert rc == 1 and not d["ok"]


def test_empty_text_and_missing_list(tmp_path):
    r = wc.Req(sharedMemory="", mappedID="", options="", windowRadix=6)
    rc, d, _ = decode_with_runner(tmp_path, wc.encode(r))
    assert rc == 0 and d["data_count"] == 0 and d["mappedID"] == ""


# ------------------------------------------------------------------ server (CPU or GPU)

@pytest.fixture()
def runner():
    p = wc.RunnerProcess()
    yield p
    p.stop()


def test_handshake_close_request(runner):
    c = runner.client()
    assert c.accept_ok, c.handshake
    assert "permessage-deflate" not in c.handshake.lower()  # declined
    c.send_frame(9, b"hi")  # ping
    op, fin, p = c.recv_frame()
    assert (op, p) == (10, b"hi")
    c.send_text("CLOSE_REQUEST")
    assert c.recv_message() == (1, b"CLOSE_COMPLETE")
    runner.proc.wait(timeout=30)
    assert runner.proc.returncode == 0


def test_unmasked_frame_is_a_protocol_error(runner):
    c = runner.client()
    c.send_frame(2, b"12345678", mask=False)
    op, p = c.recv_message()
    assert op == 8 and p[:2] == (1002).to_bytes(2, "big")


def test_error_reply_echoes_the_request(runner):
    """Undecodable input, exp < 6 and a bad Gaussian sigma come back as
    sharedMemory == "ERR" with the request's fields (the client matches replies
    by mappedID, StandAlone/src/STFTProxy.cpp:40-47)."""
    c = runner.client()
    c.send_binary(b"\x00" * 24)
    op, p = c.recv_message()
    assert op == 2 and wc.decode(p).sharedMemory == "ERR"
    x = np.ones(1000, np.float32)
    for rid, exp, opts in (("7", 5, ""), ("8", 10, "--gaussian_window=3<<sigma")):
        c.send_binary(wc.encode(wc.Req(mappedID=rid, windowRadix=exp, overlapRatio=0.5, dataLength=1000,
                                       options=opts, data=x)))
        op, p = c.recv_message()
        rep = wc.decode(p)
        assert rep.sharedMemory == "ERR" and rep.mappedID == rid and rep.windowRadix == exp
        assert rep.overlapRatio == 0.5 and rep.dataLength == 1000 and rep.options == ""


def test_fragmented_request(runner):
    c = runner.client()
    r, msg = sample_request(3000)
    c.send_binary(msg, fragments=7)
    op, p = c.recv_message()
    rep = wc.decode(p)
    assert op == 2 and rep.mappedID == "42"
    if gpu_available():
        assert rep.sharedMemory == ""
    else:
        assert rep.sharedMemory == "ERR"


def _still_serving(runner):
    """A fresh connection completes CLOSE_REQUEST and the process exits 0:
    the runner survived whatever came before."""
    c = runner.client()
    c.send_text("CLOSE_REQUEST")
    assert c.recv_message() == (1, b"CLOSE_COMPLETE")
    runner.proc.wait(timeout=30)
    assert runner.proc.returncode == 0
    return runner.proc.stderr.read().decode()


def test_oversized_result_is_refused_before_allocation(runner):
    """overlap 0.9999999 at window 2^20 over 2^20 samples gives ~8.4e6 frames,
    ~8.8e12 output floats: more than a reply's data list carries (< 2^29).
    The request must come back as ERR (not abort the runner on std::bad_alloc)."""
    x = np.zeros(1 << 20, np.float32)
    c = runner.client()
    c.send_binary(wc.encode(wc.Req(mappedID="big", windowRadix=20, overlapRatio=0.9999999, dataLength=len(x),
                                   data=x)))
    op, p = c.recv_message()
    rep = wc.decode(p)
    assert op == 2 and rep.sharedMemory == "ERR" and rep.mappedID == "big"
    err = _still_serving(runner)
    assert "exceeds what the reply carries" in err


def test_continuation_length_near_2_64_closes_only_that_connection(runner):
    """A continuation frame whose 64-bit length would wrap msg.size() + len is
    refused with close code 1009; the runner keeps serving."""
    c = runner.client()
    c.send_frame(2, b"\x00" * 16, fin=False)
    key = b"\x01\x02\x03\x04"
    c.s.sendall(bytes([0x80, 0x80 | 127]) + ((1 << 64) - 8).to_bytes(8, "big") + key)
    op, p = c.recv_message()
    assert op == 8 and p[:2] == (1009).to_bytes(2, "big")
    _still_serving(runner)


@pytest.mark.parametrize("data_len,olen", [(1 << 62, 1024), (16, 1 << 63), (1 << 61, (1 << 62) + 3)])
def test_shared_memory_sizes_that_overflow_are_refused(runner, data_len, olen):
    """dataLength * 4 or the page rounding of overlapdataLength * 4 would wrap
    in 64 bits: an ERR reply, not a resize that throws."""
    from multiprocessing import shared_memory

    name = f"STFTovf{os.getpid()}SHAREDMEM"
    shm = shared_memory.SharedMemory(name=name, create=True, size=4096)
    try:
        c = runner.client()
        c.send_binary(wc.encode(wc.Req(sharedMemory="/" + name, mappedID="ovf", windowRadix=10, overlapRatio=0.5,
                                       dataLength=data_len, overlapdataLength=olen)))
        op, p = c.recv_message()
        rep = wc.decode(p)
        assert op == 2 and rep.sharedMemory == "ERR" and rep.mappedID == "ovf"
        _still_serving(runner)
    finally:
        shm.close()
        shm.unlink()


def test_runner_death_is_visible_to_the_client(runner):
    """The client's fallback trigger (STFTProxy.cpp:59-72): the connection
    drops when the runner process dies."""
    c = runner.client()
    os.kill(runner.proc.pid, signal.SIGKILL)
    runner.proc.wait(timeout=30)
    with pytest.raises((ConnectionError, OSError)):
        c.send_text("x")
        c.recv_message()


# ------------------------------------------------------------------ end to end (GPU)

def _check(y, x, exp, ov, opts):
    rc, r = po.pipeline(x, exp, ov, opts)
    assert rc == 0 and y.shape == r.shape
    mx, _, zabs = frame_errors(y, r, 1 << exp)
    assert mx <= 1e-5 and zabs <= 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("far", ["none", "single", "double"])
def test_request_through_capnp_list(runner, far):
    x = np.random.default_rng(1).uniform(-1, 1, 1 << 16).astype(np.float32)
    opts = "--hanning_window --remove_dc"
    c = runner.client()
    c.send_binary(wc.encode(wc.Req(mappedID="1", windowRadix=10, overlapRatio=0.5, dataLength=len(x),
                                   options=opts, data=x), far=far))
    op, p = c.recv_message()
    rep = wc.decode(p)
    assert rep.sharedMemory == "" and rep.mappedID == "1"
    assert rep.dataLength == rep.data.size == po.geometry(len(x), 10, 0.5)["ofull"]
    _check(rep.data, x, 10, 0.5, opts)


@pytest.mark.gpu
def test_large_window_request(runner):
    """A 2^17 window (the two-pass path behind the server) through the wire
    protocol, against the float restatement with a 64-bit twiddle index."""
    exp, ov, opts = 17, 0.5, "--blackman_window --remove_dc"
    x = np.random.default_rng(3).standard_normal(3 * (1 << exp) + 77).astype(np.float32)
    c = runner.client()
    c.send_binary(wc.encode(wc.Req(mappedID="L", windowRadix=exp, overlapRatio=ov, dataLength=len(x),
                                   options=opts, data=x)))
    op, p = c.recv_message()
    rep = wc.decode(p)
    assert rep.sharedMemory == "" and rep.mappedID == "L"
    assert rep.data.size == po.geometry(len(x), exp, ov)["ofull"]
    _check(rep.data, x, exp, ov, opts)


@pytest.mark.gpu
def test_request_through_shared_memory(runner):
    x = np.random.default_rng(2).uniform(-1, 1, 300_000).astype(np.float32)
    opts = "--blackman_window --remove_dc --half_complex_return"
    s = wc.ShmRequest(f"t{os.getpid()}", x, 12, 0.75, opts)
    try:
        c = runner.client()
        c.send_binary(wc.encode(s.req))
        op, p = c.recv_message()
        rep = wc.decode(p)
        assert rep.sharedMemory == s.req.sharedMemory and rep.data.size == 0
        y = s.result()
        n = po.geometry(len(x), 12, 0.75)["ofull"]
        assert y.size >= n
        _check(y[:n], x, 12, 0.75, opts)
    finally:
        s.close()


@pytest.mark.gpu
def test_concurrent_clients(runner):
    """Two connections at once: each request keeps its own state (the
    reference shares one mapping member across connection threads)."""
    results = {}

    def one(i):
        x = np.random.default_rng(10 + i).uniform(-1, 1, 200_000).astype(np.float32)
        opts = ["--hanning_window --remove_dc", "--hamming_window"][i]
        s = wc.ShmRequest(f"c{os.getpid()}_{i}", x, 9 + i, 0.5, opts)
        try:
            c = runner.client()
            for _ in range(3):
                # the result overwrites the input in the mapping, as in the
                # reference protocol: put the input back before each request
                s.write(x)
                c.send_binary(wc.encode(s.req))
                c.recv_message()
            n = po.geometry(len(x), 9 + i, 0.5)["ofull"]
            results[i] = (s.result()[:n], x, 9 + i, opts)
        finally:
            s.close()

    ts = [threading.Thread(target=one, args=(i,)) for i in range(2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=120)
    assert set(results) == {0, 1}
    for y, x, exp, opts in results.values():
        _check(y, x, exp, 0.5, opts)
