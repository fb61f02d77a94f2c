"""Test client for `hipRun`: Cap'n Proto RequestCapnp codec, WebSocket client
and POSIX shared memory, written independently of the runner's C++ code.

Schema: StandAlone/include/FFTcapnp.capnp:1-16 (struct RequestCapnp, 6 data
words + 4 pointers; field offsets StandAlone/include/FFTcapnp.capnp.h).  The
client side of the protocol follows StandAlone/src/STFTProxy.cpp:186-211 and
StandAlone/src/FFTStruct_linux.cpp:5-66 (shm name "/STFT<id>SHAREDMEM",
page-rounded size of overlapdataLength floats).
"""
from __future__ import annotations

import base64
import hashlib
import os
import socket
import struct
import subprocess
import time
from dataclasses import dataclass, field
from multiprocessing import shared_memory

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPRUN = os.path.join(ROOT, "occastft_amd", "hipRun")


@dataclass
class Req:
    sharedMemory: str = ""
    mappedID: str = "0"
    options: str = ""
    memPTR: int = 0
    posixFileDes: int = -1
    windowsHandlePTR: int = 0
    windowRadix: int = 10
    overlapRatio: float = 0.5
    dataLength: int = 0
    overlapdataLength: int = 0
    data: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))


# ------------------------------------------------------------------ encoding

def _struct_ptr(offset: int, dwords: int, pwords: int) -> int:
    return ((offset & 0x3FFFFFFF) << 2) | (dwords << 32) | (pwords << 48)


def _list_ptr(offset: int, esize: int, count: int) -> int:
    return 1 | ((offset & 0x3FFFFFFF) << 2) | (esize << 32) | (count << 35)


def _far_ptr(pad_word: int, seg: int, double: bool = False) -> int:
    return 2 | (int(double) << 2) | (pad_word << 3) | (seg << 32)


def _text_words(s: bytes) -> int:
    return (len(s) + 1 + 7) // 8


def _data_section(r: Req) -> bytes:
    return struct.pack("<QqQIfQQ", r.memPTR, r.posixFileDes, r.windowsHandlePTR, r.windowRadix,
                       r.overlapRatio, r.dataLength, r.overlapdataLength)


def encode(r: Req, far: str = "none") -> bytes:
    """Flat-array message.  far = "none": one segment; "single": the data list
    in segment 1 behind a single-far pointer with its landing pad in front of
    it (what a MallocMessageBuilder does for a list that outgrows segment 0);
    "double": the list content in segment 2, its landing pad (far + tag) in
    segment 1."""
    texts = [r.sharedMemory.encode(), r.mappedID.encode(), r.options.encode()]
    data = np.ascontiguousarray(r.data, dtype=np.float32)
    n = data.size
    seg0 = [0] * (1 + 6 + 4)
    seg0[0] = _struct_ptr(0, 6, 4)
    ds = _data_section(r)
    for i in range(6):
        seg0[1 + i] = struct.unpack_from("<Q", ds, 8 * i)[0]
    ptr_at = 7
    blobs = {}
    for slot, t in zip((0, 2, 3), texts):
        at = len(seg0)
        seg0[ptr_at + slot] = _list_ptr(at - (ptr_at + slot) - 1, 2, len(t) + 1)
        w = _text_words(t)
        buf = t + b"\0" * (8 * w - len(t))
        seg0 += list(struct.unpack(f"<{w}Q", buf))
    data_words = (n * 4 + 7) // 8
    dbytes = data.tobytes() + b"\0" * (8 * data_words - 4 * n)
    segs = []
    if n == 0:
        segs = [seg0]
        body = [b""]
    elif far == "none":
        at = len(seg0)
        seg0[ptr_at + 1] = _list_ptr(at - (ptr_at + 1) - 1, 4, n)
        segs = [seg0]
        body = [dbytes]
    elif far == "single":
        seg0[ptr_at + 1] = _far_ptr(0, 1)
        seg1 = [_list_ptr(0, 4, n)]
        segs = [seg0, seg1]
        body = [b"", dbytes]
    elif far == "double":
        seg0[ptr_at + 1] = _far_ptr(0, 1, double=True)
        seg1 = [_far_ptr(0, 2), _list_ptr(0, 4, n)]
        segs = [seg0, seg1, []]
        body = [b"", b"", dbytes]
    else:
        raise ValueError(far)
    seg_bytes = [struct.pack(f"<{len(s)}Q", *s) + b for s, b in zip(segs, body)]
    sizes = [len(b) // 8 for b in seg_bytes]
    hdr = struct.pack("<I", len(sizes) - 1) + struct.pack(f"<{len(sizes)}I", *sizes)
    if len(hdr) % 8:
        hdr += b"\0" * (8 - len(hdr) % 8)
    return hdr + b"".join(seg_bytes)


# ------------------------------------------------------------------ decoding

def _signed30(v: int) -> int:
    v &= 0x3FFFFFFF
    return v - (1 << 30) if v & (1 << 29) else v


def decode(buf: bytes) -> Req:
    (nm1,) = struct.unpack_from("<I", buf, 0)
    n = nm1 + 1
    sizes = struct.unpack_from(f"<{n}I", buf, 4)
    off = (4 + 4 * n + 7) // 8 * 8
    segs = []
    for s in sizes:
        segs.append(buf[off: off + 8 * s])
        off += 8 * s

    def word(seg, i):
        return struct.unpack_from("<Q", segs[seg], 8 * i)[0]

    def resolve(seg, at):
        p = word(seg, at)
        if p == 0:
            return None
        kind = p & 3
        if kind == 2:
            dbl, pad, tseg = (p >> 2) & 1, (p >> 3) & 0x1FFFFFFF, p >> 32
            land = word(tseg, pad)
            if not dbl:
                return resolve(tseg, pad)
            tag = word(tseg, pad + 1)
            return tag & 3, land >> 32, (land >> 3) & 0x1FFFFFFF, tag
        return kind, seg, at + 1 + _signed30(p >> 2), p

    k, rseg, rstart, tag = resolve(0, 0)
    dwords, pwords = (tag >> 32) & 0xFFFF, tag >> 48
    ds = segs[rseg][8 * rstart: 8 * (rstart + dwords)] + b"\0" * 48
    r = Req()
    (r.memPTR, r.posixFileDes, r.windowsHandlePTR, r.windowRadix, r.overlapRatio, r.dataLength,
     r.overlapdataLength) = struct.unpack_from("<QqQIfQQ", ds, 0)

    def text(i):
        if i >= pwords:
            return ""
        t = resolve(rseg, rstart + dwords + i)
        if t is None:
            return ""
        _, s, st, tg = t
        cnt = tg >> 35
        return segs[s][8 * st: 8 * st + cnt].rstrip(b"\0").decode()

    r.sharedMemory, r.mappedID, r.options = text(0), text(2), text(3)
    t = resolve(rseg, rstart + dwords + 1) if pwords > 1 else None
    if t is not None:
        _, s, st, tg = t
        cnt = tg >> 35
        r.data = np.frombuffer(segs[s], dtype=np.float32, count=cnt, offset=8 * st).copy()
    return r


# ------------------------------------------------------------------ websocket

class WsClient:
    def __init__(self, port: int, host: str = "127.0.0.1", timeout: float = 60.0):
        self.s = socket.create_connection((host, port), timeout=timeout)
        key = base64.b64encode(os.urandom(16)).decode()
        self.s.sendall((f"GET / HTTP/1.1\r\nHost: {host}:{port}\r\nUpgrade: websocket\r\nConnection: Upgrade\r\n"
                        f"Sec-WebSocket-Key: {key}\r\nSec-WebSocket-Version: 13\r\n"
                        "Sec-WebSocket-Extensions: permessage-deflate; client_max_window_bits\r\n\r\n").encode())
        resp = b""
        while b"\r\n\r\n" not in resp:
            chunk = self.s.recv(4096)
            if not chunk:
                raise ConnectionError("closed during handshake")
            resp += chunk
        head, self.buf = resp.split(b"\r\n\r\n", 1)
        self.handshake = head.decode()
        expect = base64.b64encode(hashlib.sha1((key + "258EAFA5-E914-47DA-95CA-C5AB0DC85B11").encode()).digest())
        self.accept_ok = f"Sec-WebSocket-Accept: {expect.decode()}" in self.handshake

    def send_frame(self, opcode: int, payload: bytes, fin: bool = True, mask: bool = True):
        hdr = bytes([(0x80 if fin else 0) | opcode])
        n = len(payload)
        mbit = 0x80 if mask else 0
        if n < 126:
            hdr += bytes([mbit | n])
        elif n < 65536:
            hdr += bytes([mbit | 126]) + struct.pack(">H", n)
        else:
            hdr += bytes([mbit | 127]) + struct.pack(">Q", n)
        if mask:
            key = os.urandom(4)
            arr = np.frombuffer(payload, dtype=np.uint8)
            k = np.frombuffer((key * ((n + 3) // 4 + 1))[:n], dtype=np.uint8)
            payload = (arr ^ k).tobytes()
            hdr += key
        self.s.sendall(hdr + payload)

    def send_binary(self, payload: bytes, fragments: int = 1):
        if fragments <= 1:
            self.send_frame(2, payload)
            return
        step = max(1, len(payload) // fragments)
        parts = [payload[i: i + step] for i in range(0, len(payload), step)]
        for i, p in enumerate(parts):
            self.send_frame(2 if i == 0 else 0, p, fin=(i == len(parts) - 1))

    def send_text(self, s: str):
        self.send_frame(1, s.encode())

    def _read(self, n: int) -> bytes:
        while len(self.buf) < n:
            chunk = self.s.recv(max(65536, n - len(self.buf)))
            if not chunk:
                raise ConnectionError("connection closed")
            self.buf += chunk
        out, self.buf = self.buf[:n], self.buf[n:]
        return out

    def recv_frame(self):
        b0, b1 = self._read(2)
        n = b1 & 0x7F
        if n == 126:
            (n,) = struct.unpack(">H", self._read(2))
        elif n == 127:
            (n,) = struct.unpack(">Q", self._read(8))
        assert not (b1 & 0x80), "server frames must not be masked"
        return b0 & 0x0F, bool(b0 & 0x80), self._read(n)

    def recv_message(self):
        """(opcode, payload) of the next data message; answers nothing itself."""
        while True:
            op, fin, p = self.recv_frame()
            if op in (1, 2):
                while not fin:
                    _, fin, more = self.recv_frame()
                    p += more
                return op, p
            if op == 8:
                return 8, p

    def close(self):
        try:
            self.s.close()
        except OSError:
            pass


# ------------------------------------------------------------------ runner process

def free_port() -> int:
    with socket.socket() as s:
        s.bind(("127.0.0.1", 0))
        return s.getsockname()[1]


class RunnerProcess:
    def __init__(self, exe: str = HIPRUN):
        self.port = free_port()
        self.proc = subprocess.Popen([exe, str(self.port)], stdout=subprocess.PIPE, stderr=subprocess.PIPE)
        deadline = time.time() + 60
        while time.time() < deadline:
            if self.proc.poll() is not None:
                raise RuntimeError(f"hipRun exited: {self.proc.stderr.read().decode()}")
            try:
                with socket.create_connection(("127.0.0.1", self.port), timeout=1):
                    return
            except OSError:
                time.sleep(0.1)
        raise TimeoutError("hipRun did not start listening")

    def client(self) -> WsClient:
        return WsClient(self.port)

    def stop(self):
        if self.proc.poll() is None:
            self.proc.kill()
        self.proc.wait(timeout=30)


def overlap_length(n: int, ov: float, N: int) -> int:
    """FFTRequest::toOverlapLength (StandAlone/src/FFTStruct.cpp:146-158)."""
    if ov == 0.0:
        q = n // N + 1
    else:
        q = int(np.float32(n) / (np.float32(N) * (np.float32(1.0) - np.float32(ov)))) + 1
    return q * N


class ShmRequest:
    """A client-side request whose data travels through POSIX shared memory."""

    def __init__(self, mapped_id: str, x: np.ndarray, exp: int, ov: float, options: str):
        self.name = f"STFT{mapped_id}SHAREDMEM"
        olen = overlap_length(len(x), ov, 1 << exp)
        page = os.sysconf("SC_PAGE_SIZE")
        size = (olen * 4 + page - 1) // page * page
        self.shm = shared_memory.SharedMemory(name=self.name, create=True, size=size)
        np.ndarray(len(x), np.float32, buffer=self.shm.buf)[:] = x
        self.req = Req(sharedMemory="/" + self.name, mappedID=mapped_id, options=options, windowRadix=exp,
                       overlapRatio=ov, dataLength=len(x), overlapdataLength=olen)
        self.olen = olen

    def write(self, x: np.ndarray):
        np.ndarray(len(x), np.float32, buffer=self.shm.buf)[:] = x

    def result(self) -> np.ndarray:
        return np.ndarray(self.olen, np.float32, buffer=self.shm.buf).copy()

    def close(self):
        self.shm.close()
        self.shm.unlink()
