"""A client harness with STFTProxy's runtime-fallback semantics, and a CPU
runner speaking the same protocol, for end-to-end tests of `hipRun`.

Proxy follows StandAlone/src/STFTProxy.cpp and runtimeChecker.cpp:
  * the port is a random free port in [49152, 65535] (GeneratePortNumber);
  * a runtime is available if its runtime library loads (libamdhip64.so for
    HIP; RuntimeCheck::isAvailable) and its runner executable exists; the
    runner is started with the port as its argument and connected to with 5
    tries 500 ms apart (TryConnect);
  * a connection error or a close that is not a normal close moves to the next
    runtime of the fallback list (RuntimeFallback); an exhausted list sets the
    status "Run Out Fallback" and every later request returns None;
  * KillRunner sends "CLOSE_REQUEST" and waits up to 5 s for
    "CLOSE_COMPLETE".
A request in flight when its runner dies returns None (the reference leaves its
promise unfulfilled); later requests go to the fallback runtime.

CpuRunner is the CPU fallback: a WebSocket server answering RequestCapnp
messages with the oracle's restatement of the reference Serial runner (test
infrastructure, like the oracle itself).
"""
from __future__ import annotations

import base64
import ctypes
import hashlib
import os
import random
import socket
import struct
import subprocess
import threading
import time

import numpy as np

import wire_client as wc

import pyoracle as po

GUID = "258EAFA5-E914-47DA-95CA-C5AB0DC85B11"


# ------------------------------------------------------------------ CPU runner

class CpuRunner:
    """Serial-restatement runner on 127.0.0.1:port (one thread per connection)."""

    def __init__(self, port: int):
        self.port = port
        self.srv = socket.socket()
        self.srv.setsockopt(socket.SOL_SOCKET, socket.SO_REUSEADDR, 1)
        self.srv.bind(("127.0.0.1", port))
        self.srv.listen(8)
        self.stopped = threading.Event()
        self.served = 0
        self.thread = threading.Thread(target=self._accept, daemon=True)
        self.thread.start()

    def _accept(self):
        while not self.stopped.is_set():
            try:
                self.srv.settimeout(0.2)
                conn, _ = self.srv.accept()
            except socket.timeout:
                continue
            except OSError:
                return
            threading.Thread(target=self._serve, args=(conn,), daemon=True).start()

    @staticmethod
    def _recv_exact(conn, n, buf):
        while len(buf) < n:
            chunk = conn.recv(65536)
            if not chunk:
                raise ConnectionError("closed")
            buf += chunk
        return buf[:n], buf[n:]

    def _send(self, conn, opcode, payload):
        n = len(payload)
        hdr = bytes([0x80 | opcode])
        if n < 126:
            hdr += bytes([n])
        elif n < 65536:
            hdr += bytes([126]) + struct.pack(">H", n)
        else:
            hdr += bytes([127]) + struct.pack(">Q", n)
        conn.sendall(hdr + payload)

    def _serve(self, conn):
        buf = b""
        try:
            while b"\r\n\r\n" not in buf:
                chunk = conn.recv(4096)
                if not chunk:
                    return
                buf += chunk
            head, buf = buf.split(b"\r\n\r\n", 1)
            key = [ln.split(":", 1)[1].strip() for ln in head.decode().split("\r\n")
                   if ln.lower().startswith("sec-websocket-key:")][0]
            acc = base64.b64encode(hashlib.sha1((key + GUID).encode()).digest()).decode()
            conn.sendall(("HTTP/1.1 101 Switching Protocols\r\nUpgrade: websocket\r\nConnection: Upgrade\r\n"
                          f"Sec-WebSocket-Accept: {acc}\r\n\r\n").encode())
            msg, mop = b"", 0
            while True:
                h, buf = self._recv_exact(conn, 2, buf)
                op, n = h[0] & 0x0F, h[1] & 0x7F
                if n == 126:
                    e, buf = self._recv_exact(conn, 2, buf)
                    (n,) = struct.unpack(">H", e)
                elif n == 127:
                    e, buf = self._recv_exact(conn, 8, buf)
                    (n,) = struct.unpack(">Q", e)
                mk, buf = self._recv_exact(conn, 4, buf)
                p, buf = self._recv_exact(conn, n, buf)
                p = (np.frombuffer(p, np.uint8) ^ np.frombuffer((mk * (n // 4 + 1))[:n], np.uint8)).tobytes()
                if op == 8:
                    self._send(conn, 8, p[:2])
                    return
                if op == 9:
                    self._send(conn, 10, p)
                    continue
                if op in (1, 2):
                    msg, mop = p, op
                else:
                    msg += p
                if not h[0] & 0x80:
                    continue
                if mop == 1:
                    if msg == b"CLOSE_REQUEST":
                        self._send(conn, 1, b"CLOSE_COMPLETE")
                        self.stop()
                        return
                    continue
                self._send(conn, 2, self._answer(msg))
        except (ConnectionError, OSError):
            return
        finally:
            conn.close()

    def _answer(self, msg: bytes) -> bytes:
        try:
            r = wc.decode(msg)
        except Exception:
            return wc.encode(wc.Req(sharedMemory="ERR"))
        rc, y = po.pipeline(np.asarray(r.data, np.float32), r.windowRadix, r.overlapRatio, r.options)
        rep = wc.Req(sharedMemory=r.sharedMemory, mappedID=r.mappedID, options="", memPTR=r.memPTR,
                     posixFileDes=r.posixFileDes, windowsHandlePTR=r.windowsHandlePTR, windowRadix=r.windowRadix,
                     overlapRatio=r.overlapRatio, overlapdataLength=r.overlapdataLength)
        if rc != 0:
            rep.sharedMemory = "ERR"
            rep.dataLength = r.dataLength
        else:
            rep.data = y.astype(np.float32)
            rep.dataLength = y.size
        self.served += 1
        return wc.encode(rep)

    def stop(self):
        """Stop listening; the port is free when this returns."""
        self.stopped.set()
        try:
            self.srv.shutdown(socket.SHUT_RDWR)  # wakes a blocked accept
        except OSError:
            pass
        if threading.current_thread() is not self.thread:
            self.thread.join(timeout=5)
        try:
            self.srv.close()
        except OSError:
            pass


# ------------------------------------------------------------------ proxy

def runner_path(kind: str, path):
    """The executable a fallback entry names: a directory gets the runner's
    file name appended (runtimeChecker.cpp:139-196)."""
    if kind != "HIP" or path is None:
        return path
    return path if os.path.basename(path) == "hipRun" else os.path.join(path, "hipRun")


def runtime_available(kind: str, path) -> bool:
    """RuntimeCheck::isAvailable (runtimeChecker.cpp:138-222): the runtime
    library loads and the runner executable exists.  SERIAL is served in this
    process, so it is always available."""
    if kind == "HIP":
        try:
            ctypes.CDLL("libamdhip64.so")
        except OSError:
            return False
        return path is not None and os.path.exists(runner_path(kind, path))
    return kind == "SERIAL"


def generate_port() -> int:
    while True:
        port = random.randint(49152, 65535)
        with socket.socket() as s:
            try:
                s.bind(("127.0.0.1", port))
                return port
            except OSError:
                continue


class Proxy:
    def __init__(self, fallback, tries: int = 5, interval: float = 0.5):
        self.fallback = list(fallback)  # [(kind, path)], tried in order
        self.tries, self.interval = tries, interval  # TryConnect: 5 tries, 500 ms apart
        self.port = generate_port()
        self.status = "OK"
        self.runtime = None
        self.proc = None
        self.cpu = None
        self.client = None
        self.next_id = 0
        self._runtime_fallback()

    def _try_connect(self, kind, path) -> bool:
        if not runtime_available(kind, path):
            return False
        if kind == "HIP":
            self.proc = subprocess.Popen([runner_path(kind, path), str(self.port)], stdout=subprocess.DEVNULL,
                                         stderr=subprocess.DEVNULL)
        else:
            self.cpu = CpuRunner(self.port)
        for _ in range(self.tries):
            if self.proc is not None and self.proc.poll() is not None:
                break
            try:
                self.client = wc.WsClient(self.port, timeout=120)
                return True
            except OSError:
                time.sleep(self.interval)
        if self.proc is not None and self.proc.poll() is None:
            self.proc.kill()
            self.proc.wait(timeout=30)
        return False

    def _runtime_fallback(self):
        if self.client:
            self.client.close()
            self.client = None
        if self.proc is not None:  # the previous runner: dead, or ended here
            if self.proc.poll() is None:
                self.proc.kill()
            self.proc.wait(timeout=30)
            self.proc = None
        while self.status == "OK":
            if not self.fallback:
                self.status = "Run Out Fallback"
                break
            kind, path = self.fallback.pop(0)
            if self._try_connect(kind, path):
                self.runtime = kind
                break

    def request(self, x, exp, ov, options):
        """RequestSTFT over the data list; the result array, or None."""
        if self.status != "OK":
            return None
        x = np.ascontiguousarray(x, np.float32)
        rid = str(self.next_id)
        self.next_id += 1
        msg = wc.encode(wc.Req(mappedID=rid, windowRadix=exp, overlapRatio=ov, dataLength=x.size,
                               overlapdataLength=wc.overlap_length(x.size, ov, 1 << exp), options=options, data=x))
        try:
            self.client.send_binary(msg)
            op, p = self.client.recv_message()
            if op == 8:
                raise ConnectionError("closed by the runner")
        except (ConnectionError, OSError):
            self._runtime_fallback()
            return None
        rep = wc.decode(p)
        if rep.sharedMemory == "ERR":
            return None
        return rep.data

    def kill_runner(self) -> bool:
        """KillRunner(true): CLOSE_REQUEST, CLOSE_COMPLETE within 5 s."""
        self.status = "NOFALLBACK"
        if not self.client:
            return False
        try:
            self.client.s.settimeout(5)
            self.client.send_text("CLOSE_REQUEST")
            op, p = self.client.recv_message()
            ok = (op, p) == (1, b"CLOSE_COMPLETE")
            if ok and self.proc is not None:
                try:
                    self.proc.wait(timeout=5)  # the runner ends by itself
                except subprocess.TimeoutExpired:
                    pass
            return ok
        except (ConnectionError, OSError):
            return False
        finally:
            self.close()

    def close(self):
        if self.client:
            self.client.close()
            self.client = None
        if self.proc and self.proc.poll() is None:
            self.proc.kill()
            self.proc.wait(timeout=30)
        if self.cpu:
            self.cpu.stop()
