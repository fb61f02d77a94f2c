"""ostft_run_device_noalloc: the allocation-free variant of ostft_run_device
(no counterpart in the reference; for callers that cannot allocate device
memory).  Called through the C ABI; its output must equal ostft_run_device's
bitwise (same kernels; the two-pass windows only change the chunking of
frames, and every frame is its own transform), on two calls in a row, the
second with a smaller input.
"""
import ctypes

import numpy as np
import pytest

from occastft_amd import _native as N

pytestmark = pytest.mark.gpu

OPTS = b"--blackman_window --remove_dc --power_return"


def _ctx():
    L = N.lib()
    h = ctypes.c_void_p()
    assert L.ostft_create(0, ctypes.byref(h)) == N.OK
    return L, h


def _params(L, exp, ov, opts=OPTS):
    p = N.Params()
    assert L.ostft_parse_options(opts, exp, ov, ctypes.byref(p)) == N.OK
    return p


def _geometry(L, p, n):
    g = N.Geometry()
    assert L.ostft_get_geometry(ctypes.byref(p), n, ctypes.byref(g)) == N.OK
    return g


@pytest.mark.parametrize("exp,ov", [(10, 0.5), (20, 0.75), (16, 0.0), (17, 0.75)])
def test_noalloc_matches_run_device_twice(exp, ov):
    import torch

    L, h = _ctx()
    try:
        p = _params(L, exp, ov)
        Nw = 1 << exp
        rng = np.random.default_rng(exp)
        big = (25 * Nw if exp >= 16 else 400 * Nw) + 333  # 2^20 at 75%: 97 frames, 7 chunks of the fixed scratch
        sizes = (big, big // 3 + 17)
        x = torch.from_numpy(rng.standard_normal(big).astype(np.float32)).cuda()
        stream = torch.cuda.current_stream()
        sp = ctypes.c_void_p(stream.cuda_stream)
        outs = []
        for n in sizes:
            g = _geometry(L, p, n)
            a = torch.empty(g.out_floats + 4, dtype=torch.float32, device="cuda")
            b = torch.full((g.out_floats + 4,), float("nan"), dtype=torch.float32, device="cuda")
            outs.append((n, g, a, b))
        # the tables are built by the allocating entry point's first call
        for n, g, a, _ in outs:
            rc = L.ostft_run_device(h, ctypes.byref(p), ctypes.c_void_p(x.data_ptr()), n,
                                    ctypes.c_void_p(a.data_ptr()), a.numel(), sp)
            assert rc == N.OK, L.ostft_last_error(h)
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        for n, g, _, b in outs:  # two calls in a row, the second smaller
            rc = L.ostft_run_device_noalloc(h, ctypes.byref(p), ctypes.c_void_p(x.data_ptr()), n,
                                            ctypes.c_void_p(b.data_ptr()), b.numel(), sp)
            assert rc == N.OK, L.ostft_last_error(h)
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
        assert free1 >= free0 - (2 << 20), (free0, free1)  # no device allocation between
        for n, g, a, b in outs:
            ya = a[: g.out_floats].cpu().numpy()
            yb = b[: g.out_floats].cpu().numpy()
            assert np.array_equal(ya.view(np.uint32), yb.view(np.uint32)), (exp, ov, n)
            assert np.all(np.isnan(b[g.out_floats:].cpu().numpy()))  # nothing written past the result
    finally:
        L.ostft_destroy(h)


def test_noalloc_needs_resident_tables():
    import torch

    L, h = _ctx()
    try:
        p = _params(L, 18, 0.5, b"--nuttall_window --power_return")
        x = torch.zeros(1 << 19, dtype=torch.float32, device="cuda")
        g = _geometry(L, p, x.numel())
        out = torch.empty(g.out_floats, dtype=torch.float32, device="cuda")
        rc = L.ostft_run_device_noalloc(h, ctypes.byref(p), ctypes.c_void_p(x.data_ptr()), x.numel(),
                                        ctypes.c_void_p(out.data_ptr()), out.numel(), None)
        assert rc == N.ERR_NOT_PREPARED
        assert b"not resident" in L.ostft_last_error(h)
    finally:
        L.ostft_destroy(h)
