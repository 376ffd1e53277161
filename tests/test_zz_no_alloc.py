"""scft_spline_derivs_noalloc: the allocation-free entry point, through the C-ABI.

It must give what scft_spline_derivs (the chunked host-row entry point it varies) gives on
the same rows, on two calls in a row that share the library's static scratch, the second
with a smaller input (fewer replicas, fewer points, other nodes), and refuse more replicas
than the scratch holds.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from scft_string_amd import Context, capi

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _reference(m, W, x, doiis, n):
    """scft_spline_derivs on host rows (W [R][2M], W- first) over points 0 .. n-1"""
    R = W.shape[0]
    out = np.full((R, n), -5.0)
    with Context(m, (4.0, 4.0, 4.0), 20, 0.5, 0.1, 1, 0.8, 15.0, nrep=R) as ctx:
        ctx.spline_setup(n)
        ctx.spline_set_x(x)
        ctx.spline_derivs(np.ascontiguousarray(W), 0, n, doiis, out)
    return out


def _noalloc(W_dev, x, doiis, n, wstride, out_dev):
    import torch

    st = torch.cuda.current_stream()
    xs = np.ascontiguousarray(x, dtype=np.float64)
    d = np.ascontiguousarray(doiis, dtype=np.int32)
    rc = capi.lib().scft_spline_derivs_noalloc(len(xs), n, capi._dp(xs), d.ctypes.data_as(C.POINTER(C.c_int)),
                                               C.c_void_p(W_dev.data_ptr()), wstride, C.c_void_p(out_dev.data_ptr()),
                                               out_dev.shape[1], C.c_void_p(st.cuda_stream))
    st.synchronize()
    return rc


def test_noalloc_matches_chunked_entry_point_twice():
    import torch

    m = (16, 16, 16)
    M = int(np.prod(m))
    rng = np.random.default_rng(21)
    calls = [(9, M, np.array([1, 1, 1, 0, 1, 1, 1, 1, 1])), (5, 1000, np.array([0, 1, 1, 1, 1]))]
    for c, (R, n, doiis) in enumerate(calls):
        W = rng.standard_normal((R, 2 * M))
        x = np.concatenate([[0.0], np.sort(rng.uniform(0.02, 0.98, R - 2)), [1.0]])
        _say(f"noalloc call {c}: {R} replicas, {n} points")
        want = _reference(m, W, x, doiis, n)
        W_dev = torch.from_numpy(W).cuda()
        out_dev = torch.full((R, n), -5.0, dtype=torch.float64, device="cuda")
        assert _noalloc(W_dev, x, doiis, n, 2 * M, out_dev) == capi.SCFT_OK, capi.lib().scft_last_error(None)
        got = out_dev.cpu().numpy()
        act = doiis == 1
        err = np.max(np.abs(got[act] - want[act])) / np.max(np.abs(want[act]))
        _say(f"noalloc call {c}: rel diff against scft_spline_derivs {err:.2e}")
        assert err <= 1e-12
        assert np.all(got[~act] == -5.0)


def test_noalloc_refuses_what_scratch_cannot_hold():
    import torch

    R, n = 65, 256
    W_dev = torch.zeros((R, n), dtype=torch.float64, device="cuda")
    out_dev = torch.zeros((R, n), dtype=torch.float64, device="cuda")
    x = np.linspace(0.0, 1.0, R)
    rc = _noalloc(W_dev, x, np.ones(R), n, n, out_dev)
    assert rc == capi.SCFT_ERR_UNSUPPORTED
    assert b"static scratch" in capi.lib().scft_last_error(None)
    x2 = np.array([0.0, 0.5, 0.5, 1.0])
    assert _noalloc(W_dev[:4], x2, np.ones(4), n, n, out_dev[:4]) == capi.SCFT_ERR_INVALID
