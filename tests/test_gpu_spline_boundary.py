"""The string's spline at the boundary: class density's chunked host-row methods
(density.h:462-528) as C-ABI entry points, and a caller written like scft.cu against the
C++ mirror of density.

- scft_spline_derivs / scft_spline_redist against the oracle's per-point spline
  (orc_spline_deriv_nodes / orc_spline_value_at, themselves pinned by scipy's natural
  CubicSpline), over ragged chunks, with rows that doiis leaves alone untouched;
- the argument checks;
- scft_dropin_example (scft.cu's solve_field0 / solve_field body in host loops over W[ii]
  rows, with props and the spline calls behind scft_host::density) against the oracle's
  string iteration, iteration by iteration.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import orc
from scft_string_amd import Context
from scft_string_amd.capi import ScftError

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXAMPLE = os.path.join(ROOT, "scft_string_amd", "scft_dropin_example")


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _columns(fn, x, ww, *args):
    return np.stack([fn(x, ww[:, r].copy(), *args) for r in range(ww.shape[1])], axis=1)


def test_chunked_spline_calls_match_oracle():
    m, R = (16, 16, 16), 7
    M = int(np.prod(m))
    rng = np.random.default_rng(11)
    ww = np.ascontiguousarray(rng.standard_normal((R, M + 5)))   # rows longer than M, as W[ii] are
    x = np.concatenate([[0.0], np.sort(rng.uniform(0.05, 0.95, R - 2)), [1.0]])
    xref = np.arange(R) / (R - 1.0)
    doiis = np.array([1, 1, 0, 1, 1, 1, 0], dtype=np.int32)
    _say("chunked spline: 16^3, 7 replicas, chunks of 1000")
    with Context(m, (4.0, 4.0, 4.0), 20, 0.5, 0.1, 1, 0.8, 15.0, nrep=R) as ctx:
        with pytest.raises(ScftError):
            ctx.spline_derivs(ww, 0, 10, doiis, np.zeros((R, M)))      # no staging yet
        ctx.spline_setup(1000)
        with pytest.raises(ScftError):
            ctx.spline_derivs(ww, 0, 10, doiis, np.zeros((R, M)))      # no nodes yet
        with pytest.raises(ScftError):
            ctx.spline_set_x(np.array([0.0, 0.2, 0.2, 0.5, 0.6, 0.8, 1.0]))
        ctx.spline_set_x(x)
        ctx.spline_set_xref(xref)
        with pytest.raises(ScftError):
            ctx.spline_derivs(ww, M - 10, 20, doiis, np.zeros((R, M)))  # past the grid
        with pytest.raises(ScftError):
            ctx.spline_derivs(ww, 0, 1001, doiis, np.zeros((R, M)))     # past the staging
        out = np.full((R, M), 7.0)
        for r0 in range(0, M, 1000):
            ctx.spline_derivs(ww, r0, min(1000, M - r0), doiis, out)
        w2 = ww.copy()
        for r0 in range(0, M, 1000):
            ctx.spline_redist(w2, r0, min(1000, M - r0), doiis, 0.05)
    want_d = _columns(orc.spline_deriv_nodes, x, ww[:, :M])
    want_v = _columns(orc.spline_value_at, x, ww[:, :M], xref)
    act = doiis == 1
    np.testing.assert_allclose(out[act], want_d[act], rtol=1e-12, atol=1e-12)
    assert np.all(out[~act] == 7.0)
    inner = act.copy()
    inner[0] = inner[-1] = False
    want_w = ww[:, :M] + 0.05 * (want_v - ww[:, :M])
    np.testing.assert_allclose(w2[inner, :M], want_w[inner], rtol=1e-13, atol=1e-13)
    assert np.array_equal(w2[~inner], ww[~inner])       # ends and doiis = 0 rows untouched
    assert np.array_equal(w2[:, M:], ww[:, M:])         # nothing past the chunk


def test_scft_cu_shaped_caller_matches_oracle(tmp_path):
    if not os.path.exists(EXAMPLE):
        pytest.fail(f"{EXAMPLE} is not built (python -m scft_string_amd.build --driver)")
    m, D, N, R = (16, 16, 16), (4.0, 4.0, 4.0), 20, 5
    f, alpha, ens, phidb, chi, springsteps, iters, chunk = 0.5, 0.1, 1, 0.8, 15.0, 1, 3, 1500
    M = int(np.prod(m))
    rng = np.random.default_rng(4)
    x = (np.arange(m[0]) * D[0] / m[0])[:, None, None] * np.ones(m)
    y = (np.arange(m[1]) * D[1] / m[1])[None, :, None] * np.ones(m)
    a = 0.8 * np.cos(2 * np.pi * x / D[0])
    b = 0.8 * np.cos(2 * np.pi * x / D[0]) * np.cos(2 * np.pi * y / D[1])
    W0 = np.zeros((R, 2, M))
    for i in range(R):
        t = i / (R - 1)
        W0[i, 0] = ((1 - t) * a + t * b).reshape(-1) + 0.02 * rng.standard_normal(M)
    cfg = tmp_path / "config"
    cfg.write_text(f"{m[0]} {m[1]} {m[2]} {D[0]} {D[1]} {D[2]} {N} {f} {alpha} {ens} {phidb} {chi} {R} 1 1 "
                   f"{springsteps} {iters} {chunk}\n")
    W0.tofile(tmp_path / "W0.bin")
    _say("scft.cu-shaped caller: 16^3, 5 replicas, 3 iterations (one spring step)")
    r = subprocess.run([EXAMPLE, str(cfg), str(tmp_path / "W0.bin"), str(tmp_path / "W1.bin")], capture_output=True,
                       text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-1500:] + r.stderr[-1500:]
    rows = np.array([[float(v) for v in l.split()] for l in r.stdout.strip().splitlines()])
    W1 = np.fromfile(tmp_path / "W1.bin").reshape(R, 2, M)

    g = orc.Grid(m, D)
    st = orc.StringState(g, dict(chi=chi, f=f, alpha=alpha, phidb=phidb, N=N, ens=ens, dostring=1, fix=(1, 1),
                                 springsteps=springsteps), W0)
    for k in range(iters):
        rep = st.iterate(k, 1e-3)
        got = rows[k]
        assert int(got[0]) == k
        for j, key in enumerate(("errs", "errs3", "x")):
            np.testing.assert_allclose(got[1 + j * R:1 + (j + 1) * R], rep[key], rtol=1e-10, atol=1e-12,
                                       err_msg=f"iteration {k} {key}")
    Wref = st.W[:, :2 * M].reshape(R, 2, M)
    err = np.max(np.abs(W1 - Wref)) / np.max(np.abs(Wref))
    _say(f"scft.cu-shaped caller vs oracle after {iters} iterations: rel |dW| = {err:.2e}")
    assert err <= 1e-10
