"""Checkpointed q store (scft_params.ckpt = K): props, densities, the box gradient and the
string iteration against the full store (ckpt = 0), through the C-ABI.

With K > 0 the forward contour loop keeps q at every K-th contour index (plus N and Nh), and
the densities and the box gradient recompute the rest in windows of K from those
checkpoints. The recomputed propagators come from the same kernels on the same inputs, so
only the order of the FP64 Simpson and gradient sums differs: the bar is 1e-12 relative, in
FP64 and in FP32 (the FP32 propagators are recomputed bit for bit too). The full store is
itself pinned against the oracle by the parity tests.
"""
import re
import sys

import numpy as np
import pytest

from scft_string_amd import Context

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _fields(m, D, R, seed):
    rng = np.random.default_rng(seed)
    M = int(np.prod(m))
    x = (np.arange(m[0]) * D[0] / m[0])[:, None, None] * np.ones(m)
    z = (np.arange(m[2]) * D[2] / m[2])[None, None, :] * np.ones(m)
    W = np.zeros((R, 2, M))
    for i in range(R):
        W[i, 0] = (1.5 * np.cos(2 * np.pi * x / D[0]) + 0.4 * np.sin(2 * np.pi * z / D[2])).reshape(-1)
        W[i, 0] += 0.2 * rng.standard_normal(M)
        W[i, 1] = 0.1 * rng.standard_normal(M) + 0.2 * i
    return W


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


CASES = [
    # m, D, N, f, alpha, ens, phidb, phihA, prec, nrep, K, env
    ((32, 32, 32), (4.0, 4.0, 4.0), 40, 0.45, 0.1, 1, 0.9, 0.0, "fp64", 3, 6, {}),
    ((16, 16, 16), (3.0, 3.0, 3.0), 50, 0.5, 0.2, 1, 0.7, 0.0, "fp32", 4, 7, {}),
    ((16, 16, 16), (3.0, 3.0, 3.0), 30, 0.5, 0.3, 2, 0.4, 0.0, "fp64", 2, 5, {"SCFT_LEGACY_FFT": "1"}),
    ((20, 15, 15), (2.0, 1.5, 1.5), 24, 0.55, 0.4, 1, 0.5, 0.2, "fp64", 2, 5, {}),
    ((20, 15, 15), (2.0, 1.5, 1.5), 24, 0.55, 0.4, 2, 0.5, 0.3, "fp32", 2, 4, {}),
    ((12, 8, 16), (3.0, 2.5, 4.0), 21, 0.4, 0.1, 1, 0.7, 0.0, "fp64", 2, 1, {}),
    ((12, 8, 16), (3.0, 2.5, 4.0), 21, 0.4, 0.1, 1, 0.7, 0.0, "fp64", 2, 30, {}),
    # persistent plane_mid (96^2 FP32) and the 4 x 12 z split of 48 inside the recompute windows
    ((96, 96, 96), (6.0, 6.0, 6.0), 24, 0.45, 0.1, 1, 0.9, 0.0, "fp32", 3, 5, {}),
    ((48, 48, 48), (4.0, 4.0, 4.0), 30, 0.5, 0.1, 2, 0.6, 0.0, "fp64", 2, 6, {}),
]


def _ids(c):
    return "x".join(map(str, c[0])) + f"_N{c[2]}_K{c[10]}_{c[8]}_ens{c[5]}" + ("_hA" if c[7] else "") + (
        "_legacy" if c[11] else "")


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_checkpointed_props_and_gradient_match_full_store(case, monkeypatch):
    m, D, N, f, alpha, ens, phidb, phihA, prec, R, K, env = case
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    _say("checkpointed store", case)
    W = _fields(m, D, R, seed=11)
    out = {}
    for ck in (0, K):
        with Context(m, D, N, f, alpha, ens, phidb, chi=15.0, nrep=R, precision=prec, phihA=phihA,
                     ckpt=ck) as ctx:
            info = ctx.engine_info
            phi, lnQ, phis = ctx.props_batch(W)
            lq, dq = ctx.dlnq_dbox(W)
            out[ck] = (phi, lnQ, phis, lq, dq, info)
    info = out[K][5]
    _say(f"  {info}")
    mt = re.search(r"q checkpointed every (\d+) steps \((\d+) of (\d+) entries", info)
    assert mt and int(mt.group(1)) == K and int(mt.group(3)) == N + 1
    assert int(mt.group(2)) == N // K + 3 + 2 * (K + 1)
    p0, p1 = out[0], out[K]
    d_phi = _rel(p1[0], p0[0])
    d_lnQ = float(np.max(np.abs(p1[1] - p0[1])) / max(1.0, np.max(np.abs(p0[1]))))
    d_phis = _rel(p1[2], p0[2])
    d_grad = _rel(p1[4], p0[4])
    _say(f"  rel diff: phi {d_phi:.1e}, lnQ {d_lnQ:.1e}, phis {d_phis:.1e}, dlnQ/dD {d_grad:.1e}")
    assert np.array_equal(p1[1], p0[1])  # lnQ comes from the same stored q[N], q[Nh]
    assert np.array_equal(p1[3], p0[3])
    assert d_phi <= TOL and d_phis <= TOL and d_grad <= TOL


def test_checkpointed_string_iterations_match_full_store():
    """Two string iterations (16^3, 4 replicas, N = 32) and a free energy with K = 6."""
    m, D = (16, 16, 16), (4.0, 4.0, 4.0)
    W = _fields(m, D, 4, seed=5)
    res = {}
    for ck in (0, 6):
        with Context(m, D, 32, 0.45, 0.1, 1, 0.9, 15.0, nrep=4, precision="fp64", fix=(1, 1), ckpt=ck) as ctx:
            ctx.set_fields(W)
            for k in range(2):
                ctx.string_iterate(k)
            res[ck] = (ctx.get_fields(), ctx.free_energy()["FEs"])
    dW = _rel(res[6][0], res[0][0])
    dF = float(np.max(np.abs(res[6][1] - res[0][1])))
    _say(f"string with K = 6: rel dW {dW:.1e}, |dFEs| {dF:.1e}")
    assert dW <= 1e-11 and dF <= 1e-11


def test_invalid_checkpoint_interval_rejected():
    with pytest.raises(Exception):
        Context((16, 16, 16), (3.0, 3.0, 3.0), 20, 0.5, 0.1, 1, 0.8, 15.0, ckpt=-1)


def test_failed_allocation_does_not_poison_the_next_context():
    """A context whose q store cannot be allocated (512^3 FP32, N = 2000: 2 TB) fails with
    an error, and a context created after it runs props normally: the failed hipMalloc is
    not left behind as the error a later launch check reports."""
    with pytest.raises(Exception):
        Context((512, 512, 512), (8.0, 8.0, 8.0), 2000, 0.5, 0.1, 1, 0.8, 15.0, precision="fp32")
    m, D = (16, 16, 16), (3.0, 3.0, 3.0)
    with Context(m, D, 20, 0.5, 0.1, 1, 0.8, 15.0, precision="fp64", ckpt=4) as ctx:
        _, lnQ, _ = ctx.props_batch(_fields(m, D, 1, seed=1))
    assert np.isfinite(lnQ[0])
