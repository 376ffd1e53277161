"""The two-pass contour step (scft_schedule family "two-pass": one plane pass and one column
pass per step, the kinetic factor split by axis) against the FP64 oracle and against the
four-launch two-stage step, through the C-ABI."""
import sys

import numpy as np
import pytest

from oracle import orc
from scft_string_amd import Context

pytestmark = pytest.mark.gpu

TP = {"family": "two-pass"}


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _fields(m, D, seed, amp=1.5):
    rng = np.random.default_rng(seed)
    M = int(np.prod(m))
    x = (np.arange(m[0]) * D[0] / m[0])[:, None, None] * np.ones(m)
    y = (np.arange(m[1]) * D[1] / m[1])[None, :, None] * np.ones(m)
    Wm = amp * (np.cos(2 * np.pi * x / D[0]) * np.cos(2 * np.pi * y / D[1])).reshape(-1) + 0.3 * rng.standard_normal(M)
    Wp = 0.2 * rng.standard_normal(M)
    return np.stack([Wm, Wp])


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


CASES = [
    # m, D, N, f, alpha, ens, phidb, precision, tol (relative, against the FP64 oracle)
    ((24, 24, 24), (3.0, 3.0, 3.0), 20, 0.5, 0.1, 1, 0.8, "fp64", 1e-10),
    ((32, 32, 32), (4.0, 4.0, 4.0), 24, 0.45, 0.125, 2, 0.7, "fp64", 1e-10),
    ((48, 48, 48), (5.0, 5.0, 5.0), 12, 0.5, 0.1, 1, 0.8, "fp64", 1e-10),
    ((64, 64, 64), (6.0, 6.0, 6.0), 16, 0.45, 0.125, 1, 0.9, "fp64", 1e-10),
    ((64, 64, 64), (6.0, 6.0, 6.0), 16, 0.45, 0.125, 1, 0.9, "fp32", 2e-5),
    ((96, 96, 96), (8.0, 8.0, 8.0), 8, 0.5, 0.125, 1, 0.8, "fp32", 2e-5),
    ((128, 128, 128), (8.0, 8.0, 8.0), 8, 0.5, 0.125, 1, 0.8, "fp32", 2e-5),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[0])) + f"_{c[7]}_ens{c[5]}")
def test_two_pass_props_vs_oracle(case):
    m, D, N, f, alpha, ens, phidb, prec, tol = case
    _say("two-pass props", m, prec)
    W = np.stack([_fields(m, D, seed=60 + i) for i in range(2)])
    g = orc.Grid(m, D)
    with Context(m, D, N, f, alpha, ens, phidb, 15.0, nrep=2, precision=prec, schedule=TP) as ctx:
        assert ctx.fft_path == "two-pass", ctx.fft_path
        phi, lnQ, phis = ctx.props_batch(W)
        qq = ctx.qq_invariant(1)
    for i in range(2):
        ref = g.props(W[i], N=N, alpha=alpha, phic=phidb, f=f, ens=ens)
        dl = abs(lnQ[i] - ref["lnQ"]) / max(1.0, abs(ref["lnQ"]))
        dphi = max(_rel(phi[i, c], ref["phi"][c]) for c in range(4))
        _say(f"  row {i}: rel dlnQ {dl:.2e}, rel dphi {dphi:.2e}")
        assert dl <= tol and dphi <= tol
    # the symmetric-operator invariant <q1[s] q2[N-s]> = QC for every s
    dev = float(np.max(np.abs(qq - qq[0])) / abs(qq[0]))
    _say(f"  invariant spread {dev:.2e}")
    assert dev <= (1e-12 if prec == "fp64" else 1e-5)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-5)])
def test_two_pass_agrees_with_two_stage_at_64cube(prec, tol):
    m, D = (64, 64, 64), (6.0, 6.0, 6.0)
    W = np.stack([_fields(m, D, seed=70 + i) for i in range(3)])
    out = {}
    for fam in ("two-stage", "two-pass"):
        with Context(m, D, 40, 0.45, 0.1, 1, 0.9, 18.0, nrep=3, precision=prec, schedule={"family": fam}) as ctx:
            out[fam] = ctx.props_batch(W)
    dl = float(np.max(np.abs(out["two-pass"][1] - out["two-stage"][1])))
    dphi = _rel(out["two-pass"][0], out["two-stage"][0])
    _say(f"two-pass vs two-stage {prec}: |dlnQ| {dl:.2e}, rel dphi {dphi:.2e}")
    assert dl <= tol and dphi <= tol


@pytest.mark.parametrize("sched", [{"family": "two-pass", "coltile": 32, "streams": 3},
                                   {"family": "two-pass", "coltile": 64, "streams": 1}])
def test_two_pass_schedules_and_checkpointed_store(sched):
    """Column tiles and replica-group streams on the two-pass step, and the checkpointed q
    store (recompute windows resume from stored q through the column pass) against the full
    store: FP64, equal to rounding."""
    m, D = (32, 32, 32), (4.0, 4.0, 4.0)
    W = np.stack([_fields(m, D, seed=80 + i) for i in range(4)])
    with Context(m, D, 30, 0.45, 0.1, 1, 0.8, 15.0, nrep=4, precision="fp64") as c0:
        ref = c0.props_batch(W)
    for ck in (0, 5):
        with Context(m, D, 30, 0.45, 0.1, 1, 0.8, 15.0, nrep=4, precision="fp64", schedule=sched, ckpt=ck) as c1:
            assert c1.fft_path.startswith("two-pass"), c1.fft_path
            phi, lnQ, _ = c1.props_batch(W)
        dl = float(np.max(np.abs(lnQ - ref[1])))
        dphi = _rel(phi, ref[0])
        _say(f"two-pass {sched} ckpt {ck}: |dlnQ| {dl:.2e}, rel dphi {dphi:.2e}")
        assert dl <= 1e-12 * max(1.0, float(np.max(np.abs(ref[1])))) and dphi <= 1e-12


def test_two_pass_string_iteration_matches_two_stage_fp64():
    """Two solve_field iterations of a 6-replica string on each step: FP64 results equal to
    rounding (errs, the fields)."""
    m, D = (32, 32, 32), (4.0, 4.0, 4.0)
    R = 6
    W = np.stack([_fields(m, D, seed=90 + i, amp=0.5) for i in range(R)])
    res = {}
    for fam in ("two-stage", "two-pass"):
        with Context(m, D, 32, 0.45, 0.1, 1, 0.85, 16.0, nrep=R, precision="fp64", fix=(1, 1),
                     schedule={"family": fam}) as ctx:
            ctx.set_fields(W)
            reps = [ctx.string_iterate(k) for k in range(2)]
            res[fam] = (np.array([r["errs"] for r in reps]), ctx.get_fields())
    de = float(np.max(np.abs(res["two-pass"][0] - res["two-stage"][0])))
    dw = _rel(res["two-pass"][1], res["two-stage"][1])
    _say(f"string iterations two-pass vs two-stage: |d errs| {de:.2e}, rel dW {dw:.2e}")
    assert de <= 1e-11 and dw <= 1e-11
