"""The propagator autotuner (scft_autotune, and scft_params.autotune = 1 that runs it inside
scft_ctx_create): every schedule the grid can run (kernel family; column tile x plane
workgroup size for the two-stage kernels; replica-group streams) is timed in a context of
its own, and the fastest is returned. Every schedule computes the same props (FP64: to
rounding), so a context on any of them agrees with a default one.
"""
import re
import sys

import numpy as np
import pytest

from scft_string_amd import Context
from scft_string_amd.capi import SCFT_ERR_INVALID, ScftError, autotune

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _fields(m, D, R, seed=2):
    rng = np.random.default_rng(seed)
    M = int(np.prod(m))
    x = (np.arange(m[0]) * D[0] / m[0])[:, None, None] * np.ones(m)
    W = np.zeros((R, 2, M))
    for i in range(R):
        W[i, 0] = 1.5 * np.cos(2 * np.pi * x / D[0]).reshape(-1) + 0.3 * rng.standard_normal(M)
    return W


def _cands(report):
    return re.findall(r"([a-z-]+(?: c\d+)?(?: nt\d+)? s\d) ([0-9.]+);", report)


@pytest.mark.parametrize("m,R,prec,nfam", [((64, 64, 64), 3, "fp32", 9), ((32, 32, 32), 3, "fp64", 7),
                                           ((16, 16, 16), 2, "fp64", 3)])
def test_autotune_reports_every_candidate_and_picks_the_fastest(m, R, prec, nfam):
    """64^3 has the two-stage kernels at two column tiles x two plane workgroup sizes, the
    two-pass step at two column tiles, the Stockham kernels and the pencil and split line
    steps (9 kernel choices); 32^3 the two-stage kernels at two column tiles and one plane
    size, the two-pass step at two column tiles, Stockham, two line steps (7); 16^3, whose
    axes are a single register DFT and so have no two-stage set, Stockham and two line steps
    (3)."""
    _say("autotune candidates", m, prec, R)
    D = (4.0, 4.0, 4.0)
    best, rep = autotune(m, D, 30, 0.45, 0.1, 1, 0.8, 15.0, nrep=R, precision=prec, dostring=0,
                         W=_fields(m, D, R))
    _say(f"  {rep}")
    cands = _cands(rep)
    assert "failed" not in rep
    assert len(cands) == nfam * min(3, R), rep  # every kernel choice x streams 1 .. min(3, R)
    name, _ = min(cands, key=lambda c: float(c[1]))
    assert rep.endswith("chosen: " + name)
    assert 1 <= best["streams"] <= min(3, R)


@pytest.mark.parametrize("m,R", [((32, 32, 32), 3), ((16, 16, 16), 2)])
def test_autotuned_context_matches_default(m, R):
    D = (4.0, 4.0, 4.0)
    W = _fields(m, D, R)
    with Context(m, D, 30, 0.45, 0.1, 1, 0.8, 15.0, nrep=R, precision="fp64") as c0:
        phi0, lnQ0, _ = c0.props_batch(W)
    with Context(m, D, 30, 0.45, 0.1, 1, 0.8, 15.0, nrep=R, precision="fp64", autotune=1) as c1:
        phi1, lnQ1, _ = c1.props_batch(W)
        info = c1.engine_info
        phi2, lnQ2, _ = c1.props_batch(W)   # the chosen schedule, again
    _say(f"autotune {m}: {info}")
    assert "(autotuned)" in info
    assert np.max(np.abs(lnQ1 - lnQ0)) <= 1e-12 * np.max(np.abs(lnQ0))
    assert np.max(np.abs(phi1 - phi0)) <= 1e-12 * np.max(np.abs(phi0))
    assert np.array_equal(phi1, phi2)
    name, _ = min(_cands(info), key=lambda c: float(c[1]))
    fam = name.split(" ")[0]
    assert info.startswith(fam), info
    assert info.split(",")[1].strip().startswith(name[-1] + " replica-group stream"), info


@pytest.mark.parametrize("sched", [
    {"family": "two-stage", "coltile": 32, "plane_nt": 256, "streams": 3},
    {"family": "two-pass", "coltile": 32, "streams": 2},
    {"family": "two-stage", "coltile": 64, "plane_nt": 512, "streams": 1},
    {"family": "stockham", "streams": 2},
    {"family": "pencil", "streams": 3},
])
def test_explicit_schedules_agree_at_64cube_fp32(sched):
    """Each schedule on the 64^3 FP32 grid computes the default context's props to FP32
    rounding (the kernels' arithmetic per item is the same; contraction may differ)."""
    m, D = (64, 64, 64), (6.0, 6.0, 6.0)
    W = _fields(m, D, 3, seed=5)
    with Context(m, D, 24, 0.45, 0.1, 1, 0.9, 18.0, nrep=3, precision="fp32") as c0:
        phi0, lnQ0, _ = c0.props_batch(W)
    with Context(m, D, 24, 0.45, 0.1, 1, 0.9, 18.0, nrep=3, precision="fp32", schedule=sched) as c1:
        phi1, lnQ1, _ = c1.props_batch(W)
        info = c1.engine_info
    dl = float(np.max(np.abs(lnQ1 - lnQ0)))
    dphi = float(np.max(np.abs(phi1 - phi0)) / np.max(np.abs(phi0)))
    _say(f"schedule {sched}: {info}; |dlnQ| {dl:.2e}, rel dphi {dphi:.2e}")
    assert info.startswith(sched["family"])
    assert f"{sched['streams']} replica-group stream" in info
    assert dl <= 2e-5 and dphi <= 2e-5


@pytest.mark.parametrize("sched", [{"family": "two-stage", "plane_nt": 384}, {"family": "two-stage", "coltile": 16},
                                   {"family": "pencil", "coltile": 32}])
def test_schedule_the_grid_cannot_run_is_refused(sched):
    m, D = (64, 64, 64), (6.0, 6.0, 6.0)
    with pytest.raises(ScftError) as ei:
        Context(m, D, 24, 0.45, 0.1, 1, 0.9, 18.0, nrep=1, precision="fp32", schedule=sched)
    assert ei.value.code == SCFT_ERR_INVALID
