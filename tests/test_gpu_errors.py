"""Error behaviour of the C-ABI. The reference exits the process on any CUDA error
(HANDLE_ERROR, GPU_kernels.h:16-48) and asserts when an error norm is NaN (scft.cu:539,
770); every entry point here returns a status instead, with a message in scft_last_error,
and leaves the library usable.

NaN fields only flow through arithmetic: the one search over field-derived values (the
spline interval search in spline_setup) starts at 0, advances only on a true comparison
and is clamped to the valid intervals, so NaN arc positions cannot index out of bounds.
"""
import numpy as np
import pytest

from scft_string_amd import Context, ScftError
from scft_string_amd.capi import SCFT_ERR_INVALID, SCFT_ERR_NAN, SCFT_ERR_UNSUPPORTED

pytestmark = pytest.mark.gpu

BASE = dict(m=(16, 16, 16), D=(3.0, 3.0, 3.0), N=20, f=0.5, alpha=0.1, ens=1, phidb=0.8, chi=15.0)


def _ctx(**kw):
    p = dict(BASE)
    p.update(kw)
    extra = {k: p.pop(k) for k in list(p) if k not in BASE}
    return Context(p["m"], p["D"], p["N"], p["f"], p["alpha"], p["ens"], p["phidb"], p["chi"], **extra)


@pytest.mark.parametrize("kw", [
    dict(N=1),                               # one contour step
    dict(nrep=0),
    dict(ens=3),
    dict(f=0.001),                           # n1 = round(N f) = 0
    dict(f=0.2, alpha=0.5, phihA=0.1),       # A homopolymer longer than the A block (Nh > n1)
    dict(phihA=0.3),                         # canonical: phidb + phihA > 1
    dict(anderson=5, dostring=1),            # Anderson mixing needs dostring = 0
    dict(anderson=41, dostring=0),
    dict(precision="fp64", ckpt=-2),
], ids=lambda kw: ",".join(f"{k}={v}" for k, v in kw.items()))
def test_invalid_parameters_are_rejected(kw):
    with pytest.raises(ScftError) as ei:
        _ctx(**kw)
    assert ei.value.code == SCFT_ERR_INVALID


@pytest.mark.parametrize("m", [(7, 7, 7), (2048, 1, 1), (14, 16, 16)])
def test_unsupported_grids_are_rejected(m):
    """Axes must factor into 2, 3 and 5 and be at most 1024 long."""
    with pytest.raises(ScftError) as ei:
        _ctx(m=m)
    assert ei.value.code == SCFT_ERR_UNSUPPORTED


def test_bad_row_counts_are_rejected_and_the_context_stays_usable():
    M = 16 ** 3
    with _ctx(nrep=2) as ctx:
        with pytest.raises(ScftError) as ei:
            ctx.props_batch(np.zeros((3, 2, M)))
        assert ei.value.code == SCFT_ERR_INVALID
        with pytest.raises(ScftError):
            ctx.props_batch(np.zeros((0, 2, M)))
        _, lnQ, _ = ctx.props_batch(np.zeros((2, 2, M)))
    assert np.all(np.isfinite(lnQ))


@pytest.mark.parametrize("mode", ["solve_w_plus", "string", "plain", "anderson"])
def test_nan_fields_return_nan_status(mode):
    """A NaN in the fields gives SCFT_ERR_NAN from the field iterations, where the reference
    asserts; the context then takes new fields and runs normally."""
    M = 16 ** 3
    R = 3 if mode == "string" else 1
    kw = dict(nrep=R, dostring=1 if mode == "string" else 0)
    if mode == "anderson":
        kw.update(anderson=3, anderson_start=10.0)
    rng = np.random.default_rng(1)
    W = 0.1 * rng.standard_normal((R, 2, M))
    bad = W.copy()
    bad[0, 0, 123] = np.nan
    with _ctx(**kw) as ctx:
        ctx.set_fields(bad)
        with pytest.raises(ScftError) as ei:
            if mode == "solve_w_plus":
                ctx.solve_w_plus(20, 1e-4)
            else:
                ctx.string_iterate(0)
        assert ei.value.code == SCFT_ERR_NAN and "NaN" in str(ei.value)
        ctx.set_fields(W)
        rep = ctx.string_iterate(0)
    assert np.all(np.isfinite(rep["errs"]))
