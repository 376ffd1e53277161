"""Anderson mixing (scft_params.anderson, dostring = 0) on the GPU, through the C-ABI.

The GPU iteration follows the oracle's orc_anderson_iter step by step. The 16^3 case uses
two independent replicas and two histories. With the default start it converges to the
lamella that the reference's own iteration reaches, in a fraction of the props. The box
relaxation runs on it.
"""
import sys
import time

import numpy as np
import pytest

from oracle import orc
from scft_string_amd import Context, ScftError

pytestmark = pytest.mark.gpu

LAM = dict(m=(32, 1, 1), D=(4.0, 1.0, 1.0), N=40, f=0.5, alpha=0.1, ens=1, phidb=0.8, chi=15.0)


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _lamella(m, D, amp=2.0):
    x = (np.arange(m[0]) * D[0] / m[0])[:, None, None] * np.ones(m)
    return np.stack([amp * np.cos(2 * np.pi * x / D[0]).reshape(-1), np.zeros(int(np.prod(m)))])


def _ctx(c, R=1, **kw):
    return Context(c["m"], c["D"], c["N"], c["f"], c["alpha"], c["ens"], c["phidb"], c["chi"], nrep=R,
                   dostring=0, **kw)


@pytest.mark.parametrize("case", ["1d_H20", "16cube_H5_2rep"])
def test_trajectory_matches_oracle(case):
    if case == "1d_H20":
        c, R, H, iters = LAM, 1, 20, 40
    else:
        c, R, H, iters = dict(LAM, m=(16, 16, 16), D=(4.0, 4.0, 4.0), N=20), 2, 5, 25
    _say("anderson trajectory", case)
    rng = np.random.default_rng(3)
    W0 = np.stack([_lamella(c["m"], c["D"]) + 0.05 * rng.standard_normal((2, int(np.prod(c["m"]))))
                   for _ in range(R)])
    # Anderson from the first iteration (start above any errs) exercises the history path
    with _ctx(c, R, anderson=H, anderson_lambda=0.2, anderson_start=1e30) as ctx:
        ctx.set_fields(W0)
        eg = [ctx.string_iterate(k)["errs"] for k in range(iters)]
        Wg = ctx.get_fields().reshape(R, 2, -1)
    g = orc.Grid(c["m"], c["D"])
    for i in range(R):
        W = W0[i].copy()
        a = orc.Anderson(g, H, c["chi"], c["f"], c["N"], c["alpha"], c["phidb"], c["ens"], 0.2, 1e30)
        eo = [a.iterate(W) for _ in range(iters)]
        # errs -> 0 is a difference of O(chi) terms, so its bar is absolute
        de = max(abs(eg[k][i] - eo[k]) for k in range(iters))
        dw = float(np.max(np.abs(Wg[i] - W)) / np.max(np.abs(W)))
        _say(f"  replica {i}: errs {eo[0]:.2e} -> {eo[-1]:.2e}; max |d(errs)| {de:.1e}, rel dW {dw:.1e}")
        assert de <= 1e-10 and dw <= 1e-8


def test_converges_to_the_reference_iterations_lamella():
    c = LAM
    t = time.perf_counter()
    with _ctx(c, anderson=20) as ctx:
        ctx.set_fields(_lamella(c["m"], c["D"])[None])
        ka = next(k for k in range(5000) if ctx.string_iterate(k)["errs"][0] < 1e-9)
        Wa = ctx.get_fields()[0]
    ta = time.perf_counter() - t
    t = time.perf_counter()
    props_ref = 0
    with _ctx(c) as ctx:
        ctx.set_fields(_lamella(c["m"], c["D"])[None])
        for k in range(20000):
            rep = ctx.string_iterate(k)
            props_ref += int(rep["steps"][0])
            if rep["errs"][0] < 1e-9:
                break
        Wr = ctx.get_fields()[0]
    tr = time.perf_counter() - t
    _say(f"anderson: {ka + 1} props in {ta:.2f} s; reference iteration: {props_ref} props in {tr:.2f} s")
    M = int(np.prod(c["m"]))
    assert np.max(np.abs(Wa[:M] - Wr[:M])) <= 1e-6   # W- (the W+ mean is not fixed by DEV)
    assert (ka + 1) * 4 < props_ref


def test_relax_box_with_anderson():
    c = LAM
    with _ctx(c, anderson=20) as ctx:
        ctx.set_fields(_lamella(c["m"], c["D"])[None])
        D, g, nbox = ctx.relax_box(axes=(1, 0, 0), stress_tol=1e-6, err_tol=1e-9)
    _say(f"anderson relaxed box D = {D[0]!r} after {nbox} boxes, dF/dD = {g[0]:.3e}")
    assert abs(g[0]) < 1e-6 and abs(D[0] - 4.4921980122) < 1e-6


def test_invalid_settings_are_refused():
    c = LAM
    with pytest.raises(ScftError):
        Context(c["m"], c["D"], c["N"], c["f"], c["alpha"], c["ens"], c["phidb"], c["chi"], nrep=2, dostring=1,
                anderson=5)
    with pytest.raises(ScftError):
        _ctx(c, anderson=41)
