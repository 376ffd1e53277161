"""Plain SCFT of one replica at the reference's CPU-runnable size, free-running on the GPU
against the oracle: m = 32^3, D = 4.0, N = 100, f = 0.5, chiN = 20, phidb = 0.8, canonical,
alpha = 0.1, dostring = 0 (solve_field, scft.cu:559-806, with one replica and no string
update). W- = 0.1 cos(2 pi 2x/D) + N(0, 0.01^2) from a seeded generator, W+ = 0.

The parity tests teacher-force single iterations from identical states. Here both sides run
40 iterations from the same start without being reset, so any difference that compounds
over iterations shows: the error norm of every iteration (relative 1e-9), the number of
props each iteration runs (equal), and the fields and free energy at the end (1e-9).
SCFT_C0_ITERS=200 runs the configuration's full 200 iterations (about 3 minutes, most of it
the oracle on 16 threads).
"""
import os
import sys

import numpy as np
import pytest

from oracle import orc
from scft_string_amd import Context

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_plain_scft_trajectory_matches_oracle():
    m, D, N, f, alpha, ens, phidb, chi = (32, 32, 32), (4.0, 4.0, 4.0), 100, 0.5, 0.1, 1, 0.8, 20.0
    K = int(os.environ.get("SCFT_C0_ITERS", "40"))
    M = int(np.prod(m))
    rng = np.random.default_rng(20)
    x = (np.arange(m[0]) * D[0] / m[0])[:, None, None] * np.ones(m)
    W0 = np.zeros((1, 2, M))
    W0[0, 0] = 0.1 * np.cos(2 * np.pi * 2 * x / D[0]).reshape(-1) + 0.01 * rng.standard_normal(M)

    orc.set_threads(16)
    st = orc.StringState(orc.Grid(m, D), dict(chi=chi, f=f, alpha=alpha, phidb=phidb, N=N, ens=ens, dostring=0),
                         W0)
    with Context(m, D, N, f, alpha, ens, phidb, chi, nrep=1, precision="fp64", dostring=0) as ctx:
        ctx.set_fields(W0)
        for k in range(K):
            rg = ctx.string_iterate(k)
            ro = st.iterate(k)
            eg, eo = float(rg["errs"][0]), float(ro["errs"][0])
            if k % 10 == 0 or k == K - 1:
                _say(f"k={k}: errs gpu {eg:.12e} oracle {eo:.12e}, props {int(rg['steps'][0])}/{int(ro['steps'][0])}")
            assert abs(eg - eo) <= 1e-9 * abs(eo), (k, eg, eo)
            assert int(rg["steps"][0]) == int(ro["steps"][0]), k
        Wg = ctx.get_fields()
        fg = ctx.free_energy()  # FreeE relaxes W+ first (solve_field0, scft.cu:817-818)
        Wg_fe = ctx.get_fields()
    Wo = st.W[0, :2 * M].copy()
    fo = st.free_energy()
    Wo_fe = st.W[0, :2 * M].copy()
    dW = float(np.max(np.abs(Wg[0] - Wo)) / np.max(np.abs(Wo)))
    dW_fe = float(np.max(np.abs(Wg_fe[0] - Wo_fe)) / np.max(np.abs(Wo_fe)))
    _say(f"after {K} iterations: rel dW {dW:.2e} (after FreeE {dW_fe:.2e}), "
         f"FEs gpu {fg['FEs'][0]!r} oracle {fo['FEs'][0]!r}")
    assert dW <= 1e-9 and dW_fe <= 1e-9
    assert abs(fg["FEs"][0] - fo["FEs"][0]) <= 1e-9 * abs(fo["FEs"][0])
