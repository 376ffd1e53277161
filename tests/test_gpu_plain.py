"""The benchmark's comparator (plain per-replica FP64 path with hipFFT) computes the same
props as the FP64 oracle, so the speedup it anchors compares equal work."""
import numpy as np
import pytest

from oracle import orc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("m,N,ens", [((16, 16, 16), 20, 1), ((32, 32, 32), 24, 2)])
def test_plain_fp64_matches_oracle(m, N, ens):
    from scft_string_amd.plain import PlainFP64

    D = (4.0, 4.0, 4.0)
    rng = np.random.default_rng(4)
    M = int(np.prod(m))
    W = np.stack([1.5 * rng.standard_normal(M), 0.3 * rng.standard_normal(M)])
    p = PlainFP64(m, D, N)
    phi, lnQ = p.props(W, N, 0.125, 0.8, 0.45, ens)
    p.close()
    ref = orc.Grid(m, D).props(W, N=N, alpha=0.125, phic=0.8, f=0.45, ens=ens)
    assert abs(lnQ - ref["lnQ"]) <= 1e-10 * max(1.0, abs(ref["lnQ"]))
    assert np.max(np.abs(phi - ref["phi"])) <= 1e-10 * np.max(np.abs(ref["phi"]))
