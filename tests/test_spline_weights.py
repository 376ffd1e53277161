"""scft_spline_weights (host only): the string spline as one weight matrix per launch.

At fixed nodes the natural cubic spline (GPU_kernels.h:232-329) is linear in the data, so
its derivative at the nodes and its values at xref are G @ a. The weights must reproduce the
scipy CubicSpline(bc_type='natural') fixtures and the oracle's per-point spline, keep the
strict-< interval rule, be exact on linear data and reject what the kernels cannot take.
"""
import os

import numpy as np
import pytest

from oracle import orc
from scft_string_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = np.load(os.path.join(ROOT, "tests", "golden", "fixtures.npz"))

pytestmark = pytest.mark.skipif(not os.path.exists(capi.LIB_PATH), reason="libscft_hip.so not built")


@pytest.mark.parametrize("i", range(int(GOLD["nspl"])))
def test_weights_reproduce_scipy_natural_spline(i):
    x, y, xr = GOLD[f"spl{i}_x"], GOLD[f"spl{i}_y"], GOLD[f"spl{i}_xref"]
    R = len(x)
    Gd = capi.spline_weights(x, x, np.arange(R), deriv=True)
    Gv = capi.spline_weights(x, xr, np.arange(R), deriv=False)
    np.testing.assert_allclose(Gd @ y, GOLD[f"spl{i}_deriv"], rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(Gv @ y, GOLD[f"spl{i}_value"], rtol=1e-10, atol=1e-11)


@pytest.mark.parametrize("R", [2, 3, 7, 16, 48, 64])
def test_weights_match_the_oracle_spline(R):
    rng = np.random.default_rng(R)
    x = np.concatenate([[0.0], np.sort(rng.uniform(0.01, 0.99, R - 2)), [1.0]])
    a = rng.standard_normal(R)
    xr = np.arange(R) / (R - 1.0)
    rows = np.arange(R)
    np.testing.assert_allclose(capi.spline_weights(x, x, rows, True) @ a, orc.spline_deriv_nodes(x, a),
                               rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(capi.spline_weights(x, xr, rows, False) @ a, orc.spline_value_at(x, a, xr),
                               rtol=1e-12, atol=1e-12)


def test_selected_rows_and_linear_data():
    x = np.array([0.0, 0.1, 0.35, 0.6, 0.61, 1.0])
    rows = np.array([4, 1])
    Gd = capi.spline_weights(x, x, rows, True)
    assert Gd.shape == (2, 6)
    np.testing.assert_allclose(Gd @ (2.5 * x - 1.0), 2.5, rtol=1e-12)
    # node j is evaluated on interval j - 1 (strict <): the value weights at the nodes are
    # the identity rows
    Gv = capi.spline_weights(x, x, np.arange(6), False)
    np.testing.assert_allclose(Gv, np.eye(6), atol=1e-14)


def test_bad_arguments_are_refused():
    x = np.array([0.0, 0.5, 0.5, 1.0])
    with pytest.raises(capi.ScftError):
        capi.spline_weights(x, x, np.arange(4), True)          # not increasing
    with pytest.raises(capi.ScftError):
        capi.spline_weights(np.array([0.0]), np.array([0.0]), np.arange(1), True)
    with pytest.raises(capi.ScftError):
        capi.spline_weights(np.linspace(0, 1, 4), np.linspace(0, 1, 4), np.array([5]), True)
