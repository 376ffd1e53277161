"""Writes the golden fixtures under tests/golden/ from numpy / scipy (no reference code).

The reference ships no tests, fixtures or sample outputs, so its two pieces of
third-party-style arithmetic are pinned against independent library routines:

* the 3-D real FFT pair that cuFFT D2Z/Z2D provides (density.h:67-74, 149-205):
  numpy.fft.rfftn (unnormalised forward) and numpy.fft.irfftn * M (unnormalised C2R);
* the natural cubic spline of GPU_kernels.h:232-329: scipy.interpolate.CubicSpline
  with bc_type='natural' (derivative at the nodes, value at reference positions).

Run:  python tests/golden/make_fixtures.py
"""
import os

import numpy as np
from scipy.interpolate import CubicSpline

HERE = os.path.dirname(os.path.abspath(__file__))


def fft_fixtures(rng):
    out = {}
    shapes = [(4, 4, 4), (6, 10, 12), (8, 8, 8), (5, 6, 9), (3, 4, 10), (12, 8, 16)]
    for i, shp in enumerate(shapes):
        a = rng.standard_normal(shp)
        out[f"fft{i}_shape"] = np.array(shp)
        out[f"fft{i}_in"] = a
        out[f"fft{i}_rfft"] = np.fft.rfftn(a)
        # a Hermitian half spectrum scaled by a k-symmetric real factor, for the C2R check
        b = rng.standard_normal(shp)
        kx = np.fft.fftfreq(shp[0]) * shp[0]
        ky = np.fft.fftfreq(shp[1]) * shp[1]
        kz = np.arange(shp[2] // 2 + 1)
        k2 = kx[:, None, None] ** 2 + ky[None, :, None] ** 2 + kz[None, None, :] ** 2
        fac = np.exp(-0.05 * k2)
        h = np.fft.rfftn(b) * fac
        out[f"fft{i}_hin"] = h
        out[f"fft{i}_irfft"] = np.fft.irfftn(h, s=shp, axes=(0, 1, 2)) * np.prod(shp)
    out["nfft"] = np.array(len(shapes))
    return out


def spline_fixtures(rng):
    out = {}
    cases = [3, 4, 5, 9, 16, 48]
    for i, n in enumerate(cases):
        gaps = rng.uniform(0.3, 1.7, n - 1)
        x = np.concatenate([[0.0], np.cumsum(gaps)])
        x /= x[-1]
        x[-1] = 1.0
        y = rng.standard_normal(n)
        xr = np.arange(n) / (n - 1.0)
        cs = CubicSpline(x, y, bc_type="natural")
        out[f"spl{i}_x"] = x
        out[f"spl{i}_y"] = y
        out[f"spl{i}_xref"] = xr
        out[f"spl{i}_deriv"] = cs(x, 1)
        out[f"spl{i}_value"] = cs(xr)
    out["nspl"] = np.array(len(cases))
    return out


def main():
    rng = np.random.default_rng(20240611)
    data = {}
    data.update(fft_fixtures(rng))
    data.update(spline_fixtures(rng))
    np.savez_compressed(os.path.join(HERE, "fixtures.npz"), **data)
    print("wrote", os.path.join(HERE, "fixtures.npz"), len(data), "arrays")


if __name__ == "__main__":
    main()
