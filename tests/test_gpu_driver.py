"""The C++ host driver (scft_drive) end to end: input.dat in, win / vtk / check / FEs out.

The string run repeats main()'s order (scft.cu:1102-1155):
- solve_field0 on every replica;
- maxiter iterations of solve_field, with FreeE at k = 0 and k = maxiter - 1;
- a final FreeE.
The oracle runs the same sequence from the same win files, and the FEs file (10 decimals,
scft.cu:843-850) must agree to its printed precision, with the full q store and with
`ckpt = 5`. The relaxation run checks the `relaxbox` key.
"""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from oracle import orc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "scft_string_amd", "scft_drive")


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _write_input(d, **kw):
    lines = [f"{k} = {' '.join(map(str, v)) if isinstance(v, (list, tuple)) else v}" for k, v in kw.items()]
    with open(os.path.join(d, "input.dat"), "w") as fh:
        fh.write("\n".join(lines) + "\n")


def _write_win(path, W):
    M = W.shape[1]
    with open(path, "w") as fh:
        for r in range(M):
            fh.write(f"{W[0, r]:.6f} {W[1, r]:.6f}\n")


def _run(d, timeout=240):
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} is not built (python -m scft_string_amd.build --driver)")
    out = subprocess.run([DRIVER, "input.dat"], cwd=d, capture_output=True, text=True, timeout=timeout)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    return out.stdout


@pytest.mark.parametrize("ckpt", [0, 5])
def test_string_run_matches_oracle(tmp_path, ckpt):
    m, D, N, R = (16, 16, 16), (4.0, 4.0, 4.0), 20, 4
    chi, f, alpha, phidb, ens, maxiter, prerelax = 15.0, 0.5, 0.1, 0.8, 1, 3, 100
    M = int(np.prod(m))
    rng = np.random.default_rng(5)
    x = (np.arange(m[0]) * D[0] / m[0])[:, None, None] * np.ones(m)
    y = (np.arange(m[1]) * D[1] / m[1])[None, :, None] * np.ones(m)
    a = 0.5 * np.cos(2 * np.pi * x / D[0])
    b = 0.5 * np.cos(2 * np.pi * x / D[0]) * np.cos(2 * np.pi * y / D[1])
    for i in range(R):
        t = i / (R - 1)
        W = np.zeros((2, M))
        W[0] = ((1 - t) * a + t * b).reshape(-1) + 0.01 * rng.standard_normal(M)
        _write_win(os.path.join(tmp_path, f"win{i}"), W)
    # the fields exactly as the driver parses them (6 decimals)
    W0 = np.stack([np.loadtxt(os.path.join(tmp_path, f"win{i}")).T.copy() for i in range(R)])
    _write_input(tmp_path, chi=chi, f=f, ens=ens, phidb=phidb, N=N, m=m, D=D, fix=(1, 1), readwin=1,
                 justFE=0, dostring=1, strn=R, maxiter=maxiter, prerelax=prerelax, precision=64, ckpt=ckpt)
    _say(f"driver: string run, ckpt = {ckpt}")
    log = _run(tmp_path)
    for name in ["FEs", "check"] + [f"win{i}" for i in range(R)] + [f"rhoA_{i}.vtk" for i in range(R)]:
        assert os.path.exists(os.path.join(tmp_path, name)), (name, log[-1000:])
    got = np.loadtxt(os.path.join(tmp_path, "FEs"))
    assert got.shape == (R, 3)

    # the same sequence on the oracle, from the same win-file fields
    g = orc.Grid(m, D)
    st = orc.StringState(g, dict(chi=chi, f=f, alpha=alpha, phidb=phidb, N=N, ens=ens, dostring=1, fix=(1, 1)), W0)
    for ii in range(R):
        g.solve_field0(st.W[ii], st.DEV[ii], chi, f, N, alpha, phidb, ens, prerelax, 1e-4)
    for k in range(maxiter):
        st.iterate(k, 1e-3)
        if k % 50 == 0 or k == maxiter - 1:
            st.free_energy()
    fe = st.free_energy()
    ref = np.stack([fe["alf"] / fe["alf"][-1], fe["FEs"], fe["phctot"]], axis=1)
    _say(f"driver FEs {got[:, 1]} oracle {ref[:, 1]}")
    assert np.max(np.abs(got[:, 1:] - ref[:, 1:])) <= 2e-10
    assert np.max(np.abs(got[:, 0] - ref[:, 0])) <= 2e-6   # printed with 6 decimals


def test_relaxbox_key(tmp_path):
    m, D0 = (32, 1, 1), 4.0
    x = np.arange(32) * D0 / 32
    _write_win(os.path.join(tmp_path, "win0"), np.stack([2.0 * np.cos(2 * np.pi * x / D0), np.zeros(32)]))
    _write_input(tmp_path, chi=15.0, f=0.5, ens=1, phidb=0.8, N=40, m=m, D=(D0, 1.0, 1.0), fix=(0, 0),
                 readwin=1, justFE=0, dostring=0, strn=1, maxiter=0, prerelax=200, precision=64, relaxbox=1,
                 stresstol=1e-6)
    _say("driver: relaxbox run")
    log = _run(tmp_path)
    mm = re.search(r"Box relaxed after (\d+) boxes: D = (\S+) (\S+) (\S+), dF/dD = (\S+)", log)
    assert mm, log[-2000:]
    D, g = float(mm.group(2)), float(mm.group(5))
    _say(f"driver relaxed D = {D}, dF/dD = {g}")
    assert abs(g) < 1e-6 and 4.3 < D < 4.7
