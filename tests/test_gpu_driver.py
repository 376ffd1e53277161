"""The C++ host driver (scft_drive) end to end: input.dat in, win / vtk / check / FEs out.

The string run repeats main()'s order (scft.cu:1102-1155):
- solve_field0 on every replica;
- maxiter iterations of solve_field, with FreeE at k = 0 and k = maxiter - 1;
- a final FreeE.
The oracle runs the same sequence from the same win files, and the FEs file (10 decimals,
scft.cu:843-850) must agree to its printed precision, with the full q store and with
`ckpt = 5`. The relaxation run checks the `relaxbox` key. The last test runs the reference's
shipped input.dat (input.dat:1-11) in its own layout: tabs, trailing comments, E notation and
keys written against the '='.
"""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from oracle import orc
from tests.uniform_string import Blend, replay_main

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


def test_uniform_string_files_match_closed_form(tmp_path):
    """main()'s cadence pinned by closed forms: a spatially uniform grand-canonical string
    (every quantity a scalar per replica, tests/uniform_string.py) run for 52 iterations.
    FreeE runs at k = 0, 50 and 51 and relaxes W+ of every replica, the fixed ends included,
    so the cadence shows in the check file (k = 51, errs / errs3 with 12 decimals, the FE of
    the FreeE at k = 50) and in the FEs file of the final FreeE."""
    m, D, R, maxiter, prerelax = (8, 8, 8), (4.0, 4.0, 4.0), 5, 52, 200
    N, f, alpha, phidb, chi, ens = 20, 0.5, 0.1, 0.5, 12.0, 2
    M = int(np.prod(m))
    a0 = np.array([-1.2, -0.7, -0.1, 0.3, 1.4])
    for i in range(R):
        _write_win(os.path.join(tmp_path, f"win{i}"), np.stack([np.full(M, a0[i]), np.zeros(M)]))
    _write_input(tmp_path, chi=chi, f=f, ens=ens, phidb=phidb, N=N, m=m, D=D, fix=(1, 1), readwin=1,
                 justFE=0, dostring=1, strn=R, maxiter=maxiter, prerelax=prerelax, precision=64)
    _say("driver: uniform grand-canonical string, 52 iterations")
    _run(tmp_path)
    want = replay_main(Blend(N, f, alpha, phidb, chi, ens), a0, np.zeros(R), maxiter, prerelax, fix=(1, 1), M=M)
    with open(os.path.join(tmp_path, "check")) as fh:
        lines = [l.split() for l in fh.read().strip().splitlines()]
    k, errtot, FEprev, errs, errs3 = want["check"]
    assert int(lines[0][0]) == k == maxiter - 1
    assert abs(float(lines[0][1]) - errtot) <= 1e-6 and float(lines[0][2]) == 0.05
    assert abs(float(lines[0][3]) - FEprev) <= 1e-6
    got = np.array([[float(v) for v in l[1:]] for l in lines[1:]])
    assert got.shape == (R, 2)
    assert np.max(np.abs(got[:, 0] - errs)) <= 1e-11 and np.max(np.abs(got[:, 1] - errs3)) <= 1e-11
    fes = np.loadtxt(os.path.join(tmp_path, "FEs"))
    FE, FEs, ph, alf = want["final"]
    _say(f"uniform string: FEs {fes[:, 1]} closed form {FEs}")
    assert np.max(np.abs(fes[:, 1] - FEs)) <= 2e-10 and np.max(np.abs(fes[:, 2] - ph)) <= 2e-10
    assert np.max(np.abs(fes[:, 0] - alf / alf[-1])) <= 1e-6
    for i in range(1, R - 1):   # interior rows written at the end; uniform fields
        w = np.loadtxt(os.path.join(tmp_path, f"win{i}"))
        assert np.max(np.abs(w[:, 0] - want["a"][i])) <= 1e-6 and np.max(np.abs(w[:, 1] - want["b"][i])) <= 1e-6


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


# input.dat:1-11, value for value and in the same layout (tabs, E notation, "readwin=1"),
# with comments of this file's own.
SHIPPED_INPUT = (
    "chi = \t30.00000 \t#chi N\n"
    "f = \t0.80000 \t#A fraction of the copolymer\n"
    "ens =\t2\t\t#1 canonical, 2 grand canonical\n"
    "phidb = 1.16000000E-05 \t#fugacity (GC) or volume fraction (C)\n"
    "N = \t60 \t\t#contour steps\n"
    "m = \t200 150 150 \t#grid points\n"
    "D = \t16.00000 12.00000 12.00000 #box in units of R0\n"
    "fix = \t0 1 \t\t#end replicas fixed\n"
    "readwin=1 \t\t#read win files\n"
    "justFE =1\t\t#free energy only\n"
    "dostring=1\t\t#string method\n"
)


def test_shipped_input_dat(tmp_path):
    """The shipped input.dat with no win files: every replica starts from zero fields, W+ is
    relaxed (solve_field0) and the free energy computed (justFE = 1). Two runtime keys bound
    the run: strn = 2 (the reference's strn is the compile-time 48, scft.cu:35) and
    prerelax = 30. Zero fields stay uniform, so the result does not depend on the grid: the
    oracle runs the same sequence on a 5 x 3 x 3 grid of the same box."""
    R, prerelax = 2, 30
    with open(os.path.join(tmp_path, "input.dat"), "w") as fh:
        fh.write(SHIPPED_INPUT + f"strn = {R}\nprerelax = {prerelax}\n")
    _say("driver: shipped input.dat, 200 x 150 x 150, FP64")
    log = _run(tmp_path, timeout=600)
    assert ("chiN=30 f=0.8 N=60 alpha=0.1 ens=2 phidb=1.16e-05 m=200 150 150 D=16 12 12 strn=2 fix=0 1 "
            "dostring=1 justFE=1 readwin=1 precision=fp64") in log, log[:2000]
    assert "win0 is too small" in log and "win1 is too small" in log
    assert "Solving fields" not in log   # justFE = 1: no string iterations
    got = np.loadtxt(os.path.join(tmp_path, "FEs"))
    # identical replicas have zero arc length, so the first column is 0/0 as in scft.cu:846
    assert got.shape == (R, 3) and np.all(np.isfinite(got[:, 1:])) and np.all(np.isnan(got[:, 0]))

    m, D = (5, 3, 3), (16.0, 12.0, 12.0)
    g = orc.Grid(m, D)
    W0 = np.zeros((R, 2, int(np.prod(m))))
    st = orc.StringState(g, dict(chi=30.0, f=0.8, alpha=0.1, phidb=1.16e-5, N=60, ens=2, dostring=1, fix=(0, 1)), W0)
    for ii in range(R):
        g.solve_field0(st.W[ii], st.DEV[ii], 30.0, 0.8, 60, 0.1, 1.16e-5, 2, prerelax, 1e-4)
    fe = st.free_energy()
    ref = np.stack([fe["alf"], fe["FEs"], fe["phctot"]], axis=1)
    _say(f"shipped input.dat: FEs {got[:, 1]} oracle {ref[:, 1]}, phct {got[:, 2]} oracle {ref[:, 2]}")
    assert np.max(np.abs(got[:, 1:] - ref[:, 1:])) <= 2e-10
