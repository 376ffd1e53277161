"""Writes tests/golden/c0_trajectory.npz: the oracle's free-running trajectory of the
reference's CPU-sized configuration (plain SCFT, one replica, dostring = 0: m = 32^3,
D = 4.0, N = 100, f = 0.5, chiN = 20, phidb = 0.8, canonical, alpha = 0.1, 200 iterations of
solve_field, scft.cu:559-806; W- = 0.1 cos(2 pi 2x/D) + N(0, 0.01^2) from
numpy default_rng(20), W+ = 0), then FreeE (scft.cu:812-853).

Stored: the error norm and props count of every iteration, the final fields at 256 seeded
points with their norms, and the final FEs. The GPU test runs the same 200 iterations and
compares; tests/test_c0_fixture.py re-runs the first iterations on the oracle to check the
file. Run: python tests/golden/make_c0_trajectory.py  (a few minutes on 8 threads).
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import orc  # noqa: E402

C0 = dict(m=(32, 32, 32), D=(4.0, 4.0, 4.0), N=100, f=0.5, alpha=0.1, ens=1, phidb=0.8, chi=20.0, iters=200)


def c0_initial_fields():
    m, D = C0["m"], C0["D"]
    M = int(np.prod(m))
    rng = np.random.default_rng(20)
    x = (np.arange(m[0]) * D[0] / m[0])[:, None, None] * np.ones(m)
    W0 = np.zeros((1, 2, M))
    W0[0, 0] = 0.1 * np.cos(2 * np.pi * 2 * x / D[0]).reshape(-1) + 0.01 * rng.standard_normal(M)
    return W0


def sample_points(M, n=256):
    return np.sort(np.random.default_rng(99).choice(M, size=n, replace=False))


def oracle_state():
    c = C0
    return orc.StringState(orc.Grid(c["m"], c["D"]),
                           dict(chi=c["chi"], f=c["f"], alpha=c["alpha"], phidb=c["phidb"], N=c["N"], ens=c["ens"],
                                dostring=0), c0_initial_fields())


def main():
    c = C0
    M = int(np.prod(c["m"]))
    orc.set_threads(os.cpu_count() or 1)
    st = oracle_state()
    errs, steps = [], []
    for k in range(c["iters"]):
        rep = st.iterate(k)
        errs.append(rep["errs"][0])
        steps.append(rep["steps"][0])
        if k % 20 == 0:
            print(f"k={k} errs={errs[-1]:.6e} props={steps[-1]}", flush=True)
    W = st.W[0, :2 * M].copy()
    fe = st.free_energy()
    Wfe = st.W[0, :2 * M].copy()
    idx = sample_points(M)
    out = os.path.join(ROOT, "tests", "golden", "c0_trajectory.npz")
    np.savez_compressed(out, errs=np.array(errs), steps=np.array(steps, dtype=np.int32), idx=idx,
                        Wm=W[:M][idx], Wp=W[M:][idx], norm_Wm=np.linalg.norm(W[:M]), norm_Wp=np.linalg.norm(W[M:]),
                        Wm_fe=Wfe[:M][idx], Wp_fe=Wfe[M:][idx], FEs=fe["FEs"], phctot=fe["phctot"])
    print("wrote", out)


if __name__ == "__main__":
    main()
