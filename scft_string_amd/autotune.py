"""Propagator autotuner from the command line (scft_autotune through the C-ABI).

    python -m scft_string_amd.autotune --m 64 64 64 --D 6 6 6 --N 128 --nrep 16 --prec fp32 \\
        [--fix 1 1] [--independent] [--fields string|lamella] [--reps 2]

times every schedule the grid can run (kernel family; column tile x plane workgroup size for
the two-stage kernels; replica-group streams), each in a context of its own, and prints one
JSON line: the chosen schedule and every candidate's ms per contour step. --fields string
(the default) times the seeded lamellar -> cylinder string of the benchmark scaled to the
grid; lamella lets the library build its own lamellar trial field. --independent times
independent solves (dostring = 0: every replica propagates); otherwise the string's fixed
ends are left out, as a string iteration does. Pass the schedule to Context(schedule=...) or
scft_params.sched.
"""
import argparse
import json
import time

import numpy as np


def string_fields(m, D, R, seed=2024):
    """Lamellar and cylinder-like end replicas, interior linear interpolants plus noise."""
    M = int(np.prod(m))
    X = 2 * np.pi * (np.arange(m[0]) / m[0])[:, None, None] * np.ones(m)
    Y = 2 * np.pi * (np.arange(m[1]) / m[1])[None, :, None] * np.ones(m)
    lam = 0.5 * np.cos(X)
    cyl = 0.5 * (np.cos(X) + np.cos(X / 2 + np.sqrt(3) * Y / 2) + np.cos(X / 2 - np.sqrt(3) * Y / 2)) / 3
    rng = np.random.default_rng(seed)
    W = np.zeros((R, 2, M))
    for i in range(R):
        t = i / max(1, R - 1)
        W[i, 0] = ((1 - t) * lam + t * cyl).reshape(-1)
        if 0 < i < R - 1:
            W[i, 0] += 0.01 * rng.standard_normal(M)
    return W


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--m", type=int, nargs=3, required=True)
    ap.add_argument("--D", type=float, nargs=3, required=True)
    ap.add_argument("--N", type=int, required=True)
    ap.add_argument("--f", type=float, default=0.45)
    ap.add_argument("--alpha", type=float, default=0.1)
    ap.add_argument("--ens", type=int, default=1)
    ap.add_argument("--phidb", type=float, default=0.9)
    ap.add_argument("--chi", type=float, default=18.0)
    ap.add_argument("--prec", choices=("fp32", "fp64"), default="fp32")
    ap.add_argument("--nrep", type=int, default=1)
    ap.add_argument("--fix", type=int, nargs=2, default=(1, 1))
    ap.add_argument("--independent", action="store_true")
    ap.add_argument("--fields", choices=("string", "lamella"), default="string")
    ap.add_argument("--reps", type=int, default=2)
    a = ap.parse_args(argv)
    from .capi import autotune

    W = string_fields(a.m, a.D, a.nrep) if a.fields == "string" else None
    t0 = time.perf_counter()
    best, report = autotune(a.m, a.D, a.N, a.f, a.alpha, a.ens, a.phidb, a.chi, nrep=a.nrep, precision=a.prec,
                            fix=tuple(a.fix), dostring=0 if a.independent else 1, W=W, reps=a.reps)
    print(json.dumps({"m": a.m, "N": a.N, "prec": a.prec, "nrep": a.nrep, "fields": a.fields,
                      "independent": a.independent, "schedule": best, "seconds": round(time.perf_counter() - t0, 2),
                      "candidates": report}), flush=True)


if __name__ == "__main__":
    main()
