"""Per-contour-step timing of props on one grid, through the C-ABI.

    python -m scft_string_amd.timing --m 200 150 150 --D 16 12 12 --N 60 --ens 2 \\
        --phidb 1.16e-5 --f 0.8 --chi 30 --prec fp64 --nrep 1 [--pencil] [--runtime-lengths]
        [--fields lamella] [--copy-probe] [--family two-pass] [--streams 3] [--coltile 32]

prints one JSON line: ms per contour step (the mean over reps of a props of nrep replicas,
divided by N), the kernel family, and the algorithmic HBM rate against the 8 TB/s peak.
--pencil forces the line kernels on grids that also have plane/column kernels
(SCFT_PENCIL=1); --runtime-lengths forces the runtime-planned line kernels
(SCFT_PENCIL_RT=1). --fields lamella times a lamellar field state instead of W = 0: the
step time depends on the data, and zero fields run measurably faster.

    python -m scft_string_amd.timing --stream-probe [MiB ...]

prints the streaming rates of torch's own copy and add kernels (by default 256 MiB to 4 GiB;
arrays of a few tens of MiB stay in the 256 MB last-level cache), a practical ceiling to set
the contour step's measured traffic against.
"""
import argparse
import json
import os
import sys

import numpy as np

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes_per_step(m, p):
    """12 (M p + Mk 2p): 2 chains x 6 real 3-D FFTs, each reading its input and writing its
    output once, pointwise factors fused (per contour step per replica)."""
    M = m[0] * m[1] * m[2]
    Mk = m[0] * m[1] * (m[2] // 2 + 1)
    return 12 * (M * p + Mk * 2 * p)


def lamella_fields(m, D, R, amp=1.5, seed=0):
    rng = np.random.default_rng(seed)
    M = int(np.prod(m))
    x = (np.arange(m[0]) * D[0] / m[0])[:, None, None] * np.ones(m)
    W = np.zeros((R, 2, M))
    for i in range(R):
        W[i, 0] = amp * np.cos(2 * np.pi * x / D[0]).reshape(-1) + 0.1 * rng.standard_normal(M)
    return W


def stream_probe(sizes_mib=(256, 1024, 4096), iters=20):
    """Streaming rates of torch's own elementwise kernels on this GPU, as a ceiling that
    kernels on it can reach in practice: copy (one read, one write) and add (two reads, one
    write) of FP32 arrays, timed with HIP events. Prints one JSON line per kernel and size."""
    import torch

    for mib in sizes_mib:
        n = (mib << 20) // 4
        a = torch.ones(n, device="cuda")
        b = torch.full((n,), 2.0, device="cuda")
        c = torch.empty(n, device="cuda")
        for name, fn, moved in (("copy", lambda: c.copy_(a), 2 * (mib << 20)),
                                ("add", lambda: torch.add(a, b, out=c), 3 * (mib << 20))):
            for _ in range(3):
                fn()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(iters):
                fn()
            e1.record()
            torch.cuda.synchronize()
            t = e0.elapsed_time(e1) / iters * 1e-3
            print(json.dumps({"kernel": name, "MiB": mib, "us": round(t * 1e6, 1),
                              "GBps": round(moved / t / 1e9, 1)}), flush=True)
        del a, b, c


def main(argv=None):
    args = sys.argv[1:] if argv is None else argv
    if "--stream-probe" in args:
        # --stream-probe [MiB ...]: array sizes (default 256 1024 4096)
        sizes = [int(v) for v in args[args.index("--stream-probe") + 1:] if v.isdigit()]
        stream_probe(tuple(sizes) if sizes else (256, 1024, 4096))
        return
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
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--pencil", action="store_true")
    ap.add_argument("--runtime-lengths", action="store_true")
    ap.add_argument("--fields", choices=("zero", "lamella"), default="zero")
    ap.add_argument("--family", choices=("default", "two-stage", "two-pass", "stockham", "pencil", "pencil-split"), default="default",
                    help="propagator schedule: kernel family (scft_schedule.family)")
    ap.add_argument("--streams", type=int, default=0, help="schedule: replica-group streams (0: default)")
    ap.add_argument("--coltile", type=int, default=0, help="schedule: column tile (0: default)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first a 512 MiB device copy: the known-size scale for FETCH_SIZE / WRITE_SIZE")
    a = ap.parse_args(argv)
    if a.copy_probe:
        import torch

        src = torch.ones(1 << 27, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    # read by the engine at context creation
    os.environ["SCFT_PENCIL"] = "1" if a.pencil else "0"
    os.environ["SCFT_PENCIL_RT"] = "1" if a.runtime_lengths else "0"
    from .capi import Context

    sched = {"family": a.family, "streams": a.streams, "coltile": a.coltile}
    with Context(tuple(a.m), tuple(a.D), a.N, a.f, a.alpha, a.ens, a.phidb, a.chi, nrep=a.nrep,
                 precision=a.prec, schedule=sched) as ctx:
        if a.fields == "lamella":
            ctx.set_fields(lamella_fields(a.m, a.D, a.nrep))
        ms = ctx.time_props(a.nrep, a.reps)
        path = ctx.fft_path
        engine = ctx.engine_info
    p = 4 if a.prec == "fp32" else 8
    gbs = a.nrep * algorithmic_bytes_per_step(a.m, p) / (ms * 1e-3) / 1e9
    print(json.dumps({"m": a.m, "D": a.D, "N": a.N, "prec": a.prec, "nrep": a.nrep, "path": path, "engine": engine,
                      "runtime_lengths": a.runtime_lengths, "fields": a.fields,
                      "ms_per_step": round(ms, 4), "alg_GBps": round(gbs, 1),
                      "frac": round(gbs / HBM_PEAK_GBS, 3)}), flush=True)


if __name__ == "__main__":
    main()
