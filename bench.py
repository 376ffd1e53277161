"""Benchmark: string iterations per second of SCFT_string's solve_field path on one MI355X.

Workload (the string-method configuration the metric is quoted on): 16 replicas on one GPU,
64^3 grid, N = 128 contour steps, D = 6.0 R0, chiN = 18, f = 0.45, phidb = 0.9, canonical
ensemble, both ends fixed, propagator and FFTs in FP32 with Q, free energy, norms, mixing and
the spline in FP64. End replicas: lamellar W- = 0.5 cos(2 pi x / D) and a cylinder-like
W- = 0.5 [cos X + cos(X/2 + sqrt3 Y/2) + cos(X/2 - sqrt3 Y/2)]/3; interior replicas are linear
interpolants plus N(0, 0.01^2) noise (seed 2024); W+ = 0. A step is one solve_field iteration
(scft.cu:592-806): W+ relaxation (<= 19 props) and one props per active replica, DEV, errors,
perpendicular projection, mixing, arc length, spline redistribution, with FreeE at k % 50 == 0.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-autotune]

The propagator schedule (kernel family, column tile, plane workgroup size, replica-group
streams) is the autotuner's choice for this string (scft_autotune, run before the timed
region); --no-autotune keeps the defaults.

Other modes, each printing one JSON line of its own:

    python bench.py --adjoint            box-gradient cost, full and checkpointed q store
    python bench.py --adjoint-box        box-gradient cost on the shipped input.dat box (line kernels)
    python bench.py --uneven             uneven batch: 128^3 string + 12 solves, concurrent
    python bench.py --trajectory ITERS   independent FP32 and FP64 runs of this string
                    [--trajectory-control]  (FP64 against FP64 from fields perturbed by 1e-12)
    python bench.py --pmc-probe          a 512 MiB copy, then one props: the rocprofv3 --pmc run
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from scft_string_amd.timing import HBM_PEAK_GBS, algorithmic_bytes_per_step  # noqa: E402

CFG = dict(m=(64, 64, 64), D=(6.0, 6.0, 6.0), N=128, chi=18.0, f=0.45, phidb=0.9, ens=1, alpha=0.1,
           nrep=16, fix=(1, 1))


def initial_fields(m, D, R, seed=2024):
    """End replicas: lamellar and cylinder-like W-; interior: linear interpolants + noise."""
    M = int(np.prod(m))
    xs = np.arange(m[0]) * D[0] / m[0]
    ys = np.arange(m[1]) * D[1] / m[1]
    X = 2 * np.pi * xs[:, None, None] / D[0] * np.ones(m)
    Y = 2 * np.pi * ys[None, :, None] / D[1] * np.ones(m)
    lam = 0.5 * np.cos(X)
    cyl = 0.5 * (np.cos(X) + np.cos(X / 2 + np.sqrt(3) * Y / 2) + np.cos(X / 2 - np.sqrt(3) * Y / 2)) / 3
    rng = np.random.default_rng(seed)
    W = np.zeros((R, 2, M))
    for i in range(R):
        t = i / (R - 1)
        W[i, 0] = ((1 - t) * lam + t * cyl).reshape(-1)
        if 0 < i < R - 1:
            W[i, 0] += 0.01 * rng.standard_normal(M)
    return W


def cpu_baseline(m, D, N, R_active, props_per_iter, budget_s=20.0):
    """Times the FP64 oracle (C, OpenMP) on a bounded sample: one props of one replica at the
    full grid and contour length; scaled to string iterations/s of R_active * props_per_iter.
    Also one props on a single thread, reported beside it."""
    from oracle import orc

    ncpu = os.cpu_count() or 1
    threads = min(ncpu, 16)   # the CPU share of one GPU on the measurement hosts
    orc.set_threads(threads)
    g = orc.Grid(m, D)
    W = initial_fields(m, D, 3)[1]
    t0 = time.perf_counter()
    n = 0
    while True:
        g.props(W, N=N, alpha=CFG["alpha"], phic=CFG["phidb"], f=CFG["f"], ens=CFG["ens"])
        n += 1
        if time.perf_counter() - t0 > budget_s * 0.5 or n >= 40:
            break
    dt = (time.perf_counter() - t0) / n
    orc.set_threads(1)
    t1 = time.perf_counter()
    g.props(W, N=N, alpha=CFG["alpha"], phic=CFG["phidb"], f=CFG["f"], ens=CFG["ens"])
    dt1 = time.perf_counter() - t1
    orc.set_threads(threads)
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            model = next((l.split(":", 1)[1].strip() for l in fh if l.startswith("model name")), "")
    except OSError:
        pass
    return {"value": 1.0 / (dt * R_active * props_per_iter), "unit": "string iters/s", "cores": threads,
            "kind": "port", "host": {"nproc": ncpu, "cpu": model, "threads_used": threads},
            "sample": f"{n} FP64 props of one replica at {m[0]}^3, N={N} ({dt:.2f} s each), scaled to "
                      f"{R_active} replicas x {props_per_iter} props per iteration",
            "single_thread": {"value": 1.0 / (dt1 * R_active * props_per_iter), "cores": 1,
                              "sample": f"1 FP64 props ({dt1:.2f} s)"}}


def free_energies(W, phi, lnQ, chi):
    """FEs of every row from props outputs (scft.cu:820-829), in FP64."""
    R, M = phi.shape[0], phi.shape[-1]
    W = W.reshape(R, 2, M)
    wa, wb = W[:, 0] + W[:, 1], W[:, 1] - W[:, 0]
    pa, pb = phi[:, 0], phi[:, 1]
    return -lnQ + np.mean(chi * pa * pb - wa * pa - wb * pb, axis=1)


def precision_report(W, Wshape_m, c, oracle_rows=(0,)):
    """FP32 vs FP64 GPU props on the identical (evolved) string state, plus the FP64 oracle
    on a few rows: |dF| per replica and max |d phiA|."""
    from scft_string_amd import Context

    R = W.shape[0]
    out = {}
    res = {}
    fe = {}
    for prec in ("fp32", "fp64"):
        with Context(c["m"], c["D"], c["N"], c["f"], c["alpha"], c["ens"], c["phidb"], c["chi"], nrep=R,
                     precision=prec, fix=c["fix"]) as ctx:
            phi, lnQ, _ = ctx.props_batch(W)
            ctx.set_fields(W)
            fe[prec] = ctx.free_energy()["FEs"]  # FreeE: W+ relaxation + props (F in FP64)
        res[prec] = (phi, lnQ, free_energies(W, phi, lnQ, c["chi"]))
    # the same FreeE with every props in FP32 (fe_fp64 = 0): Q, F and the sums in FP64
    with Context(c["m"], c["D"], c["N"], c["f"], c["alpha"], c["ens"], c["phidb"], c["chi"], nrep=R,
                 precision="fp32", fix=c["fix"], fe_fp64=0) as ctx:
        ctx.set_fields(W)
        fe["fp32_all"] = ctx.free_energy()["FEs"]
    # FreeE of the FP32 context (FP32 W+ relaxation, F from an FP64 props) vs the FP64 context
    dFE = np.abs(fe["fp32"] - fe["fp64"])
    rel = (fe["fp32"] - fe["fp32"][0]) - (fe["fp64"] - fe["fp64"][0])
    out["freeE_fp32_vs_fp64_max_abs_dF"] = float(dFE.max())
    out["freeE_fp32_vs_fp64_max_abs_dF_between_replicas"] = float(np.abs(rel).max())
    rel32 = (fe["fp32_all"] - fe["fp32_all"][0]) - (fe["fp64"] - fe["fp64"][0])
    out["freeE_all_fp32_vs_fp64_max_abs_dF"] = float(np.abs(fe["fp32_all"] - fe["fp64"]).max())
    out["freeE_all_fp32_vs_fp64_max_abs_dF_between_replicas"] = float(np.abs(rel32).max())
    out["FEs_fp64"] = [float(v) for v in fe["fp64"]]
    # one FP32 props against one FP64 props on identical fields (first order in FP32 roundoff)
    dF = np.abs(res["fp32"][2] - res["fp64"][2])
    out["props_fp32_vs_fp64_max_abs_dF"] = float(dF.max())
    out["props_fp32_vs_fp64_max_abs_dphiA"] = float(np.max(np.abs(res["fp32"][0][:, 0] - res["fp64"][0][:, 0])))
    try:
        from oracle import orc

        g = orc.Grid(c["m"], c["D"])
        d = []
        for i in oracle_rows:
            o = g.props(W[i].reshape(2, -1), N=c["N"], alpha=c["alpha"], phic=c["phidb"], f=c["f"], ens=c["ens"])
            Fo = free_energies(W[i:i + 1], o["phi"][None], np.array([o["lnQ"]]), c["chi"])[0]
            d.append(abs(res["fp64"][2][i] - Fo))
            d32 = abs(res["fp32"][2][i] - Fo)
            out.setdefault("fp32_vs_oracle_abs_dF", []).append(float(d32))
        out["fp64_vs_oracle_abs_dF"] = [float(v) for v in d]
        out["oracle_rows"] = list(oracle_rows)
    except Exception as e:  # the oracle library is test infrastructure; report if absent
        out["oracle"] = f"unavailable: {e}"
    return out


def trajectory_report(iters, control=False):
    """The benchmark string run independently in FP32 and in FP64 from the same seeded fields
    for `iters` string iterations (FreeE at k % 50 == 0 and at the end, as the reference's
    driver calls it, scft.cu:766-779, 1132), then per-replica |dF| of the final FreeE, max
    |d phiA| and the last iteration's error norms of the two runs. With `control`, the first
    run is FP64 too, from the same fields plus 1e-12 N(0, 1) on the interior replicas' W-:
    how far the iteration itself carries a perturbation far below FP32 roundoff."""
    from scft_string_amd import Context

    c = CFG
    R = c["nrep"]
    res = {}
    W0 = initial_fields(c["m"], c["D"], R)
    for run, prec in (("fp32", "fp64" if control else "fp32"), ("fp64", "fp64")):
        W = W0.copy()
        if control and run == "fp32":
            W[1:-1, 0] += 1e-12 * np.random.default_rng(7).standard_normal(W[1:-1, 0].shape)
        with Context(c["m"], c["D"], c["N"], c["f"], c["alpha"], c["ens"], c["phidb"], c["chi"], nrep=R,
                     precision=prec, fix=c["fix"]) as ctx:
            ctx.set_fields(W)
            t0 = time.time()
            rep = None
            for k in range(iters):
                rep = ctx.string_iterate(k)
                if k % 50 == 0:
                    ctx.free_energy()
                if k % 20 == 0:
                    print(f"trajectory {run} ({prec}): iteration {k}", file=sys.stderr, flush=True)
            fe = ctx.free_energy()["FEs"]
            phi = ctx.get_densities()
            res[run] = (fe, phi[:, 0].copy(), np.asarray(rep["errs3"]), time.time() - t0)
    d = np.abs(res["fp32"][0] - res["fp64"][0])
    rel = (res["fp32"][0] - res["fp32"][0][0]) - (res["fp64"][0] - res["fp64"][0][0])
    what = ("FP64 and FP64 from fields perturbed by 1e-12" if control else "independent FP32 and FP64 runs")
    return {"mode": "trajectory", "workload": f"benchmark string, 16 x 64^3, N=128, {what}",
            "iterations": iters,
            "max_abs_dF_per_replica": float(d.max()), "max_abs_dF_between_replicas": float(np.abs(rel).max()),
            "abs_dF": [float(v) for v in d], "max_abs_dphiA": float(np.max(np.abs(res["fp32"][1] - res["fp64"][1]))),
            "errs3_first_run_last": float(res["fp32"][2].max()), "errs3_fp64_last": float(res["fp64"][2].max()),
            "FEs_fp64": [float(v) for v in res["fp64"][0]],
            "wall_s": {"fp32": round(res["fp32"][3], 1), "fp64": round(res["fp64"][3], 1)}}


def plain_fp64_rate(W, c, props_per_iter, reps=2):
    """The plain per-replica FP64 GPU path (hipFFT, unfused, blocking host copies per props;
    scft_string_amd/plain.py): time `reps` props of one replica and scale to the string
    iteration's props count. Host-side field-update loops are not counted (generous to it)."""
    import torch

    from scft_string_amd.plain import PlainFP64

    p = PlainFP64(c["m"], c["D"], c["N"])
    Wr = W[1].reshape(2, -1)
    p.props(Wr, c["N"], c["alpha"], c["phidb"], c["f"], c["ens"])  # warm-up (plans, first touch)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        p.props(Wr, c["N"], c["alpha"], c["phidb"], c["f"], c["ens"])
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    p.close()
    return {"ms_per_props": dt * 1e3, "iters_per_s": 1.0 / (dt * props_per_iter),
            "sample": f"{reps} props of one replica at {c['m'][0]}^3, N={c['N']}, scaled to {props_per_iter:.0f} "
                      "props per string iteration"}


def pmc_probe(schedule=None):
    """Workload for rocprofv3 --pmc runs: a 512 MiB device copy of known size (the scale for
    FETCH_SIZE / WRITE_SIZE) followed by one batched props of the benchmark string, on the
    given propagator schedule (None: the defaults)."""
    import torch

    from scft_string_amd import Context

    n = 1 << 27
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    c = CFG
    R = c["nrep"]
    ctx = Context(c["m"], c["D"], c["N"], c["f"], c["alpha"], c["ens"], c["phidb"], c["chi"], nrep=R,
                  precision="fp32", fix=c["fix"], dostring=1, schedule=schedule)
    ctx.set_fields(initial_fields(c["m"], c["D"], R))
    ctx.time_props(R - sum(c["fix"]), reps=1)
    ctx.close()
    print("pmc probe done", flush=True)


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "pmc_traffic_contour_step_14x64cube_N128_fp32.json")
# the contour loop's kernels on any schedule: column / plane (two-stage, two-pass and Stockham)
# and line passes; the densities, reductions and factors are not part of a contour step
CONTOUR_KERNEL = re.compile(r"scft::k_(column2?|plane_(?:mid|end)2?p?|splane2|scol2|p[xyz])<float")


def step_kernels(engine):
    """The launches of one contour step on the engine that ran it (its family leads the
    engine description)."""
    for fam, ks in (("two-pass", "k_splane2 plane pass, k_scol2 column pass"),
                    ("two-stage", "k_column2, k_plane_mid2, k_column2, k_plane_end2"),
                    ("pencil-split", "k_ry/k_ly, k_rz/k_lz, k_ry/k_ly, k_scol2/k_lx"),
                    ("pencil", "k_pz, k_py, k_px")):
        if engine.startswith(fam):
            return ks
    return engine


def measure_traffic(N, schedule=None, timeout=150):
    """Memory traffic of one contour step of the benchmark batch, measured now: two child
    runs of `rocprofv3 --pmc FETCH_SIZE` and `--pmc WRITE_SIZE` over `bench.py --pmc-probe`
    (a 512 MiB copy, then one props of the 14 active replicas). Each counter is scaled by what
    it reads for the copy (512 MiB read and written), and the contour kernels' dispatches
    are summed and divided by N. Returns (bytes per step, details) or (None, reason)."""
    import csv
    import shutil
    import subprocess
    import tempfile

    exe = shutil.which("rocprofv3")
    if not exe:
        return None, "rocprofv3 not found"
    tot = {}
    with tempfile.TemporaryDirectory(dir="/tmp") as td:
        for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
            out = os.path.join(td, ctr)
            cmd = [exe, "--pmc", ctr, "--output-format", "csv", "-d", out, "-o", "pmc", "--", sys.executable,
                   os.path.join(ROOT, "bench.py"), "--pmc-probe", "--schedule", json.dumps(schedule or {})]
            try:
                r = subprocess.run(cmd, cwd="/tmp", env=dict(os.environ, TMPDIR="/tmp"), capture_output=True,
                                   text=True, timeout=timeout)
            except subprocess.TimeoutExpired:
                return None, f"rocprofv3 --pmc {ctr} timed out"
            files = [os.path.join(dp, f) for dp, _, fs in os.walk(out) for f in fs if f.endswith("counter_collection.csv")]
            if r.returncode != 0 or not files:
                return None, f"rocprofv3 --pmc {ctr} failed (rc {r.returncode})"
            copy_kib, kern_kib, disp = None, 0.0, 0
            for row in csv.DictReader(open(files[0])):
                if row["Counter_Name"] != ctr:
                    continue
                name, v = row["Kernel_Name"], float(row["Counter_Value"])
                if name.startswith("__amd_rocclr_copyBuffer") and int(row["Grid_Size"]) >= 1 << 16:
                    copy_kib = v
                elif CONTOUR_KERNEL.search(name):
                    kern_kib += v
                    disp += 1
            if not copy_kib:
                return None, f"no copy kernel in the {ctr} run"
            if disp < 2 * N:  # at least the two launches of every contour step
                return None, f"{disp} contour-kernel dispatches in the {ctr} run (expected at least 2 N = {2 * N})"
            scale = (512 << 20) / (copy_kib * 1024.0)
            tot[ctr] = (kern_kib * 1024.0 * scale / N, scale, disp)
    byte = tot["FETCH_SIZE"][0] + tot["WRITE_SIZE"][0]
    return byte, {"source": "measured in this run: rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE over bench.py --pmc-probe",
                  "read_bytes_per_step": tot["FETCH_SIZE"][0], "write_bytes_per_step": tot["WRITE_SIZE"][0],
                  "scale_fetch": tot["FETCH_SIZE"][1], "scale_write": tot["WRITE_SIZE"][1],
                  "dispatches": tot["FETCH_SIZE"][2],
                  "note": "L2-to-memory bytes; hits in the 256 MB Infinity Cache beyond L2 are not separated by "
                          "these counters (TCC_EA0_*_DRAM_32B read the same)"}


def batched_fp64_rate(W0, warmup, steps):
    """The same string with an FP64 propagator (every replica batched on the GPU): string
    iterations per second over `steps` iterations after `warmup`, and the contour step."""
    from scft_string_amd import Context

    c = CFG
    R = c["nrep"]
    ctx = Context(c["m"], c["D"], c["N"], c["f"], c["alpha"], c["ens"], c["phidb"], c["chi"], nrep=R,
                  precision="fp64", fix=c["fix"], dostring=1)
    ctx.set_fields(W0)
    for k in range(warmup):
        ctx.string_iterate(k)
    ctx.time_props(1, reps=1)
    t0 = time.perf_counter()
    for k in range(warmup, warmup + steps):
        ctx.string_iterate(k)
    dt = (time.perf_counter() - t0) / steps
    n_act = R - sum(c["fix"])
    ms_step = ctx.time_props(n_act, reps=2)
    ctx.close()
    bstep = algorithmic_bytes_per_step(c["m"], 8) * n_act
    return {"iters_per_s": 1.0 / dt, "ms_per_contour_step": ms_step,
            "frac": bstep / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, "sample": f"{steps} iterations after {warmup}"}


def adjoint_report():
    """Box derivative d lnQ/dD (adjoint of props) at 64^3, N = 100, chiN = 20, f = 0.5, FP32,
    on a 1-replica and an 8-replica string (lamellar -> cylinder interpolants, seed 2024),
    with D swept over 5.6 ... 6.4, with the full q store and with checkpoints every 10
    contour steps. Reports the mean wall time of the two host-buffer calls, the q-store size,
    and the largest difference of the checkpointed gradient from the full-store one:
    props_batch copies W in, runs props and copies phi, lnQ and phis out; dlnq_dbox copies W
    in, runs props and the gradient and copies lnQ and the gradient out. The copies differ
    (phi is 4M doubles per replica, the gradient 9 numbers), so the difference of the two is
    not the gradient's cost; its device time against props' is read from a kernel trace of
    this same run."""
    from scft_string_amd import Context

    m, N = (64, 64, 64), 100
    M = m[0] * m[1] * m[2]
    out = {"mode": "adjoint", "workload": "box gradient d lnQ/dD, 64^3, N=100, chiN=20, f=0.5, FP32",
           "D_sweep": [5.6 + 0.1 * i for i in range(9)]}
    for R in (1, 8):
        W = initial_fields(m, (6.0,) * 3, max(R, 2))[:R]
        grads = {}
        for ck in (0, 10):
            tp = tg = 0.0
            dmax = 0.0
            grads[ck] = []
            with Context(m, (5.6,) * 3, N, 0.5, 0.1, 1, 0.8, 20.0, nrep=R, precision="fp32", dostring=0,
                         ckpt=ck) as ctx:
                ctx.props_batch(W)
                ctx.dlnq_dbox(W)
                for D in out["D_sweep"]:
                    ctx.set_box((D, D, D))
                    t = time.perf_counter()
                    ctx.props_batch(W)
                    tp += time.perf_counter() - t
                    t = time.perf_counter()
                    _, d = ctx.dlnq_dbox(W)
                    tg += time.perf_counter() - t
                    dmax = max(dmax, float(np.max(np.abs(d))))
                    grads[ck].append(d)
            n = len(out["D_sweep"])
            entries = N + 1 if ck == 0 else N // ck + 3 + 2 * (ck + 1)
            key = f"replicas_{R}" + ("" if ck == 0 else f"_ckpt{ck}")
            out[key] = {"props_batch_call_ms": round(tp / n * 1e3, 2), "dlnq_dbox_call_ms": round(tg / n * 1e3, 2),
                        "max_abs_dlnQ_dD": dmax, "q_store_GB": round(entries * M * 8 * R / 1e9, 3)}
        g0, g1 = np.array(grads[0]), np.array(grads[10])
        out[f"replicas_{R}_ckpt10"]["max_rel_diff_vs_full_store"] = float(np.max(np.abs(g1 - g0)) / np.max(np.abs(g0)))
    print(json.dumps(out), flush=True)


def adjoint_box_report():
    """Box derivative d lnQ/dD on the reference's shipped box (input.dat: m = 200 x 150 x 150,
    D = 16 x 12 x 12, N = 60, f = 0.8, chiN = 30, ens = 2, phidb = 1.16e-5), which runs the split
    line step and the line-kernel gradient: FP64 with 1 replica and FP32 with 4, lamellar fields
    along x plus 0.1 N(0, 1) (seed 3). Mean wall time of 3 calls each of props_batch (W in,
    props, phi out) and dlnq_dbox (W in, props, gradient out); their device times are read from
    a kernel trace of this run."""
    from scft_string_amd import Context

    m, D = (200, 150, 150), (16.0, 12.0, 12.0)
    M = m[0] * m[1] * m[2]
    rng = np.random.default_rng(3)
    x = np.meshgrid(*[np.arange(m[i]) * D[i] / m[i] for i in range(3)], indexing="ij")[0]
    out = {"mode": "adjoint-box", "workload": "box gradient d lnQ/dD, 200x150x150 (input.dat), N=60, ens=2"}
    for prec, R in (("fp64", 1), ("fp32", 4)):
        W = np.zeros((R, 2, M))
        W[:, 0] = np.cos(2 * np.pi * x / D[0]).reshape(-1) + 0.1 * rng.standard_normal(M)
        with Context(m, D, 60, 0.8, 0.1, 2, 1.16e-5, 30.0, nrep=R, precision=prec, dostring=0) as ctx:
            ctx.props_batch(W)
            ctx.dlnq_dbox(W)
            tp = tg = 0.0
            for _ in range(3):
                t = time.perf_counter()
                ctx.props_batch(W)
                tp += time.perf_counter() - t
                t = time.perf_counter()
                _, d = ctx.dlnq_dbox(W)
                tg += time.perf_counter() - t
            out[f"{prec}_replicas_{R}"] = {"path": ctx.fft_path, "props_batch_call_ms": round(tp / 3 * 1e3, 1),
                                           "dlnq_dbox_call_ms": round(tg / 3 * 1e3, 1),
                                           "max_abs_dlnQ_dD": float(np.max(np.abs(d)))}
    print(json.dumps(out), flush=True)


C4 = dict(chi=22.0, f=0.4, phidb=0.85, ens=2, alpha=0.1)


def uneven_solves():
    """The 12 independent solves of the uneven batch, {(m, N): count}: m in {48^3, 64^3, 96^3}
    and N in {64, 100, 128} drawn with seed 4."""
    rng0 = np.random.default_rng(4)
    solves = {}
    for _ in range(12):
        m = [(48,) * 3, (64,) * 3, (96,) * 3][rng0.integers(3)]
        N = [64, 100, 128][rng0.integers(3)]
        solves[(m, N)] = solves.get((m, N), 0) + 1
    return solves


def uneven_specs(with_string, precision="fp32"):
    """Every context of the uneven batch as (Context keyword arguments, initial W [R][2][M]):
    the 24-replica 128^3, N = 160 grand-canonical string (chiN = 22, f = 0.4, phidb = 0.85,
    lamellar -> gyroid-like ends, D = 8) when with_string, then one context per distinct
    (m, N) of the 12 solves, its solves as independent replicas (dostring = 0, D = m/16)."""
    def grid(m):
        return np.meshgrid(*[2 * np.pi * np.arange(m[i]) / m[i] for i in range(3)], indexing="ij")

    rng = np.random.default_rng(7)
    specs = []
    if with_string:
        m, R = (128,) * 3, 24
        X, Y, Z = grid(m)
        lam = 0.5 * np.cos(X)
        gyr = 0.5 * (np.sin(X) * np.cos(Y) + np.sin(Y) * np.cos(Z) + np.sin(Z) * np.cos(X))
        W = np.zeros((R, 2, X.size))
        for i in range(R):
            t = i / (R - 1)
            W[i, 0] = ((1 - t) * lam + t * gyr).reshape(-1) + 0.01 * rng.standard_normal(X.size)
        specs.append((dict(m=m, D=(8.0,) * 3, N=160, nrep=R, precision=precision, fix=(1, 1), dostring=1, **C4), W))
    for (m, N), R in sorted(uneven_solves().items()):
        X = grid(m)[0]
        W = np.zeros((R, 2, X.size))
        for i in range(R):
            W[i, 0] = (0.5 * np.cos(X)).reshape(-1) + 0.01 * rng.standard_normal(X.size)
        specs.append((dict(m=m, D=(m[0] / 16.0,) * 3, N=N, nrep=R, precision=precision, fix=(0, 0), dostring=0, **C4),
                      W))
    return specs


def uneven_contexts(with_string, precision="fp32", tuned=False):
    """The uneven batch's contexts with their initial fields. tuned: each context on the
    schedule scft_autotune picks for it, timed alone on its own fields (returned too)."""
    from scft_string_amd import Context
    from scft_string_amd.capi import autotune

    ctxs, scheds = [], []
    for kw, W in uneven_specs(with_string, precision):
        kw = dict(kw)
        pos = [kw.pop(k) for k in ("m", "D", "N", "f", "alpha", "ens", "phidb", "chi")]
        sched = None
        if tuned:
            sched, _ = autotune(*pos, nrep=kw["nrep"], precision=kw["precision"], fix=kw["fix"],
                                dostring=kw["dostring"], W=W, reps=1)
            scheds.append(sched)
        c = Context(*pos, schedule=sched, **kw)
        c.set_fields(W)
        ctxs.append(c)
    return (ctxs, scheds) if tuned else ctxs


def uneven_report(K):
    """Uneven solves on one GPU: a 24-replica string at 128^3, N = 160, FP32, grand canonical
    (chiN = 22, f = 0.4, phidb = 0.85), lamellar -> gyroid-like ends, plus 12 independent
    single-replica solves with m in {48^3, 64^3, 96^3} and N in {64, 100, 128} drawn with
    seed 4 (same-(m, N) solves share a context as independent replicas). One round is one
    solve_field iteration of every context. Concurrent rounds use scft_iterate_concurrent;
    sequential rounds iterate the contexts one after another from identical fresh states."""
    import gc

    from scft_string_amd import iterate_concurrent

    solves = uneven_solves()

    tuned_scheds = {}

    def build(with_string, tuned):
        if not tuned:
            return uneven_contexts(with_string)
        ctxs, scheds = uneven_contexts(with_string, tuned=True)
        tuned_scheds[with_string] = scheds
        return ctxs

    def run(with_string, concurrent, tuned=False):
        ctxs = build(with_string, tuned)
        if concurrent:
            iterate_concurrent(ctxs, [0] * len(ctxs))
        else:
            for c in ctxs:
                c.string_iterate(0)
        t = time.perf_counter()
        props = 0
        for k in range(1, K + 1):
            reps = iterate_concurrent(ctxs, [k] * len(ctxs)) if concurrent else [c.string_iterate(k) for c in ctxs]
            props += sum(int(r["steps"].sum()) for r in reps)
        dt = (time.perf_counter() - t) / K
        for c in ctxs:
            c.close()
        gc.collect()
        return dt, props // K

    out = {"mode": "uneven", "rounds": K,
           "independent_solves": {f"{m[0]}^3 N{N}": n for (m, N), n in sorted(solves.items())}}
    for with_string in (False, True):
        tc, pc = run(with_string, True)
        ts, ps = run(with_string, False)
        tt, pt = run(with_string, True, tuned=True)
        key = "string_24x128cube_N160_plus_12_solves" if with_string else "12_independent_solves"
        out[key] = {"concurrent_ms_per_round": round(tc * 1e3, 1), "sequential_ms_per_round": round(ts * 1e3, 1),
                    "props_per_round": [pc, ps], "speedup": round(ts / tc, 3),
                    "concurrent_autotuned_ms_per_round": round(tt * 1e3, 1), "autotuned_props_per_round": pt,
                    "autotuned_schedules": tuned_scheds[with_string]}
        print(f"uneven: {key} done", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the batched FP64 rate of the same string")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--pmc-probe", action="store_true")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 --pmc traffic measurement")
    ap.add_argument("--no-autotune", action="store_true",
                    help="run the default propagator schedule instead of the autotuner's choice")
    ap.add_argument("--schedule", default=None, help=argparse.SUPPRESS)  # --pmc-probe: JSON schedule
    ap.add_argument("--adjoint", action="store_true",
                    help="box-gradient timing: 64^3, N=100, FP32, D swept over 5.6-6.4, 1 and 8 replicas")
    ap.add_argument("--adjoint-box", action="store_true",
                    help="box-gradient timing on the shipped 200x150x150 box (line kernels): FP64 x 1, FP32 x 4")
    ap.add_argument("--uneven", action="store_true",
                    help="24 x 128^3 N=160 string + 12 independent solves, concurrent vs sequential")
    ap.add_argument("--trajectory", type=int, default=0, metavar="ITERS",
                    help="independent FP32 and FP64 runs of the benchmark string for ITERS iterations: |dF|, |d phiA|")
    ap.add_argument("--trajectory-control", action="store_true",
                    help="with --trajectory: FP64 against FP64 from fields perturbed by 1e-12")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: all string replicas share it; "
                                   "sharding replicas over GPUs is designed but not built",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    if a.pmc_probe:
        pmc_probe(json.loads(a.schedule) if a.schedule else None)
        return
    if a.adjoint:
        adjoint_report()
        return
    if a.adjoint_box:
        adjoint_box_report()
        return
    if a.uneven:
        uneven_report(a.steps)
        return
    if a.trajectory:
        print(json.dumps(trajectory_report(a.trajectory, a.trajectory_control)), flush=True)
        return

    import torch

    from scft_string_amd import Context
    from scft_string_amd.capi import autotune

    torch.cuda.init()
    c = CFG
    R = c["nrep"]
    W0 = initial_fields(c["m"], c["D"], R)
    sched, tuned = None, None
    if not a.no_autotune:
        # the propagator schedule: every candidate timed on this string, each in a context of
        # its own, before the benchmark's context exists (outside the timed region)
        t0 = time.perf_counter()
        sched, report = autotune(c["m"], c["D"], c["N"], c["f"], c["alpha"], c["ens"], c["phidb"], c["chi"], nrep=R,
                                 precision=a.precision, fix=c["fix"], dostring=1, W=W0, reps=2)
        tuned = {"schedule": sched, "candidates": report, "seconds": round(time.perf_counter() - t0, 2)}
        print(f"bench: autotune {report}", file=sys.stderr, flush=True)
    ctx = Context(c["m"], c["D"], c["N"], c["f"], c["alpha"], c["ens"], c["phidb"], c["chi"], nrep=R,
                  precision=a.precision, fix=c["fix"], dostring=1, schedule=sched)
    ctx.set_fields(W0)
    print(f"bench: {R} replicas {c['m']} N={c['N']} {a.precision}; warmup {a.warmup}", file=sys.stderr, flush=True)

    def iteration(k):
        rep = ctx.string_iterate(k)
        if k % 50 == 0:
            ctx.free_energy()
        return rep

    for k in range(a.warmup):
        iteration(k)
    torch.cuda.synchronize()
    print("bench: timing", a.steps, "iterations", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    props = 0
    n_act = R - sum(c["fix"])
    for k in range(a.warmup, a.warmup + a.steps):
        rep = iteration(k)
        props += int(np.sum(np.maximum(rep["steps"] - 1, 0)))  # steps = 1 + solve_field0's k
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    iters_per_s = a.steps / dt

    # roofline of the contour step (all active replicas, every launch of one step)
    ms_step = ctx.time_props(n_act, reps=2)
    p = 4 if a.precision == "fp32" else 8
    bstep = algorithmic_bytes_per_step(c["m"], p) * n_act
    achieved = bstep / (ms_step * 1e-3) / 1e9
    Wend = ctx.get_fields()
    engine = ctx.engine_info.split(";")[0] + (" (autotuned)" if tuned else "")
    ctx.close()

    traffic, traffic_info = None, None
    if a.precision == "fp32" and not a.no_traffic:
        traffic, traffic_info = measure_traffic(c["N"], sched)
        if traffic is None and os.path.exists(TRAFFIC_PROFILE):   # fall back to the committed profile
            why = traffic_info
            traffic = json.load(open(TRAFFIC_PROFILE)).get("hbm_bytes_per_contour_step")
            traffic_info = {"source": f"{os.path.relpath(TRAFFIC_PROFILE, ROOT)} (not measured in this run: {why})"}

    out = {
        "metric": "string iters/sec 16x64^3xN128 (batched FP32 propagator, FP64 field update)",
        "value": iters_per_s, "unit": "string iters/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": dt / a.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if a.precision == "fp32" else "f64", "data": "synthetic (seeded lamellar/cylinder string)",
        "config": {"workload": "string method, 16 replicas x 64^3, N=128, D=6.0, chiN=18, f=0.45, phidb=0.9, "
                               "canonical, fix=(1,1)", "replicas": R, "grid": list(c["m"]), "N": c["N"],
                   "props_per_iteration": props / a.steps, "parallelism": "1 GPU, all replicas batched",
                   "engine": engine},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": f"contour step ({step_kernels(engine)}) over {n_act} replicas",
                     "ms_per_contour_step": ms_step,
                     "algorithmic_bytes_per_contour_step": bstep, "traffic_info": traffic_info,
                     "time_base": "HIP events on the context stream around the contour loop of props calls "
                                  "(the schedule's replica-group streams overlapping); the one-stream kernel-trace "
                                  "averages are in profiles/kernel_trace_string_16x64cube_N128_fp32.md"},
    }
    if traffic:
        # the measured traffic's own rate: on the two-pass step the fused passes move fewer
        # bytes (11 complex fields per replica and step) than the 12-FFT algorithmic figure
        out["roofline"]["traffic_GBps"] = traffic / (ms_step * 1e-3) / 1e9
        out["roofline"]["traffic_frac"] = out["roofline"]["traffic_GBps"] / HBM_PEAK_GBS
    if tuned:
        out["autotune"] = tuned
    if a.precision == "fp32" and not a.no_fp64:
        out["batched_fp64"] = batched_fp64_rate(W0, a.warmup, 2)
    if not a.no_plain:
        pl = plain_fp64_rate(Wend, c, props / a.steps)
        pl["speedup_of_value"] = iters_per_s / pl["iters_per_s"]
        out["plain_fp64_per_replica"] = pl
    if not a.no_accuracy:
        out["accuracy"] = precision_report(Wend, c["m"], c, oracle_rows=(0, R // 2))
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(c["m"], c["D"], c["N"], n_act, props / a.steps / n_act,
                                           a.cpu_budget)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
