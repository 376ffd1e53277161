"""bench.py's command-line contract, on the CPU: more than one GPU is refused with a message
(this build runs every replica on one GPU), before anything touches a device."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_a_message():
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 2
    line = json.loads(out.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_help_lists_the_modes():
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0
    for flag in ("--steps", "--warmup", "--adjoint", "--uneven", "--trajectory", "--pmc-probe"):
        assert flag in out.stdout, flag


def test_traffic_counts_the_contour_kernels_of_every_schedule():
    """measure_traffic sums the contour loop's kernels whatever the schedule (two-stage at
    either tile or plane size, two-pass, Stockham, line passes) and none of the props' other
    kernels."""
    sys.path.insert(0, ROOT)
    import bench

    hit = ["void scft::k_column2<float, 64, 64, 64, 64>(scft::PropArgs<float>, int, int)",
           "void scft::k_column2<float, 64, 64, 64, 32>(scft::PropArgs<float>, int, int)",
           "void scft::k_plane_mid2<float, 64, 64, 0>(scft::PropArgs<float>, int)",
           "void scft::k_plane_end2<float, 64, 64, 256>(scft::PropArgs<float>, int, int)",
           "void scft::k_plane_mid2p<float, 96, 96>(scft::PropArgs<float>, int, int)",
           "void scft::k_splane2<float, 64, 64>(scft::PropArgs<float>, int)",
           "void scft::k_scol2<float, 64, 64, 64, 32>(scft::PropArgs<float>, int, int)",
           "void scft::k_column<float, 64, 64, 64, 64, 256>(scft::PropArgs<float>, int, int)",
           "void scft::k_plane_end<float, 64, 64, 256>(scft::PropArgs<float>, int, int)",
           "void scft::k_pz<float, 64>(scft::PropArgs<float>, int, int, int)",
           "void scft::k_py<float, true, 64>(scft::PropArgs<float>, int, int)",
           "void scft::k_px<float, 0>(scft::PropArgs<float>, int, int)"]
    miss = ["void scft::k_density<float>(scft::cpx<float> const*, long long, scft::ChainConst, double const*, "
            "double*, long long, int const*, double*)",
            "void scft::k_factors<float>(double const*, long long, int const*, long long, double, double const*, "
            "float*, float*)",
            "void scft::k_qsum<float>(scft::cpx<float> const*, long long, int, int, int, double*)",
            "void scft::k_plane_end2<double, 64, 64, 0>(scft::PropArgs<double>, int, int)",
            "__amd_rocclr_copyBuffer"]
    assert all(bench.CONTOUR_KERNEL.search(n) for n in hit)
    assert not any(bench.CONTOUR_KERNEL.search(n) for n in miss)
