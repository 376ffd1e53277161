"""Cross-check of the hand-written FFT propagator against rocFFT.

torch.fft on a ROCm build of PyTorch runs on rocFFT (through hipFFT). Here it drives an
independent FP64 restatement of props on the GPU: the Richardson step of density.h:129-220
(GPU_kernels.h:199-203) with R2C/C2R transforms, the contour loop and Q of density.h:344-386,
the ensembles of density.h:388-399 and the Simpson sums of density.h:401-430. It shares no
code with the HIP kernels (which pack q and q-dagger into one complex field and run their
own mixed-radix transforms) nor with the C oracle (its own FFT). rocFFT is never on the
product path; it is the cross-check only.

Bars: FP64 HIP against the rocFFT restatement at 1e-11 relative (both FP64; the parity bar
against the oracle is 1e-10). The FP32 HIP path reports its error against the same reference.
"""
import math
import sys

import numpy as np
import pytest

from scft_string_amd import Context

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _fields(m, D, seed):
    rng = np.random.default_rng(seed)
    M = int(np.prod(m))
    x = (np.arange(m[0]) * D[0] / m[0])[:, None, None] * np.ones(m)
    y = (np.arange(m[1]) * D[1] / m[1])[None, :, None] * np.ones(m)
    Wm = (1.5 * np.cos(2 * np.pi * x / D[0]) + 0.7 * np.cos(2 * np.pi * y / D[1])).reshape(-1)
    Wm += 0.2 * rng.standard_normal(M)
    Wp = 0.1 * rng.standard_normal(M) + 0.4
    return np.stack([Wm, Wp])


def _simpson(n0, n, length):
    """weights 4, 2, 4, ... for interior points n0 < s < n of a Simpson sum (density.h:405-422)"""
    w = np.zeros(length)
    wt = 4
    for s in range(n0 + 1, n):
        w[s] = wt
        wt = 6 - wt
    return w


def rocfft_props(W, m, D, N, f, alpha, phic, ens):
    dev = torch.device("cuda")
    M = int(np.prod(m))
    n1 = int(math.floor(N * f + 0.5))  # C round() for positive arguments
    n2 = N - n1
    Nh = int(alpha * N)
    ds = 1.0 / N
    k = [torch.fft.fftfreq(m[0], d=1.0 / m[0], dtype=torch.float64, device=dev),
         torch.fft.fftfreq(m[1], d=1.0 / m[1], dtype=torch.float64, device=dev),
         torch.fft.rfftfreq(m[2], d=1.0 / m[2], dtype=torch.float64, device=dev)]
    ksq = 4 * math.pi ** 2 * ((k[0][:, None, None] / D[0]) ** 2 + (k[1][None, :, None] / D[1]) ** 2
                              + (k[2][None, None, :] / D[2]) ** 2)
    k12 = torch.exp(-ksq * ds / 12)  # irfftn normalises by 1/M itself
    k6 = torch.exp(-ksq * ds / 6)
    Wt = torch.as_tensor(W, dtype=torch.float64, device=dev).reshape(2, *m)
    e4A = torch.exp(-(Wt[0] + Wt[1]) * ds / 4)
    e4B = torch.exp(-(-Wt[0] + Wt[1]) * ds / 4)
    e2A, e2B = e4A * e4A, e4B * e4B

    def sandwich(q, kf):
        return torch.fft.irfftn(torch.fft.rfftn(q) * kf, s=m)

    def step(q, e4, e2):
        h = sandwich(sandwich(q * e4, k12) * e2, k12)
        full = sandwich(q * e2, k6)
        return (h * e4 * 4 - full * e2) / 3

    one = torch.ones(m, dtype=torch.float64, device=dev)
    q1, q2 = [one], [one]
    for s in range(N):
        q1.append(step(q1[s], *((e4A, e2A) if s < n1 else (e4B, e2B))))
        q2.append(step(q2[s], *((e4B, e2B) if s < n2 else (e4A, e2A))))
    Q2 = q2[Nh].mean().item()
    QC = q1[N].mean().item()
    if ens == 1:
        phihB = 1.0 - phic
        lnQ = phic * math.log(QC) + (phihB / alpha) * math.log(Q2)
        zB, zC = phihB / (Q2 * alpha), phic / QC
    else:
        zB, zC = 1.0, phic
        lnQ = zB * Q2 + zC * QC
    Q1s = torch.stack(q1)
    Q2s = torch.stack(q2)

    def wsum(w, X):  # sum_s w[s] X[s]
        return (torch.as_tensor(w, device=dev)[:, None, None, None] * X).sum(0)

    hB = 2 * q2[Nh]
    if Nh > 1:
        hB = hB + wsum(_simpson(0, Nh, Nh + 1)[1:Nh], Q2s[1:Nh] * torch.flip(Q2s[1:Nh], dims=[0]))
    pair = Q1s * torch.flip(Q2s, dims=[0])  # q1[s] q2[N - s]
    mid = q1[n1] * q2[n2]
    cA = q2[N] + mid + wsum(_simpson(0, n1, N + 1), pair)
    cB = q1[N] + mid + wsum(_simpson(n1, N, N + 1), pair)
    hB = hB * zB * ds / 3
    cA = cA * zC * ds / 3
    cB = cB * zC * ds / 3
    phi = torch.stack([cA, hB + cB, hB, cB]).reshape(4, M)
    return phi.cpu().numpy(), lnQ


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


CASES = [
    # m, D, N, f, alpha, ens, phidb, kernel family
    ((32, 32, 32), (4.0, 4.0, 4.0), 40, 0.45, 0.1, 1, 0.9, "two-pass"),
    ((64, 64, 64), (6.0, 6.0, 6.0), 128, 0.45, 0.1, 1, 0.9, "two-pass"),
    ((20, 15, 15), (2.0, 1.5, 1.5), 30, 0.8, 0.2, 2, 0.3, "pencil-split"),
    ((12, 8, 16), (3.0, 2.5, 4.0), 25, 0.4, 0.1, 1, 0.7, None),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[0])) + f"_N{c[2]}_ens{c[5]}")
def test_props_against_rocfft(case):
    m, D, N, f, alpha, ens, phidb, fam = case
    _say("rocFFT cross-check", case)
    W = _fields(m, D, seed=3)
    ref_phi, ref_lnQ = rocfft_props(W, m, D, N, f, alpha, phidb, ens)
    with Context(m, D, N, f, alpha, ens, phidb, chi=15.0, precision="fp64") as ctx:
        if fam:
            assert ctx.fft_path == fam
        phi, lnQ, _ = ctx.props_batch(W[None])
    with Context(m, D, N, f, alpha, ens, phidb, chi=15.0, precision="fp32") as ctx:
        phi32, lnQ32, _ = ctx.props_batch(W[None])
    d64 = max(_rel(phi[0, c], ref_phi[c]) for c in range(4))
    d32 = max(_rel(phi32[0, c], ref_phi[c]) for c in range(4))
    _say(f"  FP64 vs rocFFT: dlnQ {abs(lnQ[0] - ref_lnQ):.2e}, rel dphi {d64:.2e}; "
         f"FP32: dlnQ {abs(lnQ32[0] - ref_lnQ):.2e}, rel dphi {d32:.2e}")
    assert abs(lnQ[0] - ref_lnQ) <= 1e-11 * max(1.0, abs(ref_lnQ))
    assert d64 <= 1e-11
    assert abs(lnQ32[0] - ref_lnQ) <= 1e-4 * max(1.0, abs(ref_lnQ)) and d32 <= 1e-4
