"""scft_props_batch_device: W and phi as device arrays (torch CUDA tensors).

The fields are produced by torch kernels immediately before the call, on torch's current
stream, while the library reads them on its own stream; the wrapper orders the two. The
result must equal scft_props_batch on the same rows bit for bit (same kernels, same input).
Wrong dtype, layout, device or size is refused before anything reaches the GPU.
"""
import sys

import numpy as np
import pytest

from scft_string_amd import Context

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_device_rows_match_host_rows(prec):
    import torch

    m, D, R = (32, 32, 32), (4.0, 4.0, 4.0), 3
    M = int(np.prod(m))
    _say(f"props_batch_device {m} x {R} {prec}")
    with Context(m, D, 40, 0.45, 0.1, 1, 0.8, 15.0, nrep=R, precision=prec) as ctx:
        g = torch.Generator(device="cuda:0").manual_seed(3)
        x = torch.arange(m[0], device="cuda:0", dtype=torch.float64) * D[0] / m[0]
        lam = torch.cos(2 * np.pi * x / D[0])[:, None, None].expand(m).reshape(-1)
        W = torch.zeros(R, 2, M, device="cuda:0", dtype=torch.float64)
        for i in range(R):   # a chain of small kernels right before the call
            W[i, 0] = (1.0 + 0.5 * i) * lam + 0.2 * torch.randn(M, generator=g, device="cuda:0", dtype=torch.float64)
            W[i, 1] = 0.1 * torch.randn(M, generator=g, device="cuda:0", dtype=torch.float64)
        phi = torch.full((R, 4, M), float("nan"), device="cuda:0", dtype=torch.float64)
        lnQ_d, phis_d = ctx.props_batch_device(W, phi, R)
        phi_d = phi.cpu().numpy()
        phi_h, lnQ_h, phis_h = ctx.props_batch(W.cpu().numpy())
    assert np.array_equal(lnQ_d, lnQ_h) and np.array_equal(phis_d, phis_h)
    assert np.array_equal(phi_d, phi_h)


def test_device_arguments_are_checked():
    import torch

    m, R = (16, 16, 16), 2
    M = int(np.prod(m))
    with Context(m, (4.0, 4.0, 4.0), 20, 0.5, 0.1, 1, 0.8, 15.0, nrep=R) as ctx:
        good_W = torch.zeros(R, 2 * M, device="cuda:0", dtype=torch.float64)
        good_phi = torch.zeros(R, 4 * M, device="cuda:0", dtype=torch.float64)
        bad = [
            (good_W.float(), good_phi),                               # float32
            (good_W[:, :M], good_phi),                                # short and strided
            (good_W, good_phi[:1]),                                   # too few phi rows
            (good_W.cpu(), good_phi),                                 # host tensor
            (good_W.t().contiguous().t(), good_phi),                  # not contiguous
        ]
        for W, phi in bad:
            with pytest.raises(ValueError):
                ctx.props_batch_device(W, phi, R)
        with pytest.raises(ValueError):
            ctx.props_batch_device(good_W, good_phi, R + 1)
        lnQ, _ = ctx.props_batch_device(good_W, good_phi, R)     # still usable
        assert np.all(np.abs(lnQ) < 1e-12)                          # W = 0: lnQ = 0 (canonical)
