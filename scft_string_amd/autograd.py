"""torch.autograd binding of the free-energy backward (scft_backward in include/scft_api.h).

    F = FreeEnergy.apply(ctx, D, chi)    # D: (3,) float64 tensor, chi: () float64 tensor
    F.sum().backward()                    # D.grad = sum_r dF_r/dD, chi.grad = sum_r dF_r/dchiN

forward returns the per-replica free energies FEs (nrep,) of the context's current field
state (FreeE, scft.cu:812-853). Their derivatives are those of F at a saddle point, where F
is stationary in W (the envelope theorem): the field state must be converged for the box
and chiN the context holds, so forward checks that D and chi equal them. forward calls
scft_free_energy, which like FreeE first relaxes W+ with solve_field0(100, 1e-4)
(scft.cu:817-818); at a converged saddle that changes the fields by rounding only.
"""
import numpy as np
import torch


class FreeEnergy(torch.autograd.Function):
    @staticmethod
    def forward(fctx, sctx, D, chi):
        Dh = D.detach().cpu().double().numpy().reshape(3)
        if not np.allclose(Dh, sctx.D, rtol=1e-14, atol=0):
            raise ValueError(f"FreeEnergy: D {tuple(Dh)} is not the context's box {sctx.D}; "
                             "set_box and re-converge the fields first")
        if abs(float(chi) - sctx.params.chi) > 1e-14 * abs(sctx.params.chi):
            raise ValueError(f"FreeEnergy: chi {float(chi)} is not the context's chiN {sctx.params.chi}")
        fcts = sctx.free_energy()
        fctx.sctx = sctx
        fctx.dev = D.device
        return torch.as_tensor(fcts["FEs"].copy(), dtype=torch.float64, device=D.device)

    @staticmethod
    def backward(fctx, gF):
        g = gF.detach().cpu().double().numpy()
        dD, dchi = fctx.sctx.backward(g)
        gD = torch.as_tensor(dD.sum(axis=0), dtype=torch.float64, device=fctx.dev)
        gchi = torch.as_tensor(dchi.sum(), dtype=torch.float64, device=fctx.dev)
        return None, gD, gchi
