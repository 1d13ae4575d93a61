"""Design-density sensitivities through the discrete adjoint.

`point_objective(params, lo, rho, r0, freqs, nsteps)` runs a fresh simulation
with the design block `rho` at cell `lo`, returns J (the power of the running
DFT of E at cell r0 summed over the frequencies and components) and dJ/drho
from one forward and one adjoint sweep on the GPU.

`DesignObjective` exposes the same pair to torch.autograd: the forward pass
returns J as a 0-d tensor and the backward pass scales the adjoint gradient by
the incoming gradient.  Torch is imported lazily, only for this wrapper.
"""
import numpy as np

from . import Simulation


def point_objective(params, lo, rho, r0, freqs, nsteps, checkpoint=64, spheres=()):
    sim = Simulation(params)
    try:
        for centre, radius in spheres:
            sim.add_sphere(centre, radius)
        sim.set_design(lo, np.asarray(rho, dtype=np.float32))
        return sim.adjoint(nsteps, r0, freqs, checkpoint)
    finally:
        sim.close()


def DesignObjective():
    """Returns a torch.autograd.Function class bound to point_objective."""
    import torch

    class _DesignObjective(torch.autograd.Function):
        @staticmethod
        def forward(ctx, rho, params, lo, r0, freqs, nsteps, checkpoint=64):
            J, g = point_objective(params, lo, rho.detach().cpu().numpy(), r0, freqs, nsteps, checkpoint)
            ctx.save_for_backward(torch.as_tensor(g, dtype=rho.dtype, device=rho.device))
            return torch.tensor(J, dtype=torch.float64, device=rho.device)

        @staticmethod
        def backward(ctx, gJ):
            (g,) = ctx.saved_tensors
            return (gJ.to(g.dtype) * g,) + (None,) * 6

    return _DesignObjective
