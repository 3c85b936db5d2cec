"""torch.autograd view of the SQ density with frozen Lanczos bounds.

``SQDensity.apply(veff, sq, lf_lo, lf_hi)`` returns rho (n^3) from
``SpectralQuadrature.density_fixed`` (the chi, zee of the last full density
step kept, Fermi level re-solved in [lf_lo, lf_hi], so the electron count is
fixed).  Its backward is the hand-written adjoint of the Chebyshev recurrence
(``sqg_adjoint_veff`` with the Fermi-shift term), so
``torch.autograd.gradcheck`` compares the adjoint with finite differences of
the forward step.  torch is only the harness here: both directions run in the
HIP library through the C-ABI on host buffers.
"""
from __future__ import annotations

import numpy as np
import torch


def _host(t: torch.Tensor) -> np.ndarray:
    return np.ascontiguousarray(t.detach().to("cpu", torch.float64).numpy().reshape(-1))


class SQDensity(torch.autograd.Function):
    @staticmethod
    def forward(ctx, veff: torch.Tensor, sq, lf_lo: float, lf_hi: float) -> torch.Tensor:
        out = sq.density_fixed(_host(veff), lf_lo, lf_hi)
        ctx.sq = sq
        ctx.bracket = (lf_lo, lf_hi)
        ctx.save_for_backward(veff)
        return torch.from_numpy(out["rho"]).to(veff.device, veff.dtype)

    @staticmethod
    def backward(ctx, grad_rho: torch.Tensor):
        (veff,) = ctx.saved_tensors
        sq = ctx.sq
        # the adjoint runs at the state of the last density call: put it back at veff
        sq.density_fixed(_host(veff), *ctx.bracket)
        g = sq.adjoint_veff(_host(grad_rho), fixed_n=True)
        return torch.from_numpy(g).to(veff.device, veff.dtype), None, None, None
