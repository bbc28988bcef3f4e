"""torch.autograd interface to the adjoint of libdmri.

``bloch_torrey_signal(solver, D, invT2, kappa, q, f_implicit, f_explicit, dt, gdir)`` returns the
normalised signal Re S as a 0-d float64 tensor whose backward pass applies the gradients of
``dmri_adjoint`` (checkpointed reverse sweep on the GPU): dS/dD per compartment, dS/d(1/T2) per
compartment, dS/dkappa, dS/dq and dS/df for every sample of both time profiles.

The mesh, its assembled matrices and every state vector stay in HBM inside the libdmri
context; only the handful of scalars and the time profiles cross into torch.
"""
from __future__ import annotations

import numpy as np
import torch

import dmri_native


def _np(t):
    return t.detach().to("cpu", torch.float64).numpy() if isinstance(t, torch.Tensor) else np.asarray(t, np.float64)


class BlochTorreySignal(torch.autograd.Function):
    @staticmethod
    def forward(ctx, D, invT2, kappa, q, f_implicit, f_explicit, solver, dt, gdir, theta, rtol, stride):
        if getattr(solver, "has_fields", False):
            # the scalar inputs below would replace the context's D_cell / invT2_cell / kappa_facet
            # fields and change the physics; the field multipliers' gradients are returned by
            # Context.adjoint directly
            raise ValueError("bloch_torrey_signal takes scalar coefficients; this context holds per-cell or "
                             "per-facet coefficient fields (use Context.adjoint for their multiplier gradients)")
        Dv, Tv = _np(D).reshape(2), _np(invT2).reshape(2)
        solver.set_coeffs(D=tuple(Dv), invT2=tuple(Tv), kappa=float(_np(kappa)))
        solver.assemble()
        s, g, _ = solver.adjoint(_np(f_implicit), _np(f_explicit), dt, float(_np(q)), gdir, theta=theta, rtol=rtol,
                                 checkpoint_stride=stride)
        f64 = torch.float64
        grads = [torch.as_tensor(g["D"], dtype=f64), torch.as_tensor(g["invT2"], dtype=f64),
                 torch.tensor(g["kappa"], dtype=f64), torch.tensor(g["q"], dtype=f64),
                 torch.as_tensor(g["f_implicit"], dtype=f64), torch.as_tensor(g["f_explicit"], dtype=f64)]
        ctx.save_for_backward(*[x.to(torch.float64) for x in grads])
        ctx.devices = [t.device if isinstance(t, torch.Tensor) else torch.device("cpu")
                       for t in (D, invT2, kappa, q, f_implicit, f_explicit)]
        return torch.tensor(s.real, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gout):
        gs = ctx.saved_tensors
        out = [None if not ctx.needs_input_grad[i] else (gout.to("cpu") * gs[i]).to(ctx.devices[i])
               for i in range(6)]
        return (*out, None, None, None, None, None, None)


def bloch_torrey_signal(solver: dmri_native.Context, D, invT2, kappa, q, f_implicit, f_explicit, dt, gdir,
                        theta=0.5, rtol=1e-13, checkpoint_stride=0):
    """Differentiable signal of an assembled libdmri context (its mesh must already be set)."""
    as_t = lambda v: v if isinstance(v, torch.Tensor) else torch.as_tensor(v, dtype=torch.float64)
    D = as_t(D).to(torch.float64).reshape(-1).expand(2) if as_t(D).numel() == 1 else as_t(D)
    invT2 = as_t(invT2).to(torch.float64).reshape(-1).expand(2) if as_t(invT2).numel() == 1 else as_t(invT2)
    return BlochTorreySignal.apply(D, invT2, as_t(kappa), as_t(q), as_t(f_implicit), as_t(f_explicit), solver,
                                   float(dt), tuple(float(x) for x in gdir), float(theta), float(rtol),
                                   int(checkpoint_stride))
