"""torch.autograd.Function over the C-ABI gradient (include/sw2dg.h,
sw2dg_assim_gradient): the 4D-Var misfit J as a differentiable function of
the per-cell porosity / friction coefficients and of the boundary-value
table, so that torch code can compose it with its own parametrisation
(zones, hydrograph knots, a network's outputs) and let autograd do the
chain rule.

The forward calls sw2dg_assim_gradient once (dry run with sqrt(N)
checkpoints, segment recomputation and the hand-written adjoint on the
GPU) and keeps the gradients; the backward scales them by the incoming
gradient.  This mirrors how the reference exposes its Enzyme gradient
(assim/swSpE_checkpoint.cpp:722-1281 returns J and dJ/d(controls) from one
call) rather than re-running anything in backward.

Boundary values: the table `bc_v` [n_series, n_times] at times `bc_t` is
installed on the solver before the dry run (SpSolver.set_bc_series); its
gradient is the per-step boundary gradient mapped onto the table by the
_E time interpolation weights of the recorded step times (the same
linear map the forward applies, assim/swSpE.h:138-153).
"""
import numpy as np
import torch

from .assim import interp_weights


class GaugeMisfit(torch.autograd.Function):
    """J = misfit(phi, K1, K2, bc_v) of an Assim (a twin of its .gradient)."""

    @staticmethod
    def forward(ctx, assim, phi, K1, K2, bc_v, bc_t, bc_types):
        def host(x):
            return None if x is None else x.detach().to("cpu", torch.float64).numpy()

        if bc_v is not None:
            assim.s.set_bc_series(np.asarray(bc_t, dtype=np.float64), host(bc_v),
                                  types=None if bc_types is None else np.asarray(bc_types))
        G = assim.gradient(phi=host(phi), K1=host(K1), K2=host(K2))
        ref = next((x for x in (phi, K1, K2, bc_v) if x is not None), None)
        dev = ref.device if ref is not None else torch.device("cpu")
        outs = []
        for x, g in ((phi, G["phi"]), (K1, G["K1"]), (K2, G["K2"])):
            outs.append(None if x is None else torch.as_tensor(g, dtype=x.dtype).to(dev))
        if bc_v is not None:
            t_steps = assim.get_record()["t"][:-1]
            W = interp_weights(np.asarray(bc_t, dtype=np.float64), t_steps)
            outs.append(torch.as_tensor((W.T @ G["bc"]).T, dtype=bc_v.dtype).to(dev))
        else:
            outs.append(None)
        ctx.grads = outs
        ctx.n_steps = G["n_steps"]
        return torch.tensor(G["J"], dtype=torch.float64, device=dev)

    @staticmethod
    def backward(ctx, gJ):
        return (None, *[None if g is None else gJ * g for g in ctx.grads], None, None)


def gauge_misfit(assim, phi=None, K1=None, K2=None, bc_v=None, bc_t=None, bc_types=None):
    """J of `assim` at the given controls (torch tensors or None = the
    solver's own); differentiable in every tensor argument."""
    if bc_v is not None and bc_t is None:
        raise ValueError("bc_v needs its knot times bc_t")
    return GaugeMisfit.apply(assim, phi, K1, K2, bc_v, bc_t, bc_types)
