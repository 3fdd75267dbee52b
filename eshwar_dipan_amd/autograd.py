"""torch.autograd binding of the adjoint (gp_adjoint_* through gpstep.Adjoint).

    loss = phi_mismatch_loss(theta, phi0, mu0, target, inp, nsteps)

theta is a float64 tensor (dab, Tau, GAMMA, D_liquid) of the pair (solid 0,
liquid) of a two-phase binary parameter set `inp`; phi0 [2, MX, MY] and
mu0 [1, MX, MY] are the initial state, target [MX, MY] the solid fraction to
match.  The forward pass runs `nsteps` explicit steps on the GPU and returns
L = 1/2 sum (phi_0(T) - target)^2; the backward pass returns the adjoint's
gradients for theta, phi0 and mu0, computed in the same call.  Tensors live on
the host: the HIP library owns the device memory.
"""
import copy

import numpy as np
import torch

from . import gpstep
from ._abi import GpInput


def with_theta(params, theta):
    """Copy of GpParams with (dab, Tau, GAMMA, D_liquid) of pair 0-1 set from theta."""
    p = copy.deepcopy(params.p if isinstance(params, GpInput) else params)
    dab, tau, gam, dl = (float(v) for v in theta)
    p.dab[0][1] = p.dab[1][0] = dab
    p.tau_ab[0][1] = p.tau_ab[1][0] = tau
    p.gamma[0][1] = p.gamma[1][0] = gam
    p.diffusivity[p.nphases - 1][0][0] = dl
    return p


def theta_of(params):
    p = params.p if isinstance(params, GpInput) else params
    return torch.tensor([p.dab[0][1], p.tau_ab[0][1], p.gamma[0][1], p.diffusivity[p.nphases - 1][0][0]],
                        dtype=torch.float64)


class PhiMismatchLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, phi0, mu0, target, params, nsteps, dtype, nsnap):
        p = with_theta(params, theta.detach().cpu())
        s = gpstep.Solver(p, dtype)
        adj = gpstep.Adjoint(s, nsteps, nsnap)
        want_state = phi0.requires_grad or mu0.requires_grad
        loss, grad, sg = adj.run(phi0.detach().cpu().numpy(), mu0.detach().cpu().numpy(),
                                 target.detach().cpu().numpy(), state_grad=want_state)
        adj.close()
        s.close()
        ctx.g_theta = torch.tensor([grad[k] for k in gpstep.Adjoint.PARAMS], dtype=torch.float64)
        ctx.g_phi = torch.from_numpy(sg[0]) if want_state else None
        ctx.g_mu = torch.from_numpy(sg[1]) if want_state else None
        return torch.tensor(loss, dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        gp = g * ctx.g_phi if ctx.g_phi is not None else None
        gm = g * ctx.g_mu if ctx.g_mu is not None else None
        return g * ctx.g_theta, gp, gm, None, None, None, None, None


def phi_mismatch_loss(theta, phi0, mu0, target, params, nsteps, dtype="f64", nsnap=0):
    """Differentiable L(theta, phi0, mu0) of `nsteps` steps (see module doc)."""
    as_t = lambda a: a if isinstance(a, torch.Tensor) else torch.from_numpy(np.asarray(a, dtype=np.float64))  # noqa: E731
    return PhiMismatchLoss.apply(as_t(theta), as_t(phi0), as_t(mu0), as_t(target), params, int(nsteps), dtype,
                                 int(nsnap))
