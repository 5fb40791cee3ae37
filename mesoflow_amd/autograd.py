"""torch.autograd.Function over the discrete adjoint of the C-ABI.

Thin wrapper used by the gradient checks: theta = (log_A[0..nreact-1], in_vel)
of a porous-reactor level -> J, the time-integrated outlet species mass flux of
mf_adjoint_forward. The forward runs the fixed-dt steps from a given initial
state and the reverse sweep at once (mf_adjoint_forward + mf_adjoint_backward),
so the gradient belongs to that forward even when several forwards run before
backward is called (torch.autograd.gradcheck does that).
"""
import ctypes as C

import numpy as np
import torch

from . import abi


def plugin_with(base, theta):
    """Copy of a PluginParams with log_A[r] = theta[r] (r < nreact) and in_vel = theta[nreact]."""
    p = abi.PluginParams()
    C.memmove(C.byref(p), C.byref(base), C.sizeof(p))
    nr = int(base.nreact)
    for r in range(nr):
        p.log_A[r] = float(theta[r])
    p.in_vel = float(theta[nr])
    return p


def theta_of(plugin):
    nr = int(plugin.nreact)
    return np.array([plugin.log_A[r] for r in range(nr)] + [plugin.in_vel], dtype=np.float64)


def run_objective(level, theta, state0, nsteps, ncheck, spec, with_grad=True):
    """J(theta) and, with with_grad, dJ/dtheta (numpy) from the initial state state0."""
    ctx = level.ctx
    ctx.set_plugin_params(plugin_with(ctx.case.plugin, theta))
    level.set_state(state0)
    level.set_time(0.0, 1.0e100, 0)
    J = level.adjoint_forward(nsteps, ncheck, spec)
    if not with_grad:
        return J, None
    gA, gv = level.adjoint_backward()
    return J, np.concatenate([gA, [gv]])


class OutletFluxObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, level, state0, nsteps, ncheck, spec):
        J, g = run_objective(level, theta.detach().cpu().numpy(), state0, nsteps, ncheck, spec)
        ctx.save_for_backward(torch.from_numpy(g))
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ):
        (g,) = ctx.saved_tensors
        return gJ * g, None, None, None, None, None
