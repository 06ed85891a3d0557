"""Gradients of a final-time kinetic energy or enstrophy with respect to the
initial Fourier coefficients and the viscosity, through every RK4 step.

``final_objective(u_hat0, nu, nsteps, dt, objective)`` returns J as a 0-dim
tensor inside torch autograd: its backward runs the hand-written adjoint
kernels (vector-Jacobian products of the right-hand side, libtgdns
tg_adjoint_backward) with sqrt(n) checkpointing: the forward sweep keeps a
checkpoint every ceil(sqrt(nsteps)) steps; the backward sweep recomputes one
segment at a time, keeps its stage states and reverses it.

u_hat0 is the compact spectrum of a hydro plan (3, K, K, KZ); its gradient
follows torch's complex convention dJ/dRe + i dJ/dIm per stored entry.  The
reference has no gradient code; its jax.grad through advance_one_step
(jax/taylor_green_jax.py:188-196) is what this replaces.
"""
import math

import torch

from ._lib import supported_sizes
from .solver import get_plan

OBJECTIVES = {"energy": 0, "enstrophy": 1}


def sqrt_segment(nsteps):
    return max(1, int(math.ceil(math.sqrt(nsteps))))


def choose_segment(plan, nsteps, budget_bytes=None):
    """Checkpoint segment length: every stage of the whole trajectory kept
    (seg = nsteps: nothing recomputed) when that adjoint workspace fits in
    `budget_bytes` (default: a quarter of the free device memory), else the
    sqrt(n) schedule."""
    if budget_bytes is None:
        free, _ = torch.cuda.mem_get_info(plan.device)
        budget_bytes = free // 4
    return nsteps if plan.adjoint_workspace_bytes(nsteps, nsteps) <= budget_bytes else sqrt_segment(nsteps)


class FinalTimeObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u_hat0, nu, plan, nsteps, dt, objective, seg):
        aws = plan.adjoint_workspace(nsteps, seg)
        nu_f = float(nu)
        J = plan.adjoint_forward(u_hat0.detach().contiguous(), nsteps, dt, nu_f, objective, seg, aws)
        ctx.plan, ctx.aws = plan, aws
        ctx.args = (nsteps, dt, nu_f, seg)
        ctx.nu_is_tensor = torch.is_tensor(nu)
        return torch.tensor(J, dtype=torch.float64, device=u_hat0.device)

    @staticmethod
    def backward(ctx, gJ):
        nsteps, dt, nu, seg = ctx.args
        grad, dnu = ctx.plan.adjoint_backward(nsteps, dt, nu, seg, ctx.aws)
        ctx.aws = None
        g = gJ.to(torch.float64)
        gu = grad * g.to(grad.real.dtype)
        gnu = (g * dnu) if ctx.nu_is_tensor else None
        return gu, gnu, None, None, None, None, None


def final_objective(u_hat0, nu, nsteps, dt, objective="enstrophy", seg=None, L=2.0 * math.pi):
    """J = E(T) or Z(T) after `nsteps` RK4 steps from the compact spectrum
    u_hat0 (complex128 -> FP64 plan, complex64 -> FP32 plan).  `seg`:
    checkpoint interval; None picks choose_segment(); "sqrt" forces the
    sqrt(n) schedule."""
    dtype = torch.float64 if u_hat0.dtype == torch.complex128 else torch.float32
    M = (u_hat0.shape[-1] - 1)
    N = None
    for n in supported_sizes():
        if n // 3 == M and 2 * M + 1 == u_hat0.shape[-2]:
            N = n
            break
    if N is None:
        raise ValueError(f"no supported grid for a compact spectrum of shape {tuple(u_hat0.shape)}")
    plan = get_plan(N, dtype, False, L, u_hat0.device)
    if seg is None:
        seg = choose_segment(plan, nsteps)
    elif seg == "sqrt":
        seg = sqrt_segment(nsteps)
    return FinalTimeObjective.apply(u_hat0, nu, plan, int(nsteps), float(dt), OBJECTIVES[objective], int(seg))
