"""Persistent on-GPU solvers: the compact spectral state stays in HBM across
steps; the run loops and benchmarks drive these.

The hydro solver is what ``run_simulation`` (jax/taylor_green_jax.py:204-317)
loops over; the MHD solver is what ``TaylorGreenMhdSimulation.run``
(cuda/taylor_green_mhd/simulation.py:49-100) loops over.
"""
import math

import torch

from ._lib import Plan

_PLANS = {}


def get_plan(N, dtype=torch.float64, mhd=False, L=2.0 * math.pi, device=None):
    """Cached Plan per (N, dtype, equations, L, device)."""
    dev = torch.device(device if device is not None else "cuda")
    if dev.type == "cuda" and dev.index is None:
        dev = torch.device("cuda", torch.cuda.current_device())
    key = (int(N), dtype, bool(mhd), float(L), str(dev))
    p = _PLANS.get(key)
    if p is None:
        p = Plan(N, dtype=dtype, mhd=mhd, domain_length=L, device=dev)
        _PLANS[key] = p
    return p


def clear_plans():
    _PLANS.clear()


class TaylorGreenSolver:
    """Hydro state u_hat (3, K, K, KZ) on one GPU, advanced by classic RK4."""

    def __init__(self, N, nu, dtype=torch.float64, L=2.0 * math.pi, device=None):
        self.plan = get_plan(N, dtype, False, L, device)
        self.N = N
        self.nu = float(nu)
        self.u_hat = self.plan.empty_spec()
        self.u_hat.zero_()

    def set_velocity(self, u):
        u = torch.as_tensor(u).to(self.plan.device, self.plan.dtype).contiguous()
        self.plan.forward(u, out=self.u_hat)

    def set_spectrum(self, u_hat):
        self.u_hat.copy_(u_hat)

    def velocity(self):
        return self.plan.inverse(self.u_hat)

    def step(self, nsteps=1, dt=0.005):
        self.plan.rk4(self.u_hat, nsteps=int(nsteps), dt=float(dt), nu=self.nu)

    def rhs(self):
        return self.plan.rhs(self.u_hat, nu=self.nu)

    def diagnostics(self, divmax=True):
        return self.plan.diagnostics(self.u_hat, divmax=divmax)

    def save(self, path, step=0, time=0.0, dt=0.0):
        """Write the compact state and (step, time, dt, nu) to a checkpoint file."""
        self.plan.checkpoint_write(path, self.u_hat, step=step, time=time, dt=dt, nu=self.nu)

    @classmethod
    def load(cls, path, device=None):
        """A solver restored from a checkpoint file, and the file's metadata."""
        from ._lib import checkpoint_info
        n, prec, mhd, L, meta = checkpoint_info(path)
        if mhd:
            raise ValueError(f"{path} holds an MHD state: use MhdSolver.load")
        s = cls(n, meta["nu"], dtype=torch.float64 if prec == 64 else torch.float32, L=L, device=device)
        return s, s.plan.checkpoint_read(path, s.u_hat)


class MhdSolver:
    """MHD state (u_hat, b_hat), each (3, K, K, KZ), advanced by classic RK4.

    lorentz_sign = -1 reproduces the reference code (rhs_u = -P(adv) -
    P(j x b), cuda/taylor_green_mhd/operators.py:88-92); +1 is the documented
    momentum equation (cuda/README_MHD.md:9).
    """

    def __init__(self, N, nu, eta, dtype=torch.float64, L=2.0 * math.pi, device=None, lorentz_sign=-1.0):
        self.plan = get_plan(N, dtype, True, L, device)
        self.N = N
        self.nu = float(nu)
        self.eta = float(eta)
        self.lorentz_sign = float(lorentz_sign)
        self.u_hat = self.plan.empty_spec()
        self.b_hat = self.plan.empty_spec()
        self.u_hat.zero_()
        self.b_hat.zero_()

    def set_fields(self, u, b):
        p = self.plan
        p.forward(torch.as_tensor(u).to(p.device, p.dtype).contiguous(), out=self.u_hat)
        p.forward(torch.as_tensor(b).to(p.device, p.dtype).contiguous(), out=self.b_hat)

    def fields(self):
        return self.plan.inverse(self.u_hat), self.plan.inverse(self.b_hat)

    def step(self, nsteps=1, dt=0.005):
        self.plan.rk4(self.u_hat, self.b_hat, nsteps=int(nsteps), dt=float(dt), nu=self.nu, eta=self.eta,
                      lorentz_sign=self.lorentz_sign)

    def rhs(self):
        return self.plan.rhs(self.u_hat, self.b_hat, nu=self.nu, eta=self.eta, lorentz_sign=self.lorentz_sign)

    def diagnostics(self, divmax=True):
        return self.plan.diagnostics(self.u_hat, self.b_hat, divmax=divmax)

    def save(self, path, step=0, time=0.0, dt=0.0):
        """Write (u_hat, b_hat) and (step, time, dt, nu, eta, lorentz_sign) to a checkpoint file."""
        self.plan.checkpoint_write(path, self.u_hat, self.b_hat, step=step, time=time, dt=dt, nu=self.nu,
                                   eta=self.eta, lorentz_sign=self.lorentz_sign)

    @classmethod
    def load(cls, path, device=None):
        """A solver restored from a checkpoint file, and the file's metadata."""
        from ._lib import checkpoint_info
        n, prec, mhd, L, meta = checkpoint_info(path)
        if not mhd:
            raise ValueError(f"{path} holds a hydro state: use TaylorGreenSolver.load")
        s = cls(n, meta["nu"], meta["eta"], dtype=torch.float64 if prec == 64 else torch.float32, L=L,
                device=device, lorentz_sign=meta["lorentz_sign"])
        return s, s.plan.checkpoint_read(path, s.u_hat, s.b_hat)
