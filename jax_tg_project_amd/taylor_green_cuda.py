"""Drop-in for the reference CuPy hydro solver: the functional shim
cuda/taylor_green_cuda.py:12-79 and the classes of the taylor_green_cuda
package (config.py, grid.py, initial_conditions.py, operators.py,
integrator.py, diagnostics.py, simulation.py, plotting.py, cli.py).

The reference keeps full-cube complex spectra (3, N, N, N) on the GPU; so
does this module, on torch tensors.  At the edge each spectrum X is split
into

* its Hermitian part restricted to the kept modes (the compact spectrum
  that the gfx950 kernels advance: every transform, right-hand side, RK4
  step and diagnostic runs in libtgdns.so), and
* a remainder R = X - compact_to_full(compact): the anti-Hermitian part
  of the kept modes and everything outside the 2/3 mask.

For a spectrum of a real field that the reference itself produced
(to_spectral, dealias, a step) R is exactly zero.  Otherwise R is carried
exactly as the reference carries it: the nonlinear terms only see
``ifftn(...).real`` (the Hermitian part), so R enters the right-hand side
only through the viscous term -nu k^2 mask R, evolves under RK4 as the
linear ODE it then obeys, and modes outside the mask pass through a step
unchanged (they are dropped where the reference dealiases).  The transforms
to physical space (to_real, diagnostics) need R to vanish outside the mask
(the pruned inverse transforms hold only kept modes) and raise otherwise.
"""
import argparse
import math
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .solver import get_plan

_TWO_PI = 2.0 * math.pi


def _dev(device):
    if device is not None:
        return torch.device(device)
    return torch.device("cuda" if torch.cuda.is_available() else "cpu")


def _cdtype_of(t):
    return torch.complex64 if t.dtype in (torch.complex64, torch.float32) else torch.complex128


def _rdtype_of(t):
    return torch.float32 if t.dtype in (torch.complex64, torch.float32) else torch.float64


# ------------------------------------------------------------------ config
@dataclass(frozen=True)
class TaylorGreenConfig:
    """cuda/taylor_green_cuda/config.py:5-38."""
    n: int = 128
    dt: float = 0.005
    reynolds: float = 1000.0
    total_time: float = 1.0
    save_every_time: float = 0.05
    domain_length: float = _TWO_PI

    def __post_init__(self):
        if self.n < 4:
            raise ValueError("n must be at least 4")
        if self.dt <= 0.0:
            raise ValueError("dt must be positive")
        if self.reynolds <= 0.0:
            raise ValueError("reynolds must be positive")
        if self.total_time <= 0.0:
            raise ValueError("total_time must be positive")
        if self.save_every_time <= 0.0:
            raise ValueError("save_every_time must be positive")
        if self.domain_length <= 0.0:
            raise ValueError("domain_length must be positive")

    @property
    def viscosity(self):
        return self.domain_length / self.reynolds

    @property
    def steps(self):
        return int(self.total_time / self.dt)

    @property
    def save_every(self):
        return max(1, int(self.save_every_time / self.dt))


# ------------------------------------------------------------------ grid
class SpectralGrid:
    """cuda/taylor_green_cuda/grid.py:4-40 on torch tensors.  to_spectral /
    to_real run the pruned HIP transforms; the compact <-> full-cube
    conversions and the split described in the module docstring are edge
    plumbing."""

    def __init__(self, n, domain_length, device=None):
        self.n = int(n)
        self.domain_length = domain_length
        k = 2.0 * np.pi * np.fft.fftfreq(n, d=domain_length / n)
        cutoff = (n // 3) * (2.0 * np.pi / domain_length)
        keep = np.abs(k) <= cutoff
        dev = _dev(device)
        kt = torch.tensor(k, dtype=torch.float64, device=dev)
        kx, ky, kz = torch.meshgrid(kt, kt, kt, indexing="ij")
        kp = torch.tensor(keep, device=dev)
        self._set_wavenumbers(kx, ky, kz, kp[:, None, None] & kp[None, :, None] & kp[None, None, :])

    @classmethod
    def from_arrays(cls, kx, ky, kz, dealias_mask):
        """grid.py:15-22.  The arrays must be the canonical wavenumber grids
        and 2/3 cube mask of some (N, L): the kernels implement exactly that
        truncation (ValueError otherwise)."""
        from .taylor_green_jax import grid_from_arrays
        n, L = grid_from_arrays(kx, ky, kz, dealias_mask)
        grid = cls.__new__(cls)
        grid.n = n
        grid.domain_length = None
        grid._L = L
        grid._set_wavenumbers(torch.as_tensor(kx), torch.as_tensor(ky), torch.as_tensor(kz),
                              torch.as_tensor(dealias_mask))
        return grid

    def _set_wavenumbers(self, kx, ky, kz, dealias_mask):
        self.kx, self.ky, self.kz = kx, ky, kz
        self.k_squared = kx * kx + ky * ky + kz * kz
        self.dealias_mask = dealias_mask

    @property
    def length(self):
        if self.domain_length is not None:
            return float(self.domain_length)
        return float(getattr(self, "_L", _TWO_PI))

    def plan(self, dtype=torch.float64, mhd=False):
        _lib._require_gpu()
        return get_plan(self.n, dtype, mhd, self.length)

    def _on(self, t, device):
        return t if t.device == device else t.to(device)

    # compact <-> full-cube conversions --------------------------------------
    def _kept(self, device):
        n, M = self.n, self.n // 3
        return torch.cat([torch.arange(0, M + 1), torch.arange(n - M, n)]).to(device)

    def full_to_compact(self, f_hat):
        """Kept modes with kz >= 0 of a full-cube spectrum, as stored."""
        idx = self._kept(f_hat.device)
        M = self.n // 3
        return f_hat.index_select(-3, idx).index_select(-2, idx)[..., : M + 1].contiguous()

    def hermitian_compact(self, f_hat):
        """Kept modes (kz >= 0) of the Hermitian part (X(k) + conj X(-k)) / 2,
        the part that ifftn(X).real depends on.  Equals full_to_compact for
        the spectrum of a real field, exactly."""
        n, M = self.n, self.n // 3
        idx = self._kept(f_hat.device)
        nidx = (-idx) % n
        iz = torch.arange(M + 1, device=f_hat.device)
        pos = f_hat.index_select(-3, idx).index_select(-2, idx).index_select(-1, iz)
        neg = f_hat.index_select(-3, nidx).index_select(-2, nidx).index_select(-1, (-iz) % n)
        return (0.5 * (pos + torch.conj(neg))).contiguous()

    def compact_to_full(self, c_hat):
        n, M = self.n, self.n // 3
        idx = self._kept(c_hat.device)
        full = torch.zeros(c_hat.shape[:-3] + (n, n, n), dtype=c_hat.dtype, device=c_hat.device)
        ix = idx[:, None, None]
        iy = idx[None, :, None]
        iz = torch.arange(M + 1, device=c_hat.device)[None, None, :]
        full[..., ix, iy, iz] = c_hat
        nx = (-idx) % n
        izn = (-torch.arange(1, M + 1, device=c_hat.device)) % n
        full[..., nx[:, None, None], nx[None, :, None], izn[None, None, :]] = torch.conj(c_hat[..., 1:])
        return full

    def split(self, f_hat):
        """(compact Hermitian part, remainder or None when it is zero)."""
        c = self.hermitian_compact(f_hat)
        rest = f_hat - self.compact_to_full(c)
        if not bool(torch.any(rest != 0)):
            rest = None
        return c, rest

    def require_in_mask(self, f_hat, rest):
        """The pruned inverse transforms hold only kept modes: a remainder
        outside the mask (beyond round-off) cannot be transformed."""
        if rest is None:
            return
        outside = rest * (~self._on(self.dealias_mask, rest.device))
        scale = float(f_hat.abs().max()) if f_hat.numel() else 0.0
        if float(outside.abs().max()) > 1e-12 * max(scale, 1e-300):
            raise ValueError("spectrum has modes outside the 2/3 dealiasing mask; apply grid.dealias first "
                             "(the pruned transforms hold only the kept modes)")

    def mask_like(self, t):
        return self._on(self.dealias_mask, t.device).to(_rdtype_of(t))

    def k2_like(self, t):
        return self._on(self.k_squared, t.device).to(_rdtype_of(t))

    # reference methods -------------------------------------------------------
    def to_spectral(self, field):
        """mask * fftn(field) as a full cube (grid.py:31-34); exact for any
        real field (the pruned forward transform computes exactly the kept
        modes).  float32 input gives complex64, as fftn does."""
        f = torch.as_tensor(field)
        dtype = torch.float32 if f.dtype == torch.float32 else torch.float64
        p = self.plan(dtype)
        f = f.to(p.device, dtype).contiguous()
        lead = f.shape[:-3]
        flat = f.reshape((-1, self.n, self.n, self.n))
        outs = []
        for i in range(0, flat.shape[0], 3):  # the plan transforms vectors of 3 components
            chunk = flat[i:i + 3]
            m = chunk.shape[0]
            if m < 3:
                chunk = torch.cat([chunk, torch.zeros((3 - m,) + chunk.shape[1:], dtype=dtype, device=p.device)])
            outs.append(self.compact_to_full(p.forward(chunk.contiguous()))[:m])
        return torch.cat(outs).reshape(lead + (self.n, self.n, self.n))

    def to_real(self, field_hat):
        """ifftn(field_hat).real (grid.py:36-37) for a spectrum inside the mask."""
        fh = torch.as_tensor(field_hat)
        dtype = _rdtype_of(fh)
        p = self.plan(dtype)
        fh = fh.to(p.device, p.cdtype)
        c, rest = self.split(fh)
        self.require_in_mask(fh, rest)
        lead = c.shape[:-3]
        flat = c.reshape((-1,) + c.shape[-3:])
        outs = []
        for i in range(0, flat.shape[0], 3):
            chunk = flat[i:i + 3]
            m = chunk.shape[0]
            if m < 3:
                chunk = torch.cat([chunk, torch.zeros((3 - m,) + chunk.shape[1:], dtype=c.dtype, device=c.device)])
            outs.append(p.inverse(chunk.contiguous())[:m])
        return torch.cat(outs).reshape(lead + (self.n, self.n, self.n))

    def dealias(self, field_hat):
        return field_hat * self._on(self.dealias_mask, field_hat.device)


def rk4_linear(x, lam, dt):
    """Classic RK4 of x' = lam x with the reference's combination order
    (integrator.py:7-12): the evolution of the remainder of a step."""
    r1 = lam * x
    r2 = lam * (x + 0.5 * dt * r1)
    r3 = lam * (x + 0.5 * dt * r2)
    r4 = lam * (x + dt * r3)
    return x + (dt / 6.0) * (r1 + 2.0 * r2 + 2.0 * r3 + r4)


# ------------------------------------------------------------------ IC
class TaylorGreenInitialCondition:
    """cuda/taylor_green_cuda/initial_conditions.py:4-13 (float32 fields)."""

    def __init__(self, domain_length=_TWO_PI):
        self.domain_length = domain_length

    def create(self, n, device=None):
        dev = _dev(device)
        x = torch.tensor(np.linspace(0.0, self.domain_length, n, endpoint=False, dtype=np.float64),
                         dtype=torch.float32, device=dev)
        X, Y, Z = torch.meshgrid(x, x, x, indexing="ij")
        u = torch.sin(X) * torch.cos(Y) * torch.cos(Z)
        v = -torch.cos(X) * torch.sin(Y) * torch.cos(Z)
        return torch.stack((u, v, torch.zeros_like(u)))


# ------------------------------------------------------------------ operator / integrator
class CupySpectralOperator:
    """cuda/taylor_green_cuda/operators.py:4-48: right-hand side of the
    full-cube spectrum in the fused HIP passes.  The reference's float64
    wavenumber arrays promote every spectrum to complex128, so the FP64
    kernels run and complex128 comes back whatever the input precision;
    ``precision="fp32"`` selects the FP32 kernels instead (complex64)."""

    def __init__(self, grid, precision="fp64"):
        if precision not in ("fp64", "fp32"):
            raise ValueError("precision must be 'fp64' or 'fp32'")
        self.grid = grid
        self.precision = precision

    def gradients(self, field_hat):
        """operators.py:8-14: dealiased i k_d field_hat[c] (elementwise)."""
        g = self.grid
        waves = (g.kx, g.ky, g.kz)
        return tuple(tuple(g.dealias(1j * g._on(w, field_hat.device) * field_hat[c]) for w in waves)
                     for c in range(3))

    def spectral_curl_hat(self, field_hat):
        """operators.py:16-22 (elementwise)."""
        gr = self.gradients(field_hat)
        return self.grid.dealias(torch.stack((gr[2][1] - gr[1][2], gr[0][2] - gr[2][0], gr[1][0] - gr[0][1])))

    def _plan_for(self, t=None):
        return self.grid.plan(torch.float64 if self.precision == "fp64" else torch.float32)

    def rhs(self, velocity_hat, viscosity):
        """-P mask F(u . grad u) - nu k^2 mask u_hat (operators.py:24-48)."""
        vh = torch.as_tensor(velocity_hat)
        p = self._plan_for(vh)
        vh = vh.to(p.device, p.cdtype)
        c, rest = self.grid.split(vh)
        out = self.grid.compact_to_full(p.rhs(c, nu=float(viscosity)))
        if rest is not None:
            out = out - viscosity * self.grid.k2_like(vh) * self.grid.mask_like(vh) * rest
        return out


class RK4Integrator:
    """cuda/taylor_green_cuda/integrator.py:1-12: one classic RK4 step of the
    full-cube state (all four stages in the fused GPU passes); the result is
    not dealiased, as there."""

    def __init__(self, operator, dt, viscosity):
        self.operator = operator
        self.dt = dt
        self.viscosity = viscosity

    def _step(self, state, dealias):
        op = self.operator
        g = op.grid
        st = torch.as_tensor(state)
        p = op._plan_for(st)
        st = st.to(p.device, p.cdtype)
        c, rest = g.split(st)
        p.rk4(c, nsteps=1, dt=float(self.dt), nu=float(self.viscosity))
        out = g.compact_to_full(c)
        if rest is not None:
            m = g.mask_like(st)
            kept = rk4_linear(m * rest, -self.viscosity * g.k2_like(st) * m, self.dt)
            out = out + (kept if dealias else kept + (1.0 - m) * rest)
        return out

    def step(self, state):
        return self._step(state, dealias=False)


class Diagnostics:
    """cuda/taylor_green_cuda/diagnostics.py:4-25: E = 1/2 <|u|^2> and
    Z = 1/2 <|curl u|^2> as FP64 Parseval sums of the compact state."""

    def __init__(self, operator):
        self.operator = operator
        self.times = []
        self.energies = []
        self.enstrophies = []

    def record_values(self, time_value, energy, enstrophy):
        self.times.append(time_value)
        self.energies.append(energy)
        self.enstrophies.append(enstrophy)
        return energy, enstrophy

    def record(self, time, state_hat):
        g = self.operator.grid
        st = torch.as_tensor(state_hat)
        p = self.operator._plan_for()
        st = st.to(p.device, p.cdtype)
        c, rest = g.split(st)
        g.require_in_mask(st, rest)
        d = p.diagnostics(c, divmax=False)
        return self.record_values(time, d["kinetic_energy"], d["enstrophy"])


def check_finite(d, step, time_value):
    """Raise FloatingPointError when the diagnostics report NaN or infinite
    values in the state (the reference would continue with NaN fields)."""
    if d.get("nonfinite", 0.0) > 0.0 or not math.isfinite(d["kinetic_energy"]):
        raise FloatingPointError(
            f"the state holds {int(d.get('nonfinite', 0))} non-finite values at step {step} "
            f"(t = {time_value:g}): the run has diverged")


# ------------------------------------------------------------------ simulation
class TaylorGreenSimulation:
    """cuda/taylor_green_cuda/simulation.py:13-52.  The compact state stays on
    the GPU between saves and the steps between two saves run as one call
    into the library; ``state`` gives the full-cube spectrum."""

    def __init__(self, config=None, make_plots=True, precision="fp64", check_finite=True):
        from .solver import TaylorGreenSolver
        self.config = config or TaylorGreenConfig()
        c = self.config
        _lib._require_gpu()
        self.grid = SpectralGrid(c.n, c.domain_length)
        self.initial_condition = TaylorGreenInitialCondition(c.domain_length)
        self.operator = CupySpectralOperator(self.grid)
        self.integrator = RK4Integrator(self.operator, c.dt, c.viscosity)
        dtype = torch.float64 if precision == "fp64" else torch.float32
        self.solver = TaylorGreenSolver(c.n, c.viscosity, dtype=dtype, L=c.domain_length)
        self.solver.set_velocity(self.initial_condition.create(c.n, device=self.solver.plan.device))
        self.diagnostics = Diagnostics(self.operator)
        self.make_plots = make_plots
        self.check_finite = check_finite

    @property
    def state(self):
        return self.grid.compact_to_full(self.solver.u_hat)

    def _record(self, step):
        d = self.solver.diagnostics(divmax=False)
        if self.check_finite:
            check_finite(d, step, step * self.config.dt)
        return self.diagnostics.record_values(step * self.config.dt, d["kinetic_energy"], d["enstrophy"])

    def record(self, step):
        energy, value = self._record(step)
        print(f"Save time step: {step}, time: {step * self.config.dt:.6f}, enstrophy: {value:.8e}")
        return energy, value

    def run(self):
        c = self.config
        self._record(0)
        start = time.perf_counter()
        step = 0
        while step < c.steps:
            nxt = min(c.steps, (step // c.save_every + 1) * c.save_every)
            self.solver.step(nxt - step, c.dt)
            step = nxt
            self.record(step)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - start
        print(f"GPU device: {torch.cuda.get_device_name()}")
        print(f"Reynolds number: {c.reynolds:.1f}")
        print(f"Simulated time: {c.total_time:.3f} ({c.steps} steps) in {elapsed:.3f} s")
        print(f"Energy trace: {self.diagnostics.energies[0]:.6f} -> {self.diagnostics.energies[-1]:.6f}")
        state = self.state
        if self.make_plots:
            save_plots(state, self.diagnostics, c.n, self.grid)
        return state, self.diagnostics


def save_plots(state_hat, diagnostics, n, grid):
    """cuda/taylor_green_cuda/plotting.py:7-28 (PNG side effect)."""
    try:
        import matplotlib
        matplotlib.use("Agg")
        import matplotlib.pyplot as plt
    except ImportError:  # pragma: no cover
        print("matplotlib unavailable: plots skipped")
        return
    slice_data = grid.to_real(state_hat)[:, :, :, n // 2].cpu().numpy()
    figure, axes = plt.subplots(1, 3, figsize=(12, 4))
    for component, axis in enumerate(axes):
        image = axis.imshow(slice_data[component].T, origin="lower", cmap="viridis")
        axis.set_title(f"u{component + 1} slice")
        figure.colorbar(image, ax=axis)
    figure.tight_layout()
    figure.savefig("taylor_green_cuda_slice.png", dpi=200)
    plt.close(figure)
    figure, (energy_axis, enstrophy_axis) = plt.subplots(2, 1, figsize=(10, 8))
    energy_axis.plot(diagnostics.times, diagnostics.energies, "b-o")
    energy_axis.set_ylabel("Kinetic energy")
    energy_axis.grid(True, alpha=0.3)
    enstrophy_axis.plot(diagnostics.times, diagnostics.enstrophies, "r-s")
    enstrophy_axis.set_xlabel("Time")
    enstrophy_axis.set_ylabel("Enstrophy")
    enstrophy_axis.grid(True, alpha=0.3)
    figure.tight_layout()
    figure.savefig("taylor_green_cuda_diagnostics.png", dpi=200)
    plt.close(figure)


# ------------------------------------------------------------------ functional shim
def make_wavenumbers(n, length):
    """cuda/taylor_green_cuda.py:12-14."""
    g = SpectralGrid(n, length)
    return g.kx, g.ky, g.kz


def make_dealias_mask(n, length):
    """cuda/taylor_green_cuda.py:17-18."""
    return SpectralGrid(n, length).dealias_mask


def initial_taylor_green(n, length=_TWO_PI):
    """cuda/taylor_green_cuda.py:21-22 (float32 physical field)."""
    return TaylorGreenInitialCondition(length).create(n)


def initial_taylor_green_spectral(n, length=_TWO_PI):
    """cuda/taylor_green_cuda.py:25-27 (complex64, as fftn of float32)."""
    grid = SpectralGrid(n, length)
    return grid.to_spectral(TaylorGreenInitialCondition(length).create(n))


def spectral_rhs(velocity_hat, nu, kx, ky, kz, dealias_mask):
    """cuda/taylor_green_cuda.py:34-37."""
    return CupySpectralOperator(SpectralGrid.from_arrays(kx, ky, kz, dealias_mask)).rhs(velocity_hat, nu)


def advance_one_step(velocity_hat, dt, nu, kx, ky, kz, dealias_mask):
    """cuda/taylor_green_cuda.py:40-46 (the result is dealiased, as there)."""
    grid = SpectralGrid.from_arrays(kx, ky, kz, dealias_mask)
    return RK4Integrator(CupySpectralOperator(grid), dt, nu)._step(velocity_hat, dealias=True)


def kinetic_energy(velocity):
    """cuda/taylor_green_cuda.py:49-50."""
    v = torch.as_tensor(velocity)
    return 0.5 * torch.mean(torch.sum(v * v, dim=0))


def enstrophy(velocity, kx, ky, kz):
    """cuda/taylor_green_cuda.py:53-60: the enstrophy of mask * F(velocity)
    (the reference transforms with the 2/3-masked grid of side 2 pi and
    takes the curl without a mask), as a 0-dim tensor."""
    from .taylor_green_jax import grid_from_arrays
    n, _ = grid_from_arrays(kx, ky, kz)
    grid = SpectralGrid(n, _TWO_PI)
    v = torch.as_tensor(velocity)
    dtype = torch.float32 if v.dtype == torch.float32 else torch.float64
    p = grid.plan(dtype)
    c = p.forward(v.to(p.device, dtype).contiguous())
    return torch.tensor(p.diagnostics(c, divmax=False)["enstrophy"], dtype=torch.float64)


def run_simulation(n=128, dt=0.005, reynolds=1000.0, total_time=1.0, save_every_time=0.05, make_plots=True):
    """cuda/taylor_green_cuda.py:63-69: (physical velocity, times, energies,
    enstrophies)."""
    config = TaylorGreenConfig(n, dt, reynolds, total_time, save_every_time)
    simulation = TaylorGreenSimulation(config, make_plots=make_plots)
    state_hat, diagnostics = simulation.run()
    return simulation.grid.to_real(state_hat), diagnostics.times, diagnostics.energies, diagnostics.enstrophies


def main(argv=None):
    """cuda/taylor_green_cuda/cli.py:7-23 (python -m jax_tg_project_amd.taylor_green_cuda)."""
    parser = argparse.ArgumentParser(description="3D Taylor-Green vortex on the MI355X HIP kernels")
    parser.add_argument("--n", type=int, default=128)
    parser.add_argument("--dt", type=float, default=0.005)
    parser.add_argument("--reynolds", type=float, default=1000.0)
    parser.add_argument("--total-time", type=float, default=1.0)
    parser.add_argument("--save-every-time", type=float, default=0.05)
    parser.add_argument("--no-plots", action="store_true")
    args = parser.parse_args(argv)
    config = TaylorGreenConfig(n=args.n, dt=args.dt, reynolds=args.reynolds, total_time=args.total_time,
                               save_every_time=args.save_every_time)
    return TaylorGreenSimulation(config, make_plots=not args.no_plots).run()


if __name__ == "__main__":
    main()
