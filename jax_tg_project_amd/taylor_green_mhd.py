"""Drop-in for the reference MHD solver: the functional shim
cuda/taylor_green_mhd.py:12-106 and the classes of the taylor_green_mhd
package it re-exports (config.py, initial_conditions.py, operators.py,
integrator.py, diagnostics.py, simulation.py) plus SpectralGrid
(cuda/taylor_green_cuda/grid.py).

The reference keeps full-cube complex spectra (3, N, N, N); this module
accepts and returns the same, converting at the edge to the compact
2/3-truncated half spectrum that the gfx950 kernels advance.  Every
transform, right-hand side, RK4 step and diagnostic runs in libtgdns.so.

Lorentz sign: the reference code subtracts P(j x b) (operators.py:88-92)
while its README states + (curl B) x B (README_MHD.md:9).  ``lorentz_sign``
defaults to -1.0, the code's behaviour, so results match the reference;
+1.0 selects the documented equation.
"""
import math
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .solver import MhdSolver, get_plan

_TWO_PI = 2.0 * math.pi


# ------------------------------------------------------------------ config
@dataclass(frozen=True)
class TaylorGreenMhdConfig:
    """cuda/taylor_green_mhd/config.py:7-50 (+ lorentz_sign, initial_condition,
    dtype selectors with reference-compatible defaults)."""
    n: int = 128
    dt: float = 0.005
    reynolds: float = 1000.0
    magnetic_reynolds: float = 1000.0
    magnetic_amplitude: float = 0.5
    total_time: float = 1.0
    save_every_time: float = 0.05
    domain_length: float = _TWO_PI
    lorentz_sign: float = -1.0
    initial_condition: str = "reference"
    precision: str = "fp64"

    def __post_init__(self):
        if self.n < 4:
            raise ValueError("n must be at least 4")
        if self.dt <= 0.0:
            raise ValueError("dt must be positive")
        if self.reynolds <= 0.0:
            raise ValueError("reynolds must be positive")
        if self.magnetic_reynolds <= 0.0:
            raise ValueError("magnetic_reynolds must be positive")
        if self.magnetic_amplitude < 0.0:
            raise ValueError("magnetic_amplitude must be nonnegative")
        if self.total_time <= 0.0:
            raise ValueError("total_time must be positive")
        if self.save_every_time <= 0.0:
            raise ValueError("save_every_time must be positive")
        if self.domain_length <= 0.0:
            raise ValueError("domain_length must be positive")
        if self.initial_condition not in ("reference", "config"):
            raise ValueError("initial_condition must be 'reference' or 'config'")
        if self.precision not in ("fp64", "fp32"):
            raise ValueError("precision must be 'fp64' or 'fp32'")

    @property
    def viscosity(self):
        return self.domain_length / self.reynolds

    @property
    def resistivity(self):
        return self.domain_length / self.magnetic_reynolds

    @property
    def steps(self):
        return int(self.total_time / self.dt)

    @property
    def save_every(self):
        return max(1, int(self.save_every_time / self.dt))


def magnetic_reynolds_to_resistivity(magnetic_reynolds, velocity_amplitude=1.0, domain_length=_TWO_PI):
    """cuda/taylor_green_mhd/config.py:53-60."""
    if magnetic_reynolds <= 0.0:
        raise ValueError("magnetic_reynolds must be positive")
    return float((velocity_amplitude * domain_length) / magnetic_reynolds)


def resistivity_to_magnetic_reynolds(resistivity, velocity_amplitude=1.0, domain_length=_TWO_PI):
    """cuda/taylor_green_mhd/config.py:63-72."""
    if resistivity < 0.0:
        raise ValueError("resistivity must be nonnegative")
    if resistivity == 0.0:
        return float("inf")
    return float((velocity_amplitude * domain_length) / resistivity)


# ------------------------------------------------------------------ grid
def _dev(device):
    if device is not None:
        return torch.device(device)
    return torch.device("cuda" if torch.cuda.is_available() else "cpu")


class SpectralGrid:
    """cuda/taylor_green_cuda/grid.py:4-40 on torch tensors; to_spectral /
    to_real run the pruned HIP transforms."""

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
        from .taylor_green_jax import grid_from_arrays
        n, L = grid_from_arrays(kx, ky, kz, dealias_mask)
        grid = cls.__new__(cls)
        grid.n = n
        grid.domain_length = L
        grid._set_wavenumbers(torch.as_tensor(kx), torch.as_tensor(ky), torch.as_tensor(kz),
                              torch.as_tensor(dealias_mask))
        return grid

    def _set_wavenumbers(self, kx, ky, kz, dealias_mask):
        self.kx, self.ky, self.kz = kx, ky, kz
        self.k_squared = kx * kx + ky * ky + kz * kz
        self.dealias_mask = dealias_mask

    @property
    def length(self):
        return self.domain_length if self.domain_length is not None else _TWO_PI

    def plan(self, dtype=torch.float64):
        _lib._require_gpu()
        return get_plan(self.n, dtype, False, self.length)

    # compact <-> full-cube conversions (edge plumbing)
    def _kept(self, device):
        n, M = self.n, self.n // 3
        idx = torch.cat([torch.arange(0, M + 1), torch.arange(n - M, n)]).to(device)
        return idx

    def full_to_compact(self, f_hat):
        idx = self._kept(f_hat.device)
        M = self.n // 3
        c = f_hat.index_select(-3, idx).index_select(-2, idx)[..., : M + 1]
        return c.contiguous()

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

    def to_spectral(self, field):
        """mask * fftn(field) as a full cube (grid.py:31-34)."""
        f = torch.as_tensor(field)
        dtype = torch.float64 if f.dtype not in (torch.float32,) else torch.float32
        p = self.plan(torch.float64)
        f = f.to(p.device, torch.float64).contiguous()
        return self.compact_to_full(p.forward(f))

    def to_real(self, field_hat):
        """ifftn(field_hat).real (grid.py:36-37) for a masked spectrum."""
        p = self.plan(torch.float64)
        fh = torch.as_tensor(field_hat).to(p.device, torch.complex128)
        c = self.full_to_compact(fh)
        _check_masked(fh, c)
        return p.inverse(c)

    def dealias(self, field_hat):
        return field_hat * self.dealias_mask.to(field_hat.device)


def _check_masked(full, compact):
    """The full-cube spectrum must have nothing outside the kept modes."""
    tot = float((full.abs() ** 2).sum())
    kz0 = compact[..., 0].abs() ** 2
    kept = float(kz0.sum() + 2.0 * (compact[..., 1:].abs() ** 2).sum())
    if abs(tot - kept) > 1e-20 + 1e-9 * tot:
        raise ValueError("spectrum has modes outside the 2/3 dealiasing mask (or is not Hermitian); "
                         "apply grid.dealias / pass a real field's spectrum")


def validate_timestep(dt, viscosity, resistivity, grid, max_velocity=1.0, max_field=0.5, cfl=0.5):
    """cuda/taylor_green_mhd/config.py:75-118."""
    if dt <= 0.0:
        raise ValueError("dt must be positive")
    if viscosity < 0.0:
        raise ValueError("viscosity must be nonnegative")
    if resistivity < 0.0:
        raise ValueError("resistivity must be nonnegative")
    if cfl <= 0.0:
        raise ValueError("cfl must be positive")
    if max_velocity <= 0.0:
        raise ValueError("max_velocity must be positive")
    if max_field <= 0.0:
        raise ValueError("max_field must be positive")
    kmax = float(max(grid.kx.abs().max(), grid.ky.abs().max(), grid.kz.abs().max()))
    if kmax == 0.0:
        raise ValueError("wavenumber arrays must contain nonzero modes")
    h = math.pi / kmax
    k2max = float(grid.k_squared.max())
    lims = [cfl * h / max_velocity, cfl * h / max_field,
            math.inf if viscosity == 0.0 else 2.0 / (viscosity * k2max),
            math.inf if resistivity == 0.0 else 2.0 / (resistivity * k2max)]
    limit = float(min(lims))
    if dt > limit:
        raise ValueError(f"dt={dt:g} exceeds stability limit {limit:g}; reduce dt or increase cfl")
    return limit


# ------------------------------------------------------------------ IC
class TaylorGreenMhdInitialCondition:
    """cuda/taylor_green_mhd/initial_conditions.py:9-21 (float32 fields, as
    the reference builds them); kind='config' gives b = -omega_TG =
    (cos x sin y sin z, sin x cos y sin z, -2 sin x sin y cos z)."""

    def __init__(self, domain_length=_TWO_PI, magnetic_amplitude=0.5, kind="reference", dtype=torch.float32):
        self.domain_length = domain_length
        self.magnetic_amplitude = magnetic_amplitude
        self.kind = kind
        self.dtype = dtype

    def create(self, n, device=None):
        dev = _dev(device)
        x = torch.tensor(np.linspace(0.0, self.domain_length, n, endpoint=False, dtype=np.float64),
                         dtype=self.dtype, device=dev)
        X, Y, Z = torch.meshgrid(x, x, x, indexing="ij")
        u = torch.sin(X) * torch.cos(Y) * torch.cos(Z)
        v = -torch.cos(X) * torch.sin(Y) * torch.cos(Z)
        vel = torch.stack((u, v, torch.zeros_like(u)))
        if self.kind == "reference":
            a = torch.tensor(self.magnetic_amplitude, dtype=self.dtype)
            mag = torch.stack((a * u, a * v, torch.zeros_like(u)))
        else:
            mag = torch.stack((torch.cos(X) * torch.sin(Y) * torch.sin(Z),
                               torch.sin(X) * torch.cos(Y) * torch.sin(Z),
                               -2.0 * torch.sin(X) * torch.sin(Y) * torch.cos(Z)))
        return vel, mag


# ------------------------------------------------------------------ operator / integrator
class MhdSpectralOperator:
    """cuda/taylor_green_mhd/operators.py:4-107 on full-cube spectra."""

    def __init__(self, grid, lorentz_sign=-1.0):
        self.grid = grid
        self.lorentz_sign = float(lorentz_sign)

    def _mhd_plan(self):
        _lib._require_gpu()
        return get_plan(self.grid.n, torch.float64, True, self.grid.length)

    def rhs(self, velocity_hat, magnetic_field_hat, viscosity, resistivity):
        p = self._mhd_plan()
        g = self.grid
        uh = g.full_to_compact(torch.as_tensor(velocity_hat).to(p.device, torch.complex128))
        bh = g.full_to_compact(torch.as_tensor(magnetic_field_hat).to(p.device, torch.complex128))
        ru, rb = p.rhs(uh, bh, nu=viscosity, eta=resistivity, lorentz_sign=self.lorentz_sign)
        return g.compact_to_full(ru), g.compact_to_full(rb)

    def spectral_curl_hat(self, field_hat):
        g = self.grid
        return g.dealias(torch.stack((1j * (g.ky * field_hat[2] - g.kz * field_hat[1]),
                                      1j * (g.kz * field_hat[0] - g.kx * field_hat[2]),
                                      1j * (g.kx * field_hat[1] - g.ky * field_hat[0]))))

    def spectral_curl(self, field_hat):
        p = self.grid.plan(torch.float64)
        c = self.grid.full_to_compact(torch.as_tensor(field_hat).to(p.device, torch.complex128))
        return p.inverse(c, op=_lib.TG_OP_CURL)

    def leray_project(self, vector_hat):
        g = self.grid
        k2 = torch.where(g.k_squared == 0.0, torch.ones_like(g.k_squared), g.k_squared)
        kd = g.kx * vector_hat[0] + g.ky * vector_hat[1] + g.kz * vector_hat[2]
        return vector_hat - torch.stack((g.kx * kd / k2, g.ky * kd / k2, g.kz * kd / k2))

    def divergence_hat(self, vector_hat):
        g = self.grid
        return 1j * (g.kx * vector_hat[0] + g.ky * vector_hat[1] + g.kz * vector_hat[2])

    def divergence_from_hat(self, vector_hat):
        p = self.grid.plan(torch.float64)
        c = self.grid.full_to_compact(torch.as_tensor(vector_hat).to(p.device, torch.complex128))
        return p.inverse(c, op=_lib.TG_OP_DIV)


class MhdRK4Integrator:
    """cuda/taylor_green_mhd/integrator.py:1-32: one classic RK4 step of
    (u_hat, b_hat), all four stages in the fused GPU passes."""

    def __init__(self, operator, dt, viscosity, resistivity):
        self.operator = operator
        self.dt = dt
        self.viscosity = viscosity
        self.resistivity = resistivity

    def step(self, state):
        op = self.operator
        p = op._mhd_plan()
        g = op.grid
        uh = g.full_to_compact(torch.as_tensor(state[0]).to(p.device, torch.complex128))
        bh = g.full_to_compact(torch.as_tensor(state[1]).to(p.device, torch.complex128))
        p.rk4(uh, bh, nsteps=1, dt=self.dt, nu=self.viscosity, eta=self.resistivity,
              lorentz_sign=op.lorentz_sign)
        return g.compact_to_full(uh), g.compact_to_full(bh)


class MhdDiagnostics:
    """cuda/taylor_green_mhd/diagnostics.py:4-49 (FP64 Parseval sums of the
    compact state; max |div| on the grid)."""

    def __init__(self, operator=None):
        self.operator = operator
        self.times = []
        self.kinetic_energies = []
        self.magnetic_energies = []
        self.total_energies = []
        self.cross_helicities = []
        self.enstrophies = []
        self.divergence_u_max = []
        self.divergence_b_max = []

    def record_values(self, time_value, d):
        self.times.append(time_value)
        self.kinetic_energies.append(d["kinetic_energy"])
        self.magnetic_energies.append(d["magnetic_energy"])
        self.total_energies.append(d["kinetic_energy"] + d["magnetic_energy"])
        self.cross_helicities.append(d["cross_helicity"])
        self.enstrophies.append(d["enstrophy"])
        self.divergence_u_max.append(d["div_u_max"])
        self.divergence_b_max.append(d["div_b_max"])
        return d["kinetic_energy"], d["magnetic_energy"], d["cross_helicity"], d["enstrophy"]

    def record(self, time_value, velocity_hat, magnetic_field_hat):
        g = self.operator.grid
        p = get_plan(g.n, torch.float64, True, g.length)
        uh = g.full_to_compact(torch.as_tensor(velocity_hat).to(p.device, torch.complex128))
        bh = g.full_to_compact(torch.as_tensor(magnetic_field_hat).to(p.device, torch.complex128))
        return self.record_values(time_value, p.diagnostics(uh, bh))


class TaylorGreenMhdSimulation:
    """cuda/taylor_green_mhd/simulation.py:15-100; the state stays compact on
    the GPU between saves and the steps between saves run as one call."""

    def __init__(self, config=None, make_plots=True):
        self.config = config or TaylorGreenMhdConfig()
        c = self.config
        _lib._require_gpu()
        self.grid = SpectralGrid(c.n, c.domain_length)
        self.initial_condition = TaylorGreenMhdInitialCondition(c.domain_length, c.magnetic_amplitude,
                                                                kind=c.initial_condition)
        self.operator = MhdSpectralOperator(self.grid, c.lorentz_sign)
        self.integrator = MhdRK4Integrator(self.operator, c.dt, c.viscosity, c.resistivity)
        dtype = torch.float64 if c.precision == "fp64" else torch.float32
        self.solver = MhdSolver(c.n, c.viscosity, c.resistivity, dtype=dtype, L=c.domain_length,
                                lorentz_sign=c.lorentz_sign)
        vel, mag = self.initial_condition.create(c.n, device=self.solver.plan.device)
        self.solver.set_fields(vel, mag)
        self.diagnostics = MhdDiagnostics(self.operator)
        self.make_plots = make_plots

    @property
    def velocity_hat(self):
        return self.grid.compact_to_full(self.solver.u_hat.to(torch.complex128))

    @property
    def magnetic_field_hat(self):
        return self.grid.compact_to_full(self.solver.b_hat.to(torch.complex128))

    def record(self, step):
        k, m, h, z = self.diagnostics.record_values(step * self.config.dt, self.solver.diagnostics())
        print(f"Save time step: {step}, time: {step * self.config.dt:.6f}, "
              f"E_k: {k:.8e}, E_m: {m:.8e}, H_c: {h:.8e}, enstrophy: {z:.8e}")
        return k, m, h, z

    def run(self):
        c = self.config
        max_field = c.magnetic_amplitude if c.initial_condition == "reference" else 2.0
        validate_timestep(c.dt, c.viscosity, c.resistivity, self.grid, max_velocity=1.0, max_field=max_field)
        self.diagnostics.record_values(0.0, self.solver.diagnostics())
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
        print(f"Reynolds number: Re = {c.reynolds:.1f}")
        print(f"Magnetic Reynolds number: Rm = {c.magnetic_reynolds:.1f}")
        print(f"Simulated time: {c.total_time:.3f} ({c.steps} steps) in {elapsed:.3f} s")
        d = self.diagnostics
        print(f"Kinetic energy trace: {d.kinetic_energies[0]:.6f} -> {d.kinetic_energies[-1]:.6f}")
        print(f"Magnetic energy trace: {d.magnetic_energies[0]:.6f} -> {d.magnetic_energies[-1]:.6f}")
        print(f"Max divergence u: {max(d.divergence_u_max):.3e}, B: {max(d.divergence_b_max):.3e}")
        if self.make_plots:
            _save_plots(self)
        return self.velocity_hat, self.magnetic_field_hat, self.diagnostics


def _save_plots(sim):
    """cuda/taylor_green_mhd/plotting.py:9-47 (PNG side effect)."""
    try:
        import matplotlib
        matplotlib.use("Agg")
        import matplotlib.pyplot as plt
    except ImportError:  # pragma: no cover
        return
    n = sim.config.n
    u, b = sim.solver.fields()
    us, bs = u[:, :, :, n // 2].cpu().numpy(), b[:, :, :, n // 2].cpu().numpy()
    fig, axes = plt.subplots(2, 3, figsize=(12, 8))
    for i in range(3):
        for row, (f, lab) in enumerate(((us, "u"), (bs, "B"))):
            im = axes[row, i].imshow(f[i].T, origin="lower", cmap="viridis")
            axes[row, i].set_title(f"{lab}{i + 1} slice")
            fig.colorbar(im, ax=axes[row, i])
    plt.tight_layout()
    plt.savefig("taylor_green_mhd_slice.png", dpi=150)
    plt.close(fig)
    d = sim.diagnostics
    fig, (a1, a2) = plt.subplots(2, 1, figsize=(10, 8))
    a1.plot(d.times, d.kinetic_energies, label="E_k")
    a1.plot(d.times, d.magnetic_energies, label="E_m")
    a1.plot(d.times, d.total_energies, label="E_total")
    a1.legend()
    a1.grid(True, alpha=0.3)
    a2.plot(d.times, d.cross_helicities, label="H_c")
    a2.legend()
    a2.grid(True, alpha=0.3)
    plt.tight_layout()
    plt.savefig("taylor_green_mhd_diagnostics.png", dpi=150)
    plt.close(fig)
    print("Saved: taylor_green_mhd_slice.png")
    print("Saved: taylor_green_mhd_diagnostics.png")


# ------------------------------------------------------------------ functional shim
def make_wavenumbers(n, length):
    """cuda/taylor_green_mhd.py:12-15."""
    g = SpectralGrid(n, length)
    return g.kx, g.ky, g.kz


def make_dealias_mask(n, length):
    """cuda/taylor_green_mhd.py:18-20."""
    return SpectralGrid(n, length).dealias_mask


def initial_taylor_green_mhd(n, length=_TWO_PI, magnetic_amplitude=0.5):
    """cuda/taylor_green_mhd.py:23-25 (float32 physical fields)."""
    return TaylorGreenMhdInitialCondition(length, magnetic_amplitude).create(n)


def initial_taylor_green_mhd_spectral(n, length=_TWO_PI, magnetic_amplitude=0.5):
    """cuda/taylor_green_mhd.py:28-33."""
    grid = SpectralGrid(n, length)
    vel, mag = TaylorGreenMhdInitialCondition(length, magnetic_amplitude).create(n)
    return grid.to_spectral(vel), grid.to_spectral(mag)


def spectral_rhs_mhd(velocity_hat, magnetic_field_hat, viscosity, resistivity, kx, ky, kz, dealias_mask,
                     lorentz_sign=-1.0):
    """cuda/taylor_green_mhd.py:40-46."""
    grid = SpectralGrid.from_arrays(kx, ky, kz, dealias_mask)
    return MhdSpectralOperator(grid, lorentz_sign).rhs(velocity_hat, magnetic_field_hat, viscosity, resistivity)


def advance_one_step(velocity_hat, magnetic_field_hat, dt, viscosity, resistivity, kx, ky, kz, dealias_mask,
                     lorentz_sign=-1.0):
    """cuda/taylor_green_mhd.py:49-59 (the result is dealiased, as there)."""
    grid = SpectralGrid.from_arrays(kx, ky, kz, dealias_mask)
    integ = MhdRK4Integrator(MhdSpectralOperator(grid, lorentz_sign), dt, viscosity, resistivity)
    return integ.step((velocity_hat, magnetic_field_hat))


def kinetic_energy(velocity):
    """cuda/taylor_green_mhd.py:62-64."""
    v = torch.as_tensor(velocity).double()
    return float(0.5 * torch.mean(torch.sum(v * v, dim=0)))


def magnetic_energy(magnetic_field):
    """cuda/taylor_green_mhd.py:67-69."""
    return kinetic_energy(magnetic_field)


def run_simulation(n=128, dt=0.005, reynolds=1000.0, magnetic_reynolds=1000.0, magnetic_amplitude=0.5,
                   total_time=1.0, save_every_time=0.05, make_plots=True, **extra):
    """cuda/taylor_green_mhd.py:72-99: returns (u, b, diagnostics) with u, b
    physical fields."""
    config = TaylorGreenMhdConfig(n=n, dt=dt, reynolds=reynolds, magnetic_reynolds=magnetic_reynolds,
                                  magnetic_amplitude=magnetic_amplitude, total_time=total_time,
                                  save_every_time=save_every_time, **extra)
    sim = TaylorGreenMhdSimulation(config, make_plots=make_plots)
    _, _, diagnostics = sim.run()
    u, b = sim.solver.fields()
    return u, b, diagnostics
