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
from .taylor_green_cuda import SpectralGrid, check_finite, rk4_linear

_TWO_PI = 2.0 * math.pi


# ------------------------------------------------------------------ config
@dataclass(frozen=True)
class TaylorGreenMhdConfig:
    """cuda/taylor_green_mhd/config.py:7-50 (+ lorentz_sign, initial_condition,
    dtype selectors with reference-compatible defaults).

    With lorentz_sign = -1 (the reference code's sign) the state
    initial_condition="config" (b0 = -curl u_TG) is linearly unstable: at
    n=128, nu = eta = 1/200, dt = 2e-3 a run from the float32 initial arrays
    diverges within 300 steps in FP32 and in FP64, and the run loop raises
    FloatingPointError.  lorentz_sign = +1 (the documented equation) keeps it
    finite."""
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
    """cuda/taylor_green_mhd/operators.py:4-107 on full-cube spectra.  rhs
    runs the fused MHD passes on the Hermitian kept part of (u_hat, b_hat);
    a remainder (anti-Hermitian or outside the mask; zero for the spectra
    of real fields) enters only the viscous / resistive terms, exactly as
    in the reference (see taylor_green_cuda.SpectralGrid.split)."""

    def __init__(self, grid, lorentz_sign=-1.0):
        self.grid = grid
        self.lorentz_sign = float(lorentz_sign)

    def _mhd_plan(self):
        _lib._require_gpu()
        return get_plan(self.grid.n, torch.float64, True, self.grid.length)

    def gradients(self, field_hat):
        """operators.py:8-14: dealiased i k_d field_hat[c] (elementwise)."""
        g = self.grid
        waves = (g.kx, g.ky, g.kz)
        return tuple(tuple(g.dealias(1j * g._on(w, field_hat.device) * field_hat[c]) for w in waves)
                     for c in range(3))

    def rhs(self, velocity_hat, magnetic_field_hat, viscosity, resistivity):
        p = self._mhd_plan()
        g = self.grid
        uf = torch.as_tensor(velocity_hat).to(p.device, torch.complex128)
        bf = torch.as_tensor(magnetic_field_hat).to(p.device, torch.complex128)
        uh, ur = g.split(uf)
        bh, br = g.split(bf)
        ru, rb = p.rhs(uh, bh, nu=viscosity, eta=resistivity, lorentz_sign=self.lorentz_sign)
        ru, rb = g.compact_to_full(ru), g.compact_to_full(rb)
        if ur is not None:
            ru = ru - viscosity * g.k2_like(ur) * g.mask_like(ur) * ur
        if br is not None:
            rb = rb - resistivity * g.k2_like(br) * g.mask_like(br) * br
        return ru, rb

    def spectral_curl_hat(self, field_hat):
        """operators.py:16-22 (elementwise)."""
        gr = self.gradients(field_hat)
        return self.grid.dealias(torch.stack((gr[2][1] - gr[1][2], gr[0][2] - gr[2][0], gr[1][0] - gr[0][1])))

    def spectral_curl(self, field_hat):
        """operators.py:24-25: to_real(spectral_curl_hat(field_hat))."""
        p = self.grid.plan(torch.float64)
        fh = torch.as_tensor(field_hat).to(p.device, torch.complex128)
        c, rest = self.grid.split(self.grid.dealias(fh))
        return p.inverse(c, op=_lib.TG_OP_CURL)

    def leray_project(self, vector_hat):
        """operators.py:27-38 (elementwise)."""
        g = self.grid
        dev = vector_hat.device
        kx, ky, kz, k2 = (g._on(t, dev) for t in (g.kx, g.ky, g.kz, g.k_squared))
        k2s = torch.where(k2 == 0.0, torch.ones_like(k2), k2)
        kd = kx * vector_hat[0] + ky * vector_hat[1] + kz * vector_hat[2]
        return vector_hat - torch.stack((kx * kd / k2s, ky * kd / k2s, kz * kd / k2s))

    def divergence_hat(self, vector_hat):
        """operators.py:40-45 (elementwise)."""
        g = self.grid
        dev = vector_hat.device
        return 1j * (g._on(g.kx, dev) * vector_hat[0] + g._on(g.ky, dev) * vector_hat[1]
                     + g._on(g.kz, dev) * vector_hat[2])

    def divergence_clean(self, vector_hat):
        """operators.py:47-48."""
        return self.leray_project(vector_hat)

    def divergence_from_hat(self, vector_hat):
        """operators.py:50-51: to_real(i k . vector_hat) for a spectrum inside the mask."""
        p = self.grid.plan(torch.float64)
        fh = torch.as_tensor(vector_hat).to(p.device, torch.complex128)
        c, rest = self.grid.split(fh)
        self.grid.require_in_mask(fh, rest)
        return p.inverse(c, op=_lib.TG_OP_DIV)

    def spectral_divergence(self, vector):
        """operators.py:53-55: divergence of a physical vector field via its
        unmasked spectrum, for band-limited fields (ValueError otherwise:
        the pruned transforms hold only the kept modes)."""
        from .taylor_green_jax import _require_band_limited
        v = torch.as_tensor(vector)
        dtype = torch.float32 if v.dtype == torch.float32 else torch.float64
        p = self.grid.plan(dtype)
        v = v.to(p.device, dtype).contiguous()
        c = p.forward(v)
        _require_band_limited(p, v, c)
        return p.inverse(c, op=_lib.TG_OP_DIV)


class MhdRK4Integrator:
    """cuda/taylor_green_mhd/integrator.py:1-32: one classic RK4 step of
    (u_hat, b_hat), all four stages in the fused GPU passes; the result is
    not dealiased, as there (a remainder is carried as the reference
    carries it)."""

    def __init__(self, operator, dt, viscosity, resistivity):
        self.operator = operator
        self.dt = dt
        self.viscosity = viscosity
        self.resistivity = resistivity

    def _step(self, state, dealias):
        op = self.operator
        p = op._mhd_plan()
        g = op.grid
        uf = torch.as_tensor(state[0]).to(p.device, torch.complex128)
        bf = torch.as_tensor(state[1]).to(p.device, torch.complex128)
        uh, ur = g.split(uf)
        bh, br = g.split(bf)
        p.rk4(uh, bh, nsteps=1, dt=self.dt, nu=self.viscosity, eta=self.resistivity,
              lorentz_sign=op.lorentz_sign)
        out = [g.compact_to_full(uh), g.compact_to_full(bh)]
        for i, (rest, coef) in enumerate(((ur, self.viscosity), (br, self.resistivity))):
            if rest is not None:
                m = g.mask_like(rest)
                kept = rk4_linear(m * rest, -coef * g.k2_like(rest) * m, self.dt)
                out[i] = out[i] + (kept if dealias else kept + (1.0 - m) * rest)
        return out[0], out[1]

    def step(self, state):
        return self._step(state, dealias=False)


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
        uf = torch.as_tensor(velocity_hat).to(p.device, torch.complex128)
        bf = torch.as_tensor(magnetic_field_hat).to(p.device, torch.complex128)
        uh, ur = g.split(uf)
        bh, br = g.split(bf)
        g.require_in_mask(uf, ur)
        g.require_in_mask(bf, br)
        return self.record_values(time_value, p.diagnostics(uh, bh))


class TaylorGreenMhdSimulation:
    """cuda/taylor_green_mhd/simulation.py:15-100; the state stays compact on
    the GPU between saves and the steps between saves run as one call."""

    def __init__(self, config=None, make_plots=True, check_finite=True):
        self.check_finite = check_finite
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

    def _diag(self, step):
        d = self.solver.diagnostics()
        if self.check_finite:
            check_finite(d, step, step * self.config.dt)
        return d

    def record(self, step):
        k, m, h, z = self.diagnostics.record_values(step * self.config.dt, self._diag(step))
        print(f"Save time step: {step}, time: {step * self.config.dt:.6f}, "
              f"E_k: {k:.8e}, E_m: {m:.8e}, H_c: {h:.8e}, enstrophy: {z:.8e}")
        return k, m, h, z

    def run(self):
        c = self.config
        max_field = c.magnetic_amplitude if c.initial_condition == "reference" else 2.0
        validate_timestep(c.dt, c.viscosity, c.resistivity, self.grid, max_velocity=1.0, max_field=max_field)
        self.diagnostics.record_values(0.0, self._diag(0))
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
    """cuda/taylor_green_mhd/plotting.py:9-47 (PNG side effect, same files,
    panels and labels)."""
    try:
        import matplotlib
        matplotlib.use("Agg")
        import matplotlib.pyplot as plt
    except ImportError:  # pragma: no cover
        print("matplotlib unavailable: plots skipped")
        return
    n = sim.config.n
    u, b = sim.solver.fields()
    velocity_slice = u[:, :, :, n // 2].double().cpu().numpy()
    magnetic_slice = b[:, :, :, n // 2].double().cpu().numpy()
    figure, axes = plt.subplots(2, 3, figsize=(14, 8))
    for component in range(3):
        image = axes[0, component].imshow(velocity_slice[component].T, origin="lower", cmap="viridis")
        axes[0, component].set_title(f"u{component + 1} slice")
        figure.colorbar(image, ax=axes[0, component])
        image = axes[1, component].imshow(magnetic_slice[component].T, origin="lower", cmap="plasma")
        axes[1, component].set_title(f"B{component + 1} slice")
        figure.colorbar(image, ax=axes[1, component])
    figure.tight_layout()
    figure.savefig("taylor_green_mhd_slice.png", dpi=200)
    plt.close(figure)
    d = sim.diagnostics
    figure, axes = plt.subplots(2, 2, figsize=(12, 10))
    panels = ((0, 0, d.kinetic_energies, "b-o", "Kinetic energy"), (0, 1, d.magnetic_energies, "r-o", "Magnetic energy"),
              (1, 0, d.total_energies, "g-o", "Total energy"), (1, 1, d.cross_helicities, "m-s", "Cross helicity"))
    for r, c, values, style, label in panels:
        axes[r, c].plot(d.times, values, style, markersize=4)
        axes[r, c].set_ylabel(label)
        axes[r, c].grid(True, alpha=0.3)
        if r == 1:
            axes[r, c].set_xlabel("Time")
    figure.tight_layout()
    figure.savefig("taylor_green_mhd_diagnostics.png", dpi=200)
    plt.close(figure)
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
    return integ._step((velocity_hat, magnetic_field_hat), dealias=True)


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


def main(argv=None):
    """cuda/taylor_green_mhd/cli.py:7-29 (python -m jax_tg_project_amd.taylor_green_mhd), plus
    --lorentz-sign, --initial-condition and --precision selectors."""
    import argparse
    parser = argparse.ArgumentParser(description="3D Taylor-Green MHD vortex on the MI355X HIP kernels")
    parser.add_argument("--n", type=int, default=128)
    parser.add_argument("--dt", type=float, default=0.005)
    parser.add_argument("--reynolds", type=float, default=1000.0)
    parser.add_argument("--magnetic-reynolds", type=float, default=1000.0)
    parser.add_argument("--magnetic-amplitude", type=float, default=0.5)
    parser.add_argument("--total-time", type=float, default=1.0)
    parser.add_argument("--save-every-time", type=float, default=0.05)
    parser.add_argument("--no-plots", action="store_true")
    parser.add_argument("--lorentz-sign", type=float, default=-1.0)
    parser.add_argument("--initial-condition", choices=("reference", "config"), default="reference")
    parser.add_argument("--precision", choices=("fp64", "fp32"), default="fp64")
    args = parser.parse_args(argv)
    config = TaylorGreenMhdConfig(n=args.n, dt=args.dt, reynolds=args.reynolds,
                                  magnetic_reynolds=args.magnetic_reynolds,
                                  magnetic_amplitude=args.magnetic_amplitude, total_time=args.total_time,
                                  save_every_time=args.save_every_time, lorentz_sign=args.lorentz_sign,
                                  initial_condition=args.initial_condition, precision=args.precision)
    return TaylorGreenMhdSimulation(config, make_plots=not args.no_plots).run()


if __name__ == "__main__":
    main()
