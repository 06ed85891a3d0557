"""FP64 numpy/scipy restatement of the reference pseudo-spectral solver.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  Every function names
the reference lines it restates.  Physical layout is ``(3, N, N, N)`` with
axes (component, x, y, z), C order, exactly as the reference.
"""
import math
import os

import numpy as np

try:  # scipy.fft supports a thread count; numpy.fft is the fallback.
    import scipy.fft as _sfft
except ImportError:  # pragma: no cover
    _sfft = None

_AXES = (-3, -2, -1)
_WORKERS = [1]


def set_workers(n):
    """Number of threads scipy.fft uses for the oracle's 3D transforms."""
    _WORKERS[0] = max(1, int(n))


def _fftn(a):
    if _sfft is not None:
        return _sfft.fftn(a, axes=_AXES, workers=_WORKERS[0])
    return np.fft.fftn(a, axes=_AXES)


def _ifftn(a):
    if _sfft is not None:
        return _sfft.ifftn(a, axes=_AXES, workers=_WORKERS[0])
    return np.fft.ifftn(a, axes=_AXES)


# ---------------------------------------------------------------- hydro (JAX)
def reynolds_to_viscosity(reynolds, velocity_amplitude=1.0, domain_length=2.0 * np.pi):
    """jax/taylor_green_jax.py:17-33."""
    if reynolds <= 0:
        raise ValueError("Reynolds number must be positive")
    return float((velocity_amplitude * domain_length) / reynolds)


def viscosity_to_reynolds(nu, velocity_amplitude=1.0, domain_length=2.0 * np.pi):
    """jax/taylor_green_jax.py:36-52."""
    if nu < 0.0:
        raise ValueError("Viscosity must be nonnegative")
    if nu == 0.0:
        return float("inf")
    return float((velocity_amplitude * domain_length) / nu)


def make_wavenumbers(N, L):
    """jax/taylor_green_jax.py:55-59: 2*pi*fftfreq(N, L/N), meshgrid 'ij'."""
    k = 2.0 * np.pi * np.fft.fftfreq(N, d=L / N)
    return np.meshgrid(k, k, k, indexing="ij")


def make_dealias_mask(N, L):
    """jax/taylor_green_jax.py:62-67: inclusive cube mask |k_i| <= (N//3)*2*pi/L."""
    cutoff = (N // 3) * (2.0 * np.pi / L)
    k = 2.0 * np.pi * np.fft.fftfreq(N, d=L / N)
    m = np.abs(k) <= cutoff
    return m[:, None, None] * m[None, :, None] * m[None, None, :]


def initial_taylor_green(N, L=2.0 * np.pi, amp=1.0):
    """jax/taylor_green_jax.py:70-79."""
    x = np.linspace(0.0, L, N, endpoint=False)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    u = amp * np.sin(X) * np.cos(Y) * np.cos(Z)
    v = -amp * np.cos(X) * np.sin(Y) * np.cos(Z)
    return np.stack([u, v, np.zeros_like(u)], axis=0)


def spectral_gradient(field_hat, kx, ky, kz):
    """jax/taylor_green_jax.py:82-87."""
    return (_ifftn(1j * kx * field_hat).real,
            _ifftn(1j * ky * field_hat).real,
            _ifftn(1j * kz * field_hat).real)


def spectral_divergence(u, kx, ky, kz):
    """jax/taylor_green_jax.py:90-96."""
    u_hat = _fftn(u)
    return _ifftn(1j * (kx * u_hat[0] + ky * u_hat[1] + kz * u_hat[2])).real


def vorticity(u, kx, ky, kz):
    """jax/taylor_green_jax.py:99-109 (unmasked F(u))."""
    u_hat = _fftn(u)
    du = spectral_gradient(u_hat[0], kx, ky, kz)
    dv = spectral_gradient(u_hat[1], kx, ky, kz)
    dw = spectral_gradient(u_hat[2], kx, ky, kz)
    return np.stack([dw[1] - dv[2], du[2] - dw[0], dv[0] - du[1]], axis=0)


def enstrophy(u, kx, ky, kz):
    """jax/taylor_green_jax.py:112-115."""
    om = vorticity(u, kx, ky, kz)
    return 0.5 * np.mean(np.sum(om * om, axis=0))


def viscous_dissipation(u, nu, kx, ky, kz):
    """jax/taylor_green_jax.py:118-128 (normalised by N**6)."""
    u_hat = _fftn(u)
    g2 = np.sum(np.abs(1j * kx * u_hat) ** 2 + np.abs(1j * ky * u_hat) ** 2
                + np.abs(1j * kz * u_hat) ** 2, axis=0)
    return nu * np.sum(g2) / (u.shape[-1] ** 6)


def validate_timestep(dt, nu, kx, ky, kz, cfl=0.5, max_velocity=1.0):
    """jax/taylor_green_jax.py:131-155."""
    if dt <= 0.0:
        raise ValueError("dt must be positive")
    if nu < 0.0:
        raise ValueError("nu must be nonnegative")
    if cfl <= 0.0:
        raise ValueError("cfl must be positive")
    if max_velocity <= 0.0:
        raise ValueError("max_velocity must be positive")
    kmax = float(max(np.max(np.abs(kx)), np.max(np.abs(ky)), np.max(np.abs(kz))))
    if kmax == 0.0:
        raise ValueError("wavenumber arrays must contain nonzero modes")
    adv = cfl * (np.pi / kmax) / max_velocity
    k2max = float(np.max(kx * kx + ky * ky + kz * kz))
    diff = np.inf if nu == 0.0 else 2.0 / (nu * k2max)
    limit = float(min(adv, diff))
    if dt > limit:
        raise ValueError(f"dt={dt:g} exceeds stability limit {limit:g}; "
                         "reduce dt or increase cfl")
    return limit


def leray(v_hat, kx, ky, kz):
    """Projection n - k (k.n)/k2 with k2==0 -> 1: jax/taylor_green_jax.py:162-182."""
    k2 = kx * kx + ky * ky + kz * kz
    k2s = np.where(k2 == 0.0, 1.0, k2)
    kd = kx * v_hat[0] + ky * v_hat[1] + kz * v_hat[2]
    return v_hat - np.stack([kx * kd / k2s, ky * kd / k2s, kz * kd / k2s], axis=0)


def spectral_rhs(u, nu, kx, ky, kz, dealias_mask, form="advective"):
    """jax/taylor_green_jax.py:158-185.

    ``form="advective"`` is the reference (u_raw . grad(F^-1 mask F u), 18
    complex FFTs).  ``form="rotational"`` replaces the product by
    ``omega x u`` of the masked field (it equals the advective form after
    truncation and projection when 3*(N//3) < N); it is the form the
    ensemble sizes N=48, 96 are compared in.
    """
    u_hat = _fftn(u) * dealias_mask
    k2 = kx * kx + ky * ky + kz * kz
    if form == "advective":
        ux, uy, uz = u
        du = spectral_gradient(u_hat[0], kx, ky, kz)
        dv = spectral_gradient(u_hat[1], kx, ky, kz)
        dw = spectral_gradient(u_hat[2], kx, ky, kz)
        nlin = np.stack([ux * du[0] + uy * du[1] + uz * du[2],
                         ux * dv[0] + uy * dv[1] + uz * dv[2],
                         ux * dw[0] + uy * dw[1] + uz * dw[2]], axis=0)
    elif form == "rotational":
        ul = _ifftn(u_hat).real
        om = _ifftn(np.stack([1j * (ky * u_hat[2] - kz * u_hat[1]),
                              1j * (kz * u_hat[0] - kx * u_hat[2]),
                              1j * (kx * u_hat[1] - ky * u_hat[0])])).real
        nlin = np.cross(om, ul, axis=0)
    else:
        raise ValueError(form)
    nlin_hat = _fftn(nlin) * dealias_mask
    rhs_hat = -leray(nlin_hat, kx, ky, kz) - nu * k2 * u_hat
    return _ifftn(rhs_hat).real


def advance_one_step(u, dt, nu, kx, ky, kz, dealias_mask, form="advective"):
    """jax/taylor_green_jax.py:188-196 (classic RK4 on the physical state)."""
    r1 = spectral_rhs(u, nu, kx, ky, kz, dealias_mask, form)
    r2 = spectral_rhs(u + 0.5 * dt * r1, nu, kx, ky, kz, dealias_mask, form)
    r3 = spectral_rhs(u + 0.5 * dt * r2, nu, kx, ky, kz, dealias_mask, form)
    r4 = spectral_rhs(u + dt * r3, nu, kx, ky, kz, dealias_mask, form)
    return u + (dt / 6.0) * (r1 + 2.0 * r2 + 2.0 * r3 + r4)


def kinetic_energy(u):
    """jax/taylor_green_jax.py:199-201."""
    return 0.5 * np.mean(u[0] ** 2 + u[1] ** 2 + u[2] ** 2)


def run_simulation(N=16, dt=0.005, nu=None, reynolds=None, total_time=0.5,
                   save_every_time=0.05, cfl=0.5, return_diagnostics=False,
                   u0=None, form="advective", verbose=False):
    """jax/taylor_green_jax.py:204-317 without the plotting side effects.

    ``u0`` optionally replaces the Taylor-Green initial condition.
    """
    if N < 4 or total_time <= 0.0 or save_every_time <= 0.0:
        raise ValueError("N must be at least 4, total_time and save_every_time must be positive")
    if dt <= 0.0:
        raise ValueError("dt must be positive")
    n_steps = int(total_time / dt)
    save_every = max(1, int(save_every_time / dt))
    L = 2.0 * np.pi
    if reynolds is not None:
        nu = reynolds_to_viscosity(reynolds, 1.0, L)
    elif nu is None:
        nu = reynolds_to_viscosity(100, 1.0, L)
    kx, ky, kz = make_wavenumbers(N, L)
    mask = make_dealias_mask(N, L)
    u = initial_taylor_green(N, L=L, amp=1.0) if u0 is None else np.array(u0, dtype=np.float64)
    validate_timestep(dt, nu, kx, ky, kz, cfl, max_velocity=float(np.max(np.abs(u))))
    times = [0.0]
    energies = [float(kinetic_energy(u))]
    div = [float(np.max(np.abs(spectral_divergence(u, kx, ky, kz))))]
    ens = [float(enstrophy(u, kx, ky, kz))]
    dis = [float(viscous_dissipation(u, nu, kx, ky, kz))]
    for step in range(n_steps):
        u = advance_one_step(u, dt, nu, kx, ky, kz, mask, form)
        if (step + 1) % save_every == 0 or step == n_steps - 1:
            times.append((step + 1) * dt)
            energies.append(float(kinetic_energy(u)))
            div.append(float(np.max(np.abs(spectral_divergence(u, kx, ky, kz)))))
            ens.append(float(enstrophy(u, kx, ky, kz)))
            dis.append(float(viscous_dissipation(u, nu, kx, ky, kz)))
            if verbose:
                print(f"Save time step: {step + 1}, Enstrophy: {ens[-1]:.6e}")
    diag = {"time": times, "energy": energies, "divergence_max": div,
            "enstrophy": ens, "dissipation": dis}
    if return_diagnostics:
        return u, energies, diag
    return u, energies


# ------------------------------------------------------------------ MHD (CuPy)
class SpectralGrid:
    """cuda/taylor_green_cuda/grid.py:4-40."""

    def __init__(self, n, domain_length):
        self.n = n
        k = 2.0 * np.pi * np.fft.fftfreq(n, d=domain_length / n)
        self.kx, self.ky, self.kz = np.meshgrid(k, k, k, indexing="ij")
        cutoff = (n // 3) * (2.0 * np.pi / domain_length)
        keep = np.abs(k) <= cutoff
        self.dealias_mask = keep[:, None, None] & keep[None, :, None] & keep[None, None, :]
        self.k_squared = self.kx ** 2 + self.ky ** 2 + self.kz ** 2

    def to_spectral(self, field):
        return _fftn(field) * self.dealias_mask

    def to_real(self, field_hat):
        return _ifftn(field_hat).real

    def dealias(self, field_hat):
        return field_hat * self.dealias_mask


def cupy_rhs(velocity_hat, viscosity, grid):
    """cuda/taylor_green_cuda/operators.py:8-48: on the full-cube spectrum,
    dealias, u = ifftn(u_hat).real, gradients i k_d u_hat_c dealiased,
    n_c = sum_d u_d d_d u_c (the product of the DEALIASED velocity, unlike
    the JAX path's raw u), mask F(n), Leray projection with k2 == 0 -> 1,
    -P n_hat - nu k^2 u_hat."""
    kx, ky, kz = grid.kx, grid.ky, grid.kz
    u_hat = grid.dealias(velocity_hat)
    u = grid.to_real(u_hat)
    waves = (kx, ky, kz)
    nl = np.stack([sum(u[d] * grid.to_real(grid.dealias(1j * waves[d] * u_hat[c])) for d in range(3))
                   for c in range(3)])
    n_hat = _fftn(nl) * grid.dealias_mask
    return -leray(n_hat, kx, ky, kz) - viscosity * grid.k_squared * u_hat


def cupy_rk4_step(state, dt, viscosity, grid):
    """cuda/taylor_green_cuda/integrator.py:7-12 (no dealias afterwards)."""
    r1 = cupy_rhs(state, viscosity, grid)
    r2 = cupy_rhs(state + 0.5 * dt * r1, viscosity, grid)
    r3 = cupy_rhs(state + 0.5 * dt * r2, viscosity, grid)
    r4 = cupy_rhs(state + dt * r3, viscosity, grid)
    return state + (dt / 6.0) * (r1 + 2.0 * r2 + 2.0 * r3 + r4)


def cupy_diagnostics(state_hat, grid):
    """cuda/taylor_green_cuda/diagnostics.py:15-25 -> (E, Z): E from
    to_real(state), Z from the dealiased spectral curl (operators.py:16-22)."""
    u = grid.to_real(state_hat)
    om = grid.to_real(_curl_hat(grid, state_hat))
    return (float(0.5 * np.mean(np.sum(u * u, axis=0))), float(0.5 * np.mean(np.sum(om * om, axis=0))))


def cupy_run(n, dt, viscosity, total_time, save_every_time, u0=None, length=2.0 * np.pi):
    """cuda/taylor_green_cuda/simulation.py:28-52 without plots and prints:
    record(0), then per step RK4 + dealias, record when step % save_every
    == 0 or step == steps at time step * dt.  Returns (state, times,
    energies, enstrophies)."""
    steps = int(total_time / dt)
    save_every = max(1, int(save_every_time / dt))
    grid = SpectralGrid(n, length)
    if u0 is None:
        u0 = initial_taylor_green(n, length)
    state = grid.to_spectral(np.asarray(u0, dtype=np.float64))
    times, energies, enstrophies = [], [], []

    def record(t):
        e, z = cupy_diagnostics(state, grid)
        times.append(t)
        energies.append(e)
        enstrophies.append(z)

    record(0.0)
    for step in range(1, steps + 1):
        state = grid.dealias(cupy_rk4_step(state, dt, viscosity, grid))
        if step % save_every == 0 or step == steps:
            record(step * dt)
    return state, times, energies, enstrophies


def mhd_initial_condition(n, length=2.0 * np.pi, magnetic_amplitude=0.5, kind="reference"):
    """Physical (u, b).

    ``kind="reference"``: cuda/taylor_green_mhd/initial_conditions.py:9-21
    (b = amp * (u_TG, v_TG, 0)); the grid is built in float64 here so the
    oracle and the GPU get identical inputs.  ``kind="config"``: the MHD
    benchmark field b = (cos x sin y sin z, sin x cos y sin z,
    -2 sin x sin y cos z) (= -omega_TG).
    """
    x = np.linspace(0.0, length, n, endpoint=False)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    u = np.sin(X) * np.cos(Y) * np.cos(Z)
    v = -np.cos(X) * np.sin(Y) * np.cos(Z)
    vel = np.stack([u, v, np.zeros_like(u)])
    if kind == "reference":
        mag = np.stack([magnetic_amplitude * u, magnetic_amplitude * v, np.zeros_like(u)])
    elif kind == "config":
        mag = np.stack([np.cos(X) * np.sin(Y) * np.sin(Z),
                        np.sin(X) * np.cos(Y) * np.sin(Z),
                        -2.0 * np.sin(X) * np.sin(Y) * np.cos(Z)])
    else:
        raise ValueError(kind)
    return vel, mag


def _curl_hat(grid, f_hat):
    """cuda/taylor_green_mhd/operators.py:8-22 (every derivative dealiased)."""
    kx, ky, kz, d = grid.kx, grid.ky, grid.kz, grid.dealias
    return d(np.stack([d(1j * ky * f_hat[2]) - d(1j * kz * f_hat[1]),
                       d(1j * kz * f_hat[0]) - d(1j * kx * f_hat[2]),
                       d(1j * kx * f_hat[1]) - d(1j * ky * f_hat[0])]))


def _div_hat(grid, f_hat):
    """cuda/taylor_green_mhd/operators.py:40-45."""
    return 1j * (grid.kx * f_hat[0] + grid.ky * f_hat[1] + grid.kz * f_hat[2])


def mhd_rhs(u_hat, b_hat, viscosity, resistivity, grid, lorentz_sign=-1.0):
    """cuda/taylor_green_mhd/operators.py:57-107.

    ``lorentz_sign=-1`` is the reference code (rhs_u = -P(adv) - P(j x b),
    operators.py:88-92); ``+1`` is the documented equation
    cuda/README_MHD.md:9 (``+ (curl B) x B``).
    """
    kx, ky, kz = grid.kx, grid.ky, grid.kz
    u_hat = grid.dealias(u_hat)
    b_hat = grid.dealias(b_hat)
    u = grid.to_real(u_hat)
    b = grid.to_real(b_hat)
    waves = (kx, ky, kz)
    nl = np.stack([sum(u[d] * grid.to_real(grid.dealias(1j * waves[d] * u_hat[c]))
                       for d in range(3)) for c in range(3)])
    adv = leray(_fftn(nl) * grid.dealias_mask, kx, ky, kz)
    j = grid.to_real(_curl_hat(grid, b_hat))
    lor = leray(_fftn(np.cross(j, b, axis=0)) * grid.dealias_mask, kx, ky, kz)
    rhs_u = -adv + lorentz_sign * lor - viscosity * grid.k_squared * u_hat
    cr_hat = _fftn(np.cross(u, b, axis=0)) * grid.dealias_mask
    ind = _curl_hat(grid, cr_hat) * grid.dealias_mask
    rhs_b = leray(ind, kx, ky, kz) - resistivity * grid.k_squared * b_hat
    return rhs_u, rhs_b


def mhd_rk4_step(u_hat, b_hat, dt, viscosity, resistivity, grid, lorentz_sign=-1.0, dealias=True):
    """cuda/taylor_green_mhd/integrator.py:14-32 plus the shim's final dealias
    (cuda/taylor_green_mhd.py:58-59); ``dealias=False`` is the integrator's
    step alone."""
    def f(a, b):
        return mhd_rhs(a, b, viscosity, resistivity, grid, lorentz_sign)
    k1u, k1b = f(u_hat, b_hat)
    k2u, k2b = f(u_hat + 0.5 * dt * k1u, b_hat + 0.5 * dt * k1b)
    k3u, k3b = f(u_hat + 0.5 * dt * k2u, b_hat + 0.5 * dt * k2b)
    k4u, k4b = f(u_hat + dt * k3u, b_hat + dt * k3b)
    un = u_hat + (dt / 6.0) * (k1u + 2.0 * k2u + 2.0 * k3u + k4u)
    bn = b_hat + (dt / 6.0) * (k1b + 2.0 * k2b + 2.0 * k3b + k4b)
    if not dealias:
        return un, bn
    return grid.dealias(un), grid.dealias(bn)


def mhd_diagnostics(u_hat, b_hat, grid):
    """cuda/taylor_green_mhd/diagnostics.py:20-49 -> (E_k, E_m, H_c, Z, div_u, div_b)."""
    u = grid.to_real(u_hat)
    b = grid.to_real(b_hat)
    ek = 0.5 * np.mean(np.sum(u * u, axis=0))
    em = 0.5 * np.mean(np.sum(b * b, axis=0))
    hc = np.mean(u[0] * b[0] + u[1] * b[1] + u[2] * b[2])
    om = grid.to_real(_curl_hat(grid, u_hat))
    z = 0.5 * np.mean(np.sum(om * om, axis=0))
    du = float(np.max(np.abs(grid.to_real(_div_hat(grid, u_hat)))))
    db = float(np.max(np.abs(grid.to_real(_div_hat(grid, b_hat)))))
    return float(ek), float(em), float(hc), float(z), du, db


def mhd_run(n, dt, viscosity, resistivity, total_time, save_every_time, magnetic_amplitude=0.5,
            kind="reference", lorentz_sign=-1.0, u0=None, b0=None, length=2.0 * np.pi, max_field=None):
    """cuda/taylor_green_mhd/simulation.py:49-100 (TaylorGreenMhdSimulation.run
    without plots and prints), with the step count and save cadence of
    config.py:44-50: validate the time step, record at t = 0, then per step
    RK4 + dealias (integrator.py:14-32, simulation.py:63-66) and record when
    ``step % save_every == 0 or step == steps`` at time ``step * dt``.

    ``u0``/``b0`` (physical) replace the initial condition so the GPU and
    the oracle start from identical arrays; ``max_field`` defaults to the
    magnetic amplitude as the reference passes it.  Returns (u_hat, b_hat,
    records) with the MhdDiagnostics lists (diagnostics.py:4-49).
    """
    steps = int(total_time / dt)
    save_every = max(1, int(save_every_time / dt))
    grid = SpectralGrid(n, length)
    if u0 is None:
        u0, b0 = mhd_initial_condition(n, length, magnetic_amplitude, kind)
    uh = grid.to_spectral(np.asarray(u0, dtype=np.float64))
    bh = grid.to_spectral(np.asarray(b0, dtype=np.float64))
    mhd_validate_timestep(dt, viscosity, resistivity, grid, max_velocity=1.0,
                          max_field=magnetic_amplitude if max_field is None else max_field)
    rec = {k: [] for k in ("times", "kinetic_energies", "magnetic_energies", "total_energies",
                           "cross_helicities", "enstrophies", "divergence_u_max", "divergence_b_max")}

    def record(t):
        ek, em, hc, z, du, db = mhd_diagnostics(uh, bh, grid)
        for k, v in zip(("times", "kinetic_energies", "magnetic_energies", "total_energies",
                         "cross_helicities", "enstrophies", "divergence_u_max", "divergence_b_max"),
                        (t, ek, em, ek + em, hc, z, du, db)):
            rec[k].append(v)

    record(0.0)
    for step in range(1, steps + 1):
        uh, bh = mhd_rk4_step(uh, bh, dt, viscosity, resistivity, grid, lorentz_sign)
        if step % save_every == 0 or step == steps:
            record(step * dt)
    return uh, bh, rec


def mhd_validate_timestep(dt, viscosity, resistivity, grid, max_velocity=1.0,
                          max_field=0.5, cfl=0.5):
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
    kmax = float(np.max(np.abs(np.stack((grid.kx, grid.ky, grid.kz)))))
    if kmax == 0.0:
        raise ValueError("wavenumber arrays must contain nonzero modes")
    h = np.pi / kmax
    k2max = float(np.max(grid.k_squared))
    lims = [cfl * h / max_velocity, cfl * h / max_field,
            np.inf if viscosity == 0.0 else 2.0 / (viscosity * k2max),
            np.inf if resistivity == 0.0 else 2.0 / (resistivity * k2max)]
    limit = float(min(lims))
    if dt > limit:
        raise ValueError(f"dt={dt:g} exceeds stability limit {limit:g}; reduce dt or increase cfl")
    return limit


# ------------------------------------------------------ half-spectrum helpers
def kept_counts(N):
    """(M, K, KZ): cutoff N//3, kept modes per x/y axis, kept kz >= 0 modes."""
    M = N // 3
    return M, 2 * M + 1, M + 1


def kept_index(N):
    """Full FFT indices of the kept modes in compact order (0..M, N-M..N-1)."""
    M = N // 3
    return np.concatenate([np.arange(0, M + 1), np.arange(N - M, N)])


def full_to_compact(f_hat):
    """Full-cube spectrum (..., N, N, N) -> compact kept half spectrum (..., K, K, KZ)."""
    N = f_hat.shape[-1]
    idx = kept_index(N)
    M = N // 3
    return np.ascontiguousarray(f_hat[..., idx, :, :][..., :, idx, :][..., :, :, :M + 1])


def compact_to_full(c_hat, N):
    """Compact kept half spectrum -> full Hermitian cube (zeros outside the mask)."""
    M = N // 3
    idx = kept_index(N)
    lead = c_hat.shape[:-3]
    full = np.zeros(lead + (N, N, N), dtype=np.complex128)
    ix = idx[:, None, None]
    iy = idx[None, :, None]
    iz = np.arange(M + 1)[None, None, :]
    full[..., ix, iy, iz] = c_hat
    # Hermitian completion for kz < 0: X[-kx,-ky,-kz] = conj X[kx,ky,kz]
    nx = (-idx) % N
    izn = (-np.arange(1, M + 1)) % N
    full[..., nx[:, None, None], nx[None, :, None], izn[None, None, :]] = np.conj(c_hat[..., 1:])
    return full


def compact_wavenumbers(N, L=2.0 * np.pi):
    """Integer-valued wavenumbers of the compact layout, scaled by 2*pi/L."""
    M = N // 3
    k1 = np.concatenate([np.arange(0, M + 1), np.arange(-M, 0)]).astype(np.float64)
    kz1 = np.arange(0, M + 1).astype(np.float64)
    s = 2.0 * np.pi / L
    return np.meshgrid(k1 * s, k1 * s, kz1 * s, indexing="ij")


def shear_wave(N, A=1.0, L=2.0 * np.pi):
    """u = (0, A sin x, 0): an exact Navier-Stokes solution with zero nonlinear term."""
    x = np.linspace(0.0, L, N, endpoint=False)
    X, _, _ = np.meshgrid(x, x, x, indexing="ij")
    z = np.zeros((N, N, N))
    return np.stack([z, A * np.sin(X), z])


def rk4_amplification(z):
    """Classic RK4 stability polynomial R(z) = 1 + z + z^2/2 + z^3/6 + z^4/24."""
    return 1.0 + z + z * z / 2.0 + z ** 3 / 6.0 + z ** 4 / 24.0


def random_perturbation(N, rng, rms=0.05, kmax=4.0, L=2.0 * np.pi):
    """Seeded divergence-free perturbation on 0 < |k| <= kmax (complex Gaussian,
    Hermitian via a real field, Leray-projected), scaled to rms(du)=rms."""
    kx, ky, kz = make_wavenumbers(N, L)
    kk = np.sqrt(kx * kx + ky * ky + kz * kz)
    sel = (kk > 0) & (kk <= kmax)
    spec = (rng.standard_normal((3, N, N, N)) + 1j * rng.standard_normal((3, N, N, N))) * sel
    f = _ifftn(spec).real                     # real field: Hermitian part of the draw
    f_hat = leray(_fftn(f), kx, ky, kz) * sel
    du = _ifftn(f_hat).real
    cur = math.sqrt(float(np.mean(np.sum(du * du, axis=0)) / 3.0))
    return du * (rms / cur)


def cpu_count():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1
