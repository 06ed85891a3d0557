"""Drop-in for the reference JAX hydro module (jax/taylor_green_jax.py).

Same function names, signatures, return conventions and ValueError messages
as the reference; arrays are torch tensors (float64 unless the input is
float32) on the ROCm GPU.  Setup helpers (wavenumbers, mask, initial
condition, Re <-> nu, timestep validation) are host/plumbing code; every
transform, right-hand side, RK4 step and spectral diagnostic runs in the
gfx950 HIP library (libtgdns.so) and raises if it is unavailable.

Inside, the state is the compact 2/3-truncated half spectrum.  The physical
API entry points (spectral_rhs, advance_one_step, spectral_divergence,
vorticity, enstrophy, viscous_dissipation) transform their physical input
to it, which is exact for band-limited inputs (every state of the
reference's run loop: the TG initial condition and all RK4 updates lie
inside the mask).  An input whose part outside the mask exceeds round-off
(relative amplitude BAND_TOL) raises ValueError instead of silently
differing from the reference, which would transform it unmasked.

The compiled kernels form the nonlinear term as u x omega.  After the 2/3
truncation and the projection this equals the reference's u . grad u
whenever 3 (N // 3) < N.  Where 3 divides a compiled N (12, 24, 48, 96, 192,
384) the inclusive mask keeps the shell |k_i| = N/3 that the product aliases
into, and there the two forms differ, by about 1e-5 relative on a random
band-limited field.  Sizes without compiled kernels run the direct-sum path,
which forms the reference's advective term exactly.
"""
import math
import time

import numpy as np
import torch

from . import _lib
from .solver import TaylorGreenSolver, get_plan

_TWO_PI = 2.0 * math.pi


def _device(device=None):
    if device is not None:
        return torch.device(device)
    return torch.device("cuda" if torch.cuda.is_available() else "cpu")


def reynolds_to_viscosity(reynolds, velocity_amplitude=1.0, domain_length=_TWO_PI):
    """nu = U L / Re (jax/taylor_green_jax.py:17-33)."""
    if reynolds <= 0:
        raise ValueError("Reynolds number must be positive")
    return float((velocity_amplitude * domain_length) / reynolds)


def viscosity_to_reynolds(nu, velocity_amplitude=1.0, domain_length=_TWO_PI):
    """Re = U L / nu, inf for nu = 0 (jax/taylor_green_jax.py:36-52)."""
    if nu < 0.0:
        raise ValueError("Viscosity must be nonnegative")
    if nu == 0.0:
        return float("inf")
    return float((velocity_amplitude * domain_length) / nu)


def _k1d(N, L):
    # 2 pi * fftfreq(N, d=L/N) exactly as numpy computes it
    return 2.0 * np.pi * np.fft.fftfreq(N, d=L / N)


def make_wavenumbers(N, L, device=None):
    """(kx, ky, kz) meshgrid 'ij' of 2 pi fftfreq(N, L/N) (jax/taylor_green_jax.py:55-59)."""
    k = torch.tensor(_k1d(N, L), dtype=torch.float64, device=_device(device))
    kx, ky, kz = torch.meshgrid(k, k, k, indexing="ij")
    return kx, ky, kz


def make_dealias_mask(N, L, device=None):
    """Inclusive cube 2/3 mask |k_i| <= (N//3) 2 pi / L (jax/taylor_green_jax.py:62-67)."""
    cutoff = (N // 3) * (2.0 * np.pi / L)
    m = torch.tensor(np.abs(_k1d(N, L)) <= cutoff, device=_device(device))
    return m[:, None, None] & m[None, :, None] & m[None, None, :]


def initial_taylor_green(N, L=_TWO_PI, amp=1.0, device=None):
    """u = amp sin x cos y cos z, v = -amp cos x sin y cos z, w = 0 on the
    endpoint-free grid (jax/taylor_green_jax.py:70-79)."""
    x = torch.tensor(np.linspace(0.0, L, N, endpoint=False), dtype=torch.float64, device=_device(device))
    X, Y, Z = torch.meshgrid(x, x, x, indexing="ij")
    u = amp * torch.sin(X) * torch.cos(Y) * torch.cos(Z)
    v = -amp * torch.cos(X) * torch.sin(Y) * torch.cos(Z)
    return torch.stack([u, v, torch.zeros_like(u)], dim=0)


def validate_timestep(dt, nu, kx, ky, kz, cfl=0.5, max_velocity=1.0):
    """CFL and diffusive limits (jax/taylor_green_jax.py:131-155)."""
    if dt <= 0.0:
        raise ValueError("dt must be positive")
    if nu < 0.0:
        raise ValueError("nu must be nonnegative")
    if cfl <= 0.0:
        raise ValueError("cfl must be positive")
    if max_velocity <= 0.0:
        raise ValueError("max_velocity must be positive")
    kx, ky, kz = (torch.as_tensor(k) for k in (kx, ky, kz))
    max_wavenumber = float(max(kx.abs().max(), ky.abs().max(), kz.abs().max()))
    if max_wavenumber == 0.0:
        raise ValueError("wavenumber arrays must contain nonzero modes")
    grid_spacing = math.pi / max_wavenumber
    advective_limit = cfl * grid_spacing / max_velocity
    k2_max = float((kx * kx + ky * ky + kz * kz).max())
    diffusive_limit = math.inf if nu == 0.0 else 2.0 / (nu * k2_max)
    limit = float(min(advective_limit, diffusive_limit))
    if dt > limit:
        raise ValueError(f"dt={dt:g} exceeds stability limit {limit:g}; reduce dt or increase cfl")
    return limit


# ------------------------------------------------------------ physical API
def grid_from_arrays(kx, ky, kz, dealias_mask=None):
    """(N, L) of canonical wavenumber arrays (and mask); ValueError otherwise.
    N is kx.shape[0] and L = 2 pi / kx[1,0,0], as SpectralGrid.from_arrays
    (cuda/taylor_green_cuda/grid.py:15-22) takes them."""
    kx = torch.as_tensor(kx)
    N = int(kx.shape[0])
    if tuple(kx.shape) != (N, N, N):
        raise ValueError("wavenumber arrays must have shape (N, N, N)")
    k1 = float(kx[1, 0, 0]) if N > 1 else 0.0
    if not k1 > 0.0:
        raise ValueError("wavenumber arrays must contain nonzero modes")
    L = _TWO_PI / k1
    want = torch.tensor(_k1d(N, L), dtype=torch.float64)
    for arr, sl in ((kx, (slice(None), 0, 0)), (torch.as_tensor(ky), (0, slice(None), 0)),
                    (torch.as_tensor(kz), (0, 0, slice(None)))):
        got = arr[sl].detach().to("cpu", torch.float64)
        if not torch.allclose(got, want, rtol=1e-12, atol=1e-12 * abs(k1)):
            raise ValueError("wavenumber arrays are not the canonical 2*pi*fftfreq(N, L/N) grids")
    if dealias_mask is not None:
        m = torch.as_tensor(dealias_mask)
        M = N // 3
        want_m = torch.tensor(np.abs(np.fft.fftfreq(N) * N) <= M)
        if tuple(m.shape) != (N, N, N) or not (
                torch.equal(m[:, 0, 0].cpu().bool(), want_m) and torch.equal(m[0, :, 0].cpu().bool(), want_m)
                and torch.equal(m[0, 0, :].cpu().bool(), want_m)
                and int(m.sum()) == int(want_m.sum()) ** 3):
            raise ValueError("dealias_mask is not the 2/3-rule cube mask of this grid")
    return N, L


def _as_state(u):
    t = torch.as_tensor(u)
    if t.dtype not in (torch.float64, torch.float32):
        t = t.to(torch.float64)
    if t.device.type != "cuda":
        _lib._require_gpu()
        t = t.to("cuda")
    return t.contiguous()


# Relative L2 amplitude of an input's part outside the 2/3 mask above which
# the physical-API functions refuse it: the round-off of a forward + inverse
# transform pair stays ~100x below these.
BAND_TOL = {torch.float64: 1e-11, torch.float32: 2e-6}


def _require_band_limited(plan, u, u_hat):
    """The field must equal the inverse transform of its kept modes: compare
    amplitudes, || u - F^-1(mask F u) || <= BAND_TOL || u ||."""
    back = plan.inverse(u_hat) if u.dim() == 4 else None
    if back is None:
        raise ValueError("expected a (3, N, N, N) field")
    num = float(torch.linalg.vector_norm((u - back).double()))
    den = float(torch.linalg.vector_norm(u.double()))
    if num > BAND_TOL[plan.dtype] * max(den, 1e-300):
        raise ValueError(
            "input field has modes outside the 2/3 dealiasing mask (relative amplitude "
            f"{num / max(den, 1e-300):.3e}); the pruned transforms advance band-limited states only")


def _to_spectral(u, kx, ky, kz, dealias_mask=None, check=True):
    N, L = grid_from_arrays(kx, ky, kz, dealias_mask)
    u = _as_state(u)
    if tuple(u.shape) != (3, N, N, N):
        raise ValueError(f"velocity must have shape (3, {N}, {N}, {N})")
    plan = get_plan(N, u.dtype, False, L, u.device)
    u_hat = plan.forward(u)
    if check:
        _require_band_limited(plan, u, u_hat)
    return plan, u_hat


def spectral_gradient(field_hat, kx, ky, kz):
    """(ifftn(i kx f).real, ifftn(i ky f).real, ifftn(i kz f).real) of one
    scalar full-cube spectrum (jax/taylor_green_jax.py:82-87).  Only the
    Hermitian part of f enters .real; its kept modes go through the HIP
    gradient transform (TG_OP_GRAD).  A spectrum with content outside the
    2/3 mask raises ValueError (the pruned transforms hold kept modes only)."""
    from .taylor_green_cuda import SpectralGrid
    N, L = grid_from_arrays(kx, ky, kz)
    f = torch.as_tensor(field_hat)
    if tuple(f.shape) != (N, N, N):
        raise ValueError(f"field_hat must have shape ({N}, {N}, {N})")
    dtype = torch.float32 if f.dtype == torch.complex64 else torch.float64
    _lib._require_gpu()
    plan = get_plan(N, dtype, False, L)
    grid = SpectralGrid(N, L, device=plan.device)
    f = f.to(plan.device, plan.cdtype)
    c, rest = grid.split(f)
    grid.require_in_mask(f, rest)
    g = plan.inverse(c, op=_lib.TG_OP_GRAD)
    return g[0], g[1], g[2]


def spectral_divergence(u, kx, ky, kz):
    """Physical divergence field (jax/taylor_green_jax.py:90-96)."""
    plan, u_hat = _to_spectral(u, kx, ky, kz)
    return plan.inverse(u_hat, op=_lib.TG_OP_DIV)


def vorticity(u, kx, ky, kz):
    """Physical vorticity (jax/taylor_green_jax.py:99-109)."""
    plan, u_hat = _to_spectral(u, kx, ky, kz)
    return plan.inverse(u_hat, op=_lib.TG_OP_CURL)


def enstrophy(u, kx, ky, kz):
    """1/2 <|curl u|^2> (jax/taylor_green_jax.py:112-115), FP64 Parseval sum."""
    plan, u_hat = _to_spectral(u, kx, ky, kz)
    return plan.diagnostics(u_hat, divmax=False)["enstrophy"]


def viscous_dissipation(u, nu, kx, ky, kz):
    """nu <|grad u|^2> (jax/taylor_green_jax.py:118-128), FP64 Parseval sum."""
    plan, u_hat = _to_spectral(u, kx, ky, kz)
    return nu * plan.diagnostics(u_hat, divmax=False)["grad_sq"]


def spectral_rhs(u, nu, kx, ky, kz, dealias_mask):
    """-P(mask F(u.grad u)) - nu k^2 mask F(u), returned in physical space
    (jax/taylor_green_jax.py:158-185)."""
    plan, u_hat = _to_spectral(u, kx, ky, kz, dealias_mask)
    return plan.inverse(plan.rhs(u_hat, nu=float(nu)))


def advance_one_step(u, dt, nu, kx, ky, kz, dealias_mask):
    """One classic RK4 step (jax/taylor_green_jax.py:188-196)."""
    plan, u_hat = _to_spectral(u, kx, ky, kz, dealias_mask)
    plan.rk4(u_hat, nsteps=1, dt=float(dt), nu=float(nu))
    return plan.inverse(u_hat)


def kinetic_energy(u):
    """1/2 <u^2 + v^2 + w^2> (jax/taylor_green_jax.py:199-201)."""
    u = torch.as_tensor(u)
    return float(0.5 * torch.mean(u[0].double() ** 2 + u[1].double() ** 2 + u[2].double() ** 2))


def _save_plots(u, times, energies, enstrophies, N):
    """The reference's two PNGs (jax/taylor_green_jax.py:283-306)."""
    try:
        import matplotlib
        matplotlib.use("Agg")
        import matplotlib.pyplot as plt
    except ImportError:  # pragma: no cover
        print("matplotlib unavailable: plots skipped")
        return
    sl = u[:, :, N // 2].cpu().numpy()
    fig, axes = plt.subplots(1, 3, figsize=(12, 4))
    for i, ax in enumerate(axes):
        im = ax.imshow(sl[i].T, origin="lower", cmap="viridis")
        ax.set_title(f"u{i + 1} slice")
        fig.colorbar(im, ax=ax)
    plt.tight_layout()
    plt.savefig("taylor_green_slice.png", dpi=200)
    plt.close(fig)
    print("Saved: taylor_green_slice.png")
    fig, (ax1, ax2) = plt.subplots(2, 1, figsize=(10, 8))
    ax1.plot(times, energies, "b-o", linewidth=2, markersize=4)
    ax1.set_xlabel("Time")
    ax1.set_ylabel("Kinetic Energy")
    ax1.set_title("Energy Decay")
    ax1.grid(True, alpha=0.3)
    ax2.plot(times, enstrophies, "r-s", linewidth=2, markersize=4)
    ax2.set_xlabel("Time")
    ax2.set_ylabel("Enstrophy")
    ax2.set_title("Enstrophy Evolution (Turbulence Growth/Decay)")
    ax2.grid(True, alpha=0.3)
    plt.tight_layout()
    plt.savefig("taylor_green_diagnostics.png", dpi=200)
    plt.close(fig)
    print("Saved: taylor_green_diagnostics.png")


def run_simulation(N=16, dt=0.005, nu=None, reynolds=None, total_time=0.5, save_every_time=0.05,
                   cfl=0.5, return_diagnostics=False, dtype=torch.float64, make_plots=True, check_finite=True):
    """Decaying 3D Taylor-Green run (jax/taylor_green_jax.py:204-317).

    Same argument checks, step count int(total_time/dt), save cadence
    max(1, int(save_every_time/dt)), record times (step+1)*dt, diagnostic
    dict keys and prints.  The steps between two saves run as one call into
    the library; diagnostics are FP64 Parseval sums of the on-GPU state.
    ``dtype`` (default float64, the reference's precision) selects the
    FP32 variant; ``make_plots`` switches the reference's PNG side effect;
    with ``check_finite`` a save point whose state holds NaN or infinite
    values raises FloatingPointError (the reference would carry on with NaN).
    Where 3 divides a compiled N the nonlinear term's aliased shell differs
    from the reference's advective form (module docstring).
    """
    from .taylor_green_cuda import check_finite as _finite
    if N < 4 or total_time <= 0.0 or save_every_time <= 0.0:
        raise ValueError("N must be at least 4, total_time and save_every_time must be positive")
    if dt <= 0.0:
        raise ValueError("dt must be positive")
    n_steps = int(total_time / dt)
    save_every = max(1, int(save_every_time / dt))
    L = _TWO_PI
    velocity_amplitude = 1.0
    if reynolds is not None:
        nu = reynolds_to_viscosity(reynolds, velocity_amplitude, L)
    elif nu is None:
        reynolds = 100
        nu = reynolds_to_viscosity(reynolds, velocity_amplitude, L)
    actual_reynolds = viscosity_to_reynolds(nu, velocity_amplitude, L)
    _lib._require_gpu()
    dev = torch.device("cuda", torch.cuda.current_device())
    kx, ky, kz = make_wavenumbers(N, L, device="cpu")
    u0 = initial_taylor_green(N, L=L, amp=1.0, device=dev)
    validate_timestep(dt, nu, kx, ky, kz, cfl, max_velocity=float(u0.abs().max()))
    solver = TaylorGreenSolver(N, nu=nu, dtype=dtype, L=L, device=dev)
    solver.set_velocity(u0)
    del u0

    d = solver.diagnostics()
    if check_finite:
        _finite(d, 0, 0.0)
    times = [0.0]
    energies = [d["kinetic_energy"]]
    divergence_max = [d["div_u_max"]]
    enstrophies = [d["enstrophy"]]
    dissipations = [nu * d["grad_sq"]]
    t0 = time.time()
    step = 0
    while step < n_steps:
        nxt = min(n_steps, ((step // save_every) + 1) * save_every)
        solver.step(nxt - step, dt)
        step = nxt
        d = solver.diagnostics()
        if check_finite:
            _finite(d, step, step * dt)
        times.append(step * dt)
        energies.append(d["kinetic_energy"])
        divergence_max.append(d["div_u_max"])
        enstrophies.append(d["enstrophy"])
        dissipations.append(nu * d["grad_sq"])
        print(f"Save time step: {step}, Enstrophy: {enstrophies[-1]:.6e}")
    elapsed = time.time() - t0
    print(f"Reynolds number: Re = {actual_reynolds:.1f}")
    print(f"Simulated time: {total_time:.3f} ({n_steps} steps) in {elapsed:.3f} s")
    print(f"Energy trace: {energies[0]:.6f} -> {energies[-1]:.6f}")
    print(f"Max divergence: {max(divergence_max):.3e}")
    u = solver.velocity()
    if make_plots:
        _save_plots(u, times, energies, enstrophies, N)
    diagnostics = {"time": times, "energy": energies, "divergence_max": divergence_max,
                   "enstrophy": enstrophies, "dissipation": dissipations}
    if return_diagnostics:
        return u, energies, diagnostics
    return u, energies


def main(argv=None):
    """jax/taylor_green_jax.py:320-333 (python -m jax_tg_project_amd.taylor_green_jax): report the
    device and run the reference's default case (N=256, dt=0.005, Re=1000,
    total time 10 * 2 pi * 1000, saves every 0.05), each overridable."""
    import argparse
    parser = argparse.ArgumentParser(description="3D decaying Taylor-Green vortex on the MI355X HIP kernels")
    parser.add_argument("--n", type=int, default=256)
    parser.add_argument("--dt", type=float, default=0.005)
    parser.add_argument("--reynolds", type=float, default=1000.0)
    parser.add_argument("--total-time", type=float, default=10 * 2 * math.pi * 1000)
    parser.add_argument("--save-every-time", type=float, default=0.05)
    parser.add_argument("--precision", choices=("fp64", "fp32"), default="fp64")
    parser.add_argument("--no-plots", action="store_true")
    args = parser.parse_args(argv)
    if torch.cuda.is_available():
        print("GPU devices:", [torch.cuda.get_device_name(i) for i in range(torch.cuda.device_count())])
    else:
        print("WARNING: No GPU detected by torch. Check the ROCm setup before scaling up.")
    return run_simulation(N=args.n, dt=args.dt, reynolds=args.reynolds, total_time=args.total_time,
                          save_every_time=args.save_every_time, make_plots=not args.no_plots,
                          dtype=torch.float64 if args.precision == "fp64" else torch.float32)


if __name__ == "__main__":
    main()
