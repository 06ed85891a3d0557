"""Pins for the FP64 oracle (CPU only).

Each check is chosen so that a plausible mistake in the restatement (a sign,
a dropped term, a wrong normalisation, a transposed cross product) fails it:
the notebook fixture (the only reference-held numbers), the reference's own
unit-test invariants, closed forms and brute force.
"""
import json
import os

import numpy as np
import pytest

from oracle import tg_oracle as o
from tests.helpers import band_limited_field, rel_l2

L = 2 * np.pi
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "notebook_tgv_n128.json")


def _fixture():
    with open(GOLDEN) as f:
        return json.load(f)


def _compare_to_notebook(diag, energies):
    fx = _fixture()
    assert diag["time"] == pytest.approx(fx["times"], abs=1e-15)
    zs = [float(s) for s in fx["enstrophy_str"]]
    dz = max(abs(a - b) for a, b in zip(diag["enstrophy"], zs))
    de = abs(energies[-1] - float(fx["final_energy_str"]))
    return dz, de


def test_oracle_reproduces_notebook_histories():
    """jax/taylor_green_colab.ipynb:134-161,180-200: N=128, dt=0.005,
    Re=1000, 200 steps.  TGV energy at t<=1 lives in low modes, so the
    oracle at N=32 must match every printed enstrophy to its 7 significant
    digits (|dZ| < 5e-8) and the final energy to 1e-10 relative."""
    o.set_workers(min(8, o.cpu_count()))
    u, energies, diag = o.run_simulation(N=32, dt=0.005, reynolds=1000, total_time=1.0,
                                         save_every_time=0.05, return_diagnostics=True)
    dz, de = _compare_to_notebook(diag, energies)
    assert dz < 5e-8, dz
    assert de < 1e-10 * energies[-1], de
    assert max(diag["divergence_max"]) < 1e-12


def test_notebook_fixture_rejects_planted_sign_error(monkeypatch):
    """Plant a mistake in the oracle (nonlinear term with the wrong sign) and
    show the notebook comparison fails; otherwise the pin above is vacuous."""
    real = o.spectral_rhs

    def wrong(u, nu, kx, ky, kz, mask, form="advective"):
        r = real(u, nu, kx, ky, kz, mask, form)
        uh = np.fft.fftn(u, axes=(-3, -2, -1)) * mask
        visc = np.fft.ifftn(-nu * (kx * kx + ky * ky + kz * kz) * uh, axes=(-3, -2, -1)).real
        return 2 * visc - r  # = +P(n) - nu k^2 u
    monkeypatch.setattr(o, "spectral_rhs", wrong)
    u, energies, diag = o.run_simulation(N=16, dt=0.005, reynolds=1000, total_time=1.0,
                                         save_every_time=0.05, return_diagnostics=True)
    dz, de = _compare_to_notebook(diag, energies)
    assert dz > 1e-5 or de > 1e-8


def test_reference_unit_invariants():
    """jax/test_taylor_green_jax.py:22-75 on the oracle."""
    kx, ky, kz = o.make_wavenumbers(16, L)
    v = o.initial_taylor_green(16, L)
    assert np.max(np.abs(o.spectral_divergence(v, kx, ky, kz))) < 1e-12
    m = o.make_dealias_mask(16, L)
    assert m.shape == (16, 16, 16) and int(m.sum()) == 11 ** 3
    assert not m[8, 0, 0] and m[5, 0, 0]
    with pytest.raises(ValueError, match="stability limit"):
        o.validate_timestep(1.0, 0.01, kx, ky, kz)
    with pytest.raises(ValueError, match="positive"):
        o.validate_timestep(0.0, 0.01, kx, ky, kz)
    k8 = o.make_wavenumbers(8, L)
    m8 = o.make_dealias_mask(8, L)
    assert np.max(np.abs(o.spectral_rhs(np.ones((3, 8, 8, 8)), 0.01, *k8, m8))) < 1e-12
    u, energies, diag = o.run_simulation(N=8, dt=0.005, nu=0.01, total_time=0.02, save_every_time=0.01,
                                         return_diagnostics=True)
    assert np.all(np.isfinite(u))
    assert diag["time"] == [0.0, 0.01, 0.02]
    assert max(diag["divergence_max"]) < 1e-10
    assert all(a >= b for a, b in zip(energies, energies[1:]))
    with pytest.raises(ValueError, match="Reynolds number must be positive"):
        o.reynolds_to_viscosity(-100)
    assert abs(o.viscosity_to_reynolds(o.reynolds_to_viscosity(100.0)) - 100.0) < 1e-10


def test_closed_form_t0_rhs_and_diagnostics():
    """spectral_rhs(u_TG) = -P(n) - 3 nu u_TG; E0 = 1/8, Z0 = 3/8, eps0 = 3 nu / 4."""
    N, nu = 16, 0.037
    kx, ky, kz = o.make_wavenumbers(N, L)
    u0 = o.initial_taylor_green(N)
    x = np.linspace(0, L, N, endpoint=False)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    Pn = np.stack([np.sin(2 * X) * np.cos(2 * Z), np.sin(2 * Y) * np.cos(2 * Z),
                   -(np.cos(2 * X) + np.cos(2 * Y)) * np.sin(2 * Z)]) / 8
    r = o.spectral_rhs(u0, nu, kx, ky, kz, o.make_dealias_mask(N, L))
    assert np.max(np.abs(r - (-Pn - 3 * nu * u0))) < 1e-14
    assert abs(o.kinetic_energy(u0) - 0.125) < 1e-15
    assert abs(o.enstrophy(u0, kx, ky, kz) - 0.375) < 1e-14
    assert abs(o.viscous_dissipation(u0, nu, kx, ky, kz) - 0.75 * nu) < 1e-15


def test_shear_wave_rk4_amplification():
    """u = (0, A sin x, 0) has zero nonlinear term; RK4 multiplies it by
    R(-nu dt) per step exactly, so E_n = E0 R^(2n)."""
    N, nu, dt, n = 8, 0.3, 0.05, 7
    kx, ky, kz = o.make_wavenumbers(N, L)
    m = o.make_dealias_mask(N, L)
    u = o.shear_wave(N, 0.7)
    e0 = o.kinetic_energy(u)
    for _ in range(n):
        u = o.advance_one_step(u, dt, nu, kx, ky, kz, m)
    R = o.rk4_amplification(-nu * dt)
    assert abs(o.kinetic_energy(u) - e0 * R ** (2 * n)) < 1e-15


@pytest.mark.parametrize("N", [8, 16, 32])
def test_rotational_form_equals_advective_when_dealiasing_is_exact(N):
    u = band_limited_field(N, N)
    kx, ky, kz = o.make_wavenumbers(N, L)
    m = o.make_dealias_mask(N, L)
    a = o.spectral_rhs(u, 0.01, kx, ky, kz, m, "advective")
    r = o.spectral_rhs(u, 0.01, kx, ky, kz, m, "rotational")
    assert rel_l2(r, a) < 1e-14


def test_rotational_form_differs_on_the_aliased_shell_n48():
    """3*(N//3) == N: the inclusive mask keeps |k| = N/3 where the product
    aliases, so the two forms differ (why the ensemble compares like forms)."""
    N = 48
    u = band_limited_field(N, 1)
    kx, ky, kz = o.make_wavenumbers(N, L)
    m = o.make_dealias_mask(N, L)
    a = o.spectral_rhs(u, 0.01, kx, ky, kz, m, "advective")
    r = o.spectral_rhs(u, 0.01, kx, ky, kz, m, "rotational")
    assert rel_l2(r, a) > 1e-9


def test_brute_force_dft_conventions():
    """numpy.fft (the oracle's transform) against the O(N^6) definition."""
    N = 4
    rng = np.random.default_rng(0)
    f = rng.standard_normal((N, N, N))
    n = np.arange(N)
    F = np.zeros((N, N, N), complex)
    for a in range(N):
        for b in range(N):
            for c in range(N):
                ph = np.exp(-2j * np.pi * (a * n[:, None, None] + b * n[None, :, None] + c * n[None, None, :]) / N)
                F[a, b, c] = np.sum(f * ph)
    assert np.max(np.abs(F - np.fft.fftn(f))) < 1e-12
    assert np.max(np.abs(np.fft.ifftn(F).real - f)) < 1e-14


@pytest.mark.parametrize("N", [4, 8, 16, 48])
def test_compact_layout_round_trip(N):
    u = band_limited_field(N, 3, divfree=False)
    m = o.make_dealias_mask(N, L)
    full = np.fft.fftn(u, axes=(-3, -2, -1)) * m
    c = o.full_to_compact(full)
    M, K, KZ = o.kept_counts(N)
    assert c.shape == (3, K, K, KZ)
    assert np.max(np.abs(o.compact_to_full(c, N) - full)) < 1e-12 * np.max(np.abs(full))


def test_mhd_closed_form_t0_config_ic():
    """b0 = -omega_TG: j = -3 u0, P(j x b0) = -3 P(n) and
    curl(u0 x b0) = (1/2 sin2y sin2z, -1/2 sin2x sin2z, 0); so
    rhs_u = +2 P(n) - 3 nu u0 (reference sign), -4 P(n) - 3 nu u0 (physical),
    rhs_b = curl(u0 x b0) - 3 eta b0; E_m(0) = 3/8, H_c(0) = 0."""
    N, nu, eta = 16, 0.02, 0.03
    grid = o.SpectralGrid(N, L)
    u0, b0 = o.mhd_initial_condition(N, kind="config")
    uh, bh = grid.to_spectral(u0), grid.to_spectral(b0)
    x = np.linspace(0, L, N, endpoint=False)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    Pn = np.stack([np.sin(2 * X) * np.cos(2 * Z), np.sin(2 * Y) * np.cos(2 * Z),
                   -(np.cos(2 * X) + np.cos(2 * Y)) * np.sin(2 * Z)]) / 8
    ind = np.stack([0.5 * np.sin(2 * Y) * np.sin(2 * Z), -0.5 * np.sin(2 * X) * np.sin(2 * Z), 0 * X])
    for sign, coef in ((-1.0, 2.0), (1.0, -4.0)):
        ru, rb = o.mhd_rhs(uh, bh, nu, eta, grid, lorentz_sign=sign)
        assert np.max(np.abs(grid.to_real(ru) - (coef * Pn - 3 * nu * u0))) < 1e-14
        assert np.max(np.abs(grid.to_real(rb) - (ind - 3 * eta * b0))) < 1e-14
    ek, em, hc, z, du, db = o.mhd_diagnostics(uh, bh, grid)
    assert abs(ek - 0.125) < 1e-15 and abs(em - 0.375) < 1e-14 and abs(hc) < 1e-15
    assert du < 1e-13 and db < 1e-13


def test_mhd_reference_ic_has_no_induction_at_t0():
    """Reference IC b = 0.5 u_TG: u x b = 0, so rhs_b = -eta k^2 b_hat."""
    N, eta = 8, 0.05
    grid = o.SpectralGrid(N, L)
    u0, b0 = o.mhd_initial_condition(N, kind="reference")
    uh, bh = grid.to_spectral(u0), grid.to_spectral(b0)
    _, rb = o.mhd_rhs(uh, bh, 0.01, eta, grid)
    assert np.max(np.abs(rb + eta * grid.k_squared * bh)) < 1e-12


def test_mhd_ideal_energy_conservation_depends_on_lorentz_sign():
    """With nu = eta = 0 the physical sign conserves E_k + E_m (Galerkin
    truncation preserves the triad cancellation when 3*(N//3) < N); the
    reference code's sign does not."""
    N, dt = 16, 0.01
    grid = o.SpectralGrid(N, L)
    u0, b0 = o.mhd_initial_condition(N, kind="config")
    out = {}
    for sign in (1.0, -1.0):
        uh, bh = grid.to_spectral(u0), grid.to_spectral(b0)
        e0 = sum(o.mhd_diagnostics(uh, bh, grid)[:2])
        for _ in range(10):
            uh, bh = o.mhd_rk4_step(uh, bh, dt, 0.0, 0.0, grid, lorentz_sign=sign)
        out[sign] = abs(sum(o.mhd_diagnostics(uh, bh, grid)[:2]) - e0) / e0
    assert out[1.0] < 1e-9, out
    assert out[-1.0] > 1e-4, out


def test_mhd_reference_unit_invariants():
    """cuda/test_taylor_green_mhd.py:23-88 on the oracle (float64 grid)."""
    grid = o.SpectralGrid(16, L)
    u, b = o.mhd_initial_condition(16)
    uh, bh = grid.to_spectral(u), grid.to_spectral(b)
    *_, du, db = o.mhd_diagnostics(uh, bh, grid)
    assert du < 1e-5 and db < 1e-5
    with pytest.raises(ValueError, match="stability limit"):
        o.mhd_validate_timestep(1.0, L / 100, L / 100, grid)
    with pytest.raises(ValueError, match="positive"):
        o.mhd_validate_timestep(0.0, L / 100, L / 100, grid)
    g8 = o.SpectralGrid(8, L)
    u, b = o.mhd_initial_condition(8)
    uh, bh = g8.to_spectral(u), g8.to_spectral(b)
    for _ in range(4):
        uh, bh = o.mhd_rk4_step(uh, bh, 0.005, L / 100, L / 100, g8)
    assert np.all(np.isfinite(uh)) and np.all(np.isfinite(bh))
    *_, du, db = o.mhd_diagnostics(uh, bh, g8)
    assert du < 1e-4 and db < 1e-4


def test_mhd_reference_ic_closed_form_diagnostics():
    """b0 = 0.5 u_TG: E_k = 1/8, E_m = 1/4 * 1/8 = 1/32, H_c = <u.b> =
    0.5 <|u|^2> = 1/8 (no 1/2, diagnostics.py:25-29: a 1/2 factor would give
    1/16), Z = 3/8."""
    grid = o.SpectralGrid(16, L)
    u0, b0 = o.mhd_initial_condition(16, kind="reference")
    ek, em, hc, z, du, db = o.mhd_diagnostics(grid.to_spectral(u0), grid.to_spectral(b0), grid)
    assert abs(ek - 0.125) < 1e-15 and abs(em - 0.03125) < 1e-15
    assert abs(hc - 0.125) < 1e-15
    assert abs(z - 0.375) < 1e-14
    assert du < 1e-13 and db < 1e-13


def test_mhd_config_ic_enstrophy():
    """Z(0) = 3/8 for u_TG whatever b is (the kinetic enstrophy)."""
    grid = o.SpectralGrid(16, L)
    u0, b0 = o.mhd_initial_condition(16, kind="config")
    z = o.mhd_diagnostics(grid.to_spectral(u0), grid.to_spectral(b0), grid)[3]
    assert abs(z - 0.375) < 1e-14


def test_mhd_run_cadence_times_and_records():
    """The reference's short-run test (cuda/test_taylor_green_mhd.py:50-66):
    n=8, dt=0.005, total 0.02, saves every 0.01 -> records at steps 0, 2, 4,
    times step*dt exactly; the first record is the t=0 closed form, the last
    the diagnostics of the returned state; a non-divisible total time records
    the last step as well."""
    uh, bh, rec = o.mhd_run(8, 0.005, L / 100.0, L / 100.0, 0.02, 0.01)
    assert rec["times"] == [0.0, 2 * 0.005, 4 * 0.005]
    assert abs(rec["kinetic_energies"][0] - 0.125) < 1e-15
    assert abs(rec["cross_helicities"][0] - 0.125) < 1e-15
    assert rec["total_energies"][-1] == rec["kinetic_energies"][-1] + rec["magnetic_energies"][-1]
    grid = o.SpectralGrid(8, L)
    last = o.mhd_diagnostics(uh, bh, grid)
    assert last[0] == rec["kinetic_energies"][-1] and last[4] == rec["divergence_u_max"][-1]
    assert max(rec["divergence_u_max"]) < 1e-4 and max(rec["divergence_b_max"]) < 1e-4
    _, _, rec2 = o.mhd_run(8, 0.005, L / 100.0, L / 100.0, 0.035, 0.01)
    assert rec2["times"] == [0.0, 2 * 0.005, 4 * 0.005, 6 * 0.005, 7 * 0.005]
    with pytest.raises(ValueError, match="stability limit"):
        o.mhd_run(8, 1.0, L / 100.0, L / 100.0, 2.0, 1.0)


def test_cupy_rhs_equals_jax_rhs_on_band_limited_states_and_closed_form():
    """The CuPy operator (operators.py:24-48) multiplies the dealiased u, the
    JAX one the raw u (jax/taylor_green_jax.py:165): equal for band-limited
    states, so the notebook-pinned JAX restatement pins it; plus the t=0
    closed form -P(n) - 3 nu u0."""
    N, nu = 16, 0.02
    grid = o.SpectralGrid(N, L)
    kx, ky, kz = o.make_wavenumbers(N, L)
    m = o.make_dealias_mask(N, L)
    u = band_limited_field(N, 7)
    r_cupy = grid.to_real(o.cupy_rhs(grid.to_spectral(u), nu, grid))
    r_jax = o.spectral_rhs(u, nu, kx, ky, kz, m)
    assert rel_l2(r_cupy, r_jax) < 1e-13
    u0 = o.initial_taylor_green(N)
    x = np.linspace(0, L, N, endpoint=False)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    Pn = np.stack([np.sin(2 * X) * np.cos(2 * Z), np.sin(2 * Y) * np.cos(2 * Z),
                   -(np.cos(2 * X) + np.cos(2 * Y)) * np.sin(2 * Z)]) / 8
    r0 = grid.to_real(o.cupy_rhs(grid.to_spectral(u0), nu, grid))
    assert np.max(np.abs(r0 - (-Pn - 3 * nu * u0))) < 1e-14


def test_cupy_run_matches_reference_test_and_jax_run():
    """cuda/test_taylor_green_cuda.py:25-37: n=8, dt=0.005, Re=100,
    total 0.02, saves every 0.01 -> 3 records, finite; and the same run as
    the JAX restatement's E history (both advance band-limited states)."""
    nu = L / 100.0
    state, times, energies, enstrophies = o.cupy_run(8, 0.005, nu, 0.02, 0.01)
    assert times == [0.0, 0.01, 0.02]
    assert np.all(np.isfinite(state))
    assert abs(energies[0] - 0.125) < 1e-15 and abs(enstrophies[0] - 0.375) < 1e-14
    _, e_jax, d_jax = o.run_simulation(N=8, dt=0.005, nu=nu, total_time=0.02, save_every_time=0.01,
                                       return_diagnostics=True)
    assert np.allclose(energies, e_jax, rtol=1e-13, atol=0)
    assert np.allclose(enstrophies, d_jax["enstrophy"], rtol=1e-12, atol=0)
