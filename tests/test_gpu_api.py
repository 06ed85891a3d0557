"""The reference's own unit tests (jax/test_taylor_green_jax.py,
cuda/test_taylor_green_mhd.py) run against the drop-in modules, plus
API-level parity with the oracle and the notebook fixture at N=128."""
import json
import math
import os

import numpy as np
import pytest
import torch

from oracle import tg_oracle as o
from tests.helpers import band_limited_field, rel_l2

pytestmark = pytest.mark.gpu
LENGTH = 2.0 * math.pi


@pytest.fixture(scope="module")
def tg():
    from jax_tg_project_amd import taylor_green_jax
    return taylor_green_jax


@pytest.fixture(scope="module")
def mhd():
    from jax_tg_project_amd import taylor_green_mhd
    return taylor_green_mhd


# ---------------------------------------------------- jax/test_taylor_green_jax.py
def test_initial_condition_is_divergence_free(tg):
    kx, ky, kz = tg.make_wavenumbers(16, LENGTH)
    velocity = tg.initial_taylor_green(16, LENGTH)
    divergence = tg.spectral_divergence(velocity, kx, ky, kz)
    assert float(torch.max(torch.abs(divergence))) < 1e-12


def test_dealias_mask_keeps_two_thirds_of_each_axis(tg):
    mask = tg.make_dealias_mask(16, LENGTH)
    assert tuple(mask.shape) == (16, 16, 16)
    assert int(mask.sum()) == 11 ** 3
    assert not bool(mask[8, 0, 0])
    assert bool(mask[5, 0, 0])


def test_timestep_validation_rejects_unstable_values(tg):
    kx, ky, kz = tg.make_wavenumbers(16, LENGTH)
    with pytest.raises(ValueError, match="stability limit"):
        tg.validate_timestep(1.0, 0.01, kx, ky, kz)
    with pytest.raises(ValueError, match="positive"):
        tg.validate_timestep(0.0, 0.01, kx, ky, kz)


def test_zero_mode_is_not_viscously_damped(tg):
    N = 8
    kx, ky, kz = tg.make_wavenumbers(N, LENGTH)
    mask = tg.make_dealias_mask(N, LENGTH)
    rhs = tg.spectral_rhs(torch.ones((3, N, N, N), dtype=torch.float64), 0.01, kx, ky, kz, mask)
    assert float(torch.max(torch.abs(rhs))) < 1e-12


def test_short_run_is_finite_divergence_free_and_decays(tg):
    velocity, energies, diagnostics = tg.run_simulation(N=8, dt=0.005, nu=0.01, total_time=0.02,
                                                        save_every_time=0.01, return_diagnostics=True,
                                                        make_plots=False)
    assert bool(torch.all(torch.isfinite(velocity)))
    assert diagnostics["time"] == [0.0, 0.01, 0.02]
    assert max(diagnostics["divergence_max"]) < 1e-10
    assert all(a >= b for a, b in zip(energies, energies[1:]))


def test_step_preserves_finite_values(tg):
    N = 8
    kx, ky, kz = tg.make_wavenumbers(N, LENGTH)
    mask = tg.make_dealias_mask(N, LENGTH)
    velocity = tg.initial_taylor_green(N, LENGTH)
    updated = tg.advance_one_step(velocity, 0.005, 0.01, kx, ky, kz, mask)
    assert bool(torch.all(torch.isfinite(updated)))


def test_run_with_reynolds_number(tg):
    u, energies = tg.run_simulation(N=8, dt=0.005, reynolds=100, total_time=0.02, save_every_time=0.01,
                                    make_plots=False)
    assert bool(torch.all(torch.isfinite(u)))
    assert all(a >= b for a, b in zip(energies, energies[1:]))


# ---------------------------------------------------- API-level parity
def test_api_rhs_and_step_match_oracle(tg):
    N, nu, dt = 16, 0.02, 0.01
    u = o.initial_taylor_green(N) + 0.2 * band_limited_field(N, 4)
    kx, ky, kz = tg.make_wavenumbers(N, LENGTH)
    mask = tg.make_dealias_mask(N, LENGTH)
    okx, oky, okz = o.make_wavenumbers(N, LENGTH)
    om = o.make_dealias_mask(N, LENGTH)
    ut = torch.tensor(u)
    assert rel_l2(tg.spectral_rhs(ut, nu, kx, ky, kz, mask).cpu().numpy(),
                  o.spectral_rhs(u, nu, okx, oky, okz, om)) < 1e-12
    assert rel_l2(tg.advance_one_step(ut, dt, nu, kx, ky, kz, mask).cpu().numpy(),
                  o.advance_one_step(u, dt, nu, okx, oky, okz, om)) < 1e-12
    assert rel_l2(tg.vorticity(ut, kx, ky, kz).cpu().numpy(), o.vorticity(u, okx, oky, okz)) < 1e-12
    assert abs(tg.enstrophy(ut, kx, ky, kz) - o.enstrophy(u, okx, oky, okz)) < 1e-13
    assert abs(tg.viscous_dissipation(ut, nu, kx, ky, kz) - o.viscous_dissipation(u, nu, okx, oky, okz)) < 1e-14
    assert abs(tg.kinetic_energy(ut) - o.kinetic_energy(u)) < 1e-15


def test_api_rejects_input_outside_the_mask(tg):
    N = 16
    kx, ky, kz = tg.make_wavenumbers(N, LENGTH)
    mask = tg.make_dealias_mask(N, LENGTH)
    rng = np.random.default_rng(1)
    u = torch.tensor(rng.standard_normal((3, N, N, N)))
    with pytest.raises(ValueError, match="outside the 2/3 dealiasing mask"):
        tg.advance_one_step(u, 0.001, 0.01, kx, ky, kz, mask)


def test_notebook_run_n128(tg):
    """The notebook run itself (jax/taylor_green_colab.ipynb cells 3-5): N=128,
    dt=0.005, Re=1000, T=1: every printed enstrophy to 7 digits, the final
    energy to 1e-12 relative, and divergence at round-off."""
    with open(os.path.join(os.path.dirname(__file__), "golden", "notebook_tgv_n128.json")) as f:
        fx = json.load(f)
    u, energies, diag = tg.run_simulation(N=128, dt=0.005, reynolds=1000, total_time=1.0,
                                          save_every_time=0.05, return_diagnostics=True, make_plots=False)
    assert diag["time"] == pytest.approx(fx["times"], abs=1e-15)
    zs = [float(s) for s in fx["enstrophy_str"]]
    assert max(abs(a - b) for a, b in zip(diag["enstrophy"], zs)) < 5e-8
    e_ref = float(fx["final_energy_str"])
    assert abs(energies[-1] - e_ref) < 1e-12 * e_ref, (energies[-1], e_ref)
    assert max(diag["divergence_max"]) < 1e-12


def test_fp32_run_energy_error(tg):
    N = 64
    _, e64 = tg.run_simulation(N=N, dt=0.005, reynolds=1000, total_time=0.5, make_plots=False)
    _, e32 = tg.run_simulation(N=N, dt=0.005, reynolds=1000, total_time=0.5, make_plots=False,
                               dtype=torch.float32)
    assert max(abs(a - b) / a for a, b in zip(e64, e32)) < 1e-5


# ---------------------------------------------------- cuda/test_taylor_green_mhd.py
def test_mhd_initial_condition_is_divergence_free(mhd):
    grid = mhd.SpectralGrid(16, LENGTH)
    operator = mhd.MhdSpectralOperator(grid)
    velocity, magnetic_field = mhd.TaylorGreenMhdInitialCondition(LENGTH, 0.5).create(16)
    div_u = operator.divergence_from_hat(grid.to_spectral(velocity))
    div_b = operator.divergence_from_hat(grid.to_spectral(magnetic_field))
    assert float(torch.max(torch.abs(div_u))) < 1e-5
    assert float(torch.max(torch.abs(div_b))) < 1e-5


def test_mhd_dealias_mask(mhd):
    mask = mhd.SpectralGrid(16, LENGTH).dealias_mask
    assert tuple(mask.shape) == (16, 16, 16)
    assert int(mask.sum()) == 11 ** 3
    assert not bool(mask[8, 0, 0]) and bool(mask[5, 0, 0])


def test_mhd_timestep_validation(mhd):
    grid = mhd.SpectralGrid(16, LENGTH)
    with pytest.raises(ValueError, match="stability limit"):
        mhd.validate_timestep(1.0, LENGTH / 100.0, LENGTH / 100.0, grid)
    with pytest.raises(ValueError, match="positive"):
        mhd.validate_timestep(0.0, LENGTH / 100.0, LENGTH / 100.0, grid)


def test_mhd_short_run_is_finite_and_divergence_free(mhd):
    config = mhd.TaylorGreenMhdConfig(n=8, dt=0.005, reynolds=100.0, magnetic_reynolds=100.0,
                                      magnetic_amplitude=0.5, total_time=0.02, save_every_time=0.01)
    velocity_hat, magnetic_field_hat, diagnostics = mhd.TaylorGreenMhdSimulation(config, make_plots=False).run()
    assert bool(torch.all(torch.isfinite(velocity_hat)))
    assert bool(torch.all(torch.isfinite(magnetic_field_hat)))
    assert len(diagnostics.times) == 3
    assert max(diagnostics.divergence_u_max) < 1e-4
    assert max(diagnostics.divergence_b_max) < 1e-4


def test_mhd_step_preserves_finite_values(mhd):
    grid = mhd.SpectralGrid(8, LENGTH)
    operator = mhd.MhdSpectralOperator(grid)
    integrator = mhd.MhdRK4Integrator(operator, 0.005, LENGTH / 100.0, LENGTH / 100.0)
    velocity, magnetic_field = mhd.TaylorGreenMhdInitialCondition(LENGTH, 0.5).create(8)
    uh, bh = integrator.step((grid.to_spectral(velocity), grid.to_spectral(magnetic_field)))
    assert bool(torch.all(torch.isfinite(uh))) and bool(torch.all(torch.isfinite(bh)))


def test_mhd_shim_matches_oracle(mhd):
    N, nu, eta, dt = 16, 0.03, 0.02, 0.004
    kx, ky, kz = mhd.make_wavenumbers(N, LENGTH)
    mask = mhd.make_dealias_mask(N, LENGTH)
    grid = o.SpectralGrid(N, LENGTH)
    u0, b0 = o.mhd_initial_condition(N, kind="config")
    u0 = u0 + 0.1 * band_limited_field(N, 8)
    uh, bh = grid.to_spectral(u0), grid.to_spectral(b0)
    ru, rb = mhd.spectral_rhs_mhd(torch.tensor(uh), torch.tensor(bh), nu, eta, kx, ky, kz, mask)
    wu, wb = o.mhd_rhs(uh, bh, nu, eta, grid)
    assert rel_l2(ru.cpu().numpy(), wu) < 1e-12 and rel_l2(rb.cpu().numpy(), wb) < 1e-12
    su, sb = mhd.advance_one_step(torch.tensor(uh), torch.tensor(bh), dt, nu, eta, kx, ky, kz, mask)
    wu, wb = o.mhd_rk4_step(uh, bh, dt, nu, eta, grid)
    assert rel_l2(su.cpu().numpy(), wu) < 1e-12 and rel_l2(sb.cpu().numpy(), wb) < 1e-12
