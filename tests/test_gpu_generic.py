"""Grid sizes with no compiled kernels (any N in [4, 2048] outside
tg_supported_sizes) run the direct-sum path: transforms against numpy.fft,
the hydro RHS and RK4 steps against the FP64 oracle's advective form
(jax/taylor_green_jax.py:158-196 restated) on the whole kept band, which
the direct-sum path reproduces also where 3 divides N, and the MHD RHS
against cuda/taylor_green_mhd/operators.py:57-107 restated."""
import math

import numpy as np
import pytest
import torch

from oracle import tg_oracle as o
from tests.helpers import band_limited_field, rel_l2

pytestmark = pytest.mark.gpu
L = 2.0 * math.pi
SIZES = [5, 7, 9, 10, 20, 36, 100]
TOL = {torch.float64: 1e-12, torch.float32: 5e-6}


def _plan(N, dtype=torch.float64, mhd=False):
    from jax_tg_project_amd._lib import Plan
    p = Plan(N, dtype=dtype, mhd=mhd)
    assert p.generic
    return p


def test_compiled_sizes_are_not_generic():
    from jax_tg_project_amd._lib import Plan, supported_sizes
    assert 20 not in supported_sizes() and 24 in supported_sizes()
    assert not Plan(24).generic


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("N", [5, 7, 10, 20, 36])
def test_transforms_vs_numpy(N, dtype):
    from jax_tg_project_amd._lib import TG_OP_CURL, TG_OP_DIV, TG_OP_GRAD
    rng = np.random.default_rng(N)
    u = rng.standard_normal((3, N, N, N))
    p = _plan(N, dtype)
    got = p.forward(torch.tensor(u, dtype=dtype, device="cuda")).cpu().numpy()
    want = o.full_to_compact(np.fft.fftn(u, axes=(-3, -2, -1)))
    assert rel_l2(got, want) < TOL[dtype]
    full = o.compact_to_full(want, N)
    ct = torch.tensor(want, dtype=p.cdtype, device="cuda")
    inv = lambda f: np.fft.ifftn(f, axes=(-3, -2, -1)).real
    assert rel_l2(p.inverse(ct).cpu().numpy(), inv(full)) < TOL[dtype]
    kx, ky, kz = o.make_wavenumbers(N, L)
    curl = inv(np.stack([1j * (ky * full[2] - kz * full[1]), 1j * (kz * full[0] - kx * full[2]),
                         1j * (kx * full[1] - ky * full[0])]))
    assert rel_l2(p.inverse(ct, op=TG_OP_CURL).cpu().numpy(), curl) < TOL[dtype]
    div = inv(1j * (kx * full[0] + ky * full[1] + kz * full[2]))
    assert rel_l2(p.inverse(ct, op=TG_OP_DIV).cpu().numpy(), div) < TOL[dtype]
    grad = inv(np.stack([1j * kx * full[1], 1j * ky * full[1], 1j * kz * full[1]]))
    assert rel_l2(p.inverse(ct[1].contiguous(), op=TG_OP_GRAD).cpu().numpy(), grad) < TOL[dtype]


@pytest.mark.parametrize("N", SIZES)
def test_hydro_rhs_full_band_vs_oracle_advective(N):
    u = band_limited_field(N, 40 + N)
    nu = 0.013
    kx, ky, kz = o.make_wavenumbers(N, L)
    want = o.spectral_rhs(u, nu, kx, ky, kz, o.make_dealias_mask(N, L))
    p = _plan(N)
    got = p.inverse(p.rhs(p.forward(torch.tensor(u, device="cuda")), nu=nu)).cpu().numpy()
    err = rel_l2(got, want)
    print(f"generic hydro rhs N={N}: {err:.2e}", flush=True)
    assert err < 1e-12, err


@pytest.mark.parametrize("N", [9, 20])
def test_rk4_steps_vs_oracle(N):
    u = o.initial_taylor_green(N) + 0.1 * band_limited_field(N, 3)
    nu, dt = 0.01, 0.01
    kx, ky, kz = o.make_wavenumbers(N, L)
    mask = o.make_dealias_mask(N, L)
    p = _plan(N)
    uh = p.forward(torch.tensor(u, device="cuda"))
    ref = u.copy()
    for _ in range(5):
        ref = o.advance_one_step(ref, dt, nu, kx, ky, kz, mask)
        p.rk4(uh, nsteps=1, dt=dt, nu=nu)
        err = rel_l2(p.inverse(uh).cpu().numpy(), ref)
        assert err < 1e-12, err


@pytest.mark.parametrize("N", [7, 10, 20])
def test_mhd_rhs_vs_oracle(N):
    grid = o.SpectralGrid(N, L)
    u = band_limited_field(N, 1 + N)
    b = 0.7 * band_limited_field(N, 2 + N)
    nu, eta = 0.011, 0.017
    ru, rb = o.mhd_rhs(grid.to_spectral(u), grid.to_spectral(b), nu, eta, grid, lorentz_sign=-1.0)
    p = _plan(N, mhd=True)
    uh = p.forward(torch.tensor(u, device="cuda"))
    bh = p.forward(torch.tensor(b, device="cuda"))
    gu, gb = p.rhs(uh, bh, nu=nu, eta=eta, lorentz_sign=-1.0)
    assert rel_l2(p.inverse(gu).cpu().numpy(), grid.to_real(ru)) < 1e-12
    assert rel_l2(p.inverse(gb).cpu().numpy(), grid.to_real(rb)) < 1e-12
    un, bn = o.mhd_rk4_step(grid.to_spectral(u), grid.to_spectral(b), 0.01, nu, eta, grid)
    p.rk4(uh, bh, nsteps=1, dt=0.01, nu=nu, eta=eta, lorentz_sign=-1.0)
    assert rel_l2(p.inverse(uh).cpu().numpy(), grid.to_real(un)) < 1e-12
    assert rel_l2(p.inverse(bh).cpu().numpy(), grid.to_real(bn)) < 1e-12


def test_diagnostics_and_drop_in_entry_points_at_n20():
    """Taylor-Green closed forms (E = 1/8, Z = 3/8, div u = 0), the drop-in
    run loop, and the clear errors of what the direct-sum path lacks."""
    from jax_tg_project_amd import adjoint
    from jax_tg_project_amd import taylor_green_jax as tj
    from jax_tg_project_amd._lib import TgError
    N = 20
    p = _plan(N)
    uh = p.forward(torch.tensor(o.initial_taylor_green(N), device="cuda"))
    d = p.diagnostics(uh)
    assert abs(d["kinetic_energy"] - 0.125) < 1e-14 and abs(d["enstrophy"] - 0.375) < 1e-14
    assert d["div_u_max"] < 1e-14 and d["nonfinite"] == 0.0
    with pytest.raises(TgError, match="no compiled plan"):
        p.diagnostics_noalloc(uh)
    with pytest.raises(TgError, match="no compiled plan"):
        p.vjp(uh, uh, 0.01)
    with pytest.raises(ValueError):
        adjoint.final_objective(uh, 0.01, 2, 0.01)
    u, energies, diag = tj.run_simulation(N=N, dt=0.01, nu=0.05, total_time=0.05, save_every_time=0.01,
                                          make_plots=False, return_diagnostics=True)
    _, ref = o.run_simulation(N=N, dt=0.01, nu=0.05, total_time=0.05, save_every_time=0.01)
    assert tuple(u.shape) == (3, N, N, N) and len(energies) == len(ref) == 6
    assert np.allclose(energies, ref, rtol=1e-12, atol=0)
