"""GPU hydro parity against the FP64 oracle (jax/taylor_green_jax.py restated):
RHS (rel L2 <= 1e-12), RK4 steps per step (rel L2 <= 1e-12) and diagnostics
histories (<= 1e-10), through the C ABI."""
import numpy as np
import pytest
import torch

from oracle import tg_oracle as o
from tests.helpers import band_limited_field, rel_l2, to_compact

pytestmark = pytest.mark.gpu
L = 2 * np.pi


def _plan(N, dtype=torch.float64):
    from jax_tg_project_amd._lib import Plan
    return Plan(N, dtype=dtype)


@pytest.mark.parametrize("N", [4, 8, 16, 32, 64, 128])
def test_rhs_matches_oracle_advective(N):
    u = band_limited_field(N, 11 + N)
    nu = 0.013
    kx, ky, kz = o.make_wavenumbers(N, L)
    want = o.spectral_rhs(u, nu, kx, ky, kz, o.make_dealias_mask(N, L))
    p = _plan(N)
    uh = p.forward(torch.tensor(u, device="cuda"))
    got = p.inverse(p.rhs(uh, nu=nu)).cpu().numpy()
    assert rel_l2(got, want) < 1e-12, rel_l2(got, want)


@pytest.mark.parametrize("N", [48, 96])
def test_rhs_matches_oracle_rotational_form(N):
    # 3*(N//3) == N: the inclusive mask keeps the aliased shell, so the
    # reference's advective form and u x omega differ; compare like with like.
    u = band_limited_field(N, 5 + N)
    nu = 0.02
    kx, ky, kz = o.make_wavenumbers(N, L)
    want = o.spectral_rhs(u, nu, kx, ky, kz, o.make_dealias_mask(N, L), form="rotational")
    p = _plan(N)
    got = p.inverse(p.rhs(p.forward(torch.tensor(u, device="cuda")), nu=nu)).cpu().numpy()
    assert rel_l2(got, want) < 1e-12, rel_l2(got, want)


def test_rhs_closed_form_t0_256():
    """spectral_rhs(u_TG) = -P(n) - 3 nu u_TG with P(n) = 1/8 (sin2x cos2z,
    sin2y cos2z, -(cos2x+cos2y) sin2z), at the full 256^3 size."""
    N, nu = 256, 1.0 / 1600
    u0 = o.initial_taylor_green(N)
    x = np.linspace(0.0, L, N, endpoint=False)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    Pn = np.stack([np.sin(2 * X) * np.cos(2 * Z), np.sin(2 * Y) * np.cos(2 * Z),
                   -(np.cos(2 * X) + np.cos(2 * Y)) * np.sin(2 * Z)]) / 8.0
    want = -Pn - 3.0 * nu * u0
    del X, Y, Z, Pn
    p = _plan(N)
    got = p.inverse(p.rhs(p.forward(torch.tensor(u0, device="cuda")), nu=nu)).cpu().numpy()
    assert rel_l2(got, want) < 1e-13, rel_l2(got, want)


@pytest.mark.parametrize("N,dt,nu,steps", [(8, 0.005, 0.01, 4), (16, 0.01, 0.01, 20)])
def test_rk4_steps_match_oracle_each_step(N, dt, nu, steps):
    u = o.initial_taylor_green(N) + 0.1 * band_limited_field(N, 3)
    kx, ky, kz = o.make_wavenumbers(N, L)
    mask = o.make_dealias_mask(N, L)
    p = _plan(N)
    uh = p.forward(torch.tensor(u, device="cuda"))
    ref = u.copy()
    for _ in range(steps):
        ref = o.advance_one_step(ref, dt, nu, kx, ky, kz, mask)
        p.rk4(uh, nsteps=1, dt=dt, nu=nu)
        got = p.inverse(uh).cpu().numpy()
        assert rel_l2(got, ref) < 1e-12, rel_l2(got, ref)


def test_config1_n32_100_steps_histories():
    """Hydro TGV N=32, nu=0.01, dt=0.01, 100 steps: per-step fields (rel L2
    <= 1e-12) and energy/enstrophy/dissipation histories (<= 1e-10)."""
    N, dt, nu, steps = 32, 0.01, 0.01, 100
    kx, ky, kz = o.make_wavenumbers(N, L)
    mask = o.make_dealias_mask(N, L)
    u = o.initial_taylor_green(N)
    p = _plan(N)
    uh = p.forward(torch.tensor(u, device="cuda"))
    worst = 0.0
    for s in range(steps):
        u = o.advance_one_step(u, dt, nu, kx, ky, kz, mask)
        p.rk4(uh, nsteps=1, dt=dt, nu=nu)
        got = p.inverse(uh).cpu().numpy()
        worst = max(worst, rel_l2(got, u))
        if (s + 1) % 5 == 0:
            d = p.diagnostics(uh)
            assert abs(d["kinetic_energy"] - o.kinetic_energy(u)) <= 1e-10 * o.kinetic_energy(u)
            z = o.enstrophy(u, kx, ky, kz)
            assert abs(d["enstrophy"] - z) <= 1e-10 * z
            eps = o.viscous_dissipation(u, nu, kx, ky, kz)
            assert abs(nu * d["grad_sq"] - eps) <= 1e-10 * eps
            assert d["div_u_max"] < 1e-12
    assert worst < 1e-12, worst
