"""Adjoint parity: the hand-written vector-Jacobian product of the RHS and
the sqrt(n)-checkpointed gradient through RK4, against the dot-product test,
the torch-CPU autograd restatement, central finite differences (eps = 1e-6,
16 seeded directions, relative 1e-5) and the shear-wave closed form."""
import math

import numpy as np
import pytest
import torch

from oracle import tg_oracle as o
from oracle import tg_torch_ad as ad
from tests.helpers import band_limited_field, to_compact

pytestmark = pytest.mark.gpu
L = 2 * np.pi


def _plan(N, dtype=torch.float64):
    from jax_tg_project_amd._lib import Plan
    return Plan(N, dtype=dtype)


def _weights(N):
    M = N // 3
    w = np.ones(M + 1)
    w[1:] = 2.0
    return w


def _ip(a, b, N):
    """Physical L2 inner product of two compact spectra."""
    return float(np.sum(_weights(N) * np.real(np.conj(a) * b)) / N ** 3)


def _state(N, seed, amp=0.2):
    return to_compact(o.initial_taylor_green(N) + amp * band_limited_field(N, seed))


def test_vjp_dot_product_test():
    N, nu = 16, 0.02
    p = _plan(N)
    x = _state(N, 1)
    v = to_compact(band_limited_field(N, 2, divfree=False))
    wv = to_compact(band_limited_field(N, 3, divfree=False))
    X = lambda a: torch.tensor(a, device="cuda")
    eps = 1e-6
    jv = (p.rhs(X(x + eps * v), nu=nu) - p.rhs(X(x - eps * v), nu=nu)).cpu().numpy() / (2 * eps)
    jtw = p.vjp(X(x), X(wv), nu).cpu().numpy()
    lhs, rhs = _ip(jv, wv, N), _ip(v, jtw, N)
    assert abs(lhs - rhs) <= 1e-8 * abs(lhs), (lhs, rhs)


@pytest.mark.parametrize("N", [8, 16, 32])
def test_vjp_matches_torch_autograd(N):
    nu = 0.013
    x = _state(N, 4)
    a = to_compact(band_limited_field(N, 5, divfree=False))
    c = torch.tensor(x, requires_grad=True)
    r = ad.rhs_full(c, nu, N)
    w = torch.tensor(_weights(N))
    Lval = (w * torch.real(torch.conj(torch.tensor(a)) * r)).sum() / N ** 3
    (g,) = torch.autograd.grad(Lval, c)
    p = _plan(N)
    mu = p.vjp(torch.tensor(x, device="cuda"), torch.tensor(a, device="cuda"), nu).cpu().numpy()
    ours = mu * _weights(N) / N ** 3
    err = np.linalg.norm(ours - g.numpy()) / np.linalg.norm(g.numpy())
    assert err < 1e-12, err


@pytest.mark.parametrize("kind", ["energy", "enstrophy"])
def test_gradient_matches_torch_autograd(kind):
    N, nu, dt, n, seg = 16, 0.01, 0.01, 12, 2  # six segments: the first two recomputed
    x = _state(N, 6)
    c = torch.tensor(x, requires_grad=True)
    nut = torch.tensor(nu, dtype=torch.float64, requires_grad=True)
    J = ad.final_objective(c, nut, n, dt, N, kind)
    gc, gnu = torch.autograd.grad(J, (c, nut))
    from jax_tg_project_amd.adjoint import final_objective
    u0 = torch.tensor(x, device="cuda", requires_grad=True)
    nug = torch.tensor(nu, dtype=torch.float64, device="cuda", requires_grad=True)
    Jg = final_objective(u0, nug, n, dt, kind, seg=seg)
    Jg.backward()
    assert abs(Jg.item() - J.item()) < 1e-13 * abs(J.item())
    err = (np.linalg.norm(u0.grad.cpu().numpy() - gc.numpy()) / np.linalg.norm(gc.numpy()))
    assert err < 1e-11, err
    assert abs(nug.grad.item() - gnu.item()) < 1e-10 * abs(gnu.item()), (nug.grad.item(), gnu.item())


def test_gradient_radix12_grid_matches_torch_autograd():
    """N = 48 (radix-12 plans, not a power of two): the VJPs transform the
    state and the cotangent (no kept stage fields), the workspace vectors
    use the line-padded pitch (KZ = 17 -> 24), and seg = 1 over 6 steps
    recomputes two segments on the overlapping stream (the forward sweep
    keeps the last four)."""
    N, nu, dt, n, seg = 48, 0.02, 0.01, 6, 1
    x = _state(N, 11)
    c = torch.tensor(x, requires_grad=True)
    nut = torch.tensor(nu, dtype=torch.float64, requires_grad=True)
    J = ad.final_objective(c, nut, n, dt, N, "energy")
    gc, gnu = torch.autograd.grad(J, (c, nut))
    p = _plan(N)
    aws = p.adjoint_workspace(n, seg)
    Jg = p.adjoint_forward(torch.tensor(x, device="cuda"), n, dt, nu, 0, seg, aws)
    g, dnu = p.adjoint_backward(n, dt, nu, seg, aws)
    assert abs(Jg - J.item()) < 1e-13 * abs(J.item())
    err = np.linalg.norm(g.cpu().numpy() - gc.numpy()) / np.linalg.norm(gc.numpy())
    assert err < 1e-11, err
    assert abs(dnu - gnu.item()) < 1e-10 * abs(gnu.item()), (dnu, gnu.item())


@pytest.mark.parametrize("seg", [2, 10])
def test_gradient_with_kept_stage_fields_matches_torch_autograd(seg):
    """N = 32 runs the register adjoint z pass with each stage's kept
    inverse-y output (u, w), so the VJPs transform only the cotangent; seg =
    2 also overlaps the segment recompute with the reverse sweep, seg = 10
    keeps every stage.  Both match the torch autograd restatement, and each
    other bitwise."""
    N, nu, dt, n = 32, 0.01, 0.01, 10
    x = _state(N, 9)
    c = torch.tensor(x, requires_grad=True)
    nut = torch.tensor(nu, dtype=torch.float64, requires_grad=True)
    J = ad.final_objective(c, nut, n, dt, N, "enstrophy")
    gc, gnu = torch.autograd.grad(J, (c, nut))
    p = _plan(N)
    xg = torch.tensor(x, device="cuda")
    aws = p.adjoint_workspace(n, seg)
    Jg = p.adjoint_forward(xg, n, dt, nu, 1, seg, aws)
    g, dnu = p.adjoint_backward(n, dt, nu, seg, aws)
    assert abs(Jg - J.item()) < 1e-13 * abs(J.item())
    err = np.linalg.norm(g.cpu().numpy() - gc.numpy()) / np.linalg.norm(gc.numpy())
    assert err < 1e-11, err
    assert abs(dnu - gnu.item()) < 1e-10 * abs(gnu.item()), (dnu, gnu.item())
    other = 10 if seg == 2 else 2
    aws2 = p.adjoint_workspace(n, other)
    p.adjoint_forward(xg, n, dt, nu, 1, other, aws2)
    g2, dnu2 = p.adjoint_backward(n, dt, nu, other, aws2)
    assert torch.equal(g, g2) and dnu == dnu2


def test_shear_wave_nu_gradient_closed_form():
    """u = (0, A sin x, 0): E_n = E0 R(z)^(2n), z = -nu dt, so
    dE/dnu = 2n E0 R^(2n-1) R'(z) (-dt)."""
    N, nu, dt, n, A = 16, 0.3, 0.05, 9, 0.7
    p = _plan(N)
    x = to_compact(o.shear_wave(N, A))
    aws = p.adjoint_workspace(n, 3)
    J = p.adjoint_forward(torch.tensor(x, device="cuda"), n, dt, nu, 0, 3, aws)
    _, dnu = p.adjoint_backward(n, dt, nu, 3, aws)
    z = -nu * dt
    R = o.rk4_amplification(z)
    dR = 1 + z + z * z / 2 + z ** 3 / 6
    E0 = A * A / 4
    assert abs(J - E0 * R ** (2 * n)) < 1e-15
    want = 2 * n * E0 * R ** (2 * n - 1) * dR * (-dt)
    assert abs(dnu - want) < 1e-13 * abs(want), (dnu, want)


def test_adjoint_config_fd_16_directions():
    """Adjoint config: N = 128, nu = 1/400, dt = 5e-3, 200 steps, J = Z(T),
    seeded TG + |k| <= 4 perturbation (rms 0.05), sqrt(200) checkpointing:
    <grad, v> against central differences (eps = 1e-6, 16 seeded unit
    directions) and dJ/dnu against its central difference, relative 1e-5."""
    from jax_tg_project_amd import ensemble
    from jax_tg_project_amd.adjoint import sqrt_segment
    N, nu, dt, n = 128, 1.0 / 400, 5e-3, 200
    p = _plan(N)
    u0 = ensemble.perturbed_taylor_green(N, 2024)
    x = p.forward(torch.tensor(u0, device="cuda"))
    seg = sqrt_segment(n)
    aws = p.adjoint_workspace(n, seg)
    J = p.adjoint_forward(x, n, dt, nu, 1, seg, aws)
    g, dnu = p.adjoint_backward(n, dt, nu, seg, aws)

    def Jof(xx, nuv):
        y = xx.clone()
        p.rk4(y, nsteps=n, dt=dt, nu=nuv)
        return p.diagnostics(y, divmax=False)["enstrophy"]

    rng = np.random.default_rng(16)
    eps = 1e-6
    xn = torch.linalg.vector_norm(x).item()
    for d in range(16):
        v = torch.tensor(rng.standard_normal(x.shape) + 1j * rng.standard_normal(x.shape), device="cuda")
        v *= xn / torch.linalg.vector_norm(v).item()
        fd = (Jof(x + eps * v, nu) - Jof(x - eps * v, nu)) / (2 * eps)
        an = float(torch.sum(torch.real(torch.conj(g) * v)))
        assert abs(an - fd) <= 1e-5 * abs(fd), (d, an, fd)
    h = 1e-6 * nu
    fdn = (Jof(x, nu + h) - Jof(x, nu - h)) / (2 * h)
    assert abs(dnu - fdn) <= 1e-5 * abs(fdn), (dnu, fdn)
    assert abs(J - Jof(x, nu)) < 1e-14 * J


def test_fp32_adjoint_vs_fp64():
    N, nu, dt, n = 64, 1.0 / 400, 5e-3, 50
    res = {}
    u0 = o.initial_taylor_green(N) + 0.05 * band_limited_field(N, 11)
    for dtype in (torch.float64, torch.float32):
        p = _plan(N, dtype)
        x = p.forward(torch.tensor(u0, dtype=dtype, device="cuda"))
        seg = 7
        aws = p.adjoint_workspace(n, seg)
        J = p.adjoint_forward(x, n, dt, nu, 1, seg, aws)
        g, dnu = p.adjoint_backward(n, dt, nu, seg, aws)
        res[dtype] = (J, g.to(torch.complex128).cpu().numpy(), dnu)
    J64, g64, n64 = res[torch.float64]
    J32, g32, n32 = res[torch.float32]
    assert abs(J32 - J64) <= 1e-5 * J64
    assert np.linalg.norm(g32 - g64) <= 1e-4 * np.linalg.norm(g64)
    assert abs(n32 - n64) <= 1e-4 * abs(n64)


def test_fp32_adjoint_config_128_vs_fp64():
    """The adjoint config (N = 128, nu = 1/400, dt = 5e-3, 200 steps, J = Z(T),
    sqrt(200) checkpointing) in FP32 against the FP64 gradient, which the
    central-difference test above pins: J to 1e-5, the gradient and dJ/dnu to
    1e-4 relative."""
    from jax_tg_project_amd import ensemble
    from jax_tg_project_amd.adjoint import sqrt_segment
    N, nu, dt, n = 128, 1.0 / 400, 5e-3, 200
    u0 = ensemble.perturbed_taylor_green(N, 2024)
    seg = sqrt_segment(n)
    res = {}
    for dtype in (torch.float64, torch.float32):
        p = _plan(N, dtype)
        x = p.forward(torch.tensor(u0, dtype=dtype, device="cuda"))
        aws = p.adjoint_workspace(n, seg)
        J = p.adjoint_forward(x, n, dt, nu, 1, seg, aws)
        g, dnu = p.adjoint_backward(n, dt, nu, seg, aws)
        res[dtype] = (J, g.to(torch.complex128), dnu)
        del aws
        torch.cuda.empty_cache()
    J64, g64, n64 = res[torch.float64]
    J32, g32, n32 = res[torch.float32]
    eg = (torch.linalg.vector_norm(g32 - g64) / torch.linalg.vector_norm(g64)).item()
    print(f"FP32 vs FP64 at 128^3: J {abs(J32 - J64) / J64:.2e}, grad {eg:.2e}, "
          f"dJ/dnu {abs(n32 - n64) / abs(n64):.2e}", flush=True)
    assert abs(J32 - J64) <= 1e-5 * J64
    assert eg <= 1e-4
    assert abs(n32 - n64) <= 1e-4 * abs(n64)


@pytest.mark.parametrize("N", [24, 64, 96, 128, 192, 256])
def test_vjp_matches_composed_transforms(N):
    """The fused VJP against the same formula composed from the separately
    verified transforms (tg_inverse / tg_forward) and torch elementwise
    products, at every large grid (regression: N = 128 is the size where the
    6 adjoint z-pass outputs need more room than the 9 inverse-x fields)."""
    from jax_tg_project_amd._lib import TG_OP_CURL
    nu = 0.013
    p = _plan(N)
    x = torch.tensor(_state(N, 4), device="cuda")
    a = torch.tensor(to_compact(band_limited_field(N, 5, divfree=False)), device="cuda")
    M = N // 3
    k1 = torch.tensor(list(range(0, M + 1)) + list(range(-M, 0)), dtype=torch.float64, device="cuda")
    kz1 = torch.arange(0, M + 1, dtype=torch.float64, device="cuda")
    kx, ky, kz = torch.meshgrid(k1, k1, kz1, indexing="ij")
    k2 = kx * kx + ky * ky + kz * kz
    k2s = torch.where(k2 == 0, torch.ones_like(k2), k2)
    kd = kx * a[0] + ky * a[1] + kz * a[2]
    Pa = torch.stack([a[0] - kx * kd / k2s, a[1] - ky * kd / k2s, a[2] - kz * kd / k2s]).contiguous()
    u, w, q = p.inverse(x), p.inverse(x, op=TG_OP_CURL), p.inverse(Pa)
    P1 = p.forward(torch.cross(w, q, dim=0).contiguous())
    P2 = p.forward(torch.cross(q, u, dim=0).contiguous())
    ref = P1 + torch.stack([1j * (ky * P2[2] - kz * P2[1]), 1j * (kz * P2[0] - kx * P2[2]),
                            1j * (kx * P2[1] - ky * P2[0])]) - nu * k2 * a
    mu = p.vjp(x, a, nu)
    err = (torch.linalg.vector_norm(mu - ref) / torch.linalg.vector_norm(ref)).item()
    assert err < 1e-13, err


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_recompute_variants_change_gradient_by_round_off_only(dtype):
    """The reverse sweep's recompute may run its own tile variants (chosen by
    autotune.tune_recompute for running beside the reverse sweep): the
    gradient then differs from the plan's-own-variants gradient by round-off
    only.  N = 128, 20 steps in segments of 3 (a short first segment of 2,
    three recomputed segments), every uniform recompute set and the tuned one."""
    from jax_tg_project_amd import autotune, ensemble
    N, nu, dt, n, seg = 128, 1.0 / 400, 5e-3, 20, 3
    p = _plan(N, dtype)
    x = p.forward(torch.tensor(ensemble.perturbed_taylor_green(N, 2024), dtype=dtype, device="cuda"))
    aws = p.adjoint_workspace(n, seg)

    def grad():
        J = p.adjoint_forward(x, n, dt, nu, 1, seg, aws)
        g, dnu = p.adjoint_backward(n, dt, nu, seg, aws)
        return J, g.clone(), dnu

    p.set_recompute_variants(None)
    assert p.get_recompute_variants() == p.get_variants()
    J0, g0, n0 = grad()
    tol = 1e-12 if dtype == torch.float64 else 2e-5
    for vs in ([0] * 5, [1] * 5, [2] * 5, [1, 2, 1, 2, 0]):
        p.set_recompute_variants(vs)
        assert p.get_recompute_variants() == vs
        J, g, dnu = grad()
        assert J == J0  # the forward sweep keeps the plan's variants
        err = (torch.linalg.vector_norm(g - g0) / torch.linalg.vector_norm(g0)).item()
        assert err <= tol, (vs, err)
        assert abs(dnu - n0) <= tol * abs(n0), (vs, dnu, n0)
    best, seen = autotune.tune_recompute(p, nsteps=15, seg=3, use_cache=False)
    assert p.get_recompute_variants() == best and len(seen) >= 4
    with pytest.raises(Exception):
        p.set_recompute_variants([3, 0, 0, 0, 0])


def test_sweep_at_256_fd():
    """The reverse sweep at 256^3 (8 points per lane in the register forward x,
    the five-field forward y, the fused inverse x of q) against central
    differences: 10 steps in segments of 2 (one recomputed), J = Z(T), three
    seeded directions and nu, relative 1e-6."""
    from jax_tg_project_amd import ensemble
    N, nu, dt, n, seg = 256, 1.0 / 1600, 2e-3, 10, 2
    p = _plan(N)
    x = p.forward(torch.tensor(ensemble.perturbed_taylor_green(N, 2024), device="cuda"))
    aws = p.adjoint_workspace(n, seg)
    J = p.adjoint_forward(x, n, dt, nu, 1, seg, aws)
    g, dnu = p.adjoint_backward(n, dt, nu, seg, aws)
    del aws

    def Jof(xx, nuv):
        y = xx.clone()
        p.rk4(y, nsteps=n, dt=dt, nu=nuv)
        return p.diagnostics(y, divmax=False)["enstrophy"]

    rng = np.random.default_rng(256)
    eps = 1e-6
    xn = torch.linalg.vector_norm(x).item()
    for d in range(3):
        v = torch.tensor(rng.standard_normal(x.shape) + 1j * rng.standard_normal(x.shape), device="cuda")
        v *= xn / torch.linalg.vector_norm(v).item()
        fd = (Jof(x + eps * v, nu) - Jof(x - eps * v, nu)) / (2 * eps)
        an = float(torch.sum(torch.real(torch.conj(g) * v)))
        assert abs(an - fd) <= 1e-6 * abs(fd), (d, an, fd)
    h = 1e-5 * nu
    fdn = (Jof(x, nu + h) - Jof(x, nu - h)) / (2 * h)
    assert abs(dnu - fdn) <= 1e-6 * abs(fdn), (dnu, fdn)
    assert abs(J - Jof(x, nu)) < 1e-14 * J


def test_fused_folded_sweep_at_128_matches_torch_autograd():
    """N = 128 runs the fused and folded reverse sweep: the register forward x
    adds the kx terms of i k x p2 and does the next stage VJP's inverse x of
    its cotangent, the five-field forward y folds the ky terms, the z pass the
    kz terms.  Five steps in segments of 1 (the first one recomputed on the
    second stream) against the torch autograd restatement, the gradient to
    1e-11 and dJ/dnu to 1e-10 relative."""
    N, nu, dt, n, seg = 128, 1.0 / 400, 5e-3, 5, 1
    x = _state(N, 12, amp=0.05)
    c = torch.tensor(x, requires_grad=True)
    nut = torch.tensor(nu, dtype=torch.float64, requires_grad=True)
    J = ad.final_objective(c, nut, n, dt, N, "enstrophy")
    gc, gnu = torch.autograd.grad(J, (c, nut))
    p = _plan(N)
    xg = torch.tensor(x, device="cuda")
    aws = p.adjoint_workspace(n, seg)
    Jg = p.adjoint_forward(xg, n, dt, nu, 1, seg, aws)
    g, dnu = p.adjoint_backward(n, dt, nu, seg, aws)
    assert abs(Jg - J.item()) < 1e-13 * abs(J.item())
    err = np.linalg.norm(g.cpu().numpy() - gc.numpy()) / np.linalg.norm(gc.numpy())
    assert err < 1e-11, err
    assert abs(dnu - gnu.item()) < 1e-10 * abs(gnu.item()), (dnu, gnu.item())
