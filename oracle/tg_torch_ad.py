"""torch-CPU FP64 autograd restatement of the hydro RK4 map on the compact
spectrum -- the automatic-differentiation oracle (standing in for jax.grad
of jax/taylor_green_jax.py:188-196, which cannot run here).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Rotational form (equal
to the reference's advective form after truncation and projection when
3*(N//3) < N).  A compact spectrum c (3, K, K, KZ) is placed in the full
cube with Hermitian completion; `.real` of the inverse FFT takes the
Hermitian part, as the GPU's C2R does.
"""
import math

import numpy as np
import torch


def _kept(N):
    M = N // 3
    return torch.cat([torch.arange(0, M + 1), torch.arange(N - M, N)])


def to_full(c, N):
    M = N // 3
    idx = _kept(N)
    full = torch.zeros((3, N, N, N), dtype=torch.complex128)
    ix, iy, iz = idx[:, None, None], idx[None, :, None], torch.arange(M + 1)[None, None, :]
    full[:, ix, iy, iz] = c
    nx = (-idx) % N
    izn = (-torch.arange(1, M + 1)) % N
    full[:, nx[:, None, None], nx[None, :, None], izn[None, None, :]] = torch.conj(c[..., 1:])
    return full


def to_compact(full, N):
    M = N // 3
    idx = _kept(N)
    return full[:, idx][:, :, idx][..., : M + 1]


def wavenumbers(N):
    k = torch.tensor(np.fft.fftfreq(N) * N, dtype=torch.float64)
    return torch.meshgrid(k, k, k, indexing="ij")


def rhs_full(c, nu, N):
    """Compact spectrum -> compact RHS, differentiable."""
    kx, ky, kz = wavenumbers(N)
    full = to_full(c, N)
    u = torch.fft.ifftn(full, dim=(-3, -2, -1)).real
    om_hat = torch.stack([1j * (ky * full[2] - kz * full[1]), 1j * (kz * full[0] - kx * full[2]),
                          1j * (kx * full[1] - ky * full[0])])
    om = torch.fft.ifftn(om_hat, dim=(-3, -2, -1)).real
    n = torch.cross(u, om, dim=0)
    nh = to_compact(torch.fft.fftn(n, dim=(-3, -2, -1)), N)
    ckx, cky, ckz = (to_compact(torch.stack([k, k, k]).to(torch.complex128), N)[0].real for k in (kx, ky, kz))
    k2 = ckx * ckx + cky * cky + ckz * ckz
    k2s = torch.where(k2 == 0, torch.ones_like(k2), k2)
    kd = ckx * nh[0] + cky * nh[1] + ckz * nh[2]
    p = nh - torch.stack([ckx * kd / k2s, cky * kd / k2s, ckz * kd / k2s])
    return p - nu * k2 * c


def rk4(c, dt, nu, N):
    r1 = rhs_full(c, nu, N)
    r2 = rhs_full(c + 0.5 * dt * r1, nu, N)
    r3 = rhs_full(c + 0.5 * dt * r2, nu, N)
    r4 = rhs_full(c + dt * r3, nu, N)
    return c + (dt / 6.0) * (r1 + 2.0 * r2 + 2.0 * r3 + r4)


def objective(c, N, kind="enstrophy"):
    """E or Z of a compact spectrum (Parseval with Hermitian weights)."""
    M = N // 3
    w = torch.ones(M + 1, dtype=torch.float64)
    w[1:] = 2.0
    if kind == "energy":
        s = (w * (c.abs() ** 2)).sum()
    else:
        kx, ky, kz = wavenumbers(N)
        ckx, cky, ckz = (to_compact(torch.stack([k, k, k]).to(torch.complex128), N)[0].real for k in (kx, ky, kz))
        cr = torch.stack([cky * c[2] - ckz * c[1], ckz * c[0] - ckx * c[2], ckx * c[1] - cky * c[0]])
        s = (w * (cr.abs() ** 2)).sum()
    return 0.5 * s / float(N) ** 6


def final_objective(c0, nu, nsteps, dt, N, kind="enstrophy"):
    c = c0
    for _ in range(nsteps):
        c = rk4(c, dt, nu, N)
    return objective(c, N, kind)
