"""GPU parity of the pruned 3D real transforms against numpy.fft (FP64 to a
few ulp * log N, FP32 to ~1e-6), through the C ABI (tg_forward / tg_inverse)."""
import numpy as np
import pytest
import torch

from oracle import tg_oracle as o
from tests.helpers import random_compact, rel_l2

pytestmark = pytest.mark.gpu

SIZES = [4, 8, 16, 32, 48, 64, 96, 128]
TOL = {torch.float64: 1e-13, torch.float32: 3e-6}


def _plan(N, dtype):
    from jax_tg_project_amd._lib import Plan
    return Plan(N, dtype=dtype)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("N", SIZES)
def test_forward_matches_masked_fftn(N, dtype):
    rng = np.random.default_rng(N)
    u = rng.standard_normal((3, N, N, N))
    p = _plan(N, dtype)
    got = p.forward(torch.tensor(u, dtype=dtype, device="cuda")).cpu().numpy()
    want = o.full_to_compact(np.fft.fftn(u, axes=(-3, -2, -1)))
    assert got.shape == want.shape
    assert rel_l2(got, want) < TOL[dtype], rel_l2(got, want)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("N", SIZES)
def test_inverse_matches_ifftn(N, dtype):
    c = random_compact(N, 100 + N)
    p = _plan(N, dtype)
    ct = torch.tensor(c, dtype=p.cdtype, device="cuda")
    got = p.inverse(ct).cpu().numpy()
    full = o.compact_to_full(c, N)
    want = np.fft.ifftn(full, axes=(-3, -2, -1)).real
    assert rel_l2(got, want) < TOL[dtype], rel_l2(got, want)
    # curl and divergence outputs
    kx, ky, kz = o.make_wavenumbers(N, 2 * np.pi)
    curl = np.fft.ifftn(np.stack([1j * (ky * full[2] - kz * full[1]), 1j * (kz * full[0] - kx * full[2]),
                                  1j * (kx * full[1] - ky * full[0])]), axes=(-3, -2, -1)).real
    from jax_tg_project_amd._lib import TG_OP_CURL, TG_OP_DIV
    assert rel_l2(p.inverse(ct, op=TG_OP_CURL).cpu().numpy(), curl) < TOL[dtype]
    div = np.fft.ifftn(1j * (kx * full[0] + ky * full[1] + kz * full[2]), axes=(-3, -2, -1)).real
    assert rel_l2(p.inverse(ct, op=TG_OP_DIV).cpu().numpy(), div) < TOL[dtype]


@pytest.mark.parametrize("N", [8, 32, 256])
def test_round_trip_band_limited(N):
    c = random_compact(N, 7)
    u = np.fft.ifftn(o.compact_to_full(c, N), axes=(-3, -2, -1)).real
    p = _plan(N, torch.float64)
    ut = torch.tensor(u, device="cuda")
    back = p.inverse(p.forward(ut)).cpu().numpy()
    assert rel_l2(back, u) < 1e-14
