"""tg_diagnostics_noalloc (the allocation-free variant of tg_diagnostics)
against tg_diagnostics through the C ABI: two calls in a row on one GPU,
the second with a smaller grid, for hydro and MHD plans in FP64 and FP32,
with and without the divergence maxima; and the clear error beyond the
scratch limit.  Same kernels and summation order: equal bitwise."""
import math

import numpy as np
import pytest
import torch

from tests.helpers import band_limited_field

pytestmark = pytest.mark.gpu


def _state(plan, seed):
    u = torch.tensor(band_limited_field(plan.n, seed), dtype=plan.dtype, device=plan.device)
    uh = plan.forward(u)
    bh = plan.forward((0.3 * torch.roll(u, 1, dims=1)).contiguous()) if plan.mhd else None
    return uh, bh


@pytest.mark.parametrize("mhd", [False, True], ids=["hydro", "mhd"])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["fp64", "fp32"])
def test_noalloc_diagnostics_match_tg_diagnostics_two_calls_shrinking(dtype, mhd):
    from jax_tg_project_amd._lib import Plan, TG_DIAG_COUNT
    keys = ("kinetic_energy", "enstrophy", "grad_sq", "magnetic_energy", "cross_helicity", "div_u_sq",
            "div_b_sq", "current_sq", "div_u_max", "div_b_max", "nonfinite")
    for divmax in (True, False):
        out = torch.full((TG_DIAG_COUNT,), float("nan"), dtype=torch.float64, device="cuda")
        for n, seed in ((64, 3), (32, 4)):  # the second call on a smaller grid reuses the scratch
            p = Plan(n, dtype=dtype, mhd=mhd)
            uh, bh = _state(p, seed)
            want = p.diagnostics(uh, bh, divmax=divmax)
            p.diagnostics_noalloc(uh, bh, divmax=divmax, out=out)
            torch.cuda.synchronize()
            got = out.cpu().numpy()
            for i, k in enumerate(keys):
                assert got[i] == want[k], (n, divmax, k, got[i], want[k])
            assert math.isfinite(got[0]) and got[0] > 0


def test_noalloc_rejects_grids_beyond_the_scratch():
    from jax_tg_project_amd._lib import Plan, TgError
    from jax_tg_project_amd.taylor_green_jax import initial_taylor_green
    p = Plan(256, dtype=torch.float64)
    uh = p.forward(initial_taylor_green(256, device="cuda"))
    with pytest.raises(TgError, match="scratch"):
        p.diagnostics_noalloc(uh, divmax=True)
    d = p.diagnostics_noalloc(uh, divmax=False)
    torch.cuda.synchronize()
    assert d[0].item() == p.diagnostics(uh, divmax=False)["kinetic_energy"]
