"""State dump and restart (checkpoint files of the compact state) and the
command-line entry points: the native taylor_green_hip executable (the
argv and save-line contract of cuda/taylor_green_cuda.cu:117-125) and the
package CLIs of the three drop-in modules."""
import math
import os
import re
import subprocess

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "jax_tg_project_amd", "taylor_green_hip")
LINE = re.compile(r"Save time step: (\d+), time: ([0-9.]+), enstrophy: ([-+0-9.e]+), kinetic energy: ([-+0-9.e]+)")


def _tg(N, dtype):
    from jax_tg_project_amd.ensemble import perturbed_taylor_green
    return torch.tensor(perturbed_taylor_green(N, 77), dtype=dtype, device="cuda")


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["fp64", "fp32"])
def test_checkpoint_round_trip_and_restart_is_bitwise(tmp_path, dtype):
    from jax_tg_project_amd.solver import TaylorGreenSolver
    N, nu, dt = 32, 0.01, 0.01
    a = TaylorGreenSolver(N, nu, dtype=dtype)
    a.set_velocity(_tg(N, dtype))
    b = TaylorGreenSolver(N, nu, dtype=dtype)
    b.set_spectrum(a.u_hat)
    a.step(10, dt)
    b.step(4, dt)
    path = str(tmp_path / "hydro.tgck")
    b.save(path, step=4, time=4 * dt, dt=dt)
    c, meta = TaylorGreenSolver.load(path)
    assert meta["step"] == 4 and meta["dt"] == dt and meta["nu"] == nu and c.plan.dtype == dtype
    assert torch.equal(c.u_hat, b.u_hat)
    c.step(6, dt)
    assert torch.equal(c.u_hat, a.u_hat)


def test_mhd_checkpoint_round_trip_and_plan_mismatch(tmp_path):
    from jax_tg_project_amd.solver import MhdSolver, TaylorGreenSolver
    from jax_tg_project_amd._lib import TgError
    s = MhdSolver(16, 0.02, 0.03, lorentz_sign=1.0)
    u = _tg(16, torch.float64)
    s.set_fields(u, (0.5 * u).contiguous())
    s.step(3, 0.005)
    path = str(tmp_path / "mhd.tgck")
    s.save(path, step=3, time=0.015, dt=0.005)
    r, meta = MhdSolver.load(path)
    assert meta["eta"] == 0.03 and meta["lorentz_sign"] == 1.0 and r.lorentz_sign == 1.0
    assert torch.equal(r.u_hat, s.u_hat) and torch.equal(r.b_hat, s.b_hat)
    with pytest.raises(ValueError, match="MHD state"):
        TaylorGreenSolver.load(path)
    h = TaylorGreenSolver(16, 0.02)
    with pytest.raises(TgError, match="does not match the plan"):
        h.plan.checkpoint_read(path, h.u_hat)


def _cli(*args):
    r = subprocess.run([EXE] + [str(a) for a in args], capture_output=True, text=True, timeout=120)
    return r.returncode, r.stdout, r.stderr


def test_native_cli_matches_the_solver_and_restarts(tmp_path):
    """taylor_green_hip 16 0.01 100 0.1 0.05 --precision 64: 10 steps, saves
    at 5 and 10; the printed enstrophy and energy equal the library solver's
    (same float32 dt and nu = 2 pi / Re), and a run cut at step 5 and
    restarted from its checkpoint prints the same final line."""
    from jax_tg_project_amd.solver import TaylorGreenSolver
    from jax_tg_project_amd.taylor_green_jax import initial_taylor_green
    rc, out, err = _cli(16, 0.01, 100, 0.1, 0.05, "--precision", 64, "--checkpoint", tmp_path / "full.tgck")
    assert rc == 0, err
    assert out.splitlines()[0] == "HIP Taylor-Green: N=16 Re=100.0 steps=10"
    rows = [LINE.match(l).groups() for l in out.splitlines()[1:]]
    assert [int(r[0]) for r in rows] == [5, 10]
    dt = float(np.float32(0.01))
    s = TaylorGreenSolver(16, 2.0 * math.pi / 100.0)
    s.set_velocity(initial_taylor_green(16, device="cuda"))
    for k, row in zip((5, 10), rows):
        s.step(5, dt)
        d = s.diagnostics(divmax=False)
        assert row[1] == f"{np.float32(k * np.float32(0.01)):.6f}"
        # the executable builds the initial field with the host's sin/cos: agreement to the printed digits
        assert abs(float(row[2]) - d["enstrophy"]) <= 1e-8 * d["enstrophy"]
        assert abs(float(row[3]) - d["kinetic_energy"]) <= 1e-8 * d["kinetic_energy"]
    rc, out5, err = _cli(16, 0.01, 100, 0.05, 0.05, "--precision", 64, "--checkpoint", tmp_path / "half.tgck")
    assert rc == 0, err
    rc, outr, err = _cli(16, 0.01, 100, 0.1, 0.05, "--precision", 64, "--restart", tmp_path / "half.tgck")
    assert rc == 0, err
    assert outr.splitlines()[-1] == out.splitlines()[-1]
    assert len(outr.splitlines()) == 2  # header + the one save after the restart point
    rc, _, err = _cli(16, 0.02, 100, 0.1, 0.05, "--precision", 64, "--restart", tmp_path / "half.tgck")
    assert rc == 1 and "dt=" in err


def test_native_cli_fp32_default_generic_and_unsupported_size():
    rc, out, err = _cli(32, 0.005, 1000, 0.02, 0.01)
    assert rc == 0, err
    rows = [LINE.match(l).groups() for l in out.splitlines()[1:]]
    assert [int(r[0]) for r in rows] == [2, 4]
    assert abs(float(rows[0][3]) - 0.125) < 1e-3 and abs(float(rows[0][2]) - 0.375) < 1e-2
    # 5 divides 20: no compiled kernels, the direct-sum path runs it
    rc, out, err = _cli(20, 0.005, 1000, 0.02, 0.01)
    assert rc == 0, err
    rows = [LINE.match(l).groups() for l in out.splitlines()[1:]]
    assert [int(r[0]) for r in rows] == [2, 4] and abs(float(rows[0][3]) - 0.125) < 1e-3
    rc, _, err = _cli(4096, 0.005, 1000, 0.02, 0.01)  # above the largest grid a plan accepts
    assert rc == 2 and "N=4096" in err


def test_python_package_clis(tmp_path, monkeypatch):
    from jax_tg_project_amd import taylor_green_jax, taylor_green_mhd
    monkeypatch.chdir(tmp_path)
    u, energies = taylor_green_jax.main(["--n", "16", "--total-time", "0.02", "--save-every-time", "0.01",
                                         "--no-plots"])
    assert len(energies) == 3 and all(math.isfinite(e) for e in energies)
    _, _, d = taylor_green_mhd.main(["--n", "16", "--dt", "0.005", "--total-time", "0.02", "--save-every-time",
                                     "0.01"])
    assert len(d.times) == 3
    assert (tmp_path / "taylor_green_mhd_slice.png").exists()
    assert (tmp_path / "taylor_green_mhd_diagnostics.png").exists()
