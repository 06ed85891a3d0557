"""Benchmark: RK4 steps/s of the hydro Taylor-Green DNS at 256^3 on one MI355X.

Workload: TGV u = sin x cos y cos z, v = -cos x sin y cos z, w = 0 on
[0, 2 pi]^3, N = 256, nu = 1/1600, dt = 2e-3, classic RK4 with 2/3
dealiasing and Leray projection; FP64 (the plain drop-in precision) is the
headline, FP32 is reported beside it.  The initial state is transformed to
the compact spectrum in HBM before the timed region; the timed region is
exactly --steps RK4 steps (4 right-hand sides each, every pass on the GPU),
synchronised on both sides, diagnostics excluded.

Prints ONE JSON line.  --gpus above 1 exits: this path runs on one GPU.
"""
import argparse
import json
import math
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hydro_256_fp64_traffic.txt")
KERNEL_OF_PASS = {"inv_x": "k_inv_x<double, 256, 7, {v}>", "inv_y": "k_y_curl<double, 256, {v}>",
                  "z": ["k_zw<double, 256, 16, 4>", "k_zw<double, 256, 8, 4>", "k_z<double, 256, 0, 2>"],
                  "fwd_y": "k_y<double, 256, -1, {v}>",
                  "fwd_x": ["k_fwd_x<double, 256, 0, 0>", "k_fwd_xr<double, 256, false>", "k_fwd_x<double, 256, 0, 2>"]}
PASS_INDEX = {"inv_x": 0, "inv_y": 1, "z": 2, "fwd_y": 3, "fwd_x": 4}


def measured_traffic(pass_name, variants):
    """HBM bytes per launch of the FP64 256^3 kernel for this pass and tile
    variant, from the committed rocprofv3 counter summary (None if absent)."""
    try:
        with open(TRAFFIC_FILE) as f:
            rows = [ln.rstrip("\n").split("\t") for ln in f if not ln.startswith("#")]
    except OSError:
        return None
    v = variants[PASS_INDEX[pass_name]]
    k = KERNEL_OF_PASS[pass_name]
    want = "tg::" + (k[v] if isinstance(k, list) else k.format(v=v))
    for r in rows:
        if len(r) == 3 and r[0] == want:
            return float(r[1]) + float(r[2])
    return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def tg_initial(N, dtype, device):
    import torch
    x = torch.arange(N, dtype=torch.float64, device=device) * (2.0 * math.pi / N)
    X, Y, Z = torch.meshgrid(x, x, x, indexing="ij")
    u = torch.stack([torch.sin(X) * torch.cos(Y) * torch.cos(Z),
                     -torch.cos(X) * torch.sin(Y) * torch.cos(Z),
                     torch.zeros_like(X)])
    return u.to(dtype).contiguous()


def run_gpu(N, dtype, steps, warmup, dt, nu, tune=True):
    import torch
    from jax_tg_project_amd._lib import Plan
    from jax_tg_project_amd import autotune
    plan = Plan(N, dtype=dtype)
    tuned = autotune.tune(plan, use_cache=False)[0] if tune else plan.get_variants()
    u0 = tg_initial(N, dtype, plan.device)
    uh = plan.forward(u0)
    if warmup:
        w = uh.clone()
        plan.rk4(w, nsteps=warmup, dt=dt, nu=nu)
        del w
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan.rk4(uh, nsteps=steps, dt=dt, nu=nu)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    diag = plan.diagnostics(uh)
    # same steps again with the per-pass HIP-event hook on (kernel durations)
    uh2 = plan.forward(u0)
    plan.timing(True)
    plan.rk4(uh2, nsteps=steps, dt=dt, nu=nu)
    passes = plan.timing_read()
    plan.timing(False)
    torch.cuda.synchronize()
    return {"wall": wall, "diag": diag, "passes": passes, "plan": plan, "u_hat": uh, "variants": tuned}


def oracle_history():
    """The oracle's 500-step history of the 256^3 workload (tests/golden)."""
    try:
        with open(os.path.join(ROOT, "tests", "golden", "tgv256_history.json")) as f:
            return json.load(f)
    except OSError:
        return None


def cpu_c1(workers):
    """Config C1 in full on the host: the oracle's run_simulation(N=32,
    dt=0.01, nu=0.01, total_time=1.0): 100 RK4 steps and 21 diagnostic
    records, wall seconds."""
    from oracle import tg_oracle as o
    o.set_workers(workers)
    t0 = time.perf_counter()
    o.run_simulation(N=32, dt=0.01, nu=0.01, total_time=1.0, save_every_time=0.05, return_diagnostics=True)
    return time.perf_counter() - t0


def gpu_c1():
    """The same run through the drop-in run_simulation on the GPU (diagnostics included)."""
    import torch
    from jax_tg_project_amd.taylor_green_jax import run_simulation
    import contextlib
    import io
    with contextlib.redirect_stdout(io.StringIO()):
        run_simulation(N=32, dt=0.01, nu=0.01, total_time=1.0, save_every_time=0.05, make_plots=False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run_simulation(N=32, dt=0.01, nu=0.01, total_time=1.0, save_every_time=0.05, make_plots=False)
        torch.cuda.synchronize()
    return time.perf_counter() - t0


def cpu_baseline(N, dt, nu, gpu_one_step=None):
    """One RK4 step of the FP64 oracle (jax/taylor_green_jax.py restated:
    72 complex 3D FFTs per step) on all host cores given to this process."""
    import numpy as np
    from oracle import tg_oracle as o
    cores = o.cpu_count()
    o.set_workers(cores)
    L = 2.0 * math.pi
    kx, ky, kz = o.make_wavenumbers(N, L)
    mask = o.make_dealias_mask(N, L)
    u = o.initial_taylor_green(N)
    t0 = time.perf_counter()
    u1 = o.advance_one_step(u, dt, nu, kx, ky, kz, mask)
    sec = time.perf_counter() - t0
    out = {"value": 1.0 / sec, "unit": "RK4 steps/s", "cores": cores, "kind": "port",
           "sample": f"1 RK4 step of the {N}^3 FP64 hydro TGV (72 complex 3D FFTs, scipy.fft "
                     f"workers={cores}) on {cpu_model()}; {sec:.2f} s"}
    parity = None
    if gpu_one_step is not None:
        d = gpu_one_step - u1
        parity = float(np.linalg.norm(d.ravel()) / np.linalg.norm(u1.ravel()))
    return out, parity


def leg_roofline(passes, profile):
    """Roofline object of a leg from its HIP-event pass timings: the pass
    with the largest time per launch, its algorithmic GB/s over 8 TB/s.
    Measured traffic of the leg's kernels is in the committed summary
    `profile` (rocprofv3 counters), not re-measured here."""
    dom = max(passes, key=lambda k: passes[k]["ms_avg"])
    a = passes[dom]["GBps"]
    return {"bound": "hbm", "kernel": dom, "achieved": a, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": a / HBM_PEAK_GBS, "traffic": None, "profile": profile}


def run_mhd(N, steps, dt, nu, eta, dtype, lorentz_sign=1.0):
    """MHD TGV: u = TG, b = (cos x sin y sin z, sin x cos y sin z,
    -2 sin x sin y cos z) (= -curl u_TG), nu = eta, `steps` RK4 steps on the
    compact spectra after one untimed warm-up step; FP64 diagnostics."""
    import torch
    from jax_tg_project_amd._lib import Plan
    from jax_tg_project_amd import autotune
    plan = Plan(N, dtype=dtype, mhd=True)
    tuned = autotune.tune(plan, use_cache=False)[0]
    from jax_tg_project_amd import roofline
    x = torch.arange(N, dtype=torch.float64, device=plan.device) * (2.0 * math.pi / N)
    X, Y, Z = torch.meshgrid(x, x, x, indexing="ij")
    u = torch.stack([torch.sin(X) * torch.cos(Y) * torch.cos(Z), -torch.cos(X) * torch.sin(Y) * torch.cos(Z),
                     torch.zeros_like(X)]).to(dtype).contiguous()
    b = torch.stack([torch.cos(X) * torch.sin(Y) * torch.sin(Z), torch.sin(X) * torch.cos(Y) * torch.sin(Z),
                     -2.0 * torch.sin(X) * torch.sin(Y) * torch.cos(Z)]).to(dtype).contiguous()
    del X, Y, Z
    uh, bh = plan.forward(u), plan.forward(b)
    w, v = uh.clone(), bh.clone()
    plan.rk4(w, v, nsteps=1, dt=dt, nu=nu, eta=eta, lorentz_sign=lorentz_sign)
    del w, v
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan.rk4(uh, bh, nsteps=steps, dt=dt, nu=nu, eta=eta, lorentz_sign=lorentz_sign)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    d = plan.diagnostics(uh, bh)
    # per-pass kernel times over a few more steps (HIP events on the launch stream)
    plan.timing(True)
    plan.rk4(uh.clone(), bh.clone(), nsteps=min(steps, 20), dt=dt, nu=nu, eta=eta, lorentz_sign=lorentz_sign)
    pas = plan.timing_read()
    plan.timing(False)
    pb = roofline.pass_bytes(N, 8 if dtype == torch.float64 else 4, mhd=True)
    passes = {k: {"ms_avg": v[0] / max(1, v[1]), "launches": v[1],
                  "GBps": (pb[k] if k != "fwd_x" else sum(pb["fwd_x"]) / 4.0) / (v[0] / max(1, v[1]) * 1e-3) / 1e9}
              for k, v in pas.items()}
    step_b = roofline.step_bytes(N, 8 if dtype == torch.float64 else 4, mhd=True)
    return {"steps_per_s": steps / wall, "wall_s": wall, "tile_variants": tuned, "passes": passes,
            "hbm_GBps_algorithmic_step": step_b * steps / wall / 1e9,
            "roofline": leg_roofline(passes, "profiles/mhd_128_kernel_stats_and_traffic.md"),
            "kinetic_energy": d["kinetic_energy"], "magnetic_energy": d["magnetic_energy"],
            "cross_helicity": d["cross_helicity"], "div_u_max": d["div_u_max"], "div_b_max": d["div_b_max"]}


def run_ensemble(ninst, steps, dt, seed=0, sizes=None):
    """Seeded 64-instance FP32 ensemble (sizes {32,48,64,96}, nu log-uniform
    in [1/1600, 1/100], TG + |k|<=4 perturbation of rms 0.05): wall time of
    `steps` RK4 steps, one captured HIP graph replay per step.  `sizes`
    keeps only the instances of those grid sizes (for per-size profiles)."""
    import torch
    from jax_tg_project_amd import ensemble, roofline
    cfg = ensemble.ensemble_config(ninst, seed=seed)
    if sizes:
        cfg = [c for c in cfg if c[0] in sizes]
        ninst = len(cfg)
    ens = ensemble.build_ensemble(cfg, dtype=torch.float32)
    layout = ens.tune_launch(dt)  # times both launch layouts, keeps the faster
    ens.step(1, dt, graph=True)  # capture + warm-up
    for i, (N, _, s) in enumerate(cfg):
        ens.set_velocity(i, ensemble.perturbed_taylor_green(N, s))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.step(steps, dt, graph=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    e_eager = ensemble.build_ensemble(cfg, dtype=torch.float32)
    e_eager.set_launch(ens.mixed)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    e_eager.step(steps, dt, graph=False)
    torch.cuda.synchronize()
    wall_eager = time.perf_counter() - t1
    same = bool(torch.equal(ens.state, e_eager.state))
    nbytes = sum(roofline.step_bytes(N, 4) for N, _, _ in cfg) * steps
    counts = {n: sum(1 for c in cfg if c[0] == n) for n in ensemble.ENSEMBLE_SIZES}
    rate = nbytes / wall / 1e9
    return {"roofline": {"bound": "hbm", "kernel": "whole graph step (20 mixed launches)", "achieved": rate,
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": rate / HBM_PEAK_GBS, "traffic": None,
                         "profile": "profiles/ensemble_kernel_stats_and_traffic.md"},
            "value": ninst / wall, "unit": "solves/s", "instances": ninst, "steps": steps, "dt": dt,
            "seed": seed, "sizes": counts, "launch_layout": layout, "wall_s": wall, "wall_eager_s": wall_eager,
            "graph_equals_eager": same, "hbm_GBps_algorithmic": nbytes / wall / 1e9}


def run_adjoint(N, steps, dt, nu, dtype, all_stages=True):
    """Adjoint timing at N with J = Z(T): forward = `steps` plain RK4 steps;
    adjoint = forward sweep with sqrt(n) checkpoints + reverse sweep (and,
    with `all_stages`, the schedule that keeps every stage)."""
    import torch
    from jax_tg_project_amd._lib import Plan
    from jax_tg_project_amd import ensemble
    from jax_tg_project_amd.adjoint import sqrt_segment
    from jax_tg_project_amd import autotune
    plan = Plan(N, dtype=dtype)
    tuned = autotune.tune(plan, use_cache=False)[0]  # forward and reverse sweeps run the tuned tiles
    rtuned = autotune.tune_recompute(plan, use_cache=False)[0]  # the recompute beside the reverse sweep
    x0 = plan.forward(torch.tensor(ensemble.perturbed_taylor_green(N, 2024), dtype=dtype, device=plan.device))
    seg = sqrt_segment(steps)
    y = x0.clone()
    plan.rk4(y, nsteps=2, dt=dt, nu=nu)  # warm-up
    J = plan.adjoint_forward(x0, 2, dt, nu, 1, 1, plan.adjoint_workspace(2, 1))
    torch.cuda.synchronize()
    y = x0.clone()
    t0 = time.perf_counter()
    plan.rk4(y, nsteps=steps, dt=dt, nu=nu)
    torch.cuda.synchronize()
    fwd = time.perf_counter() - t0
    out = {"N": N, "steps": steps, "dtype": "f64" if dtype == torch.float64 else "f32", "tile_variants": tuned,
           "recompute_tile_variants": rtuned, "forward_s": fwd}
    for name, sg in (("sqrt_n", seg), ("all_stages_kept", steps))[:2 if all_stages else 1]:
        aws = plan.adjoint_workspace(steps, sg)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        J = plan.adjoint_forward(x0, steps, dt, nu, 1, sg, aws)
        g, dnu = plan.adjoint_backward(steps, dt, nu, sg, aws)
        torch.cuda.synchronize()
        adj = time.perf_counter() - t0
        out[name] = {"segment": sg, "forward_plus_adjoint_s": adj, "adjoint_over_forward": adj / fwd,
                     "workspace_GB": aws.numel() / 1e9, "J_enstrophy": J, "dJ_dnu": dnu}
        del aws
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-mhd", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--only", choices=("fp32", "ensemble", "mhd", "adjoint"),
                    help="run one leg alone (for profiling) and print its JSON object")
    ap.add_argument("--ensemble-sizes", type=int, nargs="+", default=None,
                    help="with --only ensemble: keep only the instances of these grid sizes")
    ap.add_argument("--leg-steps", type=int, default=None,
                    help="steps of the --only leg (default: the leg's workload)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="with --only: start with a 1 GiB device copy (counter calibration)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="only run a 1 GiB device copy (counter calibration) and 3 tuned FP64 steps, "
                         "for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this pseudo-spectral path runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU slab/pencil decomposition is design-only)", file=sys.stderr)
        sys.exit(2)

    import torch
    from jax_tg_project_amd import roofline
    N, dt, nu = args.size, 2e-3, 1.0 / 1600.0
    if args.traffic_probe:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        from jax_tg_project_amd._lib import Plan
        from jax_tg_project_amd import autotune
        plan = Plan(N, dtype=torch.float64)
        tuned = autotune.tune(plan, use_cache=False)[0]
        uh = plan.forward(tg_initial(N, torch.float64, plan.device))
        plan.rk4(uh, nsteps=3, dt=dt, nu=nu)
        torch.cuda.synchronize()
        print(json.dumps({"traffic_probe": True, "copy_bytes_read": 1 << 30, "copy_bytes_written": 1 << 30,
                          "tile_variants": tuned}), flush=True)
        return
    K, W = args.steps, args.warmup
    if args.only:
        if args.calibrate_copy:
            src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
            dst = torch.empty_like(src)
            for _ in range(3):
                dst.copy_(src)
            torch.cuda.synchronize()
            del src, dst
        ls = args.leg_steps
        if args.only == "fp32":
            r = run_gpu(N, torch.float32, ls or K, W, dt, nu, tune=not args.no_tune)
            out = {"fp32_steps_per_s": (ls or K) / r["wall"], "tile_variants_fp32": r["variants"],
                   "passes_fp32": {k: {"ms_avg": v[0] / max(1, v[1]), "launches": v[1]}
                                   for k, v in r["passes"].items()}}
        elif args.only == "ensemble":
            out = run_ensemble(64, ls or 200, 5e-3, sizes=args.ensemble_sizes)
        elif args.only == "mhd":
            out = {"fp64": run_mhd(128, ls or 300, 2e-3, 1.0 / 200, 1.0 / 200, torch.float64),
                   "fp32": run_mhd(128, ls or 300, 2e-3, 1.0 / 200, 1.0 / 200, torch.float32)}
        else:
            out = [run_adjoint(128, ls or 200, 5e-3, 1.0 / 400, d_, all_stages=False)
                   for d_ in (torch.float64, torch.float32)]
        print(json.dumps({"only": args.only, "result": out}), flush=True)
        return

    r64 = run_gpu(N, torch.float64, K, W, dt, nu, tune=not args.no_tune)
    steps_s = K / r64["wall"]
    pb = roofline.pass_bytes(N, 8)
    step_b = roofline.step_bytes(N, 8)
    # dominant kernel = largest total time over the timed steps
    pas = r64["passes"]
    dom = max(pas, key=lambda k: pas[k][0])
    dom_ms_avg = pas[dom][0] / max(1, pas[dom][1])
    dom_bytes = pb[dom] if dom != "fwd_x" else sum(pb["fwd_x"]) / 4.0
    achieved = dom_bytes / (dom_ms_avg * 1e-3) / 1e9
    per_pass = {k: {"ms_avg": v[0] / max(1, v[1]), "launches": v[1],
                    "GBps": ((pb[k] if k != "fwd_x" else sum(pb["fwd_x"]) / 4.0)
                             / (v[0] / max(1, v[1]) * 1e-3) / 1e9)} for k, v in pas.items()}

    res = {
        "metric": "RK4 steps/s + HBM GB/s at 256^3 (hydro TGV, FP64); adjoint/forward time; ensemble solves/s",
        "value": steps_s,
        "unit": "RK4 steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": 1e3 * r64["wall"] / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Taylor-Green initial condition, generated on the GPU)",
        "config": {"workload": f"hydro TGV N={N}^3 FP64, nu=1/1600, dt=2e-3, classic RK4, 2/3 dealiasing, "
                               f"Leray projection, {K} steps", "N": N, "nu": nu, "dt": dt},
        "hbm_GBps_algorithmic_step": step_b / (r64["wall"] / K) / 1e9,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(dom, r64["variants"]) if N == 256 else None,
                     "bytes_per_launch": dom_bytes, "ms_per_launch": dom_ms_avg},
        "passes_fp64": per_pass,
        "tile_variants_fp64": r64["variants"],
        "energy_fp64_final": r64["diag"]["kinetic_energy"],
        "enstrophy_fp64_final": r64["diag"]["enstrophy"],
        "div_max_fp64_final": r64["diag"]["div_u_max"],
    }
    e64 = r64["diag"]["kinetic_energy"]
    hist = oracle_history()
    if hist is not None and N == 256 and K == 500 and abs(dt - 2e-3) < 1e-15:
        res["energy_fp64_rel_err_vs_oracle"] = abs(e64 - hist["energy"][-1]) / hist["energy"][-1]
        res["enstrophy_fp64_rel_err_vs_oracle"] = abs(r64["diag"]["enstrophy"] - hist["enstrophy"][-1]) / \
            hist["enstrophy"][-1]
    del r64
    torch.cuda.empty_cache()

    if not args.no_fp32:
        r32 = run_gpu(N, torch.float32, K, W, dt, nu, tune=not args.no_tune)
        res["fp32_steps_per_s"] = K / r32["wall"]
        res["fp32_over_fp64"] = res["fp32_steps_per_s"] / steps_s
        res["fp32_energy_rel_err_vs_fp64"] = abs(r32["diag"]["kinetic_energy"] - e64) / e64
        if "energy_fp64_rel_err_vs_oracle" in res:
            res["fp32_energy_rel_err_vs_oracle"] = abs(r32["diag"]["kinetic_energy"] - hist["energy"][-1]) / \
                hist["energy"][-1]
        pb32 = roofline.pass_bytes(N, 4)
        res["passes_fp32"] = {k: {"ms_avg": v[0] / max(1, v[1]), "launches": v[1],
                                  "GBps": ((pb32[k] if k != "fwd_x" else sum(pb32["fwd_x"]) / 4.0)
                                           / (v[0] / max(1, v[1]) * 1e-3) / 1e9)}
                              for k, v in r32["passes"].items()}
        res["roofline_fp32"] = leg_roofline(res["passes_fp32"], "profiles/hydro_256_kernel_stats.md")
        res["tile_variants_fp32"] = r32["variants"]
        del r32
        torch.cuda.empty_cache()

    if not args.no_ensemble:
        res["ensemble"] = run_ensemble(64, 200, 5e-3)
        torch.cuda.empty_cache()

    if not args.no_mhd:
        # MHD TGV at 128^3 (nu = eta = 1/200, dt = 2e-3, 300 steps) with the
        # physical Lorentz sign; FP32 errors stated against the FP64 run.
        m64 = run_mhd(128, 300, 2e-3, 1.0 / 200, 1.0 / 200, torch.float64)
        m32 = run_mhd(128, 300, 2e-3, 1.0 / 200, 1.0 / 200, torch.float32)
        for q in ("kinetic_energy", "magnetic_energy"):
            m32[q + "_rel_err_vs_fp64"] = abs(m32[q] - m64[q]) / max(abs(m64[q]), 1e-300)
        # H_c vanishes for this symmetric state: error relative to its bound 2 sqrt(E_k E_m)
        m32["cross_helicity_err_vs_fp64_over_bound"] = abs(m32["cross_helicity"] - m64["cross_helicity"]) / (
            2.0 * math.sqrt(m64["kinetic_energy"] * m64["magnetic_energy"]))
        res["mhd_128"] = {"config": "MHD TGV N=128^3, b = -curl u_TG, nu = eta = 1/200, dt = 2e-3, 300 RK4 steps, "
                                    "lorentz_sign = +1 (j x b)", "fp64": m64, "fp32": m32,
                          "fp32_over_fp64": m32["steps_per_s"] / m64["steps_per_s"]}
    if not args.no_adjoint:
        res["adjoint"] = [run_adjoint(128, 200, 5e-3, 1.0 / 400, dt_) for dt_ in (torch.float64, torch.float32)]
        res["adjoint_profile"] = "profiles/adjoint_128_kernel_stats_and_traffic.md"
        # the headline grid: 100 steps of the 256^3 TGV (keeping every stage would need ~0.27 TB in FP64)
        res["adjoint"] += [run_adjoint(256, 100, dt, nu, dt_, all_stages=False) for dt_ in (torch.float64, torch.float32)]
        torch.cuda.empty_cache()

    if not args.no_cpu_baseline:
        from jax_tg_project_amd._lib import Plan
        plan = Plan(N, dtype=torch.float64)
        plan.set_variants(res["tile_variants_fp64"])  # the tiles the timed run used
        uh = plan.forward(tg_initial(N, torch.float64, plan.device))
        plan.rk4(uh, nsteps=1, dt=dt, nu=nu)
        one = plan.inverse(uh).cpu().numpy()
        del plan, uh
        torch.cuda.empty_cache()
        cb, parity = cpu_baseline(N, dt, nu, one)
        res["cpu_baseline"] = cb
        res["parity_1step_rel_l2_vs_oracle"] = parity
        res["parity_1step_tile_variants"] = res["tile_variants_fp64"]
        from oracle import tg_oracle as o
        cores = o.cpu_count()
        res["cpu_baseline_c1"] = {
            "config": "hydro TGV N=32, nu=0.01, dt=0.01, 100 RK4 steps, 21 diagnostic records (run_simulation)",
            "oracle_all_cores_s": cpu_c1(cores), "cores": cores, "oracle_1_worker_s": cpu_c1(1),
            "gpu_s": gpu_c1()}
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
