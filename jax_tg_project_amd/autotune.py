"""Autotuner for the per-pass tile variants of a plan.

Every pass of the right-hand side (inverse x, inverse y, z, forward y,
forward x) has a few compiled tile shapes for large grids (columns per
block, points per thread, lines per block; see tg_passes.hpp).  The tuner
runs a few RK4 steps per variant with the per-pass HIP-event hook on, takes
the fastest variant of each pass independently, sets it on the plan and
stores the choice in a JSON cache keyed by (device, N, precision,
equations), so later processes skip the measurement.
"""
import json
import math
import os
import time

import torch

from ._lib import PASS_NAMES

CACHE_ENV = "TGDNS_TUNE_CACHE"


def cache_path():
    p = os.environ.get(CACHE_ENV)
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "tgdns", "tune.json")


def _key(plan):
    name = torch.cuda.get_device_name(plan.device).replace("|", "/")
    prec = "fp64" if plan.dtype == torch.float64 else "fp32"
    return f"{name}|N={plan.n}|{prec}|{'mhd' if plan.mhd else 'hydro'}"


def _load():
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(cache):
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home only loses the cache, not the tuning


def _probe_state(plan):
    """A smooth divergence-free state (TG) for timing; values do not matter."""
    N = plan.n
    x = torch.arange(N, dtype=torch.float64, device=plan.device) * (plan.domain_length / N)
    X, Y, Z = torch.meshgrid(x, x, x, indexing="ij")
    u = torch.stack([torch.sin(X) * torch.cos(Y) * torch.cos(Z), -torch.cos(X) * torch.sin(Y) * torch.cos(Z),
                     torch.zeros_like(X)]).to(plan.dtype).contiguous()
    uh = plan.forward(u)
    bh = plan.forward((0.5 * u).contiguous()) if plan.mhd else None
    return uh, bh


def tune(plan, steps=3, use_cache=True, verbose=False, rounds=3):
    """Pick the fastest tile variant per pass for `plan`; returns
    (variants, {pass: [ms per launch for each variant]}).  The variants are
    timed in `rounds` interleaved rounds of `steps` RK4 steps each; a pass
    keeps its best per-launch average over the rounds."""
    count = plan.variant_count()
    if count == 1:
        return plan.get_variants(), {}
    key = _key(plan)
    cache = _load() if use_cache else {}
    if key in cache:
        plan.set_variants(cache[key]["variants"])
        return plan.get_variants(), cache[key].get("ms", {})
    uh, bh = _probe_state(plan)
    ms = {p: [math.inf] * count for p in PASS_NAMES}
    dt, nu = 1e-4, 1e-3
    for r in range(rounds):
        for v in range(count):
            plan.set_variants([v] * len(PASS_NAMES))
            if r == 0:
                plan.rk4(uh, bh, nsteps=1, dt=dt, nu=nu, eta=nu)  # warm-up (code object load)
            plan.timing(True)
            plan.rk4(uh, bh, nsteps=steps, dt=dt, nu=nu, eta=nu)
            t = plan.timing_read()
            plan.timing(False)
            for p, (tot, n) in t.items():
                if n:
                    ms[p][v] = min(ms[p][v], tot / n)
    best = [min(range(count), key=lambda v: ms[p][v]) for p in PASS_NAMES]
    plan.set_variants(best)
    if verbose:
        print("autotune", key, best, {p: [round(x, 4) for x in ms[p]] for p in PASS_NAMES})
    if use_cache:
        cache = _load()
        cache[key] = {"variants": best, "ms": ms}
        _store(cache)
    return best, ms


def tune_recompute(plan, nsteps=64, seg=8, use_cache=True, verbose=False, rounds=2):
    """Pick the tile variants of the adjoint recompute, which runs on its own
    stream beside the reverse sweep's vector-Jacobian products: variants with
    smaller register tiles run slower alone but share the CUs with the
    reverse sweep's kernels.  Times the reverse sweep of an `nsteps`-step
    adjoint with checkpoint segments of `seg` steps for the plan's own
    variants and each uniform set, then changes one pass at a time from the
    best; sets the winner on the plan and returns (variants, {candidate:
    seconds}).  Call after tune() (the forward sweep keeps the plan's
    variants).  The trial halves seg (and nsteps with it) until its
    workspace fits a quarter of the free device memory, and tunes nothing
    when even segments of 2 steps do not fit."""
    count = plan.variant_count()
    if count == 1 or plan.mhd:
        return plan.get_recompute_variants(), {}
    key = _key(plan) + "|recompute"
    cache = _load() if use_cache else {}
    if key in cache:
        plan.set_recompute_variants(cache[key]["variants"])
        return plan.get_recompute_variants(), cache[key].get("s", {})
    budget = torch.cuda.mem_get_info(plan.device)[0] // 4
    while plan.adjoint_workspace_bytes(nsteps, seg) > budget and seg > 2:
        nsteps, seg = nsteps // 2, seg // 2
    if plan.adjoint_workspace_bytes(nsteps, seg) > budget:
        return plan.get_recompute_variants(), {}
    x0, _ = _probe_state(plan)
    aws = plan.adjoint_workspace(nsteps, seg)
    dt, nu = 1e-3, 1e-3
    seen = {}

    def sweep_s(vs):
        k = ",".join(map(str, vs))
        if k not in seen:
            plan.set_recompute_variants(vs)
            best = math.inf
            for _ in range(rounds):
                plan.adjoint_forward(x0, nsteps, dt, nu, 1, seg, aws)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                plan.adjoint_backward(nsteps, dt, nu, seg, aws)
                torch.cuda.synchronize()
                best = min(best, time.perf_counter() - t0)
            seen[k] = best
        return seen[k]

    own = plan.get_variants()
    sweep_s(own)  # warm-up: the reverse sweep's stream and events are created on first use
    seen.clear()
    best = min([own] + [[v] * len(PASS_NAMES) for v in range(count)], key=sweep_s)
    for p in range(len(PASS_NAMES)):
        for v in range(count):
            cand = list(best)
            cand[p] = v
            if sweep_s(cand) < sweep_s(best):
                best = cand
    del aws
    plan.set_recompute_variants(best)
    if verbose:
        print("autotune recompute", key, best, {k: round(s * 1e3, 2) for k, s in seen.items()})
    if use_cache:
        cache = _load()
        cache[key] = {"variants": best, "s": seen}
        _store(cache)
    return best, seen
