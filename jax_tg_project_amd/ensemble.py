"""Batched ensemble of hydro Taylor-Green instances of mixed grid sizes.

Each instance is the reference solver's problem (jax/taylor_green_jax.py:
advance_one_step on its own grid and viscosity); the ensemble advances all
of them with one batched launch per pass (libtgdns tg_batch_*), optionally
replaying one captured HIP graph per RK4 step.

The seeded workload: sizes drawn uniformly from {32, 48, 64, 96}, viscosity
log-uniform in [1/1600, 1/100], Taylor-Green velocity plus a random
divergence-free perturbation on 0 < |k| <= 4 (complex Gaussian amplitudes,
Leray-projected, scaled to rms 0.05 per component).

The batched kernels form the nonlinear term as u x omega, like the single
plans: at N = 48 and 96 (3 divides N) the aliased shell |k_i| = N/3 differs
from the reference's advective form by about 1e-5 relative (see
taylor_green_jax); at 32 and 64 the two forms agree to round-off.
"""
import ctypes
import math
import time

import numpy as np
import torch

from . import _lib
from .solver import get_plan

ENSEMBLE_SIZES = (32, 48, 64, 96)


def ensemble_config(ninst=64, seed=0, sizes=ENSEMBLE_SIZES, nu_range=(1.0 / 1600, 1.0 / 100)):
    """[(N, nu, instance seed)] drawn from numpy.random.default_rng(seed)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(ninst):
        N = int(rng.choice(sizes))
        nu = float(math.exp(rng.uniform(math.log(nu_range[0]), math.log(nu_range[1]))))
        out.append((N, nu, int(rng.integers(0, 2 ** 31 - 1))))
    return out


def perturbed_taylor_green(N, seed, rms=0.05, kmax=4.0, L=2.0 * math.pi):
    """Physical (3, N, N, N) float64: TG velocity + seeded divergence-free
    perturbation with rms(du) = rms over points and components."""
    rng = np.random.default_rng(seed)
    k1 = 2.0 * np.pi * np.fft.fftfreq(N, d=L / N)
    kx, ky, kz = np.meshgrid(k1, k1, k1, indexing="ij")
    kk = np.sqrt(kx * kx + ky * ky + kz * kz)
    sel = (kk > 0) & (kk <= kmax)
    spec = (rng.standard_normal((3, N, N, N)) + 1j * rng.standard_normal((3, N, N, N))) * sel
    f_hat = np.fft.fftn(np.fft.ifftn(spec, axes=(-3, -2, -1)).real, axes=(-3, -2, -1))
    k2 = np.where(kk == 0, 1.0, kk * kk)
    kd = kx * f_hat[0] + ky * f_hat[1] + kz * f_hat[2]
    f_hat = (f_hat - np.stack([kx * kd / k2, ky * kd / k2, kz * kd / k2])) * sel
    du = np.fft.ifftn(f_hat, axes=(-3, -2, -1)).real
    du *= rms / math.sqrt(float(np.mean(du * du)))
    x = np.linspace(0.0, L, N, endpoint=False)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    u = np.stack([np.sin(X) * np.cos(Y) * np.cos(Z), -np.cos(X) * np.sin(Y) * np.cos(Z), np.zeros_like(X)])
    return u + du


class Ensemble:
    """Instances (N_i, nu_i) advanced together on one GPU."""

    def __init__(self, sizes, nus, dtype=torch.float32, L=2.0 * math.pi, device=None):
        _lib._require_gpu()
        lib = _lib.load()
        self.sizes = [int(n) for n in sizes]
        self.nus = [float(v) for v in nus]
        if len(self.sizes) != len(self.nus) or not self.sizes:
            raise ValueError("sizes and nus must be non-empty and of equal length")
        self.dtype = dtype
        self.cdtype = torch.complex128 if dtype == torch.float64 else torch.complex64
        self.L = float(L)
        dev = torch.device(device if device is not None else "cuda")
        if dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        self._lib = lib
        n = len(self.sizes)
        h = ctypes.c_void_p()
        with torch.cuda.device(dev):
            _lib._check(lib.tg_batch_create(n, (ctypes.c_int * n)(*self.sizes), (ctypes.c_double * n)(*self.nus),
                                            64 if dtype == torch.float64 else 32, self.L, ctypes.byref(h)))
        self._h = h
        # the library's default layout: mixed launches when any size 32-96 is present
        self.mixed = any(N in (32, 48, 64, 96) for N in self.sizes)
        offs = (ctypes.c_longlong * (n + 1))()
        wsb = ctypes.c_longlong()
        _lib._check(lib.tg_batch_query(h, offs, ctypes.byref(wsb)))
        self.offsets = [int(o) for o in offs]
        self.state = torch.zeros(self.offsets[-1], dtype=self.cdtype, device=dev)
        self.ws = torch.empty(int(wsb.value), dtype=torch.uint8, device=dev)
        _lib._check(lib.tg_batch_bind(h, _lib._ptr(self.state), _lib._ptr(self.ws)))
        self.views = []
        for i, N in enumerate(self.sizes):
            M = N // 3
            shape = (3, 2 * M + 1, 2 * M + 1, M + 1)
            self.views.append(self.state[self.offsets[i]:self.offsets[i + 1]].view(shape))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.tg_batch_destroy(h)
            self._h = None

    def __len__(self):
        return len(self.sizes)

    def plan(self, i):
        """Single-instance plan of instance i's size (transforms, diagnostics)."""
        return get_plan(self.sizes[i], self.dtype, False, self.L, self.device)

    def set_velocity(self, i, u):
        p = self.plan(i)
        p.forward(torch.as_tensor(u).to(self.device, self.dtype).contiguous(), out=self.views[i])

    def velocity(self, i):
        return self.plan(i).inverse(self.views[i])

    def set_launch(self, mixed):
        """Launch layout of a step: mixed=True runs one launch per pass for
        every instance of the sizes 32, 48, 64 and 96 (their batched bodies
        share block shapes; other sizes keep their own launches); False runs
        one launch per pass and size class, the classes as parallel graph
        branches."""
        _lib._check(self._lib.tg_batch_set_launch(self._h, int(bool(mixed))))
        self.mixed = bool(mixed)

    def tune_launch(self, dt=5e-3, steps=10):
        """Time both launch layouts (graph replays, which advance the current
        state: call before setting the initial fields) and keep the faster.
        Returns {"mixed": s, "classes": s, "chosen": "mixed" | "classes"}."""
        out = {}
        for name, mixed in (("classes", False), ("mixed", True)):
            self.set_launch(mixed)
            self.step(1, dt, graph=True)  # capture
            torch.cuda.synchronize(self.device)
            t0 = time.perf_counter()
            self.step(steps, dt, graph=True)
            torch.cuda.synchronize(self.device)
            out[name] = time.perf_counter() - t0
        out["chosen"] = "mixed" if out["mixed"] <= out["classes"] else "classes"
        self.set_launch(out["chosen"] == "mixed")
        return out

    def step(self, nsteps=1, dt=5e-3, graph=True):
        with torch.cuda.device(self.device):
            _lib._check(self._lib.tg_batch_rk4_steps(self._h, int(nsteps), float(dt), int(bool(graph)),
                                                     _lib._stream()))

    def diagnostics(self, i, divmax=False):
        d = self.plan(i).diagnostics(self.views[i], divmax=divmax)
        d["dissipation"] = self.nus[i] * d["grad_sq"]
        return d


def build_ensemble(config, dtype=torch.float32, device=None):
    """Ensemble from ensemble_config() output, initial fields set."""
    ens = Ensemble([c[0] for c in config], [c[1] for c in config], dtype=dtype, device=device)
    for i, (N, _, s) in enumerate(config):
        ens.set_velocity(i, perturbed_taylor_green(N, s))
    return ens
