"""ctypes binding of libtgdns.so (C ABI in include/tgdns.h).

torch is imported before the library is loaded so that libtgdns.so binds to
the HIP runtime torch already loaded (one runtime per process; device
pointers from torch tensors are then valid in the library).  There is no
fallback: every entry point raises when the library or a GPU is missing.
"""
import ctypes
import math
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("TGDNS_LIB") or os.path.join(_HERE, "libtgdns.so")

TG_INFO_N, TG_INFO_K, TG_INFO_KZ, TG_INFO_COMPACT_ELEMS, TG_INFO_WS_BYTES = 0, 1, 2, 3, 4
TG_INFO_PRECISION, TG_INFO_MHD, TG_INFO_GENERIC, TG_INFO_COUNT = 5, 6, 7, 8
TG_OP_FIELD, TG_OP_CURL, TG_OP_DIV, TG_OP_GRAD = 0, 1, 2, 3
TG_DIAG_COUNT = 11

_lib = None

_VP = ctypes.c_void_p


class CheckpointMeta(ctypes.Structure):
    """tg_checkpoint_meta of include/tgdns.h."""
    _fields_ = [("step", ctypes.c_longlong), ("time", ctypes.c_double), ("dt", ctypes.c_double),
                ("nu", ctypes.c_double), ("eta", ctypes.c_double), ("lorentz_sign", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_SIGS = {
    "tg_version": ([], ctypes.c_int),
    "tg_supported_sizes": ([ctypes.POINTER(ctypes.c_int), ctypes.c_int], ctypes.c_int),
    "tg_plan_create": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                        ctypes.POINTER(_VP)], ctypes.c_int),
    "tg_plan_destroy": ([_VP], ctypes.c_int),
    "tg_plan_query": ([_VP, ctypes.POINTER(ctypes.c_longlong)], ctypes.c_int),
    "tg_forward": ([_VP, _VP, _VP, _VP, _VP], ctypes.c_int),
    "tg_inverse": ([_VP, _VP, ctypes.c_int, _VP, _VP, _VP], ctypes.c_int),
    "tg_rhs": ([_VP, _VP, _VP, ctypes.c_double, ctypes.c_double, ctypes.c_double, _VP, _VP,
                _VP, _VP], ctypes.c_int),
    "tg_rk4_steps": ([_VP, _VP, _VP, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                      ctypes.c_double, ctypes.c_double, _VP, _VP], ctypes.c_int),
    "tg_diagnostics": ([_VP, _VP, _VP, ctypes.c_int, ctypes.POINTER(ctypes.c_double), _VP, _VP],
                       ctypes.c_int),
    "tg_diagnostics_noalloc": ([_VP, _VP, _VP, ctypes.c_int, _VP, _VP], ctypes.c_int),
    "tg_plan_variants": ([_VP, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                          ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "tg_plan_recompute_variants": ([_VP, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)],
                                   ctypes.c_int),
    "tg_timing_enable": ([_VP, ctypes.c_int], ctypes.c_int),
    "tg_timing_read": ([_VP, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_longlong)],
                       ctypes.c_int),
    "tg_batch_create": ([ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double),
                         ctypes.c_int, ctypes.c_double, ctypes.POINTER(_VP)], ctypes.c_int),
    "tg_batch_destroy": ([_VP], ctypes.c_int),
    "tg_batch_query": ([_VP, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)],
                       ctypes.c_int),
    "tg_batch_bind": ([_VP, _VP, _VP], ctypes.c_int),
    "tg_batch_set_launch": ([_VP, ctypes.c_int], ctypes.c_int),
    "tg_batch_rk4_steps": ([_VP, ctypes.c_int, ctypes.c_double, ctypes.c_int, _VP], ctypes.c_int),
    "tg_vjp": ([_VP, _VP, _VP, ctypes.c_double, _VP, _VP, _VP], ctypes.c_int),
    "tg_adjoint_workspace": ([_VP, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong)], ctypes.c_int),
    "tg_adjoint_forward": ([_VP, _VP, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                            ctypes.POINTER(ctypes.c_double), _VP, _VP, _VP], ctypes.c_int),
    "tg_adjoint_backward": ([_VP, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, _VP,
                             ctypes.POINTER(ctypes.c_double), _VP, _VP, _VP], ctypes.c_int),
    "tg_checkpoint_write": ([_VP, ctypes.c_char_p, _VP, _VP, ctypes.POINTER(CheckpointMeta), _VP], ctypes.c_int),
    "tg_checkpoint_info": ([ctypes.c_char_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                            ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double),
                            ctypes.POINTER(CheckpointMeta)], ctypes.c_int),
    "tg_checkpoint_read": ([_VP, ctypes.c_char_p, _VP, _VP, ctypes.POINTER(CheckpointMeta), _VP], ctypes.c_int),
    "tg_last_error": ([ctypes.c_char_p, ctypes.c_size_t], ctypes.c_int),
}
PASS_NAMES = ("inv_x", "inv_y", "z", "fwd_y", "fwd_x")
EXPORTED = tuple(_SIGS)


class TgError(RuntimeError):
    """Error reported by libtgdns.so."""


def load():
    """Load libtgdns.so (once).  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "or `python -m jax_tg_project_amd.build`")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib
    return _lib


def _check(rc):
    if rc != 0:
        buf = ctypes.create_string_buffer(1024)
        _lib.tg_last_error(buf, len(buf))
        raise TgError(f"libtgdns error {rc}: {buf.value.decode(errors='replace')}")


def supported_sizes():
    lib = load()
    arr = (ctypes.c_int * 32)()
    n = lib.tg_supported_sizes(arr, 32)
    return tuple(arr[i] for i in range(n))


def _check_rc(rc):
    _check(rc)


def checkpoint_info(path):
    """(n, precision_bits, mhd, domain_length, meta dict) of a checkpoint file
    (host only: no GPU needed)."""
    lib = load()
    n, prec, mhd = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    L = ctypes.c_double()
    meta = CheckpointMeta()
    _check(lib.tg_checkpoint_info(os.fsencode(path), ctypes.byref(n), ctypes.byref(prec), ctypes.byref(mhd),
                                  ctypes.byref(L), ctypes.byref(meta)))
    return int(n.value), int(prec.value), bool(mhd.value), float(L.value), meta.as_dict()


def _require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("the pseudo-spectral HIP path needs a ROCm GPU (torch.cuda.is_available() is False)")


def _stream():
    return _VP(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return _VP(t.data_ptr()) if t is not None else _VP(0)


class Plan:
    """A libtgdns plan for one grid size / precision / equation set, plus the
    device workspace it runs in (allocated once, reused by every call)."""

    def __init__(self, n, dtype=torch.float64, mhd=False, domain_length=2.0 * math.pi, device=None):
        _require_gpu()
        lib = load()
        if dtype not in (torch.float64, torch.float32):
            raise ValueError("dtype must be torch.float64 or torch.float32")
        self.n = int(n)
        self.dtype = dtype
        self.cdtype = torch.complex128 if dtype == torch.float64 else torch.complex64
        self.mhd = bool(mhd)
        self.domain_length = float(domain_length)
        dev = torch.device(device if device is not None else "cuda")
        if dev.type != "cuda":
            raise ValueError("Plan needs a cuda (ROCm) device")
        if dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        self._lib = lib
        handle = _VP()
        with torch.cuda.device(self.device):
            _check(lib.tg_plan_create(self.n, 64 if dtype == torch.float64 else 32, int(self.mhd),
                                      self.domain_length, ctypes.byref(handle)))
        self._h = handle
        info = (ctypes.c_longlong * TG_INFO_COUNT)()
        _check(lib.tg_plan_query(self._h, info))
        self.K = int(info[TG_INFO_K])
        self.KZ = int(info[TG_INFO_KZ])
        self.M = self.n // 3
        self.ws_bytes = int(info[TG_INFO_WS_BYTES])
        # no compiled kernels for n: the direct-sum path (no adjoint, no ensembles)
        self.generic = bool(info[TG_INFO_GENERIC])
        self.ws = torch.empty(self.ws_bytes, dtype=torch.uint8, device=self.device)
        self.spec_shape = (3, self.K, self.K, self.KZ)
        self.phys_shape = (3, self.n, self.n, self.n)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:
            _lib.tg_plan_destroy(h)
            self._h = None

    # -- buffers -------------------------------------------------------------
    def empty_spec(self):
        return torch.empty(self.spec_shape, dtype=self.cdtype, device=self.device)

    def empty_phys(self):
        return torch.empty(self.phys_shape, dtype=self.dtype, device=self.device)

    def _spec(self, t, name):
        if t.dtype != self.cdtype or tuple(t.shape) != self.spec_shape or not t.is_contiguous() \
                or t.device != self.device:
            raise ValueError(f"{name}: expected contiguous {self.cdtype} {self.spec_shape} on {self.device}, "
                             f"got {t.dtype} {tuple(t.shape)} on {t.device}")
        return t

    def _phys(self, t, name):
        if t.dtype != self.dtype or tuple(t.shape) != self.phys_shape or not t.is_contiguous() \
                or t.device != self.device:
            raise ValueError(f"{name}: expected contiguous {self.dtype} {self.phys_shape} on {self.device}, "
                             f"got {t.dtype} {tuple(t.shape)} on {t.device}")
        return t

    # -- entry points ----------------------------------------------------------
    def forward(self, phys, out=None):
        """Physical (3,N,N,N) -> compact masked spectrum (3,K,K,KZ)."""
        self._phys(phys, "phys")
        out = self.empty_spec() if out is None else self._spec(out, "out")
        with torch.cuda.device(self.device):
            _check(self._lib.tg_forward(self._h, _ptr(phys), _ptr(out), _ptr(self.ws), _stream()))
        return out

    def inverse(self, spec, out=None, op=TG_OP_FIELD):
        """Compact spectrum (3,K,K,KZ) -> physical real field: the field itself
        (3,N,N,N), its curl (3,N,N,N) or its divergence (N,N,N); for
        TG_OP_GRAD, one scalar component (K,K,KZ) -> its gradient (3,N,N,N)."""
        if op == TG_OP_GRAD:
            if spec.dtype != self.cdtype or tuple(spec.shape) != self.spec_shape[1:] or not spec.is_contiguous() \
                    or spec.device != self.device:
                raise ValueError(f"spec: expected contiguous {self.cdtype} {self.spec_shape[1:]} on {self.device}")
        else:
            self._spec(spec, "spec")
        if out is None:
            out = self.empty_phys() if op != TG_OP_DIV else torch.empty(
                self.phys_shape[1:], dtype=self.dtype, device=self.device)
        else:
            want = self.phys_shape if op != TG_OP_DIV else self.phys_shape[1:]
            if out.dtype != self.dtype or tuple(out.shape) != want or not out.is_contiguous():
                raise ValueError(f"out: expected contiguous {self.dtype} {want}")
        with torch.cuda.device(self.device):
            _check(self._lib.tg_inverse(self._h, _ptr(spec), int(op), _ptr(out), _ptr(self.ws), _stream()))
        return out

    def rhs(self, u_hat, b_hat=None, nu=0.0, eta=0.0, lorentz_sign=-1.0):
        self._spec(u_hat, "u_hat")
        ru = self.empty_spec()
        rb = None
        if self.mhd:
            self._spec(b_hat, "b_hat")
            rb = self.empty_spec()
        with torch.cuda.device(self.device):
            _check(self._lib.tg_rhs(self._h, _ptr(u_hat), _ptr(b_hat), float(nu), float(eta),
                                    float(lorentz_sign), _ptr(ru), _ptr(rb), _ptr(self.ws), _stream()))
        return (ru, rb) if self.mhd else ru

    def rk4(self, u_hat, b_hat=None, nsteps=1, dt=0.0, nu=0.0, eta=0.0, lorentz_sign=-1.0):
        """nsteps RK4 steps in place."""
        self._spec(u_hat, "u_hat")
        if self.mhd:
            self._spec(b_hat, "b_hat")
        with torch.cuda.device(self.device):
            _check(self._lib.tg_rk4_steps(self._h, _ptr(u_hat), _ptr(b_hat), int(nsteps), float(dt),
                                          float(nu), float(eta), float(lorentz_sign), _ptr(self.ws),
                                          _stream()))

    # -- adjoint ---------------------------------------------------------------
    def vjp(self, x_hat, a_hat, nu):
        """J_rhs(x)^T a (physical-L2 cotangents as unnormalised spectra)."""
        self._spec(x_hat, "x_hat")
        self._spec(a_hat, "a_hat")
        out = self.empty_spec()
        with torch.cuda.device(self.device):
            _check(self._lib.tg_vjp(self._h, _ptr(x_hat), _ptr(a_hat), float(nu), _ptr(out), _ptr(self.ws),
                                    _stream()))
        return out

    def adjoint_workspace_bytes(self, nsteps, seg):
        """Bytes of adjoint workspace for (nsteps, seg)."""
        b = ctypes.c_longlong()
        _check(self._lib.tg_adjoint_workspace(self._h, int(nsteps), int(seg), ctypes.byref(b)))
        return int(b.value)

    def adjoint_workspace(self, nsteps, seg):
        return torch.empty(self.adjoint_workspace_bytes(nsteps, seg), dtype=torch.uint8, device=self.device)

    def adjoint_forward(self, x0_hat, nsteps, dt, nu, objective, seg, aws):
        self._spec(x0_hat, "x0_hat")
        J = ctypes.c_double()
        with torch.cuda.device(self.device):
            _check(self._lib.tg_adjoint_forward(self._h, _ptr(x0_hat), int(nsteps), float(dt), float(nu),
                                                int(objective), int(seg), ctypes.byref(J), _ptr(self.ws),
                                                _ptr(aws), _stream()))
        return float(J.value)

    def adjoint_backward(self, nsteps, dt, nu, seg, aws):
        grad = self.empty_spec()
        d = ctypes.c_double()
        with torch.cuda.device(self.device):
            _check(self._lib.tg_adjoint_backward(self._h, int(nsteps), float(dt), float(nu), int(seg),
                                                 _ptr(grad), ctypes.byref(d), _ptr(self.ws), _ptr(aws),
                                                 _stream()))
        return grad, float(d.value)

    # -- state dump / restart ---------------------------------------------------
    def checkpoint_write(self, path, u_hat, b_hat=None, step=0, time=0.0, dt=0.0, nu=0.0, eta=0.0,
                         lorentz_sign=0.0):
        self._spec(u_hat, "u_hat")
        if self.mhd:
            self._spec(b_hat, "b_hat")
        meta = CheckpointMeta(int(step), float(time), float(dt), float(nu), float(eta), float(lorentz_sign))
        with torch.cuda.device(self.device):
            _check(self._lib.tg_checkpoint_write(self._h, os.fsencode(path), _ptr(u_hat), _ptr(b_hat),
                                                 ctypes.byref(meta), _stream()))

    def checkpoint_read(self, path, u_hat, b_hat=None):
        """Load a checkpoint of this plan's grid, precision and equations into
        u_hat (and b_hat); returns its metadata."""
        self._spec(u_hat, "u_hat")
        if self.mhd:
            self._spec(b_hat, "b_hat")
        meta = CheckpointMeta()
        with torch.cuda.device(self.device):
            _check(self._lib.tg_checkpoint_read(self._h, os.fsencode(path), _ptr(u_hat), _ptr(b_hat),
                                                ctypes.byref(meta), _stream()))
        return meta.as_dict()

    def diagnostics_noalloc(self, u_hat, b_hat=None, divmax=True, out=None):
        """tg_diagnostics_noalloc: the TG_DIAG_COUNT values into the float64 device
        tensor `out` (allocated here when None), on the current stream, without
        synchronising and without library allocations."""
        self._spec(u_hat, "u_hat")
        if b_hat is not None:
            self._spec(b_hat, "b_hat")
        if out is None:
            out = torch.empty(TG_DIAG_COUNT, dtype=torch.float64, device=self.device)
        if out.dtype != torch.float64 or out.numel() < TG_DIAG_COUNT or not out.is_contiguous() \
                or out.device != self.device:
            raise ValueError(f"out: expected a contiguous float64 tensor of {TG_DIAG_COUNT} on {self.device}")
        with torch.cuda.device(self.device):
            _check(self._lib.tg_diagnostics_noalloc(self._h, _ptr(u_hat), _ptr(b_hat), int(bool(divmax)), _ptr(out),
                                                    _stream()))
        return out

    def variant_count(self):
        cnt = ctypes.c_int()
        _check(self._lib.tg_plan_variants(self._h, None, None, ctypes.byref(cnt)))
        return int(cnt.value)

    def get_variants(self):
        out = (ctypes.c_int * len(PASS_NAMES))()
        _check(self._lib.tg_plan_variants(self._h, None, out, None))
        return [int(v) for v in out]

    def set_variants(self, variants):
        arr = (ctypes.c_int * len(PASS_NAMES))(*[int(v) for v in variants])
        _check(self._lib.tg_plan_variants(self._h, arr, None, None))

    def get_recompute_variants(self):
        """Tile variants of the adjoint recompute stream (the plan's own
        unless set_recompute_variants chose others)."""
        out = (ctypes.c_int * len(PASS_NAMES))()
        _check(self._lib.tg_plan_recompute_variants(self._h, None, out))
        return [int(v) for v in out]

    def set_recompute_variants(self, variants):
        """variants: one per pass, or None for the plan's own."""
        vals = [-1] * len(PASS_NAMES) if variants is None else [int(v) for v in variants]
        arr = (ctypes.c_int * len(PASS_NAMES))(*vals)
        _check(self._lib.tg_plan_recompute_variants(self._h, arr, None))

    def timing(self, on):
        """Enable/disable the per-pass HIP-event timing hook."""
        _check(self._lib.tg_timing_enable(self._h, int(bool(on))))

    def timing_read(self):
        """{pass name: (total ms, launches)} since the last read; waits for the GPU."""
        ms = (ctypes.c_double * len(PASS_NAMES))()
        cnt = (ctypes.c_longlong * len(PASS_NAMES))()
        _check(self._lib.tg_timing_read(self._h, ms, cnt))
        return {n: (float(ms[i]), int(cnt[i])) for i, n in enumerate(PASS_NAMES)}

    def diagnostics(self, u_hat, b_hat=None, divmax=True):
        """Dict of FP64 diagnostics of the compact state (synchronises)."""
        self._spec(u_hat, "u_hat")
        if b_hat is not None:
            self._spec(b_hat, "b_hat")
        out = (ctypes.c_double * TG_DIAG_COUNT)()
        with torch.cuda.device(self.device):
            _check(self._lib.tg_diagnostics(self._h, _ptr(u_hat), _ptr(b_hat), int(bool(divmax)), out,
                                            _ptr(self.ws), _stream()))
        keys = ("kinetic_energy", "enstrophy", "grad_sq", "magnetic_energy", "cross_helicity",
                "div_u_sq", "div_b_sq", "current_sq", "div_u_max", "div_b_max", "nonfinite")
        return {k: float(out[i]) for i, k in enumerate(keys)}
