"""MI355X-native pseudo-spectral Taylor-Green DNS (hydro + MHD).

Hand-written gfx950 HIP kernels (libtgdns.so, C ABI in include/tgdns.h)
behind the reference's Python entry points:

* ``jax_tg_project_amd.taylor_green_jax`` mirrors ``jax/taylor_green_jax.py``;
* ``jax_tg_project_amd.taylor_green_mhd`` mirrors ``cuda/taylor_green_mhd.py``
  and the ``taylor_green_mhd`` package classes;
* ``jax_tg_project_amd.solver`` holds the persistent on-GPU state used by the
  run loops and benchmarks.
"""
from . import _lib  # noqa: F401

__all__ = ["_lib"]
