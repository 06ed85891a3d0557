"""CPU FP64 oracle for the pseudo-spectral Taylor-Green DNS path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this
package: only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg
of ``bench.py`` use it, and only as the checker (or as the timed CPU
baseline), never as the thing measured or shipped.

``tg_oracle`` restates the reference algorithm in numpy/scipy:

* hydro: ``jax/taylor_green_jax.py:17-201`` operation for operation
  (complex ``fftn``/``ifftn`` over the full cube, advective form ``u.grad(u)``
  with the raw physical ``u``, cubic inclusive 2/3 mask, Leray projection with
  ``k2 == 0 -> 1``, classic RK4 on the physical state) and the run loop
  ``jax/taylor_green_jax.py:204-317`` without the plotting side effects;
* MHD: ``cuda/taylor_green_mhd/operators.py:8-107``,
  ``integrator.py:14-32``, ``diagnostics.py:20-49``, ``config.py:53-118`` and
  the shim ``cuda/taylor_green_mhd.py:12-99`` (spectral full-cube state).

Parity pins (each checked in ``tests/test_oracle.py``):

* notebook fixture (N=128, 200 steps): enstrophy history and final energy,
  ``jax/taylor_green_colab.ipynb:134-161,180-200``;
* the reference's own unit-test invariants, ``jax/test_taylor_green_jax.py``
  and ``cuda/test_taylor_green_mhd.py``;
* closed forms: t=0 hydro RHS ``-P(n) - 3 nu u0``, E0=1/8, Z0=3/8,
  eps0=3 nu/4, the t=0 MHD RHS for b0=-omega0, and the shear-wave RK4
  amplification factor;
* the rotational form ``u x omega`` against the advective form for sizes with
  ``3*(N//3) < N``;
* brute-force O(N^6) DFT on tiny grids for the transforms.
"""
