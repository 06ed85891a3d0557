"""Shared test helpers: seeded band-limited fields and error norms."""
import numpy as np

from oracle import tg_oracle as o


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))


def random_compact(N, seed, ncomp=3, scale=1.0):
    """Random compact (kept, kz >= 0) spectrum with Hermitian-consistent kz=0 plane."""
    rng = np.random.default_rng(seed)
    M, K, KZ = o.kept_counts(N)
    c = (rng.standard_normal((ncomp, K, K, KZ)) + 1j * rng.standard_normal((ncomp, K, K, KZ))) * scale
    # make the physical field real-consistent: take the spectrum of the real part
    full = o.compact_to_full(c, N)
    phys = np.fft.ifftn(full, axes=(-3, -2, -1)).real
    return o.full_to_compact(np.fft.fftn(phys, axes=(-3, -2, -1)))


def band_limited_field(N, seed, amp=1.0, divfree=True):
    """Random real physical field (3,N,N,N) whose spectrum lies inside the 2/3 mask."""
    c = random_compact(N, seed)
    full = o.compact_to_full(c, N)
    if divfree:
        kx, ky, kz = o.make_wavenumbers(N, 2 * np.pi)
        full = o.leray(full, kx, ky, kz)
    u = np.fft.ifftn(full, axes=(-3, -2, -1)).real
    return u * (amp / np.sqrt(np.mean(u * u)))


def to_compact(u):
    """Masked compact spectrum of a physical field (oracle side)."""
    N = u.shape[-1]
    m = o.make_dealias_mask(N, 2 * np.pi)
    return o.full_to_compact(np.fft.fftn(u, axes=(-3, -2, -1)) * m)
