"""Per-element material fields for the Q1 heat solve.

Layout and file format follow scripts/gen_field.py:5-6, 25-26, 106-113 of the
reference: element e = ix*n*n + iy*n + iz (x slowest), float64 C order in
``<out>_k.bin`` / ``<out>_rhoc.bin`` plus ``<out>_meta.json``.

Models
  homogeneous, quadrants, lognormal (FFT Gaussian field, optional harmonic /
  arithmetic upscaling) and layered restate scripts/gen_field.py:31-104.  The
  lognormal field must draw the same numbers as the reference to reproduce its
  recorded runs, so it uses the same generator (numpy PCG64 standard_normal),
  the same spectral filter and the same normalisation.
  iid_lognormal  kappa_e = exp(mu + sigma*clip(g_e, -c, c)), g_e ~ N(0,1) from a
                 counter-based generator keyed by (seed, e): the value of an
                 element does not depend on how many others are drawn.
  block_lognormal kappa = exp(mu + sigma*g_b), g_b i.i.d. per block of
                 ``block``^3 elements, same counter-based generator.
"""
from __future__ import annotations

import json
import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def _mix64(z: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser (Steele, Lea & Flood 2014), vectorised on uint64."""
    z = z.astype(np.uint64, copy=True)
    with np.errstate(over="ignore"):
        z ^= z >> np.uint64(30)
        z *= np.uint64(0xBF58476D1CE4E5B9)
        z ^= z >> np.uint64(27)
        z *= np.uint64(0x94D049BB133111EB)
        z ^= z >> np.uint64(31)
    return z


def counter_normal(seed: int, idx: np.ndarray) -> np.ndarray:
    """Standard normal deviates keyed by (seed, idx): Box-Muller on two
    53-bit uniforms drawn from SplitMix64 of the counters 2*idx and 2*idx+1."""
    idx = np.asarray(idx, dtype=np.uint64)
    key = _mix64(np.array([seed], dtype=np.uint64))[0]
    with np.errstate(over="ignore"):
        c0 = key + (np.uint64(2) * idx + np.uint64(1)) * _GOLDEN
        c1 = key + (np.uint64(2) * idx + np.uint64(2)) * _GOLDEN
    u0 = ((_mix64(c0) >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)
    u1 = (_mix64(c1) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return np.sqrt(-2.0 * np.log(u0)) * np.cos(2.0 * np.pi * u1)


def homogeneous(n, k=1.0, rhoc=1.0):
    return np.full((n, n, n), float(k)), np.full((n, n, n), float(rhoc))


def quadrants(n):
    """Four-material model evaluated at element centres
    (scripts/gen_field.py:31-39, the prototype's material_at)."""
    c = (np.arange(n) + 0.5) / n
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    k = np.full(X.shape, 400.0)
    rhoc = np.full(X.shape, 0.5 * 2.5)
    low_z = Z < 0.5
    sel = (X < 0.5) & (Y < 0.5) & low_z
    k[sel], rhoc[sel] = 100.0, 1.0 * 1.0
    sel = (X >= 0.5) & (Y < 0.5) & low_z
    k[sel], rhoc[sel] = 200.0, 0.8 * 1.5
    sel = (X < 0.5) & (Y >= 0.5) & low_z
    k[sel], rhoc[sel] = 300.0, 0.6 * 2.0
    return k, rhoc


def gaussian_field(n, L, seed):
    """Unit-variance periodic Gaussian field with covariance exp(-r^2/L^2)
    (scripts/gen_field.py:41-51): white noise filtered by the square root of
    the spectrum exp(-|k|^2 L^2 / 4), then centred and scaled to unit std."""
    white = np.random.default_rng(seed).standard_normal((n, n, n))
    freq = 2.0 * np.pi * np.fft.fftfreq(n, d=1.0 / n)
    kx, ky, kz = np.meshgrid(freq, freq, freq, indexing="ij")
    filt = np.sqrt(np.exp(-(kx ** 2 + ky ** 2 + kz ** 2) * L ** 2 / 4.0))
    g = np.fft.ifftn(np.fft.fftn(white) * filt).real
    g -= g.mean()
    g /= g.std()
    return g


def upscale(fine, r, mode):
    """Average r^3 blocks: arithmetic or harmonic (scripts/gen_field.py:54-61)."""
    n = fine.shape[0] // r
    blk = fine.reshape(n, r, n, r, n, r).transpose(0, 2, 4, 1, 3, 5).reshape(n, n, n, r ** 3)
    if mode == "arithmetic":
        return blk.mean(axis=-1)
    if mode == "harmonic":
        return 1.0 / (1.0 / blk).mean(axis=-1)
    raise ValueError(mode)


def lognormal(n, mu=0.0, sigma=1.0, rhoc=1.0, rhoc_sigma=0.0, L=0.125, upscale_r=1, seed=1):
    nf = n * upscale_r
    g = gaussian_field(nf, L, seed)
    kf = np.exp(mu + sigma * g)
    rf = np.exp(np.log(rhoc) + rhoc_sigma * g)
    if upscale_r > 1:
        return upscale(kf, upscale_r, "harmonic"), upscale(rf, upscale_r, "arithmetic")
    return kf, rf


def layered(n, k=1.0, k2=100.0, period=4, rhoc=1.0):
    ix = np.arange(n)[:, None, None]
    kk = np.where((ix // period) % 2 == 0, k, k2) * np.ones((n, n, n))
    return kk, np.full((n, n, n), float(rhoc))


def iid_lognormal(n, mu=0.0, sigma=1.0, clip=3.0, rhoc=1.0, seed=1):
    g = counter_normal(seed, np.arange(n ** 3, dtype=np.uint64)).reshape(n, n, n)
    g = np.clip(g, -clip, clip)
    return np.exp(mu + sigma * g), np.full((n, n, n), float(rhoc))


def block_lognormal(n, mu=0.0, sigma=2.0, block=4, rhoc=1.0, seed=1):
    nb = -(-n // block)
    g = counter_normal(seed, np.arange(nb ** 3, dtype=np.uint64)).reshape(nb, nb, nb)
    i = np.arange(n) // block
    gb = g[i[:, None, None], i[None, :, None], i[None, None, :]]
    return np.exp(mu + sigma * gb), np.full((n, n, n), float(rhoc))


def write_field(prefix, k, rhoc, meta):
    """Write <prefix>_k.bin, <prefix>_rhoc.bin, <prefix>_meta.json."""
    k = np.ascontiguousarray(k, dtype=np.float64)
    rhoc = np.ascontiguousarray(rhoc, dtype=np.float64)
    n = k.shape[0]
    k.ravel().tofile(f"{prefix}_k.bin")
    rhoc.ravel().tofile(f"{prefix}_rhoc.bin")
    meta = dict(meta)
    meta.update(kmin=float(k.min()), kmax=float(k.max()), contrast=float(k.max() / k.min()),
                rhocmin=float(rhoc.min()), rhocmax=float(rhoc.max()), n_elements=int(n ** 3))
    with open(f"{prefix}_meta.json", "w") as f:
        json.dump(meta, f, indent=2)
    return meta


def read_field(prefix, n):
    k = np.fromfile(f"{prefix}_k.bin", dtype=np.float64)
    rhoc = np.fromfile(f"{prefix}_rhoc.bin", dtype=np.float64)
    if k.size != n ** 3 or rhoc.size != n ** 3:
        raise ValueError(f"field {prefix} holds {k.size} elements, expected {n ** 3}")
    return k, rhoc
