"""Synthetic head-class phantoms for the N4 workloads.

Each phantom is a float32 volume in C order [x][y][z] (z fastest), the
layout of the reference's device buffers (cudaN4.cu:476).  The recipe:

* an ellipsoid head mask centred in the volume (voxel centres at i+0.5,
  inclusive test rho <= 1);
* three nested ellipsoid tissue classes at intensities 100 / 200 / 300,
  class boundaries at rho <= 1, 0.75, 0.45;
* additive Gaussian noise (or Rician-like sqrt((s+n1)^2 + n2^2));
* multiplied by exp(B), B a cubic B-spline from an n^3 lattice with
  N(0, sigma) coefficients, evaluated with the reference's span convention
  (N4.py:192, 210-217).

Seeds come from numpy.random.default_rng and are recorded by the callers.
"""
from __future__ import annotations

import math

import numpy as np

CLASS_RADII = (1.0, 0.75, 0.45)
CLASS_LEVELS = (100.0, 200.0, 300.0)


def bspline_weights(npos: int, ncpt: int, domain: float | None = None):
    """Per-axis (b, w[npos,4]) with span = domain/(ncpt-3), N4.py:192-217."""
    domain = float(npos if domain is None else domain)
    span = domain / (ncpt - 3)
    pos = np.arange(npos, dtype=np.float64) + 0.5
    b = np.floor(pos / span).astype(np.int64)
    t = (pos - b * span) / span
    t2, t3 = t * t, t * t * t
    w = np.stack([(1 - t) ** 3 / 6, (3 * t3 - 6 * t2 + 4) / 6,
                  (-3 * t3 + 3 * t2 + 3 * t + 1) / 6, t3 / 6], axis=1)
    return b, w


def dense_weights(npos: int, ncpt: int, domain: float | None = None) -> np.ndarray:
    """Banded [npos, ncpt] matrix W[i, b_i + a] = w_a(i)."""
    b, w = bspline_weights(npos, ncpt, domain)
    M = np.zeros((npos, ncpt))
    for a in range(4):
        M[np.arange(npos), b + a] = w[:, a]
    return M


def evaluate_lattice(lattice: np.ndarray, dims, domain=None) -> np.ndarray:
    """Separable evaluation of a cubic B-spline lattice on a grid (FP64)."""
    domain = dims if domain is None else domain
    Wx, Wy, Wz = (dense_weights(int(dims[d]), lattice.shape[d], domain[d]) for d in range(3))
    t = np.tensordot(Wx, lattice, axes=(1, 0))          # X, cy, cz
    t = np.tensordot(t, Wy, axes=(1, 1))                # X, cz, Y
    t = np.tensordot(t, Wz, axes=(1, 1))                # X, Y, Z
    return t


def _rho(dims, semi_axes):
    axes = [((np.arange(n) + 0.5) - n / 2.0) / a for n, a in zip(dims, semi_axes)]
    return np.sqrt(axes[0][:, None, None] ** 2 + axes[1][None, :, None] ** 2
                   + axes[2][None, None, :] ** 2)


def make_phantom(dims, semi_axes, bias_ncpt: int, bias_sigma: float, noise_sigma: float,
                 seed: int, rician: bool = False):
    """Returns (image float32, mask float32, log_bias float64)."""
    rng = np.random.default_rng(seed)
    dims = tuple(int(d) for d in dims)
    rho = _rho(dims, semi_axes)
    mask = (rho <= 1.0)
    clean = np.zeros(dims, np.float64)
    for r, v in zip(CLASS_RADII, CLASS_LEVELS):
        clean[rho <= r] = v
    if rician:
        n1 = rng.normal(0.0, noise_sigma, dims)
        n2 = rng.normal(0.0, noise_sigma, dims)
        noisy = np.sqrt((clean + n1) ** 2 + n2 ** 2)
    else:
        noisy = clean + rng.normal(0.0, noise_sigma, dims)
    lat = rng.normal(0.0, bias_sigma, (bias_ncpt,) * 3)
    logb = evaluate_lattice(lat, dims)
    im = (noisy * np.exp(logb)).astype(np.float32)
    return im, mask.astype(np.float32), logb


def c0_phantom(seed: int = 0):
    """64^3, semi-axes 28/24/26, bias lattice 4^3 N(0,0.2), noise sigma 5."""
    return make_phantom((64, 64, 64), (28.0, 24.0, 26.0), 4, 0.2, 5.0, seed)


def c1_phantom(seed: int = 1, dims=(256, 256, 176)):
    """256x256x176, semi-axes (0.49, 0.47, 0.47) x dims (mask ~45%),
    bias lattice 5^3 N(0,0.3), noise sigma 5."""
    semi = tuple(f * d for f, d in zip((0.49, 0.47, 0.47), dims))
    return make_phantom(dims, semi, 5, 0.3, 5.0, seed)


def c2_phantom(seed: int = 2, dims=(512, 512, 320)):
    """512x512x320 (7T-like), bias lattice 6^3 N(0,0.4), noise sigma 5."""
    semi = tuple(f * d for f, d in zip((0.49, 0.47, 0.47), dims))
    return make_phantom(dims, semi, 6, 0.4, 5.0, seed)


def _scale_for_fraction(dims, base, frac):
    """Bisection on k so that the ellipsoid k*base covers ~frac of the box."""
    lo, hi = 0.2, 2.0
    coarse = tuple(max(8, d // 8) for d in dims)
    for _ in range(30):
        k = 0.5 * (lo + hi)
        semi = tuple(k * b * c for b, c in zip(base, coarse))
        f = float(np.mean(_rho(coarse, semi) <= 1.0))
        if f < frac:
            lo = k
        else:
            hi = k
    return 0.5 * (lo + hi)


def batch_specs(n: int, lo: int, hi: int, seed: int):
    """Per-volume specs for the batched workloads: axes ~U{lo..hi}, mask
    fraction ~U[0.3,0.6], bias lattice ~U{4,5,6}^3 with N(0,0.3) coefficients,
    Rician-like noise sigma ~U[2,8]."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(n):
        dims = tuple(int(v) for v in rng.integers(lo, hi + 1, size=3))
        specs.append(dict(dims=dims, frac=float(rng.uniform(0.3, 0.6)),
                          bias_ncpt=int(rng.integers(4, 7)), bias_sigma=0.3,
                          noise_sigma=float(rng.uniform(2.0, 8.0)),
                          seed=int(rng.integers(0, 2**31 - 1))))
    return specs


def make_batch_volume(spec):
    base = (0.5, 0.45, 0.48)
    k = _scale_for_fraction(spec["dims"], base, spec["frac"])
    semi = tuple(k * b * d for b, d in zip(base, spec["dims"]))
    return make_phantom(spec["dims"], semi, spec["bias_ncpt"], spec["bias_sigma"],
                        spec["noise_sigma"], spec["seed"], rician=True)


def mask_fraction(mask) -> float:
    return float(np.count_nonzero(mask)) / mask.size


def analytic_ellipsoid_count(semi_axes) -> float:
    a, b, c = semi_axes
    return 4.0 / 3.0 * math.pi * a * b * c


def batch_shrinks(n: int, seed: int, choices=(2, 3, 4)):
    """Per-volume shrink factors for the mixed-shrink batch, drawn uniformly
    from `choices` with their own seeded generator."""
    rng = np.random.default_rng(seed)
    return [int(v) for v in rng.choice(np.asarray(choices), size=n)]


def generate_batch(specs, workers: int = 16, mask_dtype=np.uint8):
    """make_batch_volume over specs on a thread pool (numpy releases the GIL
    in the heavy array work).  Returns [(image float32, mask)] with the mask
    cast to mask_dtype."""
    from concurrent.futures import ThreadPoolExecutor

    def one(spec):
        im, mask, _ = make_batch_volume(spec)
        return im, mask.astype(mask_dtype)

    with ThreadPoolExecutor(max(1, int(workers))) as ex:
        return list(ex.map(one, specs))
