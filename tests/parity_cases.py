"""Parity checks of the library against the FP64 oracle, shared by the GPU
tests (product library) and the CPU tests of the kernel logic (host
emulation build).  Each check takes a loaded n4amd.Library.

Tolerances (stated once, used everywhere):
  * mask voxel counts and histogram bin indices: bit-exact;
  * histogram mass: 1e-9 relative (2^-32 fixed-point accumulation);
  * E table: 1e-11 relative (FP64 FFT on both sides);
  * FP32 stage outputs (sharpened, lattice, field): FP32 rounding bounds;
  * log bias field of the whole correction: 1e-4 absolute over masked voxels
    (the FP32 path's bar against the FP64 oracle), at equal iteration counts.
"""
from __future__ import annotations

import numpy as np

import oracle as O
from n4amd import phantom as P

LOGBIAS_TOL = 1e-4


def f32(a):
    return np.asarray(a, dtype=np.float32)


def check_histogram_exact(lib, x, mask, nbins=200):
    x = f32(x)
    H, idx, (mn, mx, slope) = lib.histogram(x, mask, nbins)
    n, omn, omx = O.masked_minmax(x.astype(np.float64), mask)
    oslope = (omx - omn) / (nbins - 1)
    assert (mn, mx, slope) == (omn, omx, oslope)
    Ho, io = O.histogram(x.astype(np.float64), mask, nbins, omn, oslope, with_idx=True)
    assert np.array_equal(idx, io), "bin indices differ"
    assert np.allclose(H, Ho, rtol=1e-9, atol=1e-9 * max(1.0, Ho.max()))
    return idx


def check_sharpen(lib, x, mask, n4py_out=None):
    x = f32(x)
    y, E = lib.sharpen(x, mask)
    o, _, Eo, _ = O.sharpen(x.astype(np.float64), mask)
    m = np.asarray(mask) != 0
    assert np.abs(E - Eo).max() <= 1e-11 * np.abs(Eo).max()
    # FP32 output of an FP64 computation: within 2 ulp of |E|
    assert np.abs(y - o)[m].max() <= 4e-7 * max(1.0, np.abs(Eo).max())
    if n4py_out is not None:
        # N4.py smooths a complex64 numerator (single precision): ~1e-6
        assert np.abs(y - n4py_out)[m].max() <= 5e-6 * max(1.0, np.abs(Eo).max())


def check_fit_eval_upsample(lib, dims, ncpt, seed=0):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal(dims).astype(np.float32)
    m = rng.random(dims) < 0.7
    lat = lib.fit_bspline3d(f, m, ncpt)
    lo = O.fit_bspline3d(f.astype(np.float64), m, ncpt)
    assert np.abs(lat - lo).max() <= 2e-6 * max(1.0, np.abs(lo).max())
    lat32 = lo.astype(np.float32)
    ev = lib.evaluate_bspline3d(lat32, dims)
    eo = O.evaluate_bspline3d(lat32.astype(np.float64), dims)
    assert np.abs(ev - eo).max() <= 1e-6 * max(1.0, np.abs(eo).max())
    up = lib.upsample_lattice3d(lat32)
    uo = O.upsample_lattice3d(lat32.astype(np.float64))
    assert np.abs(up - uo).max() <= 1e-6 * max(1.0, np.abs(uo).max())


def check_convergence_and_reduce(lib, n=5000, seed=1):
    rng = np.random.default_rng(seed)
    f1 = (0.01 * rng.standard_normal(n)).astype(np.float32)
    f2 = (0.01 * rng.standard_normal(n)).astype(np.float32)
    m = rng.random(n) < 0.5
    cv = lib.convergence(f1, f2, m)
    e = np.exp(f1.astype(np.float64) - f2.astype(np.float64))[m]
    assert abs(cv - e.std() / e.mean()) <= 1e-6 * (e.std() / e.mean())
    s, mx, mn = lib.reduce(f1)
    assert abs(s - f1.astype(np.float64).sum()) <= 1e-9
    assert mx == f1.max() and mn == f1.min()


def sensitivity(im, mask, n_trials=2, seed=3, **kw):
    """Max change of the oracle's log bias under 1-ulp FP32 input changes:
    the conditioning floor below which no FP32 path can be judged."""
    r0 = O.n4(im, mask, **kw)
    rng = np.random.default_rng(seed)
    worst = 0.0
    mm = (np.asarray(mask) != 0) & (im > kw.get("low_value", 10.0)) if mask is not None else im > 10
    for _ in range(n_trials):
        im2 = im.copy()
        sel = rng.random(im.shape) < 0.5
        im2[sel] = np.nextafter(im2[sel], np.float32(np.inf))
        r1 = O.n4(im2, mask, **kw)
        worst = max(worst, float(np.abs(r1["log_bias"] - r0["log_bias"])[mm].max()))
    return worst


def check_n4_vs_oracle(lib, im, mask, levels, max_iter, thr, shrink=1, tol=LOGBIAS_TOL, ncpt=4,
                       fixed=False, ref=None, precision="fp32", same_iterations=True):
    """Whole correction vs the FP64 oracle.  Returns (max err, rms err, its)."""
    r = lib.correct(im, mask, levels=levels, max_iter=max_iter, threshold=thr, shrink=shrink, ncpt=ncpt,
                    fixed_iterations=fixed, want_bias=True, precision=precision)
    ro = ref if ref is not None else O.n4(im, mask, levels=levels, max_iter=max_iter, conv_threshold=thr,
                                          shrink=shrink, ncpt=ncpt)
    assert r["n_mask_full"] == ro["n_mask_full"]
    assert r["n_mask_work"] == ro["n_mask_work"]
    if same_iterations:
        assert r["iterations"] == ro["iterations"], (r["iterations"], ro["iterations"])
    mm = ((np.asarray(mask) != 0) if mask is not None else np.ones(im.shape, bool)) & (im > 10)
    err = np.abs(r["log_bias"].astype(np.float64) - ro["log_bias"])[mm]
    assert err.max() <= tol, f"log bias max err {err.max():.3g} > {tol:.3g}"
    # normalized image and bias field follow from the log bias
    assert np.allclose(r["bias"], np.exp(r["log_bias"].astype(np.float64)), rtol=1e-6)
    thr_im = np.where(im > 10, im, 0).astype(np.float64)
    assert np.allclose(r["normalized"], thr_im / np.exp(r["log_bias"].astype(np.float64)), rtol=2e-6, atol=1e-6)
    return float(err.max()), float(np.sqrt(np.mean(err ** 2))), r["iterations"]


def small_phantom(dims=(24, 22, 20), seed=5, bias=0.3):
    semi = tuple(0.42 * d for d in dims)
    return P.make_phantom(dims, semi, 4, bias, 5.0, seed)
