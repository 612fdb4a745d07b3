"""Confidence-weighted masks (N4_FLAG_WEIGHTED_MASK).

The reference takes binary masks only (cudaN4.h:346).  Under the flag a
float mask holds confidence weights: a voxel is foreground where its weight
is > 0 and finite, the fit's numerator and denominator terms of a voxel
(N4.py:230-236) are scaled by its weight, histogram and convergence measure
stay unweighted, and a shrunk working voxel carries the mean weight of its
block's foreground voxels.

The oracle's weighted fit is pinned below by linearity against its binary
fit (a weight taking two values splits into two binary fits), its weighted
shrink by a brute-force block loop, and its whole weighted correction by
the reduction to the binary case when every weight is 0 or 1.  The
emulated kernels (CPU) and the HIP library (GPU) are then compared with the
weighted oracle at the same tolerances as the binary parity tests.
"""
import numpy as np
import pytest

import oracle as O
import parity_cases as PC


def weights_for(mask, seed, lo=0.2):
    """Confidence weights in [lo, 1] inside the mask, a few zeros inside it
    too, 0 outside."""
    rng = np.random.default_rng(seed)
    w = (lo + (1.0 - lo) * rng.random(mask.shape)).astype(np.float32)
    w[rng.random(mask.shape) < 0.05] = 0.0
    return np.where(np.asarray(mask) != 0, w, 0.0).astype(np.float32)


# ----------------------------------------------------------- oracle pins --
def test_oracle_weighted_fit_is_linear_in_the_weights():
    rng = np.random.default_rng(11)
    dims, nc = (13, 11, 9), (5, 6, 4)
    f = rng.standard_normal(dims)
    m = rng.random(dims) < 0.8
    a = rng.random(dims) < 0.5
    conf = np.where(a, 0.5, 2.0)
    num, den = O.fit_numden(f, m, nc, conf=conf)
    na, da = O.fit_numden(f, m & a, nc)
    nb, db = O.fit_numden(f, m & ~a, nc)
    assert np.allclose(num, 0.5 * na + 2.0 * nb, rtol=1e-12, atol=1e-12)
    assert np.allclose(den, 0.5 * da + 2.0 * db, rtol=1e-12, atol=1e-12)
    # unit weights are the binary fit, bit for bit
    n1, d1 = O.fit_numden(f, m, nc, conf=np.ones(dims))
    n0, d0 = O.fit_numden(f, m, nc)
    assert np.array_equal(n1, n0) and np.array_equal(d1, d0)


@pytest.mark.parametrize("s", [1, 2, 3])
def test_oracle_weighted_shrink_brute_force(s):
    rng = np.random.default_rng(5 + s)
    dims = (10, 9, 7)
    im = (20.0 + 80.0 * rng.random(dims)).astype(np.float32)
    im[rng.random(dims) < 0.1] = 5.0  # below low_value: background
    w = rng.random(dims).astype(np.float32)
    w[rng.random(dims) < 0.2] = 0.0
    w[0, 0, 0] = -1.0
    w[1, 1, 1] = np.inf
    imlog, mw, nf, nw, cw = O.threshold_log_shrink(im, w, shrink=s, weighted=True)
    W = [(d + s - 1) // s for d in dims]
    fg_full = (w > 0) & np.isfinite(w) & (im > 10.0)
    assert nf == int(fg_full.sum())
    for I in range(W[0]):
        for J in range(W[1]):
            for K in range(W[2]):
                blk = (slice(I * s, (I + 1) * s), slice(J * s, (J + 1) * s), slice(K * s, (K + 1) * s))
                sel = fg_full[blk]
                nm, nb = int(sel.sum()), sel.size
                fg = nm > 0 and 2 * nm >= nb
                assert mw[I, J, K] == fg
                if fg:
                    assert abs(cw[I, J, K] - w[blk][sel].astype(np.float64).mean()) <= 1e-15
                    assert abs(imlog[I, J, K] - np.log(im[blk][sel].astype(np.float64)).mean()) <= 1e-12
                else:
                    assert cw[I, J, K] == 0.0
    # binary read of the same mask: foreground where nonzero
    _, mb, nfb, _ = O.threshold_log_shrink(im, w, shrink=s)
    assert nfb == int(((w != 0) & (im > 10.0)).sum())


def test_oracle_weighted_binary_weights_reduce_to_binary():
    im, m, _ = PC.small_phantom((22, 20, 18), seed=9)
    wb = (np.asarray(m) != 0).astype(np.float32)
    kw = dict(levels=2, max_iter=[4, 4], conv_threshold=0.0, shrink=2)
    rb = O.n4(im, wb, **kw)
    rw = O.n4(im, wb, weighted=1, **kw)
    assert np.array_equal(rb["log_bias"], rw["log_bias"])
    # and real weights change the answer
    rr = O.n4(im, weights_for(m, 3), weighted=1, **kw)
    assert np.abs(rr["log_bias"] - rb["log_bias"]).max() > 1e-4


# ------------------------------------------------------ library vs oracle --
def check_weighted(lib, im, w, levels, max_iter, thr, shrink=1, precision="fp32", tol=PC.LOGBIAS_TOL, ncpt=4):
    r = lib.correct(im, w, levels=levels, max_iter=max_iter, threshold=thr, shrink=shrink, ncpt=ncpt,
                    precision=precision, weighted_mask=True)
    ro = O.n4(im, w, levels=levels, max_iter=max_iter, conv_threshold=thr, shrink=shrink, ncpt=ncpt, weighted=1)
    assert r["n_mask_full"] == ro["n_mask_full"] and r["n_mask_work"] == ro["n_mask_work"]
    assert r["iterations"] == ro["iterations"], (r["iterations"], ro["iterations"])
    mm = (w > 0) & (im > 10)
    err = np.abs(r["log_bias"].astype(np.float64) - ro["log_bias"])[mm]
    assert err.max() <= tol, f"log bias max err {err.max():.3g} > {tol:.3g}"
    return float(err.max())


@pytest.mark.parametrize("shrink", [1, 2])
def test_emulated_weighted_vs_oracle(emu, shrink):
    im, m, _ = PC.small_phantom((26, 24, 20), seed=31)
    check_weighted(emu, im, weights_for(m, 1), levels=3, max_iter=[6, 6, 6], thr=0.001, shrink=shrink)


def test_emulated_unit_weights_equal_binary(emu):
    im, m, _ = PC.small_phantom((24, 22, 20), seed=8)
    wb = (np.asarray(m) != 0).astype(np.float32)
    kw = dict(levels=2, max_iter=[5, 5], threshold=0.001, shrink=2)
    a = emu.correct(im, wb, **kw)
    b = emu.correct(im, wb, weighted_mask=True, **kw)
    assert a["iterations"] == b["iterations"]
    assert np.array_equal(a["log_bias"], b["log_bias"])


def test_emulated_weighted_teams_equal_steps(emu):
    vols = []
    for k, d in enumerate([(30, 26, 22), (17, 33, 25)]):
        im, m, _ = PC.small_phantom(d, seed=50 + k)
        vols.append((im, weights_for(m, 40 + k)))
    kw = dict(levels=2, max_iter=[5, 5], threshold=0.001, weighted_mask=True)
    out = {}
    for path in ("steps", "teams"):
        try:
            emu.set_tuning(path=path, step_tile_voxels=512)
            out[path] = emu.correct_batch(vols, **kw)
        finally:
            emu.set_tuning(reset=True)
    for a, b in zip(out["steps"], out["teams"]):
        assert a["iterations"] == b["iterations"]
        assert np.array_equal(a["log_bias"], b["log_bias"])


def test_unknown_flag_bits_rejected(emu):
    import n4amd
    im, m, _ = PC.small_phantom((16, 16, 16), seed=2)
    p = n4amd.default_params(levels=1, max_iter=1)
    p.flags = 0x80
    with pytest.raises(n4amd.N4Error, match="flags"):
        emu.correct(im, m, params=p)


@pytest.mark.gpu
@pytest.mark.parametrize("precision,shrink,tol", [("fp32", 1, 1e-4), ("fp32", 2, 1e-4), ("bf16", 2, 1e-3),
                                                  ("split_bf16", 2, 1e-4)])
def test_gpu_weighted_vs_oracle(gpu, precision, shrink, tol):
    im, m, _ = PC.small_phantom((40, 36, 32), seed=33)
    e = check_weighted(gpu, im, weights_for(m, 2), levels=3, max_iter=[8, 8, 8], thr=0.001, shrink=shrink,
                       precision=precision, tol=tol)
    print(f"weighted {precision} shrink {shrink}: max |d log bias| {e:.3g}")


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_gpu_unit_weights_equal_binary(gpu, precision):
    """Weights of 0 and 1 give the binary FP32 correction bit for bit (the
    MFMA precisions run the FP32 kernels for confidence weights)."""
    im, m, _ = PC.small_phantom((40, 36, 32), seed=34)
    wb = (np.asarray(m) != 0).astype(np.float32)
    kw = dict(levels=3, max_iter=[6, 6, 6], threshold=0.001, shrink=2)
    a = gpu.correct(im, wb, precision="fp32", **kw)
    b = gpu.correct(im, wb, weighted_mask=True, precision=precision, **kw)
    assert a["iterations"] == b["iterations"]
    assert np.array_equal(a["log_bias"], b["log_bias"])
