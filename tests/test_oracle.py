"""The FP64 oracle pinned against N4.py's own outputs (golden fixtures written
by tests/golden/make_golden.py) and against what the mathematics fixes:
closed forms, invariants, textbook routines and brute force on tiny inputs.
Each check is chosen so that a plausible slip (dropped term, wrong sign or
index, transposed operand) fails it."""
import numpy as np
import pytest

import oracle as O
from n4amd import phantom as P


# ------------------------------------------------------- golden (N4.py) --
@pytest.mark.parametrize("case", [0, 1, 2])
def test_sharpen_matches_n4py(golden, case):
    x, m, y = golden[f"sharpen{case}_x"], golden[f"sharpen{case}_mask"], golden[f"sharpen{case}_out"]
    o, H, E, _ = O.sharpen(x, m)
    # N4.py smooths a complex64 numerator in single precision (N4.py:75-84)
    assert np.abs(o - y)[m != 0].max() <= 5e-6 * np.abs(E).max()
    assert np.all(o[m == 0] == 0)


@pytest.mark.parametrize("case", [0, 1, 2, 3])
def test_n4_trajectory_matches_n4py(golden, case):
    im, m = golden[f"n4_{case}_im"], golden[f"n4_{case}_mask"]
    r = O.n4(im, m, levels=int(golden[f"n4_{case}_levels"]), max_iter=list(golden[f"n4_{case}_maxiter"]),
             conv_threshold=float(golden[f"n4_{case}_thr"]))
    assert r["iterations"] == [int(v) for v in golden[f"n4_{case}_iterations"]]
    mm = (m != 0) & (im > 10)
    lb = np.log(im[mm] / golden[f"n4_{case}_normalized"][mm])
    assert np.abs(r["log_bias"][mm] - lb).max() <= 2e-6
    assert np.abs(r["lattice"] - golden[f"n4_{case}_lattice"]).max() <= 2e-6
    assert np.allclose(r["normalized"], golden[f"n4_{case}_normalized"], rtol=2e-6)


# ------------------------------------------------------------- B-spline --
def test_cubic_bspline_closed_form():
    assert np.allclose(O.cubic_bspline(0.0), [1 / 6, 4 / 6, 1 / 6, 0.0], atol=1e-16)
    assert np.allclose(O.cubic_bspline(1.0), [0.0, 1 / 6, 4 / 6, 1 / 6], atol=1e-15)
    for t in np.random.default_rng(0).random(50):
        w = O.cubic_bspline(t)
        assert abs(w.sum() - 1.0) < 1e-15                      # partition of unity
        assert abs((w * np.arange(4)).sum() - (1.0 + t)) < 1e-14  # linear reproduction
        assert np.all(w >= 0)


def test_axis_table_spans():
    b, w = O.axis_table(10, 10.0, 7)   # 4 spans of 2.5
    assert list(b) == [0, 0, 1, 1, 1, 2, 2, 3, 3, 3]
    t0 = 0.5 / 2.5
    assert np.allclose(w[0], O.cubic_bspline(t0))


def test_evaluate_constant_and_linear_fields():
    dims = (9, 7, 11)
    c = O.evaluate_bspline3d(np.full((5, 6, 4), 2.5), dims)
    assert np.allclose(c, 2.5, atol=1e-14)
    # lattice linear in the control index -> field linear in position
    nc = (6, 5, 7)
    L = np.fromfunction(lambda a, b, c: 0.3 * a - 0.2 * b + 0.1 * c, nc)
    f = O.evaluate_bspline3d(L, dims)
    pos = [(np.arange(n) + 0.5) / (n / (k - 3)) + 1.0 for n, k in zip(dims, nc)]
    ref = 0.3 * pos[0][:, None, None] - 0.2 * pos[1][None, :, None] + 0.1 * pos[2][None, None, :]
    assert np.allclose(f, ref, atol=1e-12)


def test_evaluate_matches_independent_separable_form():
    rng = np.random.default_rng(1)
    L = rng.standard_normal((7, 5, 6))
    dims = (13, 8, 17)
    assert np.allclose(O.evaluate_bspline3d(L, dims), P.evaluate_lattice(L, dims), atol=1e-12)


def test_fit_single_voxel_closed_form():
    dims = (6, 5, 7)
    m = np.zeros(dims, np.uint8)
    m[2, 3, 4] = 1
    f = np.zeros(dims)
    f[2, 3, 4] = 1.7
    eps = 1e-3
    lat = O.fit_bspline3d(f, m, (5, 6, 4), eps=eps)
    ws = [P.dense_weights(n, k)[i] for n, k, i in zip(dims, (5, 6, 4), (2, 3, 4))]
    wc = ws[0][:, None, None] * ws[1][None, :, None] * ws[2][None, None, :]
    ref = 1.7 * wc ** 3 / (wc ** 2).sum() / (wc ** 2 + eps)
    assert np.allclose(lat, ref, atol=1e-14)
    # with eps = 0 a single voxel is interpolated exactly: sum_c w_c L_c = f
    lat0 = O.fit_bspline3d(f, m, (5, 6, 4), eps=0.0)
    lat0[~np.isfinite(lat0)] = 0.0
    assert abs((wc * lat0).sum() - 1.7) < 1e-12


def test_fit_numerator_denominator_separable():
    """The direct 64-term accumulation (N4.py:186-237) equals the separable
    contraction with per-axis factors w^3/sum(w^2) and w^2."""
    rng = np.random.default_rng(2)
    dims, nc = (11, 9, 10), (5, 7, 6)
    f = rng.standard_normal(dims)
    m = (rng.random(dims) < 0.6).astype(np.uint8)
    num, den = O.fit_numden(f, m, nc)
    Pw, Qw = [], []
    for n, k in zip(dims, nc):
        W = P.dense_weights(n, k)
        S = (W ** 2).sum(1, keepdims=True)
        Pw.append(W ** 3 / S)
        Qw.append(W ** 2)
    g = f * m
    num2 = np.einsum("xyz,xa,yb,zc->abc", g, *Pw)
    den2 = np.einsum("xyz,xa,yb,zc->abc", m.astype(float), *Qw)
    assert np.allclose(num, num2, atol=1e-12)
    assert np.allclose(den, den2, atol=1e-12)


@pytest.mark.parametrize("nc", [(4, 4, 4), (5, 6, 7), (7, 4, 5)])
def test_upsample_refinement_invariance(nc):
    """Dyadic B-spline subdivision leaves the represented field unchanged."""
    L = np.random.default_rng(3).standard_normal(nc)
    U = O.upsample_lattice3d(L)
    assert U.shape == tuple(2 * (n - 3) + 3 for n in nc)
    dims = (17, 12, 14)
    assert np.allclose(O.evaluate_bspline3d(U, dims), O.evaluate_bspline3d(L, dims), atol=1e-12)


# ------------------------------------------------------------ histogram --
def test_histogram_fractional_split_and_mass():
    x = np.array([0.0, 1.0, 0.25, 0.5, 0.1234, 1.0])
    m = np.ones_like(x)
    nb = 5
    slope = 1.0 / (nb - 1)
    H, idx = O.histogram(x, m, nb, 0.0, slope, with_idx=True)
    assert list(idx) == [0, 4, 1, 2, 0, 4]
    # 0.1234 -> cidx 0.4936: 0.5064 to bin 0, 0.4936 to bin 1
    ref = np.zeros(nb)
    ref[0] = 1 + (1 - 0.4936)
    ref[1] = 1 + 0.4936
    ref[2] = 1
    ref[4] = 2
    assert np.allclose(H, ref, atol=1e-12)
    # mass: every voxel contributes 1 except dropped top-edge voxels
    rng = np.random.default_rng(4)
    y = rng.normal(size=1000)
    mm = rng.random(1000) < 0.5
    n, mn, mx = O.masked_minmax(y, mm)
    H = O.histogram(y, mm, 200, mn, (mx - mn) / 199)
    assert n - 1 <= H.sum() <= n + 1e-9


def test_padded_size():
    assert O.padded_size(200) == (512, 156)
    assert O.padded_size(256) == (512, 128)
    assert O.padded_size(100) == (256, 78)


def test_fft_matches_textbook_dft():
    rng = np.random.default_rng(5)
    for n in (8, 64, 512):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        assert np.allclose(O.fft(x), np.fft.fft(x), atol=1e-12 * n)
        assert np.allclose(O.fft(x, inverse=True), np.fft.ifft(x), atol=1e-12)
    x = rng.standard_normal(16)
    k = np.arange(16)
    dft = (x[None, :] * np.exp(-2j * np.pi * k[:, None] * k[None, :] / 16)).sum(1)
    assert np.allclose(O.fft(x), dft, atol=1e-12)


def test_wiener_identity_limit():
    """Zero noise and a delta kernel (FWHM << bin width): deconvolution and
    smoothing are identities, so E maps each occupied bin to its centre."""
    H = np.zeros(200)
    H[[10, 50, 51, 120, 199]] = [3.0, 1.0, 2.0, 5.0, 1.0]
    E = O.wiener_etable(H, 2.0, 0.01, noise=0.0, fwhm=1e-6)
    occ = H > 0
    assert np.allclose(E[occ], 2.0 + np.arange(200)[occ] * 0.01, atol=1e-10)


def test_convergence_is_cv_ddof0():
    rng = np.random.default_rng(6)
    a, b = rng.normal(0, 0.02, 500), rng.normal(0, 0.02, 500)
    m = rng.random(500) < 0.7
    e = np.exp(a - b)[m]
    assert abs(O.convergence(a, b, m) - e.std() / e.mean()) < 1e-14


def test_threshold_log_shrink():
    rng = np.random.default_rng(7)
    im = rng.uniform(0, 100, (8, 6, 10)).astype(np.float32)
    mk = (rng.random(im.shape) < 0.8).astype(np.float32)
    lg, mw, nf, nw = O.threshold_log_shrink(im, mk, 10.0, 1)
    fg = (mk != 0) & (im > 10)
    assert nf == nw == fg.sum()
    assert np.array_equal(mw.astype(bool), fg)
    assert np.allclose(lg[fg], np.log(im[fg].astype(np.float64)))
    # shrink 2 on divisible dims: block mean of the masked logs, >=50% rule
    lg2, mw2, nf2, nw2 = O.threshold_log_shrink(im, mk, 10.0, 2)
    blk = lambda a: a.reshape(4, 2, 3, 2, 5, 2).transpose(0, 2, 4, 1, 3, 5).reshape(4, 3, 5, 8)
    cnt = blk(fg).sum(-1)
    s = blk(np.where(fg, np.log(np.maximum(im, 1e-30).astype(np.float64)), 0.0)).sum(-1)
    assert np.array_equal(mw2.astype(bool), (cnt > 0) & (2 * cnt >= 8))
    sel = cnt > 0
    assert np.allclose(lg2[sel], s[sel] / cnt[sel])
    assert nf2 == fg.sum()


def test_n4_scale_invariance():
    """Scaling the image (and the low threshold) by c shifts every log
    intensity by log c: the histogram, E table and sharpened image shift
    with it, so the residual and the bias field do not change."""
    im, m, _ = P.make_phantom((18, 16, 14), (7.5, 6.5, 5.5), 4, 0.3, 5.0, seed=21)
    a = O.n4(im, m, levels=2, max_iter=[6, 6], conv_threshold=0.0)
    b = O.n4((im * 4.0).astype(np.float32), m, levels=2, max_iter=[6, 6], conv_threshold=0.0, low_value=40.0)
    mm = (m != 0) & (im > 10)
    assert np.abs(a["log_bias"][mm] - b["log_bias"][mm]).max() < 1e-9


def test_n4_recovers_bias_on_phantom():
    """On the C0 recipe the estimated log bias tracks the true one up to an
    additive constant (N4 determines the field only up to a constant)."""
    im, m, true = P.c0_phantom(0)
    r = O.n4(im, m, levels=3, max_iter=50, conv_threshold=0.001, nthreads=4)
    mm = (m != 0) & (im > 10)
    d = r["log_bias"][mm] - true[mm]
    assert np.std(d) < 0.25 * np.std(true[mm])
