"""BF16 and split-BF16 MFMA forms of the separable B-spline contractions
(fit numerator and denominator, and -- when selected -- the working-grid
evaluation of the field increments) against the FP64 oracle.  The final
full-resolution reconstruction is the FP32 kernel in every mode; the
per-iteration evaluation is the FP32 kernel by default and the MFMA one
with eval_kernel="mfma".

Tolerances on the log bias field (max abs over masked voxels):
  * split-BF16 (voxel operands and basis factors as hi + lo BF16 halves,
    ~16 significant bits): 1e-4, the same bar as the FP32 path;
  * BF16 (voxel operands in BF16, basis factors hi + lo): 1e-3, the
    batched-path bar.
Where the convergence test decides the iteration count, a BF16 run may stop
one iteration earlier or later than the FP64 oracle; those cases compare
the fields without requiring equal counts."""
import numpy as np
import pytest

import oracle as O
import parity_cases as PC
from n4amd import phantom as P

pytestmark = pytest.mark.gpu

TOL = {"split_bf16": 1e-4, "bf16": 1e-3}


@pytest.fixture(params=["fp32", "mfma"])
def eval_kernel(gpu, request):
    gpu.set_tuning(eval_kernel=request.param)
    yield request.param
    gpu.set_tuning(reset=True)


@pytest.mark.parametrize("precision", ["split_bf16", "bf16"])
def test_fixed_iterations_vs_oracle(gpu, precision, eval_kernel):
    im, m, _ = PC.small_phantom((48, 40, 36), seed=2)
    print(f"{precision} ({eval_kernel} eval): 48x40x36, 3 levels x 10 fixed iterations", flush=True)
    e, rms, _ = PC.check_n4_vs_oracle(gpu, im, m, 3, [10] * 3, 0.0, fixed=True, precision=precision,
                                      tol=TOL[precision])
    print(f"{precision}: max err {e:.3g} rms {rms:.3g}")


@pytest.mark.parametrize("precision", ["split_bf16", "bf16"])
def test_c0_reference_case(gpu, precision):
    """64^3 phantom, 3 levels x 50 iterations, early exit at CV <= 0.001."""
    im, m, _ = P.c0_phantom(0)
    print(f"{precision}: C0 64^3", flush=True)
    e, rms, its = PC.check_n4_vs_oracle(gpu, im, m, 3, [50] * 3, 0.001, precision=precision, tol=TOL[precision],
                                        same_iterations=(precision == "split_bf16"))
    print(f"{precision}: C0 max err {e:.3g} rms {rms:.3g} iterations {its}")


def test_split_golden_trajectory(gpu, golden):
    im, m = golden["n4_1_im"], golden["n4_1_mask"]
    L, its, thr = int(golden["n4_1_levels"]), list(golden["n4_1_maxiter"]), float(golden["n4_1_thr"])
    print("split_bf16: golden trajectory 1", flush=True)
    PC.check_n4_vs_oracle(gpu, im, m, L, its, thr, precision="split_bf16")


@pytest.mark.parametrize("shrink", [2, 3])
def test_split_shrink_partial_tiles(gpu, shrink, eval_kernel):
    """Edge tiles everywhere: 75x62x54 leaves partial 16-row tiles and
    partial 16- and 32-wide z tiles on the working grid."""
    im, m, _ = PC.small_phantom((75, 62, 54), seed=9)
    print(f"split_bf16 ({eval_kernel} eval): shrink {shrink} on 75x62x54", flush=True)
    PC.check_n4_vs_oracle(gpu, im, m, 3, [12] * 3, 0.0, shrink=shrink, fixed=True, precision="split_bf16")


@pytest.mark.parametrize("precision", ["split_bf16", "bf16"])
def test_multi_tile_lattice(gpu, precision, eval_kernel):
    """ncpt (4, 5, 6) over 4 levels ends at 11 x 19 x 27 control points: two
    16-point MFMA tiles along y and z."""
    im, m, _ = PC.small_phantom((40, 56, 72), seed=4)
    print(f"{precision} ({eval_kernel} eval): ncpt (4,5,6), 4 levels", flush=True)
    e, rms, _ = PC.check_n4_vs_oracle(gpu, im, m, 4, [3] * 4, 0.0, ncpt=(4, 5, 6), fixed=True,
                                      precision=precision, tol=TOL[precision])
    print(f"{precision}: multi-tile max err {e:.3g} rms {rms:.3g}")


def test_split_batch_equals_single(gpu):
    vols = []
    for k, dims in enumerate([(30, 26, 22), (17, 33, 25), (41, 19, 28), (24, 24, 24), (33, 29, 18)]):
        im, m, _ = PC.small_phantom(dims, seed=40 + k)
        vols.append((im, m))
    print("split_bf16: batch of 5 uneven volumes", flush=True)
    rb = gpu.correct_batch(vols, levels=3, max_iter=[8, 8, 8], precision="split_bf16")
    for (im, m), r in zip(vols, rb):
        s = gpu.correct(im, m, levels=3, max_iter=[8, 8, 8], precision="split_bf16")
        assert r["iterations"] == s["iterations"]
        assert np.array_equal(r["log_bias"], s["log_bias"])


def test_bf16_deterministic(gpu):
    im, m, _ = PC.small_phantom((48, 40, 36), seed=3)
    a = gpu.correct(im, m, levels=3, max_iter=[10] * 3, precision="bf16")
    b = gpu.correct(im, m, levels=3, max_iter=[10] * 3, precision="bf16")
    assert np.array_equal(a["log_bias"], b["log_bias"])


def test_precision_limits(gpu):
    """The MFMA kernels cover up to 64 control points along y and z; a call
    whose final lattice has more (ncpt 5 at 6 levels: 67 along z) runs the
    FP32 kernels, bit for bit the FP32 call.  An unknown precision value is
    refused."""
    import n4amd
    im, m, _ = PC.small_phantom()
    kw = dict(levels=6, ncpt=(4, 4, 5), max_iter=3, fixed_iterations=True)
    a = gpu.correct(im, m, precision="bf16", **kw)
    b = gpu.correct(im, m, precision="fp32", **kw)
    assert a["lattice"].shape == (35, 35, 67)
    assert np.array_equal(a["log_bias"], b["log_bias"])
    with pytest.raises(n4amd.N4Error):
        gpu.correct(im, m, precision=7)


def test_split_mixed_shrink_batch_vs_oracle(gpu):
    """The 256-volume configuration's form at small size: one batch whose
    volumes carry their own shrink factors (2, 3, 4), split-BF16, each volume
    against the FP64 oracle at its own shrink."""
    dims = [(64, 56, 48), (72, 60, 54), (80, 64, 60), (50, 62, 44)]
    shr = [2, 3, 4, 2]
    vols = []
    for k, d in enumerate(dims):
        im, m, _ = PC.small_phantom(d, seed=80 + k)
        vols.append((im, m))
    print("split_bf16: mixed-shrink batch of 4", flush=True)
    out = gpu.correct_batch(vols, levels=3, max_iter=[12] * 3, threshold=0.001, shrink=shr, precision="split_bf16")
    for (im, m), s, r in zip(vols, shr, out):
        ro = O.n4(im, m, levels=3, max_iter=[12] * 3, conv_threshold=0.001, shrink=s)
        assert r["iterations"] == ro["iterations"], (s, r["iterations"], ro["iterations"])
        mm = (np.asarray(m) != 0) & (im > 10)
        err = np.abs(r["log_bias"].astype(np.float64) - ro["log_bias"])[mm]
        print(f"shrink {s}: max err {err.max():.3g}")
        assert err.max() <= TOL["split_bf16"]
