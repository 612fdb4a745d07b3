"""The benchmark configurations at their full sizes on the GPU, each against
the FP64 oracle on a stated sample of volumes (the N4.py trajectory
semantics, N4.py:276-332):

  C2  one 512x512x320 phantom (bias lattice 6^3, N(0,0.4)), shrink 2
      (256x256x160 working grid), 4 levels x <=50 iterations, threshold
      0.001: FP32, BF16 and split-BF16 against one shared oracle run;
  C3  64 phantoms, axes U{128..288} (seed 0), shrink 4, 4 levels x <=50:
      FP32 and BF16 over the whole batch; volumes 0..7 against the oracle;
  C4  256 phantoms, axes U{96..320} (seed 4), shrink U{2,3,4} per volume
      (seed 5), 4 levels x <=50: split-BF16 over the whole batch; the first
      two volumes of each shrink class (in index order) against the oracle.

Bars on the log bias field (max |build - oracle| over masked voxels above
the low threshold): 1e-4 for FP32 and split-BF16, 1e-3 for BF16.  Iteration
counts per level: equal for FP32 and split-BF16; a BF16 run may end a level
one iteration earlier or later where the convergence test decides it.
The per-output max / RMS errors are printed (-s shows them).
"""
import os

import numpy as np
import pytest

import oracle as O
from n4amd import phantom as P

pytestmark = pytest.mark.gpu

TOL = {"fp32": 1e-4, "split_bf16": 1e-4, "bf16": 1e-3}
NTHREADS = min(16, os.cpu_count() or 1)
KW = dict(levels=4, max_iter=[50] * 4)


def _oracle(im, mask, shrink):
    return O.n4(im, mask, conv_threshold=0.001, shrink=shrink, nthreads=NTHREADS, **KW)


def _errors(im, mask, r, ro):
    mm = (np.asarray(mask) != 0) & (im > 10)
    e = np.abs(r["log_bias"].astype(np.float64) - ro["log_bias"])[mm]
    # the normalised image follows from the log bias: relative error
    n_ref = np.where(im > 10, im, 0).astype(np.float64) / np.exp(ro["log_bias"])
    en = (np.abs(r["normalized"].astype(np.float64) - n_ref) / np.maximum(n_ref, 1e-30))[mm]
    return {"logbias_max": float(e.max()), "logbias_rms": float(np.sqrt(np.mean(e ** 2))),
            "normalized_rel_max": float(en.max())}


def _check_iterations(prec, got, want):
    if prec == "bf16":
        assert all(abs(a - b) <= 1 for a, b in zip(got, want)), (got, want)
    else:
        assert got == want, (got, want)


@pytest.mark.timeout(900)
def test_c2_precisions_vs_oracle(gpu):
    im, m, _ = P.c2_phantom(2)
    m = m.astype(np.uint8)
    print("C2 512x512x320 shrink 2: oracle (%d threads)" % NTHREADS, flush=True)
    ro = _oracle(im, m, 2)
    for prec in ("fp32", "bf16", "split_bf16"):
        print(f"C2 {prec}", flush=True)
        r = gpu.correct(im, m, threshold=0.001, shrink=2, precision=prec, **KW)
        assert r["n_mask_full"] == ro["n_mask_full"] and r["n_mask_work"] == ro["n_mask_work"]
        _check_iterations(prec, r["iterations"], ro["iterations"])
        err = _errors(im, m, r, ro)
        print(f"C2 {prec}: iterations {r['iterations']} (oracle {ro['iterations']}) {err}", flush=True)
        assert err["logbias_max"] <= TOL[prec]


@pytest.fixture(scope="module")
def c3_volumes():
    print("C3: generating 64 phantoms", flush=True)
    return P.generate_batch(P.batch_specs(64, 128, 288, seed=0), workers=NTHREADS)


@pytest.fixture(scope="module")
def c3_oracle(c3_volumes):
    out = []
    for i in range(8):
        im, m = c3_volumes[i]
        print(f"C3: oracle volume {i} {im.shape}", flush=True)
        out.append(_oracle(im, m, 4))
    return out


@pytest.mark.timeout(900)
@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_c3_batch_vs_oracle(gpu, c3_volumes, c3_oracle, prec):
    print(f"C3 {prec}: batch of 64, shrink 4", flush=True)
    res = gpu.correct_batch(c3_volumes, threshold=0.001, shrink=[4] * 64, precision=prec, outputs=range(8), **KW)
    worst = 0.0
    for i in range(8):
        im, m = c3_volumes[i]
        r, ro = res[i], c3_oracle[i]
        assert r["n_mask_work"] == ro["n_mask_work"]
        _check_iterations(prec, r["iterations"], ro["iterations"])
        err = _errors(im, m, r, ro)
        worst = max(worst, err["logbias_max"])
        print(f"C3 {prec} volume {i}: iterations {r['iterations']} (oracle {ro['iterations']}) {err}", flush=True)
    for r in res:
        assert 0 < sum(r["iterations"]) <= 200
    assert worst <= TOL[prec]


@pytest.mark.timeout(1200)
def test_c4_split_bf16_mixed_shrink_vs_oracle(gpu):
    specs = P.batch_specs(256, 96, 320, seed=4)
    shrinks = P.batch_shrinks(256, seed=5)
    sample = [i for s in (2, 3, 4) for i in [k for k, v in enumerate(shrinks) if v == s][:2]]
    print(f"C4: generating 256 phantoms; sample {sample}", flush=True)
    vols = P.generate_batch(specs, workers=NTHREADS)
    print("C4 split_bf16: batch of 256, shrink 2/3/4", flush=True)
    res = gpu.correct_batch(vols, threshold=0.001, shrink=shrinks, precision="split_bf16", outputs=sample, **KW)
    for r in res:
        assert 0 < sum(r["iterations"]) <= 200
    errs = []
    for i in sample:
        im, m = vols[i]
        print(f"C4: oracle volume {i} {im.shape} shrink {shrinks[i]}", flush=True)
        ro = _oracle(im, m, shrinks[i])
        r = res[i]
        assert r["n_mask_work"] == ro["n_mask_work"]
        _check_iterations("split_bf16", r["iterations"], ro["iterations"])
        err = _errors(im, m, r, ro)
        errs.append(err["logbias_max"])
        print(f"C4 volume {i} shrink {shrinks[i]}: iterations {r['iterations']} (oracle {ro['iterations']}) "
              f"{err}", flush=True)
    print(f"C4 split_bf16 log-bias max error over the sample: {errs}", flush=True)
    assert max(errs) <= TOL["split_bf16"]
