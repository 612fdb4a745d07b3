"""The allocation-free entry point n4_correct_noalloc (plan and working
buffers in the library's file-scope device scratch array) against
n4_correct on the same inputs: the outputs must be bitwise equal (same
kernels, same tiles, same reduction order), on two calls in a row, the
second with a smaller volume.  Requests beyond the scratch size fail with
N4_ERR_NO_MEMORY."""
import numpy as np
import pytest

import n4amd
import parity_cases as PC
from n4amd import phantom as P

pytestmark = pytest.mark.gpu


def _pair(gpu, im, m, **kw):
    a = gpu.correct(im, m, want_bias=True, **kw)
    b = gpu.correct(im, m, want_bias=True, noalloc=True, **kw)
    assert a["iterations"] == b["iterations"]
    for k in ("normalized", "log_bias", "bias", "lattice"):
        assert np.array_equal(a[k], b[k]), k
    return b


def test_noalloc_equals_correct_two_calls():
    gpu = n4amd.load()
    kw = dict(levels=3, max_iter=[20] * 3, threshold=0.001)
    big = PC.small_phantom((64, 60, 56), seed=11)[:2]
    small = PC.small_phantom((40, 36, 30), seed=12)[:2]
    print("noalloc: 64x60x56 then 40x36x30", flush=True)
    r1 = _pair(gpu, *big, **kw)
    r2 = _pair(gpu, *small, **kw)
    assert sum(r1["iterations"]) > 3 and sum(r2["iterations"]) > 3


def test_noalloc_c1_shape_and_precisions():
    gpu = n4amd.load()
    im, m, _ = P.c1_phantom(1)
    for prec in ("fp32", "bf16"):
        print(f"noalloc: 256x256x176 shrink 4 {prec}", flush=True)
        _pair(gpu, im, m, levels=4, max_iter=[50] * 4, threshold=0.001, shrink=4, precision=prec)


def test_noalloc_rejects_oversize():
    gpu = n4amd.load()
    im = np.full((300, 300, 300), 100.0, np.float32)
    with pytest.raises(n4amd.N4Error, match="allocation-free"):
        gpu.correct(im, None, levels=1, max_iter=[1], noalloc=True)
