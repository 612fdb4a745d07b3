"""Plan reuse across correction calls.

A call whose volume shapes, shrink factors, mask kinds, parameters and tiling
equal the previous call's reuses that call's plan (the device layout and its
upload image) and only takes the new buffer addresses.  Its results must be
the bits a freshly built plan gives, and any change of shape or parameters
(or an intervening call with another plan) must build a new plan."""
import numpy as np
import pytest

import parity_cases as PC


def _vols(seed):
    return [PC.small_phantom(d, seed=seed + k)[:2] for k, d in enumerate([(30, 26, 22), (17, 33, 25), (24, 24, 24)])]


def _same(a, b):
    for x, y in zip(a, b):
        assert x["iterations"] == y["iterations"]
        for k in ("log_bias", "normalized", "lattice"):
            assert np.array_equal(x[k], y[k]), k


def _check(lib, **kw):
    a, b = _vols(300), _vols(400)  # same shapes, other data
    other = [PC.small_phantom((20, 22, 18), seed=5)[:2]]
    lib.correct_batch(a, **kw)
    reused = lib.correct_batch(b, **kw)  # same key: the plan of the call before
    lib.correct_batch(other, **kw)  # another plan in between
    fresh = lib.correct_batch(b, **kw)  # built anew
    _same(reused, fresh)
    # a parameter change builds a new plan whose results differ
    changed = lib.correct_batch(b, **dict(kw, max_iter=[2, 2, 2]))
    assert any(x["iterations"] != y["iterations"] for x, y in zip(changed, fresh))
    # and back: equal to the fresh results again
    _same(lib.correct_batch(b, **kw), fresh)


def test_emulated_plan_reuse(emu):
    _check(emu, levels=3, max_iter=[6, 6, 6], threshold=0.001)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_gpu_plan_reuse(gpu, precision):
    _check(gpu, levels=3, max_iter=[8, 8, 8], threshold=0.001, precision=precision)
