"""CPU checks of the kernel logic: the kernel bodies of libn4amd.so compiled
for the host (libn4emu.so) against the FP64 oracle.  This is a test tool for
index arithmetic only; the product path is the GPU library (test_gpu_*)."""
import numpy as np
import pytest

import oracle as O
from n4amd import phantom as P
import parity_cases as PC


@pytest.mark.parametrize("case", [0, 1, 2])
def test_histogram_bin_indices_exact(emu, golden, case):
    PC.check_histogram_exact(emu, golden[f"sharpen{case}_x"], golden[f"sharpen{case}_mask"])


@pytest.mark.parametrize("case", [0, 1, 2])
def test_sharpen(emu, golden, case):
    PC.check_sharpen(emu, golden[f"sharpen{case}_x"], golden[f"sharpen{case}_mask"], golden[f"sharpen{case}_out"])


@pytest.mark.parametrize("dims,ncpt", [((13, 11, 9), (4, 4, 4)), ((13, 11, 9), (5, 6, 7)), ((9, 8, 7), (11, 11, 11))])
def test_fit_evaluate_upsample(emu, dims, ncpt):
    PC.check_fit_eval_upsample(emu, dims, ncpt)


def test_convergence_and_reduce(emu):
    PC.check_convergence_and_reduce(emu)


@pytest.mark.parametrize("case", [0, 1, 3])
def test_n4_golden(emu, golden, case):
    im, m = golden[f"n4_{case}_im"], golden[f"n4_{case}_mask"]
    L, its, thr = int(golden[f"n4_{case}_levels"]), list(golden[f"n4_{case}_maxiter"]), float(golden[f"n4_{case}_thr"])
    PC.check_n4_vs_oracle(emu, im, m, L, its, thr)


@pytest.mark.parametrize("shrink", [2, 3])
def test_n4_shrink_uneven(emu, shrink):
    # >= ~700 masked working voxels: below that the top-edge rule of the
    # histogram (N4.py:25-27) makes the oracle itself sensitive at 1e-4
    im, m, _ = PC.small_phantom((46, 38, 34), seed=9)
    PC.check_n4_vs_oracle(emu, im, m, 2, [6, 6], 0.0, shrink=shrink, fixed=True)


def test_batch_equals_single(emu):
    vols = [PC.small_phantom(d, seed=30 + k)[:2]
            for k, d in enumerate([(14, 12, 10), (9, 15, 11), (12, 12, 12), (10, 13, 9)])]
    rb = emu.correct_batch(vols, levels=2, max_iter=[4, 4])
    for (im, m), r in zip(vols, rb):
        s = emu.correct(im, m, levels=2, max_iter=[4, 4])
        assert np.array_equal(r["log_bias"], s["log_bias"])


def test_mfma_precisions_need_the_device_build(emu):
    """The host emulation compiles the FP32 kernel bodies only; the MFMA
    precision modes are refused there with N4_ERR_UNSUPPORTED."""
    import n4amd
    im, m, _ = PC.small_phantom()
    for prec in ("bf16", "split_bf16"):
        with pytest.raises(n4amd.N4Error, match="FP32 kernels only"):
            emu.correct(im, m, levels=2, max_iter=[2, 2], precision=prec)


def test_n4_row_tiles(emu, monkeypatch):
    """Step kernels over row tiles of the working slabs (16-row tiles on a
    46x38x34 grid: 3 tiles per slab, the last one partial): the same parity
    with the FP64 oracle as whole slabs."""
    monkeypatch.setenv("N4_STEP_TILE_VOX", "256")
    im, m, _ = PC.small_phantom((46, 38, 34), seed=9)
    PC.check_n4_vs_oracle(emu, im, m, 3, [8] * 3, 0.0, fixed=True)
