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


def _uneven_batch(seed0=40):
    vols = []
    for k, dims in enumerate([(30, 26, 22), (17, 33, 25), (41, 19, 28), (24, 24, 24), (33, 29, 18)]):
        im, m, _ = PC.small_phantom(dims, seed=seed0 + k)
        vols.append((im, m))
    return vols


def test_schedules_give_identical_results(emu):
    """Groups and chunked step enqueueing change only the launch layout:
    every such schedule returns the same fields bit for bit.  Row tiles also
    regroup the FP32 partial sums of the fit before their fixed-point
    accumulation, so a different tile size agrees to rounding only."""
    vols = _uneven_batch()
    kw = dict(levels=3, max_iter=[5, 5, 5], fixed_iterations=True)
    try:
        emu.set_tuning(reset=True)
        ref = emu.correct_batch(vols, **kw)
        for sched in [dict(groups=1, graph_steps="off"), dict(groups=2, graph_steps=3),
                      dict(groups=5, step_tile_voxels=256, graph_steps=1), dict(groups=3, graph_steps=64)]:
            emu.set_tuning(**sched)
            out = emu.correct_batch(vols, **kw)
            for a, b in zip(ref, out):
                assert a["iterations"] == b["iterations"], sched
                if "step_tile_voxels" in sched:
                    assert np.max(np.abs(a["log_bias"] - b["log_bias"])) < 1e-5, sched
                else:
                    assert np.array_equal(a["log_bias"], b["log_bias"]), sched
    finally:
        emu.set_tuning(reset=True)


def test_tuning_roundtrip_and_limits(emu):
    import n4amd
    try:
        emu.set_tuning(groups=2, step_tile_voxels=4096, graph_steps="off")
        t = emu.get_tuning().as_dict()
        assert (t["groups"], t["step_tile_voxels"], t["graph_steps"]) == (2, 4096, "off")
        for bad in [dict(groups=9), dict(step_tile_voxels=8), dict(graph_steps=65), dict(eval_kernel=3)]:
            with pytest.raises(n4amd.N4Error):
                emu.set_tuning(**bad)
        emu.set_tuning(reset=True)
        assert emu.get_tuning().as_dict()["groups"] == 0
    finally:
        emu.set_tuning(reset=True)


def test_autotune_picks_a_candidate(emu):
    """The search runs every candidate of its grid on the given volumes and
    leaves the fastest on the handle."""
    import n4amd
    vols = _uneven_batch()[:3]
    try:
        best = emu.autotune_batch(vols, levels=2, max_iter=[2, 2])
        d = best.as_dict()
        assert d["groups"] in (1, 2) and d["step_tile_voxels"] in (4096, 8192, 16384)
        assert d["graph_steps"] in ("off", 8, 16) and d["ms"] > 0
        kept = emu.get_tuning().as_dict()
        assert {k: kept[k] for k in ("groups", "step_tile_voxels", "graph_steps")} == \
            {k: d[k] for k in ("groups", "step_tile_voxels", "graph_steps")}
        with pytest.raises(n4amd.N4Error):
            emu.autotune_batch([], levels=2)
    finally:
        emu.set_tuning(reset=True)


def test_mixed_shrink_batch(emu):
    """One batch whose volumes carry their own shrink factors gives each
    volume the oracle's result at its own shrink."""
    import oracle as O
    dims, shr = [(30, 26, 22), (33, 29, 27), (36, 30, 28)], [2, 3, 4]
    vols = [PC.small_phantom(d, seed=90 + k)[:2] for k, d in enumerate(dims)]
    out = emu.correct_batch(vols, levels=2, max_iter=[4, 4], fixed_iterations=True, shrink=shr)
    for (im, m), s, r in zip(vols, shr, out):
        ro = O.n4(im, m, levels=2, max_iter=[4, 4], conv_threshold=0.0, shrink=s)
        mm = (np.asarray(m) != 0) & (im > 10)
        assert np.max(np.abs(r["log_bias"].astype(np.float64) - ro["log_bias"])[mm]) <= 1e-4


def test_prep_shrink_classes(emu):
    """Volumes of every prep shrink class (1, 2 and the generic kernel's 5 and
    6) in one batch over two stream groups: each volume gets the oracle's
    result at its own shrink (the emulated launch also checks that a prep
    launch never mixes classes)."""
    import oracle as O
    dims, shr = [(20, 18, 16), (41, 37, 33), (30, 26, 22), (22, 19, 17), (45, 40, 38)], [1, 5, 2, 1, 6]
    vols = [PC.small_phantom(d, seed=120 + k)[:2] for k, d in enumerate(dims)]
    emu.set_tuning(groups=2)
    try:
        out = emu.correct_batch(vols, levels=2, max_iter=[3, 3], fixed_iterations=True, shrink=shr)
    finally:
        emu.set_tuning(reset=True)
    for (im, m), s, r in zip(vols, shr, out):
        ro = O.n4(im, m, levels=2, max_iter=[3, 3], conv_threshold=0.0, shrink=s)
        mm = (np.asarray(m) != 0) & (im > 10)
        assert np.max(np.abs(r["log_bias"].astype(np.float64) - ro["log_bias"])[mm]) <= 1e-4, s
