"""Run records of a correction: the convergence value of every iteration
(n4_last_cv_trace; the reference prints only "Level %d, iter %d",
cudaN4.cu:1530) and each fitting level's lattice at the level's end
(n4_last_level_lattice; the reference drops even the final one,
main.cpp:122).

Pins: the oracle's trace (N4.py:121-126 per iteration, the values that
drive N4.py:301-302), and for level l's lattice the final lattice of the
oracle run truncated after level l (the first l+1 levels of a correction do
not depend on the levels after them); the last level's record equals the
call's own lattice output bit for bit.
"""
import numpy as np
import pytest

import n4amd
import oracle as O
import parity_cases as PC

CV_TOL = 1e-6    # absolute, CV values ~1e-3: FP32 path vs FP64 oracle
LAT_TOL = 1e-4   # absolute on lattice coefficients (the log-bias bar)


def check_records(lib, precision="fp32", cv_tol=CV_TOL, lat_tol=LAT_TOL, dims=(28, 26, 24), shrink=1):
    im, m, _ = PC.small_phantom(dims, seed=12)
    kw = dict(levels=3, max_iter=[6, 6, 6], ncpt=(4, 5, 4))
    r = lib.correct(im, m, threshold=0.001, want_levels=True, precision=precision, shrink=shrink, **kw)
    ro = O.n4(im, m, conv_threshold=0.001, shrink=shrink, **kw)
    assert r["iterations"] == ro["iterations"]
    assert len(r["cv_trace"]) == sum(ro["iterations"])
    assert np.abs(r["cv_trace"] - ro["cv_trace"]).max() <= cv_tol
    # each level's last value is the report's last convergence value
    ends = np.cumsum(r["iterations"]) - 1
    assert np.array_equal(r["cv_trace"][ends], np.array(r["last_cv"]))
    for l in range(3):
        lat = r["level_lattices"][l]
        assert lat.shape == tuple((c - 3) * 2 ** l + 3 for c in (4, 5, 4))
        rl = O.n4(im, m, conv_threshold=0.001, levels=l + 1, max_iter=[6] * (l + 1), ncpt=(4, 5, 4), shrink=shrink)
        assert np.abs(lat - rl["lattice"]).max() <= lat_tol
    assert np.array_equal(r["level_lattices"][-1], r["lattice"])
    return r


def test_emulated_records_vs_oracle(emu):
    check_records(emu)


def test_emulated_records_teams_path(emu):
    try:
        emu.set_tuning(path="teams", step_tile_voxels=512)
        check_records(emu, shrink=2, dims=(40, 34, 30))
    finally:
        emu.set_tuning(reset=True)


def test_emulated_batch_records_per_volume(emu):
    vols = [PC.small_phantom(d, seed=60 + k)[:2] for k, d in enumerate([(30, 26, 22), (17, 33, 25), (24, 24, 24)])]
    kw = dict(levels=2, max_iter=[5, 5], threshold=0.001)
    res = emu.correct_batch(vols, **kw)
    for i, (im, m) in enumerate(vols):
        tr = emu.last_cv_trace(i)
        ro = O.n4(im, m, levels=2, max_iter=[5, 5], conv_threshold=0.001)
        assert res[i]["iterations"] == ro["iterations"]
        assert np.abs(tr - ro["cv_trace"]).max() <= CV_TOL
        assert np.array_equal(emu.last_level_lattice(i, 1, 4), res[i]["lattice"])


def test_emulated_records_gone_after_stage_call(emu):
    im, m, _ = PC.small_phantom((20, 18, 16), seed=4)
    emu.correct(im, m, levels=1, max_iter=2)
    assert len(emu.last_cv_trace(0)) == 2
    with pytest.raises(n4amd.N4Error, match="volume"):
        emu.last_cv_trace(1)
    with pytest.raises(n4amd.N4Error, match="level"):
        emu.last_level_lattice(0, 1, 4)
    emu.fit_bspline3d(np.zeros((8, 8, 8), np.float32), np.ones((8, 8, 8), bool), (4, 4, 4))
    with pytest.raises(n4amd.N4Error, match="no run record"):
        emu.last_cv_trace(0)


def test_emulated_fixed_iterations_trace_length(emu):
    im, m, _ = PC.small_phantom((20, 18, 16), seed=4)
    r = emu.correct(im, m, levels=2, max_iter=[3, 4], threshold=1.0, fixed_iterations=True)
    assert r["iterations"] == [3, 4] and len(r["cv_trace"]) == 7


@pytest.mark.gpu
@pytest.mark.parametrize("precision,cv_tol,lat_tol", [("fp32", CV_TOL, LAT_TOL), ("bf16", 2e-5, 1e-3)])
def test_gpu_records_vs_oracle(gpu, precision, cv_tol, lat_tol):
    check_records(gpu, precision=precision, cv_tol=cv_tol, lat_tol=lat_tol)


@pytest.mark.gpu
def test_gpu_records_teams_path(gpu):
    try:
        gpu.set_tuning(path="teams")
        check_records(gpu, shrink=2, dims=(40, 34, 30))
    finally:
        gpu.set_tuning(reset=True)
