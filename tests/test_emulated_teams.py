"""CPU checks of the persistent team path (one launch runs every volume's
level/iteration loop, a team of workgroups per volume synchronising at team
barriers) against the step path, in the host emulation build, where every
workgroup of the persistent launch runs on a host thread of its own.  Both
paths run the same per-tile arithmetic, so their results must be bitwise
equal; the oracle parity of the default (team) path is covered by the other
emulated tests."""
import numpy as np
import pytest

import n4amd
import parity_cases as PC


def _run(emu, vols, path, **kw):
    try:
        emu.set_tuning(path=path, **kw.pop("tuning", {}))
        return emu.correct_batch(vols, **kw)
    finally:
        emu.set_tuning(reset=True)


def _same(a, b):
    for x, y in zip(a, b):
        assert x["iterations"] == y["iterations"]
        assert np.array_equal(x["log_bias"], y["log_bias"])
        assert np.array_equal(x["normalized"], y["normalized"])
        assert np.array_equal(x["lattice"], y["lattice"])


def _uneven(seed=20):
    return [PC.small_phantom(d, seed=seed + k)[:2]
            for k, d in enumerate([(30, 26, 22), (17, 33, 25), (41, 19, 28), (24, 24, 24)])]


@pytest.mark.parametrize("tuning", [{}, {"step_tile_voxels": 512}, {"team_wgs_per_cu": 1},
                                    {"step_tile_voxels": 256, "team_wgs_per_cu": 2}])
def test_teams_equal_steps_early_exit(emu, tuning):
    vols = _uneven()
    kw = dict(levels=3, max_iter=[8, 8, 8], threshold=0.001)
    ref = _run(emu, vols, "steps", tuning=dict(tuning, groups=1), **kw)
    _same(ref, _run(emu, vols, "teams", tuning=dict(tuning), **kw))


def test_teams_equal_steps_mixed_shrink_and_lattice(emu):
    vols = [PC.small_phantom(d, seed=70 + k)[:2] for k, d in enumerate([(40, 36, 32), (52, 48, 40), (36, 34, 30)])]
    kw = dict(levels=4, max_iter=[3, 3, 3, 3], ncpt=(4, 5, 6), fixed_iterations=True, shrink=[1, 2, 3])
    tiles = {"step_tile_voxels": 512}  # same tiles on both paths: one convergence partial per tile
    _same(_run(emu, vols, "steps", tuning=dict(tiles), **kw), _run(emu, vols, "teams", tuning=dict(tiles), **kw))


def test_teams_degenerate_and_maskless_volumes(emu):
    im, m, _ = PC.small_phantom((20, 18, 16), seed=5)
    vols = [(im, m), (im, np.zeros_like(m)), (np.full(im.shape, 100.0, np.float32), m), (im, None)]
    kw = dict(levels=2, max_iter=[5, 5], threshold=0.001)
    out = _run(emu, vols, "teams", **kw)
    _same(_run(emu, vols, "steps", **kw), out)
    assert out[1]["n_mask_work"] == 0 and np.all(out[1]["log_bias"] == 0)
    assert np.abs(out[2]["log_bias"]).max() == 0.0


def test_teams_single_volume_oracle(emu, golden):
    """The drop-in single-volume call on the team path keeps the N4.py
    golden trajectory."""
    im, m = golden["n4_0_im"], golden["n4_0_mask"]
    L, its, thr = int(golden["n4_0_levels"]), list(golden["n4_0_maxiter"]), float(golden["n4_0_thr"])
    try:
        emu.set_tuning(path="teams", step_tile_voxels=512)
        PC.check_n4_vs_oracle(emu, im, m, L, its, thr)
    finally:
        emu.set_tuning(reset=True)


def test_teams_refused_where_not_eligible(emu):
    """A lattice whose tables do not fit one workgroup's LDS cannot run as a
    team: forced, the call fails with N4_ERR_UNSUPPORTED; by default it runs
    on the step kernels."""
    im, m, _ = PC.small_phantom((24, 22, 20), seed=5)
    kw = dict(levels=2, max_iter=[2, 2], ncpt=(40, 40, 40), fixed_iterations=True)
    try:
        emu.set_tuning(path="teams")
        with pytest.raises(n4amd.N4Error, match="persistent"):
            emu.correct(im, m, **kw)
    finally:
        emu.set_tuning(reset=True)
    emu.correct(im, m, **kw)

