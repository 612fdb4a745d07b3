"""The persistent team path (one launch runs every volume's level/iteration
loop, a team of workgroups per volume synchronising at team barriers) on the
GPU against the step path.  Both run the same per-tile arithmetic on the same
tiles, so their results must be bitwise equal; here the team members are
resident on different CUs and XCDs at once, so the check covers the
agent-scope barriers and the cross-workgroup data flow that the host
emulation runs one thread per workgroup."""
import numpy as np
import pytest

import parity_cases as PC


def _run(lib, vols, path, **kw):
    try:
        lib.set_tuning(path=path, **kw.pop("tuning", {}))
        return lib.correct_batch(vols, **kw)
    finally:
        lib.set_tuning(reset=True)


def _same(a, b):
    for x, y in zip(a, b):
        assert x["iterations"] == y["iterations"]
        assert np.array_equal(x["log_bias"], y["log_bias"])
        assert np.array_equal(x["normalized"], y["normalized"])
        assert np.array_equal(x["lattice"], y["lattice"])


def _uneven(seed=20):
    return [PC.small_phantom(d, seed=seed + k)[:2]
            for k, d in enumerate([(60, 52, 44), (34, 66, 50), (82, 38, 56), (48, 48, 48)])]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp32", "bf16"])
@pytest.mark.parametrize("tuning", [{}, {"step_tile_voxels": 1024}, {"team_wgs_per_cu": 1}])
def test_gpu_teams_equal_steps(gpu, precision, tuning):
    vols = _uneven()
    kw = dict(levels=3, max_iter=[10, 10, 10], threshold=0.001, precision=precision)
    ref = _run(gpu, vols, "steps", tuning=dict(tuning, groups=1), **kw)
    _same(ref, _run(gpu, vols, "teams", tuning=dict(tuning), **kw))


@pytest.mark.gpu
def test_gpu_teams_mixed_shrink_fixed_iterations(gpu):
    vols = [PC.small_phantom(d, seed=70 + k)[:2] for k, d in enumerate([(80, 72, 64), (104, 96, 80), (72, 68, 60)])]
    kw = dict(levels=4, max_iter=[4, 4, 4, 4], ncpt=(4, 5, 6), fixed_iterations=True, shrink=[1, 2, 3])
    _same(_run(gpu, vols, "steps", **kw), _run(gpu, vols, "teams", **kw))
