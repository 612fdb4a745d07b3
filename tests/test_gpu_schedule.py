"""Launch schedules on the device: step graphs (captured chunks of
hist -> wiener -> fit -> eval steps replayed on each group stream), group
counts and the autotuner.  Schedules that only re-lay the launches give
bit-identical results, early exit included."""
import numpy as np
import pytest

import parity_cases as PC

pytestmark = pytest.mark.gpu


def _batch():
    vols = []
    for k, dims in enumerate([(30, 26, 22), (17, 33, 25), (41, 19, 28), (24, 24, 24), (33, 29, 18),
                              (48, 40, 36)]):
        im, m, _ = PC.small_phantom(dims, seed=60 + k)
        vols.append((im, m))
    return vols


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_graph_and_group_schedules_identical(gpu, precision):
    vols = _batch()
    kw = dict(levels=3, max_iter=[20, 20, 20], threshold=0.001, precision=precision)
    try:
        gpu.set_tuning(graph_steps="off", groups=1)
        ref = gpu.correct_batch(vols, **kw)
        for sched in [dict(graph_steps=1), dict(graph_steps=8, groups=2), dict(graph_steps=16, groups=6),
                      dict(graph_steps=64, groups=3)]:
            print(f"{precision}: schedule {sched}", flush=True)
            gpu.set_tuning(**sched)
            for rep in range(2):  # the second call replays the cached graphs
                out = gpu.correct_batch(vols, **kw)
                for a, b in zip(ref, out):
                    assert a["iterations"] == b["iterations"], (sched, rep)
                    assert np.array_equal(a["log_bias"], b["log_bias"]), (sched, rep)
                    assert np.array_equal(a["normalized"], b["normalized"]), (sched, rep)
    finally:
        gpu.set_tuning(reset=True)


def test_graph_recaptured_when_inputs_change(gpu):
    """A cached step graph is keyed on everything its launches take from the
    host; a different batch after it must not replay stale arguments."""
    vols = _batch()
    kw = dict(levels=2, max_iter=[6, 6], fixed_iterations=True)
    try:
        gpu.set_tuning(graph_steps="off")
        ref_a = gpu.correct_batch(vols[:3], **kw)
        ref_b = gpu.correct_batch(vols[2:], **kw)
        gpu.set_tuning(graph_steps=4)
        for _ in range(2):
            for sub, ref in ((vols[:3], ref_a), (vols[2:], ref_b)):
                out = gpu.correct_batch(sub, **kw)
                for a, b in zip(ref, out):
                    assert np.array_equal(a["log_bias"], b["log_bias"])
    finally:
        gpu.set_tuning(reset=True)


def test_graph_recaptured_when_iteration_limits_change(gpu):
    """The per-level iteration limits travel in the launches' arguments: a
    step graph captured for one set of limits must not replay for another
    (same volumes, pointers and sizes otherwise)."""
    vols = _batch()[:2]
    lims = ([3, 3], [5, 2])
    kw = dict(levels=2, fixed_iterations=True)
    try:
        gpu.set_tuning(graph_steps="off")
        ref = [gpu.correct_batch(vols, max_iter=m, **kw) for m in lims]
        gpu.set_tuning(graph_steps=4)
        for _ in range(2):
            for m, r in zip(lims, ref):
                out = gpu.correct_batch(vols, max_iter=m, **kw)
                for a, b in zip(r, out):
                    assert a["iterations"] == b["iterations"] == m
                    assert np.array_equal(a["log_bias"], b["log_bias"])
    finally:
        gpu.set_tuning(reset=True)


def test_single_volume_graph_parity(gpu):
    """The drop-in single-volume call with step graphs keeps the oracle
    parity (C0-size case with early exit)."""
    from n4amd import phantom as P
    im, m, _ = P.c0_phantom(0)
    try:
        gpu.set_tuning(graph_steps=8)
        print("fp32: C0 64^3 with 8-step graphs", flush=True)
        PC.check_n4_vs_oracle(gpu, im, m, 3, [50] * 3, 0.001)
    finally:
        gpu.set_tuning(reset=True)


def test_autotune_batch(gpu):
    vols = _batch()
    kw = dict(levels=2, max_iter=[5, 5], fixed_iterations=True)
    try:
        gpu.set_tuning(reset=True)
        ref = gpu.correct_batch(vols, **kw)
        best = gpu.autotune_batch(vols, **kw).as_dict()
        print(f"autotuned: {best}", flush=True)
        assert best["groups"] in (1, 2, 4) and best["step_tile_voxels"] in (4096, 8192, 16384)
        assert best["graph_steps"] in ("off", 8, 16) and best["ms"] > 0
        out = gpu.correct_batch(vols, **kw)
        for a, b in zip(ref, out):
            assert a["iterations"] == b["iterations"]
            assert np.max(np.abs(a["log_bias"] - b["log_bias"])) < 1e-5
    finally:
        gpu.set_tuning(reset=True)


def _group_split(xs, groups):
    """The library's volume grouping for volumes of dims (x, 16, 16), shrink
    1 (one row tile per slab, so a volume has x working slabs): largest
    first onto the least-loaded group.  Returns per group (first volume,
    volumes, first slab, slabs)."""
    order = sorted(range(len(xs)), key=lambda v: (-256 * xs[v], v))
    load, gid = [0] * groups, [0] * len(xs)
    for v in order:
        g = load.index(min(load))
        gid[v] = g
        load[g] += 256 * xs[v]
    out, v0, s0 = [], 0, 0
    for g in range(groups):
        mem = [v for v in range(len(xs)) if gid[v] == g]
        n, s = len(mem), sum(xs[v] for v in mem)
        out.append((v0, n, s0, s))
        v0, s0 = v0 + n, s0 + s
    return out


def _key_collision(groups=4, nvol=8, seed=11):
    """Two batches of equal size in which some group g >= 1 starts at the
    same slab with the same slab and volume counts but at a different first
    volume: a step graph keyed without the first volume would be replayed
    with the other batch's volume range."""
    rng = np.random.default_rng(seed)
    seen = {}
    for _ in range(20000):
        xs = [int(v) for v in rng.integers(6, 31, size=nvol)]
        for g, (v0, n, s0, s) in enumerate(_group_split(xs, groups)):
            if g == 0:
                continue
            k = (g, n, s0, s, sum(xs))
            if k in seen and seen[k][1] != v0:
                return seen[k][0], xs
            seen.setdefault(k, (xs, v0))
    return None


def test_graph_key_holds_group_first_volume(gpu):
    """Regression: batches whose group splits differ only in where a group's
    volumes start must not share a cached step graph."""
    pair = _key_collision()
    assert pair is not None
    batches = []
    for b, xs in enumerate(pair):
        vols = []
        for k, x in enumerate(xs):
            im, m, _ = PC.small_phantom((x, 16, 16), seed=300 + 10 * b + k)
            vols.append((im, m))
        batches.append(vols)
    kw = dict(levels=2, max_iter=[6, 6], fixed_iterations=True)
    try:
        gpu.set_tuning(graph_steps="off", groups=4)
        refs = [gpu.correct_batch(v, **kw) for v in batches]
        gpu.set_tuning(graph_steps=4, groups=4)
        for _ in range(2):
            for vols, ref in zip(batches, refs):
                out = gpu.correct_batch(vols, **kw)
                for a, b in zip(ref, out):
                    assert np.array_equal(a["log_bias"], b["log_bias"])
    finally:
        gpu.set_tuning(reset=True)
