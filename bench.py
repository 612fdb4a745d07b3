#!/usr/bin/env python3
"""N4 bias-field-correction throughput on one MI355X.

A step is one pass of the N4 correction over one batch of synthetic
volumes whose inputs are already resident in HBM: threshold/log/shrink,
the multi-level fitting loop (histogram sharpening with the Wiener step,
B-spline fit, field update, convergence) and the full-resolution field +
normalisation.  Default workload: 64 mixed-size head-class phantoms, each
axis drawn from U{128..288}, shrink 4, 4 levels x <= 50 iterations, one
launch per stage across all volumes.  The batched path is timed with FP32,
BF16-MFMA and split-BF16-MFMA B-spline contractions (headline: --precision),
each checked against the FP64 CPU oracle on a bounded sample of the volumes;
the plain single-volume FP32 path is timed on the same volumes.

Prints ONE JSON line (rank 0).  Progress lines go to stderr.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "cuda-n4_amd"))

METRIC = "volumes/s & Mvoxel·iter/s, batched MFMA vs plain FP32, log-field err≤1e-3"
PEAK_HBM_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_workload(name, mask_dtype=np.float32):
    from n4amd import phantom as P
    if name == "batch64":
        vols = P.generate_batch(P.batch_specs(64, 128, 288, seed=0), workers=min(16, os.cpu_count() or 1),
                                mask_dtype=mask_dtype)
        cfg = {"workload": "batch of 64 head-class phantoms, axes U{128..288} (seed 0), mask 30-60%, "
                           "bias lattice 4^3-6^3 N(0,0.3), Rician sigma U[2,8], shrink 4, "
                           "4 levels x <=50 iterations, threshold 0.001",
               "volumes": 64, "shrink": 4, "levels": 4, "max_iterations": 50}
        return vols, cfg, [4] * 64
    if name == "batch256":
        vols = P.generate_batch(P.batch_specs(256, 96, 320, seed=4), workers=min(16, os.cpu_count() or 1),
                                mask_dtype=mask_dtype)
        shrinks = P.batch_shrinks(256, seed=5)
        cfg = {"workload": "batch of 256 head-class phantoms, axes U{96..320} (seed 4), mask 30-60%, "
                           "bias lattice 4^3-6^3 N(0,0.3), Rician sigma U[2,8], shrink U{2,3,4} per volume "
                           "(seed 5), 4 levels x <=50 iterations, threshold 0.001",
               "volumes": 256, "shrink": "2-4 per volume", "levels": 4, "max_iterations": 50}
        return vols, cfg, shrinks
    if name == "c1":
        im, mask, _ = P.c1_phantom(1)
        cfg = {"workload": "256x256x176 head phantom (mask ~45%), bias lattice 5^3 N(0,0.3), shrink 4 "
                           "(64x64x44 working grid), 4 levels x <=50 iterations, threshold 0.001",
               "volumes": 1, "shrink": 4, "levels": 4, "max_iterations": 50}
        return [(im, mask)], cfg, [4]
    if name == "c2":
        im, mask, _ = P.c2_phantom(2)
        cfg = {"workload": "512x512x320 7T-like head phantom, bias lattice 6^3 N(0,0.4), shrink 2 "
                           "(256x256x160 working grid), 4 levels x <=50 iterations, threshold 0.001",
               "volumes": 1, "shrink": 2, "levels": 4, "max_iterations": 50}
        return [(im, mask)], cfg, [2]
    raise SystemExit(f"unknown workload {name}")


# Algorithmic HBM bytes per unit:
#   working-grid kernels, per working voxel and iteration executed:
#     hist 5 B (logU 4 + mask 1), fit 5 B (logU 4 + mask 1) plus 1 B (mask)
#     per level for the level's denominator, eval 9 B (mask 1 + logU read 4 +
#     logU write 4)
#   full-resolution kernels, per voxel:
#     prep 8 B (image 4 + float mask 4) + 9 B per working voxel written
#     recon 12 B (image 4 + normalised 4 + log bias 4)
def algorithmic_bytes(kernel, vols_meta, reports, levels):
    total = 0
    for meta, r in zip(vols_meta, reports):
        nw, nf = meta["n_work"], meta["n_full"]
        its = sum(int(r.iterations[l]) for l in range(levels))
        if kernel == "hist":
            total += 5 * nw * its
        elif kernel == "fit":
            total += 5 * nw * its + 1 * nw * levels
        elif kernel == "eval":
            total += 9 * nw * its
        elif kernel == "prep":
            total += 8 * nf + 9 * nw
        elif kernel == "recon":
            total += 12 * nf
    return total


def oracle_sample(vols, cfg, shrinks, cpu_seconds, nthreads):
    """FP64 oracle on the first volumes of the workload until cpu_seconds of
    wall time: the CPU baseline and the error reference."""
    import oracle as O
    out, t_cpu = [], 0.0
    for i, (im, mask) in enumerate(vols):
        t1 = time.perf_counter()
        ro = O.n4(im, mask, levels=cfg["levels"], max_iter=[cfg["max_iterations"]] * cfg["levels"],
                  conv_threshold=0.001, shrink=shrinks[i], nthreads=nthreads)
        dt = time.perf_counter() - t1
        t_cpu += dt
        out.append({"i": i, "log_bias": ro["log_bias"], "iterations": ro["iterations"]})
        log(f"oracle volume {i}: {dt:.2f}s, iterations {ro['iterations']}")
        if t_cpu >= cpu_seconds:
            break
    return out, t_cpu


def _oracle_one(job):
    import oracle as O
    im, mask, levels, iters, shrink = job
    O.n4(im, mask, levels=levels, max_iter=[iters] * levels, conv_threshold=0.001, shrink=shrink, nthreads=1)


def oracle_all_cores(vols, cfg, shrinks, n):
    """The same first n volumes through the oracle, one volume per process on
    all host cores (at most 16): wall seconds and the worker count."""
    from concurrent.futures import ProcessPoolExecutor
    workers = max(1, min(16, os.cpu_count() or 1, n))
    jobs = [(vols[i][0], vols[i][1], cfg["levels"], cfg["max_iterations"], shrinks[i]) for i in range(n)]
    t1 = time.perf_counter()
    with ProcessPoolExecutor(workers) as ex:
        list(ex.map(_oracle_one, jobs))
    return time.perf_counter() - t1, workers


def config_record(lib, name, precs, steps):
    """Compact record of one other workload: batched volumes/s per precision
    (library default schedule, 1 untimed + median of max(5, steps) timed
    runs), the plain single-volume FP32 rate (median of 5 passes) and the
    log-bias error against the FP64 oracle (16 threads) on a stated sample."""
    import n4amd
    import oracle as O
    t0 = time.time()
    vols, cfg, shrinks = make_workload(name, mask_dtype=np.uint8)
    vols = [(im, (np.asarray(m) != 0).astype(np.uint8)) for im, m in vols]
    log(f"[{name}] generated {len(vols)} volume(s) in {time.time() - t0:.1f}s")
    nvol = len(vols)
    if nvol == 1:
        sample = [0]
        sample_text = "the volume"
    else:  # the first volume of each shrink class, in index order
        sample = sorted({s: i for i, s in reversed(list(enumerate(shrinks)))}.values())
        sample_text = "first volume of each shrink class: " + ", ".join(f"#{i} (shrink {shrinks[i]})"
                                                                         for i in sample)
    levels, iters = cfg["levels"], cfg["max_iterations"]
    ref = {}
    for i in sample:
        im, m = vols[i]
        t1 = time.perf_counter()
        ref[i] = O.n4(im, m, levels=levels, max_iter=[iters] * levels, conv_threshold=0.001, shrink=shrinks[i],
                      nthreads=min(16, os.cpu_count() or 1))
        log(f"[{name}] oracle volume {i}: {time.perf_counter() - t1:.1f}s, iterations {ref[i]['iterations']}")
    kw = dict(levels=levels, max_iter=iters, threshold=0.001)
    nc = n4amd.final_ncpt(n4amd.default_params(**kw))
    arr = (n4amd.Volume * nvol)()
    dev = []
    for i, (im, m) in enumerate(vols):
        d = (lib.to_device(im), lib.to_device(m), lib.empty(im.nbytes), lib.empty(4 * nc[0] * nc[1] * nc[2]),
             lib.empty(im.nbytes) if i in ref else None)
        dev.append(d)
        v = arr[i]
        v.d_im, v.d_mask, v.mask_kind = d[0].ptr, d[1].ptr, n4amd.MASK_U8
        for k in range(3):
            v.dims[k] = im.shape[k]
        v.shrink_factor = shrinks[i]
        v.d_im_normalized, v.d_lattice, v.d_bias_field = d[2].ptr, d[3].ptr, None
        v.d_log_bias_field = d[4].ptr if d[4] is not None else None
    n_work = [int(np.prod([(s + k - 1) // k for s, k in zip(im.shape, [shrinks[i]] * 3)]))
              for i, (im, _) in enumerate(vols)]
    reps = (n4amd.Report * nvol)()
    out = {"workload": cfg["workload"], "volumes": nvol, "error_sample": sample_text, "batched": {}}
    lib.set_tuning(reset=True)
    for prec in precs:
        p = n4amd.default_params(precision=prec, **kw)
        lib._check(lib.L.n4_correct_batch(lib.h, arr, nvol, C.byref(p), reps, None))
        ts = []
        for _ in range(max(5, steps)):
            t1 = time.perf_counter()
            lib._check(lib.L.n4_correct_batch(lib.h, arr, nvol, C.byref(p), reps, None))
            ts.append(time.perf_counter() - t1)
        t = float(np.median(ts))
        its = [[int(r.iterations[l]) for l in range(levels)] for r in reps]
        errs = []
        for i in sample:
            im, m = vols[i]
            lb = dev[i][4].download(np.float32, im.shape).astype(np.float64)
            mm = (m != 0) & (im > 10)
            e = np.abs(lb - ref[i]["log_bias"])[mm]
            errs.append({"volume": i, "shrink": shrinks[i], "max": float(e.max()),
                         "rms": float(np.sqrt(np.mean(e ** 2))),
                         "iterations": its[i], "oracle_iterations": ref[i]["iterations"]})
        out["batched"][prec] = {
            "volumes_per_s": round(nvol / t, 3), "ms_median": round(1e3 * t, 4),
            "mvoxel_iter_per_s": round(sum(w * sum(i) for w, i in zip(n_work, its)) / t / 1e6, 2),
            "logbias_err_vs_fp64": {"max": max(e["max"] for e in errs), "rms_max": max(e["rms"] for e in errs),
                                    "iteration_counts_matched":
                                        f"{sum(e['iterations'] == e['oracle_iterations'] for e in errs)}/{len(errs)}",
                                    "per_volume": errs}}
        log(f"[{name}] {prec}: {nvol / t:.2f} volumes/s, err max {out['batched'][prec]['logbias_err_vs_fp64']['max']:.3g}")
    p32 = n4amd.default_params(precision="fp32", **kw)
    rep1 = n4amd.Report()

    def plain_pass():
        for i in range(nvol):
            v = arr[i]
            p32.shrink_factor = shrinks[i]
            lib._check(lib.L.n4_correct(lib.h, v.d_im, v.d_mask, v.mask_kind, v.dims[0], v.dims[1], v.dims[2],
                                        C.byref(p32), v.d_im_normalized, v.d_lattice, None, v.d_log_bias_field,
                                        C.byref(rep1), None))
    plain_pass()
    lib.synchronize()
    ts = []
    for _ in range(5):
        t1 = time.perf_counter()
        plain_pass()
        lib.synchronize()
        ts.append(time.perf_counter() - t1)
    t = float(np.median(ts))
    out["plain_fp32"] = {"volumes_per_s": round(nvol / t, 3), "protocol": "1 untimed pass, median of 5"}
    log(f"[{name}] plain FP32: {nvol / t:.2f} volumes/s")
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="batch64", choices=["batch64", "c1", "c2", "batch256"])
    ap.add_argument("--precision", default="bf16", choices=["fp32", "bf16", "split_bf16"],
                    help="precision of the headline batched figure")
    ap.add_argument("--precisions", default="fp32,bf16,split_bf16",
                    help="batched precisions timed (comma list; must include --precision)")
    ap.add_argument("--schedule", default="autotune", choices=["autotune", "default", "direct"],
                    help="launch schedule of the batched runs: the library's autotuner (per precision, before "
                         "the warmup), its defaults, or its defaults without step graphs")
    ap.add_argument("--configs", default="c1,c2,batch256",
                    help="other workloads reported as compact sub-records (comma list of c1, c2, batch256; "
                         "empty: none): volumes/s per precision, plain FP32 rate, log-bias error vs the oracle")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run device-to-device copies of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on exactly one GPU (one MI355X); multi-GPU runs are not part "
              "of it. Re-run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    precs = [q for q in args.precisions.split(",") if q]
    if args.precision not in precs:
        precs.append(args.precision)

    import n4amd
    if not args.copy_calibration_mib:
        # generate in worker processes before this process touches the GPU
        t0 = time.time()
        vols, cfg, shrinks = make_workload(args.workload)
        log(f"generated {len(vols)} volume(s) in {time.time() - t0:.1f}s")
        # one core for the batch (N4.py's execution model); all host cores (at
        # most 16) for the single large volume, which one core cannot finish
        cpu_threads = 1 if len(vols) > 1 else min(16, os.cpu_count() or 1)
        sample, t_cpu = (([], 0.0) if args.no_cpu_baseline
                         else oracle_sample(vols, cfg, shrinks, args.cpu_seconds, cpu_threads))
        # the batch's sample again with one volume per host core
        t_all, all_workers = (oracle_all_cores(vols, cfg, shrinks, len(sample))
                              if sample and len(vols) > 1 else (0.0, 0))
    lib = n4amd.load()
    log(f"library {lib.version()} ({lib.path})")
    if args.copy_calibration_mib:
        # counters calibrated on kernels of known traffic: a device copy
        # (nbytes read + nbytes written) and the library's read-only sum
        # kernel (k_flat: nbytes read, a few KB of partials written)
        nbytes = args.copy_calibration_mib << 20
        a, b = lib.empty(nbytes), lib.empty(nbytes)
        lib._check(lib.L.n4_memset(lib.h, a.ptr, 0, nbytes))
        for _ in range(3):
            lib._check(lib.L.n4_memcpy_dtod(lib.h, b.ptr, a.ptr, nbytes))
        s = C.c_double()
        for _ in range(3):
            lib._check(lib.L.n4_reduce_sum(lib.h, nbytes // 4, a.ptr, C.byref(s)))
        lib.synchronize()
        print(json.dumps({"copy_calibration_bytes": nbytes, "copies": 3, "read_only_sums": 3}), flush=True)
        return

    base_kw = dict(levels=cfg["levels"], max_iter=cfg["max_iterations"], threshold=0.001, shrink=shrinks[0])
    nc = n4amd.final_ncpt(n4amd.default_params(**base_kw))
    L = nc[0] * nc[1] * nc[2]
    # device-resident inputs and outputs
    dev, meta = [], []
    arr = (n4amd.Volume * len(vols))()
    for i, (im, mask) in enumerate(vols):
        d_im = lib.to_device(im)
        d_mask = lib.to_device(mask.astype(np.float32))
        d_out = lib.empty(im.nbytes)
        d_lb = lib.empty(im.nbytes)
        d_lat = lib.empty(4 * L)
        dev.append((d_im, d_mask, d_out, d_lb, d_lat))
        v = arr[i]
        v.d_im, v.d_mask, v.mask_kind = d_im.ptr, d_mask.ptr, n4amd.MASK_F32
        for d in range(3):
            v.dims[d] = im.shape[d]
        v.shrink_factor = shrinks[i]
        v.d_im_normalized, v.d_lattice, v.d_bias_field, v.d_log_bias_field = d_out.ptr, d_lat.ptr, None, d_lb.ptr
        W = [(s + shrinks[i] - 1) // shrinks[i] for s in im.shape]
        meta.append({"n_full": int(im.size), "n_work": int(np.prod(W)), "dims": im.shape})
    nvol = len(vols)
    reps = (n4amd.Report * nvol)()

    def batch_step(p):
        lib._check(lib.L.n4_correct_batch(lib.h, arr, nvol, C.byref(p), reps, None))

    def sample_errors(its):
        errs, matched = [], 0
        for smp in sample:
            i = smp["i"]
            im, mask = vols[i]
            lb = dev[i][3].download(np.float32, im.shape).astype(np.float64)
            mm = (mask != 0) & (im > 10)
            e = np.abs(lb - smp["log_bias"])[mm]
            errs.append((float(e.max()), float(np.sqrt(np.mean(e ** 2)))))
            matched += int(its[i] == smp["iterations"])
        if not errs:
            return None
        return {"sample_volumes": len(errs), "max": max(e[0] for e in errs), "rms_max": max(e[1] for e in errs),
                "iteration_counts_matched": f"{matched}/{len(errs)}"}

    runs = {}
    for prec in precs:
        p = n4amd.default_params(precision=prec, **base_kw)
        lib.set_tuning(reset=True)
        t_tune = 0.0
        if args.schedule == "autotune":
            best = n4amd.Tuning()
            t0 = time.perf_counter()
            lib._check(lib.L.n4_autotune_batch(lib.h, arr, nvol, C.byref(p), C.byref(best), None))
            t_tune = time.perf_counter() - t0
            log(f"[{prec}] autotuned schedule {best.as_dict()} in {t_tune:.1f}s")
        elif args.schedule == "direct":
            lib.set_tuning(graph_steps="off")
        schedule = lib.get_tuning().as_dict()
        schedule.pop("ms", None)
        # untimed warmup and the timed region, with the batch's schedule and
        # no events in the streams
        lib.profile_enable([])
        for _ in range(args.warmup):
            batch_step(p)
        lib.synchronize()
        step_s = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            t1 = time.perf_counter()
            batch_step(p)  # synchronous: returns after the batch completes
            step_s.append(time.perf_counter() - t1)
        lib.synchronize()
        dt = time.perf_counter() - t0
        # calls that plan from scratch: a call with another threshold in
        # between replaces the handle's cached plan, so each timed call
        # builds and uploads its plan again (the first call's cost)
        p_other = n4amd.default_params(precision=prec, **dict(base_kw, threshold=0.002))
        cold_s = []
        for _ in range(3):
            batch_step(p_other)
            t1 = time.perf_counter()
            batch_step(p)
            cold_s.append(time.perf_counter() - t1)
        # kernel durations, measured apart from the timed region with one
        # stream group, direct launches and the default tiles (no concurrent
        # groups sharing the CUs, no graphs; the same launches whatever the
        # schedule of the timed region): one step with events around every launch picks
        # the dominant kernel, then the same number of steps as timed with
        # events around each of its launches give its average duration
        lib.set_tuning(groups=1, graph_steps="off")  # library defaults otherwise
        lib.profile_enable(n4amd.LAUNCH_KERNELS)
        batch_step(p)
        lib.synchronize()
        breakdown = {k: lib.profile_read(k) for k in n4amd.LAUNCH_KERNELS}
        dominant = max(breakdown, key=lambda k: breakdown[k][0])
        log(f"[{prec}] one-group per-kernel device time (ms, launches): "
            + ", ".join(f"{k}={v[0]:.3f}/{v[1]}" for k, v in breakdown.items()))
        lib.profile_enable([dominant])
        for _ in range(args.steps):
            batch_step(p)
        lib.synchronize()
        dom_ms, dom_n = lib.profile_read(dominant)
        lib.profile_enable([])
        lib.set_tuning(**schedule)
        its = [[int(r.iterations[l]) for l in range(cfg["levels"])] for r in reps]
        vox_iter = sum(m["n_work"] * sum(i) for m, i in zip(meta, its))
        dom_bytes = algorithmic_bytes(dominant, meta, reps, cfg["levels"]) * args.steps
        achieved = dom_bytes / (dom_ms * 1e-3) / 1e9 if dom_ms > 0 else 0.0
        R = {"volumes_per_s": nvol * args.steps / dt, "ms_per_step": 1e3 * dt / args.steps,
             "mvoxel_iter_per_s": vox_iter * args.steps / dt / 1e6,
             "iterations_per_level_mean": [round(float(np.mean([i[l] for i in its])), 2)
                                           for l in range(cfg["levels"])],
             "dominant": dominant, "dom_ms": dom_ms, "dom_n": dom_n, "dom_bytes": dom_bytes, "achieved": achieved,
             "breakdown": {k: round(v[0], 4) for k, v in breakdown.items()},
             "err": sample_errors(its), "schedule": schedule, "autotune_s": round(t_tune, 2),
             "step_ms_median": 1e3 * float(np.median(step_s)),
             "new_plan_ms_median": 1e3 * float(np.median(cold_s))}
        runs[prec] = R
        log(f"[{prec}] batched: {R['ms_per_step']:.3f} ms/step, {R['volumes_per_s']:.1f} volumes/s; dominant "
            f"{dominant}: {dom_ms / max(1, dom_n) * 1e3:.2f} us avg over {dom_n} launches, {achieved:.1f} GB/s; "
            f"a call building its plan {R['new_plan_ms_median']:.3f} ms; err {R['err']}")

    H = runs[args.precision]
    value = H["volumes_per_s"]
    prec_text = {"fp32": "FP32 B-spline fit/evaluation/reconstruction",
                 "bf16": "MFMA B-spline fit and field-increment contractions, BF16 voxel operands x hi+lo "
                         "BF16 basis factors, FP32 accumulation; FP32 final reconstruction",
                 "split_bf16": "MFMA B-spline fit and field-increment contractions, hi+lo BF16 voxel operands "
                               "and basis factors, FP32 accumulation; FP32 final reconstruction"}
    result = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "volumes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(H["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32": "fp32", "bf16": "bf16", "split_bf16": "bf16x2"}[args.precision],
        "data": "synthetic phantoms (numpy default_rng seeds recorded in the generator), inputs resident in HBM",
        "config": dict(cfg, precision=args.precision),
        "mvoxel_iter_per_s": round(H["mvoxel_iter_per_s"], 2),
        "iterations_per_level_mean": H["iterations_per_level_mean"],
        "precision": prec_text[args.precision] + "; FP64 histogram range, Wiener step, lattice update and CV sums",
        "roofline_note": "timed steps run without events; the dominant kernel's average duration comes from the "
                         "same number of steps run again with one stream group, direct launches and the default "
                         "tiles (independent of the timed region's schedule), with events around each of its "
                         "launches",
        "plan_note": "every timed step has the shapes and parameters of the step before, so it reuses the "
                     "handle's cached plan (device layout and upload image) and uploads only descriptors and "
                     "zeroed state; batched.*.ms_per_call_new_plan_median times calls that build their plan",
        "schedule": H["schedule"],
        "schedule_source": {"autotune": "n4_autotune_batch on this workload (groups x tile voxels x graph steps), "
                                        "run before the warmup",
                            "default": "library defaults", "direct": "library defaults, no step graphs"}[
                                            args.schedule],
    }
    dominant, dom_n = H["dominant"], H["dom_n"]
    # HBM traffic per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE of the same
    # kernel-timing dispatches, scaled by a copy of known size) and the
    # profiler's own average duration of those dispatches, from the committed
    # profile of this command
    traffic, traffic_src, prof_us = None, None, None
    ppath = os.path.join(REPO, "profiles", f"{args.workload}_{args.precision}_profile.json")
    if os.path.exists(ppath):
        with open(ppath) as f:
            pj = json.load(f)
        tp = pj.get("kernel_timing_pass", {}).get(dominant)
        if tp:
            if tp.get("fetch_bytes") is not None:
                traffic = round(tp["fetch_bytes"] + (tp.get("write_bytes") or 0.0))
            prof_us = tp.get("rocprof_avg_us")
            traffic_src = os.path.relpath(ppath, REPO)
    avg_us = H["dom_ms"] / max(1, dom_n) * 1e3
    bytes_per_launch = H["dom_bytes"] / max(1, dom_n)
    result["roofline"] = {
        "bound": "hbm", "kernel": dominant, "achieved": round(H["achieved"], 2), "peak": PEAK_HBM_GBS,
        "unit": "GB/s", "frac": round(H["achieved"] / PEAK_HBM_GBS, 5), "traffic": traffic,
        "traffic_source": traffic_src,
        "algorithmic_bytes_per_launch": round(bytes_per_launch),
        "avg_launch_us": round(avg_us, 3), "launches": int(dom_n),
    }
    if prof_us:
        # HIP event pairs around a launch also hold the event markers and
        # the dispatch; the profiler times the dispatch's execution alone
        result["roofline"]["rocprof_avg_launch_us"] = round(prof_us, 3)
        result["roofline"]["achieved_at_rocprof_duration"] = round(bytes_per_launch / (prof_us * 1e-6) / 1e9, 2)
    result["kernel_ms_per_step_one_group"] = H["breakdown"]
    result["batched"] = {q: {"volumes_per_s": round(R["volumes_per_s"], 3), "ms_per_step": round(R["ms_per_step"], 4),
                             "ms_per_step_median": round(R["step_ms_median"], 4),
                             "ms_per_call_new_plan_median": round(R["new_plan_ms_median"], 4),
                             "mvoxel_iter_per_s": round(R["mvoxel_iter_per_s"], 2),
                             "logbias_err_vs_fp64": R["err"], "kernel_ms_per_step": R["breakdown"],
                             "schedule": R["schedule"], "autotune_s": R["autotune_s"]}
                         for q, R in runs.items()}
    if H["err"] is not None:
        result["logbias_err_vs_fp64"] = H["err"]
    p_head = n4amd.default_params(precision=args.precision, **base_kw)
    lib.set_tuning(**{k: v for k, v in H["schedule"].items()})

    # host-buffer rate: pageable host image + mask up, normalised + log bias down
    lib.profile_enable([])
    lib.synchronize()
    t0 = time.perf_counter()
    for (im, mask), d in zip(vols, dev):
        d[0].upload(im)
        d[1].upload(mask.astype(np.float32))
    batch_step(p_head)
    host_out = [(np.empty_like(im), np.empty_like(im)) for im, _ in vols]
    for (a, b), d in zip(host_out, dev):
        lib._check(lib.L.n4_memcpy_dtoh(lib.h, a.ctypes.data, d[2].ptr, a.nbytes))
        lib._check(lib.L.n4_memcpy_dtoh(lib.h, b.ctypes.data, d[3].ptr, b.nbytes))
    lib.synchronize()
    hdt = time.perf_counter() - t0
    del host_out
    result["host_buffers_inclusive"] = {"volumes_per_s": round(nvol / hdt, 3),
                                        "note": "one step including PCIe copies of image, mask, normalised image "
                                                "and log bias from/to pageable host memory"}
    log(f"host-buffer inclusive: {nvol / hdt:.1f} volumes/s")

    # plain single-volume FP32 path over the same volumes
    if not args.no_plain:
        p32 = n4amd.default_params(precision="fp32", **base_kw)
        rep1 = n4amd.Report()
        lib.set_tuning(reset=True)  # the drop-in as shipped: default schedule

        def plain_pass():
            for i in range(nvol):
                v = arr[i]
                p32.shrink_factor = shrinks[i]
                lib._check(lib.L.n4_correct(lib.h, v.d_im, v.d_mask, v.mask_kind, v.dims[0], v.dims[1], v.dims[2],
                                            C.byref(p32), v.d_im_normalized, v.d_lattice, None,
                                            v.d_log_bias_field, C.byref(rep1), None))
        plain_pass()
        lib.synchronize()
        passes = []
        for _ in range(max(5, args.steps)):
            t0 = time.perf_counter()
            plain_pass()
            lib.synchronize()
            passes.append(time.perf_counter() - t0)
        pdt = float(np.median(passes))
        result["plain_fp32"] = {"volumes_per_s": round(nvol / pdt, 3), "ms_per_volume": round(1e3 * pdt / nvol, 4),
                                "batched_over_plain": round(value / (nvol / pdt), 3),
                                "protocol": f"1 untimed pass, median of {len(passes)} timed passes over the "
                                            f"{nvol} volume(s)",
                                "pass_ms": [round(1e3 * t, 3) for t in passes]}
        log(f"plain single-volume path: {nvol / pdt:.1f} volumes/s (median of {len(passes)})")

    if args.configs:
        del dev, arr
        result["configs"] = {}
        for name in [c for c in args.configs.split(",") if c]:
            result["configs"][name] = config_record(lib, name, precs, args.steps)
            lib.set_tuning(reset=True)

    if sample:
        result["cpu_baseline"] = {
            "value": round(len(sample) / t_cpu, 4), "unit": "volumes/s", "cores": cpu_threads, "kind": "port",
            "sample": f"first {len(sample)} volume(s) of the workload, FP64 C restatement of N4.py (oracle/), "
                      f"{cpu_threads} thread(s), same levels/iterations/threshold/shrink",
            "host": {"cpu_model": cpu_model(), "nproc": os.cpu_count()},
        }
        if t_all > 0:
            result["cpu_baseline"]["all_cores"] = {
                "value": round(len(sample) / t_all, 4), "unit": "volumes/s", "cores": all_workers,
                "sample": "the same volumes, one per process (one thread each) on all host cores (at most 16)"}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
