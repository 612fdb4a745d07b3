"""Write golden input/output vectors from the reference's N4.py.

Runs ONLY in the build container, where the reference is mounted read-only
at /root/reference.  The fixtures it writes (n4py_golden.npz) are data:
inputs and the outputs N4.py produced for them.  Nothing under tests/ reads
the reference at test time.

Only the functions that no closed form, invariant or brute-force check pins
are run here (the two the parity plan leaves open):
  * sharpenImage (N4.py:9-118) as numpy 2.x executes it, including the
    complex64 numerator FFT of N4.py:75-84;
  * the whole N4() trajectory (N4.py:276-332): per-level iteration counts
    at early exit, the final lattice and the normalised image.

N4.py's module-level script (N4.py:335-346) loads image files the reference
does not ship and `import nibabel` (N4.py:5) names a package absent here and
unused by the functions, so only the function definitions (N4.py:1-332) are
executed, with an empty module registered under the nibabel name.
"""
from __future__ import annotations

import contextlib
import io
import os
import re
import sys
import types

import numpy as np

REF = "/root/reference/N4.py"
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "n4py_golden.npz")
sys.path.insert(0, os.path.join(HERE, "..", "..", "cuda-n4_amd"))


def load_reference_functions():
    with open(REF) as f:
        lines = f.read().splitlines()
    body = "\n".join(lines[:332])
    sys.modules.setdefault("nibabel", types.ModuleType("nibabel"))
    ns = {"__name__": "n4py_reference"}
    exec(compile(body, REF, "exec"), ns)
    return ns


def f32(a):
    """float64 array holding float32-representable values (what the GPU sees)."""
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def sharpen_cases(rng):
    cases = []
    # (a) smooth log-intensity field with an ellipsoid mask
    dims = (10, 12, 9)
    g = np.meshgrid(*[np.linspace(-1, 1, d) for d in dims], indexing="ij")
    rho = np.sqrt(g[0] ** 2 + g[1] ** 2 + g[2] ** 2)
    x = 4.6 + 0.8 * np.cos(2.5 * rho) + 0.05 * rng.standard_normal(dims)
    cases.append((f32(x), (rho <= 1.0).astype(np.float64)))
    # (b) three classes + noise, as log intensities of the C0 recipe
    dims = (16, 14, 12)
    lvl = rng.choice([np.log(100.0), np.log(200.0), np.log(300.0)], size=dims)
    x = lvl + 0.03 * rng.standard_normal(dims)
    m = (rng.random(dims) < 0.7).astype(np.float64)
    cases.append((f32(x), m))
    # (c) few voxels; duplicates at min and max
    x = f32(rng.uniform(2.0, 6.0, size=(5, 4, 3)))
    m = np.zeros_like(x)
    m.flat[[1, 5, 7, 11, 20, 33, 41, 50, 55]] = 1
    x.flat[5] = x.flat[1]
    x.flat[11] = x.flat[7]
    cases.append((x, m))
    return cases


def n4_cases():
    from n4amd import phantom as P
    cases = []
    # C0 recipe at 20x18x16 (the reference's CPU case, shrunk to fit pure-Python loops)
    im, m, _ = P.make_phantom((20, 18, 16), (8.5, 7.5, 6.5), 4, 0.2, 5.0, seed=11)
    cases.append(dict(im=im, mask=m, levels=3, iters=[6, 6, 6], thr=0.001))
    # Fixed-count case: threshold 0 never triggers, strong bias
    im, m, _ = P.make_phantom((16, 16, 14), (7.0, 6.5, 6.0), 4, 0.4, 5.0, seed=12)
    cases.append(dict(im=im, mask=m, levels=3, iters=[5, 5, 5], thr=0.0))
    # All-foreground mask with low-value background (threshold path)
    im, m, _ = P.make_phantom((14, 12, 12), (6.0, 5.0, 5.5), 4, 0.25, 5.0, seed=13)
    cases.append(dict(im=im, mask=np.ones_like(m), levels=2, iters=[5, 5], thr=0.001))
    # Tight threshold, longer levels
    im, m, _ = P.make_phantom((18, 16, 16), (8.0, 7.0, 7.0), 4, 0.3, 3.0, seed=14)
    cases.append(dict(im=im, mask=m, levels=2, iters=[12, 12], thr=1e-4))
    return cases


def main():
    ref = load_reference_functions()
    rng = np.random.default_rng(20240601)
    out = {}
    for i, (x, m) in enumerate(sharpen_cases(rng)):
        y = ref["sharpenImage"](x, m, numberOfHistogramBins=200, WienerFilterNoise=0.01,
                                BiasFieldFullWidthAtHalfMaximium=0.15)
        out[f"sharpen{i}_x"] = x
        out[f"sharpen{i}_mask"] = m
        out[f"sharpen{i}_out"] = np.asarray(y, dtype=np.float64)
        print(f"sharpen case {i}: {x.shape}, {int(m.sum())} masked")
    for i, c in enumerate(n4_cases()):
        buf = io.StringIO()
        with contextlib.redirect_stdout(buf):
            norm, lat = ref["N4"](c["im"].astype(np.float64), c["mask"].astype(np.float64),
                                  low_value=10, maximumNumberOfLevels=c["levels"],
                                  maximumNumberOfIterations=c["iters"],
                                  ConvergenceThreshold=c["thr"])
        its = [0] * c["levels"]
        for lev, _ in re.findall(r"Level (\d+), iter (\d+)", buf.getvalue()):
            its[int(lev)] += 1
        out[f"n4_{i}_im"] = c["im"]
        out[f"n4_{i}_mask"] = c["mask"].astype(np.float32)
        out[f"n4_{i}_levels"] = np.int32(c["levels"])
        out[f"n4_{i}_maxiter"] = np.asarray(c["iters"], np.int32)
        out[f"n4_{i}_thr"] = np.float64(c["thr"])
        out[f"n4_{i}_normalized"] = np.asarray(norm, np.float64)
        out[f"n4_{i}_lattice"] = np.asarray(lat, np.float64)
        out[f"n4_{i}_iterations"] = np.asarray(its, np.int32)
        print(f"N4 case {i}: {c['im'].shape}, iterations {its}")
    out["meta_numpy"] = np.array(np.__version__)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
