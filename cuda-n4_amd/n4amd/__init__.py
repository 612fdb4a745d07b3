"""Python host binding of the MI355X N4 library (libn4amd.so) over its C ABI.

The compute path is the HIP library; this module only moves arrays and calls
the entry points declared in include/n4_capi.h.  There is no Python or CPU
fallback: `load()` raises when the built library or a GPU is missing.

`load(emulate=True)` loads libn4emu.so instead: the same kernel bodies
compiled for the host.  It exists only so that CPU unit tests can check the
kernels' index arithmetic without a GPU; nothing in the product uses it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(_HERE)
BUILD_DIR = os.path.join(PKG_ROOT, "_build")
LIB_PATH = os.path.join(BUILD_DIR, "libn4amd.so")
EMU_PATH = os.path.join(BUILD_DIR, "libn4emu.so")

MAX_LEVELS = 16
MASK_F32, MASK_U8, MASK_NONE = 0, 1, 2
PREC_FP32, PREC_BF16, PREC_SPLIT_BF16 = 0, 1, 2
PRECISIONS = {"fp32": PREC_FP32, "bf16": PREC_BF16, "split_bf16": PREC_SPLIT_BF16}
FLAG_FIXED_ITERATIONS = 0x1
FLAG_WEIGHTED_MASK = 0x2

# Names every build of the library must export (include/n4_capi.h).
EXPORTS = [
    "n4_version", "n4_status_string", "n4_last_error", "n4_params_default", "n4_final_ncpt",
    "n4_create", "n4_destroy", "n4_correct", "n4_correct_noalloc", "n4_correct_host", "n4_correct_batch",
    "n4_histogram", "n4_sharpen_image", "n4_fit_bspline3d", "n4_evaluate_bspline3d",
    "n4_convergence_measurement", "n4_upsample_lattice3d", "n4_reduce_sum", "n4_reduce_max",
    "n4_reduce_min", "n4_device_count", "n4_malloc", "n4_free", "n4_memcpy_htod",
    "n4_memcpy_dtoh", "n4_memcpy_dtod", "n4_memset", "n4_synchronize", "n4_profile_enable",
    "n4_profile_read", "n4_get_tuning", "n4_set_tuning", "n4_autotune_batch", "n4_last_cv_trace",
    "n4_last_level_lattice", "n4_mask_to_u8",
]
KERNELS = ["prep", "level", "hist", "fit_den", "fit", "eval", "recon", "wiener", "lattice", "update", "teams",
           "t_wait", "t_reduce", "t_hist", "t_wiener", "t_fit", "t_update", "t_eval", "t_trans"]
# launch kernels (the t_* entries are phases timed inside the persistent launch)
LAUNCH_KERNELS = KERNELS[:11]


class Params(C.Structure):
    _fields_ = [
        ("number_of_histogram_bins", C.c_uint32),
        ("wiener_filter_noise", C.c_double),
        ("bias_field_fwhm", C.c_double),
        ("ncpt", C.c_uint32 * 3),
        ("number_of_fitting_levels", C.c_uint32),
        ("max_iterations", C.c_uint32 * MAX_LEVELS),
        ("convergence_threshold", C.c_double),
        ("low_value", C.c_double),
        ("shrink_factor", C.c_uint32),
        ("fit_eps", C.c_double),
        ("precision", C.c_uint32),
        ("flags", C.c_uint32),
    ]


class Report(C.Structure):
    _fields_ = [
        ("iterations", C.c_uint32 * MAX_LEVELS),
        ("last_convergence", C.c_double * MAX_LEVELS),
        ("n_mask_full", C.c_uint64),
        ("n_mask_work", C.c_uint64),
        ("work_dims", C.c_uint32 * 3),
        ("final_ncpt", C.c_uint32 * 3),
        ("elapsed_ms", C.c_float),
    ]


class Volume(C.Structure):
    _fields_ = [
        ("d_im", C.c_void_p),
        ("d_mask", C.c_void_p),
        ("mask_kind", C.c_uint32),
        ("dims", C.c_uint32 * 3),
        ("shrink_factor", C.c_uint32),
        ("d_im_normalized", C.c_void_p),
        ("d_lattice", C.c_void_p),
        ("d_bias_field", C.c_void_p),
        ("d_log_bias_field", C.c_void_p),
    ]


GRAPH_OFF = 0xFFFFFFFF
EVAL_KERNELS = {"fp32": 1, "mfma": 2}
EVAL_KERNELS_INV = {v: k for k, v in EVAL_KERNELS.items()}
PATHS = {"auto": 0, "steps": 1, "teams": 2}
PATHS_INV = {v: k for k, v in PATHS.items()}


class Tuning(C.Structure):
    """Launch schedule (include/n4_capi.h n4_tuning); 0 fields = defaults."""
    _fields_ = [
        ("groups", C.c_uint32),
        ("step_tile_voxels", C.c_uint32),
        ("graph_steps", C.c_uint32),
        ("eval_kernel", C.c_uint32),
        ("path", C.c_uint32),
        ("team_wgs_per_cu", C.c_uint32),
        ("ms", C.c_float),
    ]

    def as_dict(self):
        return {"groups": int(self.groups), "step_tile_voxels": int(self.step_tile_voxels),
                "graph_steps": "off" if self.graph_steps == GRAPH_OFF else int(self.graph_steps),
                "eval_kernel": EVAL_KERNELS_INV.get(int(self.eval_kernel), 0),
                "path": PATHS_INV.get(int(self.path), 0), "team_wgs_per_cu": int(self.team_wgs_per_cu),
                "ms": float(self.ms)}


class N4Error(RuntimeError):
    pass


def _declare(L):
    vp, u32, u64, i32, dbl = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int32, C.c_double
    sig = {
        "n4_version": ([], C.c_char_p),
        "n4_status_string": ([C.c_int], C.c_char_p),
        "n4_last_error": ([], C.c_char_p),
        "n4_params_default": ([C.POINTER(Params)], C.c_int),
        "n4_final_ncpt": ([C.POINTER(Params), C.POINTER(u32)], C.c_int),
        "n4_create": ([C.POINTER(vp), C.c_int], C.c_int),
        "n4_destroy": ([vp], C.c_int),
        "n4_correct": ([vp, vp, vp, u32, u32, u32, u32, C.POINTER(Params), vp, vp, vp, vp,
                        C.POINTER(Report), vp], C.c_int),
        "n4_correct_noalloc": ([vp, vp, vp, u32, u32, u32, u32, C.POINTER(Params), vp, vp, vp, vp,
                                C.POINTER(Report), vp], C.c_int),
        "n4_correct_host": ([vp, vp, vp, u32, u32, u32, C.POINTER(Params), vp, vp, vp,
                             C.POINTER(Report)], C.c_int),
        "n4_correct_batch": ([vp, C.POINTER(Volume), u32, C.POINTER(Params), C.POINTER(Report), vp], C.c_int),
        "n4_histogram": ([vp, vp, vp, u64, u32, C.POINTER(dbl), vp, vp], C.c_int),
        "n4_sharpen_image": ([vp, vp, vp, u64, u32, dbl, dbl, vp, vp], C.c_int),
        "n4_fit_bspline3d": ([vp, vp, vp, C.POINTER(u32), C.POINTER(u32), dbl, vp], C.c_int),
        "n4_evaluate_bspline3d": ([vp, vp, C.POINTER(u32), C.POINTER(u32), vp], C.c_int),
        "n4_convergence_measurement": ([vp, vp, vp, vp, u64, C.POINTER(dbl)], C.c_int),
        "n4_upsample_lattice3d": ([vp, vp, C.POINTER(u32), vp], C.c_int),
        "n4_reduce_sum": ([vp, u64, vp, C.POINTER(dbl)], C.c_int),
        "n4_reduce_max": ([vp, u64, vp, C.POINTER(C.c_float)], C.c_int),
        "n4_reduce_min": ([vp, u64, vp, C.POINTER(C.c_float)], C.c_int),
        "n4_device_count": ([C.POINTER(C.c_int)], C.c_int),
        "n4_malloc": ([vp, C.POINTER(vp), C.c_size_t], C.c_int),
        "n4_free": ([vp, vp], C.c_int),
        "n4_memcpy_htod": ([vp, vp, vp, C.c_size_t], C.c_int),
        "n4_memcpy_dtoh": ([vp, vp, vp, C.c_size_t], C.c_int),
        "n4_memcpy_dtod": ([vp, vp, vp, C.c_size_t], C.c_int),
        "n4_memset": ([vp, vp, C.c_int, C.c_size_t], C.c_int),
        "n4_synchronize": ([vp], C.c_int),
        "n4_profile_enable": ([vp, u32], C.c_int),
        "n4_profile_read": ([vp, u32, C.POINTER(dbl), C.POINTER(u64)], C.c_int),
        "n4_get_tuning": ([vp, C.POINTER(Tuning)], C.c_int),
        "n4_set_tuning": ([vp, C.POINTER(Tuning)], C.c_int),
        "n4_autotune_batch": ([vp, C.POINTER(Volume), u32, C.POINTER(Params), C.POINTER(Tuning), vp], C.c_int),
        "n4_last_cv_trace": ([vp, u32, vp, u32, C.POINTER(u32)], C.c_int),
        "n4_last_level_lattice": ([vp, u32, u32, vp, u64], C.c_int),
        "n4_mask_to_u8": ([vp, vp, u64, vp], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    return L


_loaded: dict[str, "Library"] = {}


def load(emulate: bool = False) -> "Library":
    """Load the product library (or, for CPU unit tests only, the emulation
    build).  Raises N4Error when it is missing or no GPU is visible."""
    path = EMU_PATH if emulate else LIB_PATH
    if path in _loaded:
        return _loaded[path]
    if not os.path.exists(path):
        raise N4Error(f"{path} is not built: run __graft_entry__.build() (make -C cuda-n4_amd)")
    lib = Library(_declare(C.CDLL(path)), emulate)
    if not emulate:
        n = C.c_int(0)
        st = lib.L.n4_device_count(C.byref(n))
        if st != 0 or n.value < 1:
            raise N4Error("libn4amd.so loaded but no HIP device is visible; this library has no CPU path")
    _loaded[path] = lib
    return lib


def default_params(**kw) -> Params:
    """n4_params with the reference defaults; keyword overrides:
    nbins, noise, fwhm, ncpt (int or 3-tuple), levels, max_iter (int or list),
    threshold, low_value, shrink, eps, precision ("fp32", "bf16",
    "split_bf16" or the n4_precision value), fixed_iterations,
    weighted_mask (float masks are confidence weights)."""
    p = Params()
    p.number_of_histogram_bins = 200
    p.wiener_filter_noise = 0.01
    p.bias_field_fwhm = 0.15
    for d in range(3):
        p.ncpt[d] = 4
    p.number_of_fitting_levels = 4
    for i in range(MAX_LEVELS):
        p.max_iterations[i] = 50
    p.convergence_threshold = 0.001
    p.low_value = 10.0
    p.shrink_factor = 1
    p.fit_eps = 1e-3
    p.precision = PREC_FP32
    p.flags = 0
    alias = {"nbins": "number_of_histogram_bins", "noise": "wiener_filter_noise",
             "fwhm": "bias_field_fwhm", "levels": "number_of_fitting_levels",
             "threshold": "convergence_threshold", "shrink": "shrink_factor", "eps": "fit_eps"}
    for k, v in kw.items():
        if k == "ncpt":
            v = (v, v, v) if np.isscalar(v) else v
            for d in range(3):
                p.ncpt[d] = int(v[d])
        elif k == "max_iter":
            v = [v] * MAX_LEVELS if np.isscalar(v) else list(v)
            for i in range(MAX_LEVELS):
                p.max_iterations[i] = int(v[i]) if i < len(v) else int(v[-1])
        elif k == "precision":
            p.precision = PRECISIONS[v] if isinstance(v, str) else int(v)
        elif k in ("fixed_iterations", "weighted_mask"):
            f = FLAG_FIXED_ITERATIONS if k == "fixed_iterations" else FLAG_WEIGHTED_MASK
            p.flags = (p.flags | f) if v else (p.flags & ~f)
        else:
            name = alias.get(k, k)
            if name not in {f[0] for f in Params._fields_}:
                raise TypeError(f"unknown parameter {k!r}")
            setattr(p, name, v)
    return p


def final_ncpt(p: Params):
    return tuple(int((p.ncpt[d] - 3) * 2 ** (p.number_of_fitting_levels - 1) + 3) for d in range(3))


class DeviceArray:
    """A device allocation owned by a Library handle."""

    def __init__(self, lib: "Library", nbytes: int):
        self.lib = lib
        self.nbytes = int(nbytes)
        ptr = C.c_void_p()
        lib._check(lib.L.n4_malloc(lib.h, C.byref(ptr), C.c_size_t(self.nbytes)))
        self.ptr = ptr.value

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        self.lib._check(self.lib.L.n4_memcpy_htod(self.lib.h, self.ptr, a.ctypes.data, a.nbytes))
        return self

    def download(self, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        self.lib._check(self.lib.L.n4_memcpy_dtoh(self.lib.h, out.ctypes.data, self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            self.lib.L.n4_free(self.lib.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Library:
    def __init__(self, L, emulate: bool):
        self.L = L
        self.emulate = emulate
        h = C.c_void_p()
        st = L.n4_create(C.byref(h), 0)
        if st != 0:
            raise N4Error(f"n4_create failed: {L.n4_last_error().decode()}")
        self.h = h.value

    @property
    def path(self):
        return EMU_PATH if self.emulate else LIB_PATH

    def _check(self, st):
        if st != 0:
            raise N4Error(f"{self.L.n4_status_string(st).decode()}: {self.L.n4_last_error().decode()}")

    def version(self) -> str:
        return self.L.n4_version().decode()

    def to_device(self, a: np.ndarray) -> DeviceArray:
        a = np.ascontiguousarray(a)
        return DeviceArray(self, max(a.nbytes, 1)).upload(a)

    def empty(self, nbytes: int) -> DeviceArray:
        return DeviceArray(self, max(int(nbytes), 1))

    def synchronize(self):
        self._check(self.L.n4_synchronize(self.h))

    def profile_enable(self, kernels=()):
        mask = 0
        for k in kernels:
            mask |= 1 << KERNELS.index(k)
        self._check(self.L.n4_profile_enable(self.h, mask))

    def profile_read(self, kernel: str):
        ms, n = C.c_double(), C.c_uint64()
        self._check(self.L.n4_profile_read(self.h, KERNELS.index(kernel), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    # ---------------------------------------------------- schedule ----
    def get_tuning(self) -> Tuning:
        t = Tuning()
        self._check(self.L.n4_get_tuning(self.h, C.byref(t)))
        return t

    def set_tuning(self, groups=0, step_tile_voxels=0, graph_steps=0, eval_kernel=0, path=0, team_wgs_per_cu=0,
                   reset=False, ms=None):
        """graph_steps: 0 default, "off" (or GRAPH_OFF) direct launches;
        eval_kernel: 0 default, "fp32" or "mfma"; path: 0/"auto", "steps" or
        "teams"."""
        if reset:
            self._check(self.L.n4_set_tuning(self.h, None))
            return
        g = GRAPH_OFF if graph_steps == "off" else int(graph_steps)
        e = EVAL_KERNELS[eval_kernel] if isinstance(eval_kernel, str) else int(eval_kernel)
        pth = PATHS[path] if isinstance(path, str) else int(path)
        t = Tuning(int(groups), int(step_tile_voxels), g, e, pth, int(team_wgs_per_cu), 0.0)
        self._check(self.L.n4_set_tuning(self.h, C.byref(t)))

    def autotune_batch(self, vols, params: Params | None = None, shrink=None, **kw) -> Tuning:
        """Runs the library's schedule search on these host volumes and keeps
        the winner on the handle."""
        p = params if params is not None else default_params(**kw)
        keep, arr = self._batch_arrays(vols, p, False, shrink)
        best = Tuning()
        self._check(self.L.n4_autotune_batch(self.h, arr, len(vols), C.byref(p), C.byref(best), None))
        del keep
        return best

    # ------------------------------------------------------------ N4 ----
    def last_cv_trace(self, volume=0):
        """Convergence value of every iteration of volume `volume` of the
        last correction call, level after level (n4_last_cv_trace)."""
        n = C.c_uint32(0)
        self._check(self.L.n4_last_cv_trace(self.h, volume, None, 0, C.byref(n)))
        out = np.zeros(n.value, np.float64)
        if n.value:
            self._check(self.L.n4_last_cv_trace(self.h, volume, out.ctypes.data_as(C.c_void_p), n.value,
                                                C.byref(n)))
        return out

    def last_level_lattice(self, volume, level, ncpt0):
        """FP32 lattice of fitting level `level` at that level's end
        (n4_last_level_lattice); ncpt0: the first level's control points."""
        nc0 = (ncpt0,) * 3 if np.isscalar(ncpt0) else tuple(ncpt0)
        shape = tuple((int(c) - 3) * 2 ** level + 3 for c in nc0)
        out = np.zeros(shape, np.float32)
        self._check(self.L.n4_last_level_lattice(self.h, volume, level, out.ctypes.data_as(C.c_void_p), out.size))
        return out

    def correct(self, im, mask=None, params: Params | None = None, want_bias=False, want_log_bias=True,
                mask_kind=None, noalloc=False, want_levels=False, **kw):
        """Whole correction of one volume.  Returns dict with normalized,
        log_bias (optional), bias (optional), lattice, report, cv_trace and,
        with want_levels, each level's lattice.  noalloc: the allocation-free
        entry point (n4_correct_noalloc)."""
        p = params if params is not None else default_params(**kw)
        im = np.ascontiguousarray(im, dtype=np.float32)
        dims = im.shape
        n = im.size
        d_im = self.to_device(im)
        if mask is None:
            d_mask, mk = None, MASK_NONE
        elif mask_kind == MASK_U8 or (mask_kind is None and np.asarray(mask).dtype == np.uint8):
            d_mask, mk = self.to_device(np.ascontiguousarray(mask, dtype=np.uint8)), MASK_U8
        else:
            d_mask, mk = self.to_device(np.ascontiguousarray(mask, dtype=np.float32)), MASK_F32
        nc = final_ncpt(p)
        d_out = self.empty(4 * n)
        d_lat = self.empty(4 * nc[0] * nc[1] * nc[2])
        d_bias = self.empty(4 * n) if want_bias else None
        d_lb = self.empty(4 * n) if want_log_bias else None
        rep = Report()
        entry = self.L.n4_correct_noalloc if noalloc else self.L.n4_correct
        self._check(entry(self.h, d_im.ptr, d_mask.ptr if d_mask else None, mk, dims[0], dims[1], dims[2],
                          C.byref(p), d_out.ptr, d_lat.ptr, d_bias.ptr if d_bias else None,
                          d_lb.ptr if d_lb else None, C.byref(rep), None))
        out = {
            "normalized": d_out.download(np.float32, dims),
            "lattice": d_lat.download(np.float32, nc),
            "report": rep,
            "iterations": [int(rep.iterations[i]) for i in range(p.number_of_fitting_levels)],
            "last_cv": [float(rep.last_convergence[i]) for i in range(p.number_of_fitting_levels)],
            "n_mask_full": int(rep.n_mask_full),
            "n_mask_work": int(rep.n_mask_work),
            "elapsed_ms": float(rep.elapsed_ms),
        }
        out["cv_trace"] = self.last_cv_trace(0)
        if want_levels:
            out["level_lattices"] = [self.last_level_lattice(0, l, tuple(p.ncpt))
                                     for l in range(p.number_of_fitting_levels)]
        if d_bias is not None:
            out["bias"] = d_bias.download(np.float32, dims)
        if d_lb is not None:
            out["log_bias"] = d_lb.download(np.float32, dims)
        return out

    def correct_batch(self, vols, params: Params | None = None, want_log_bias=True, shrink=None, outputs=None,
                      **kw):
        """vols: list of (im, mask) host arrays (mask may be None; a uint8
        mask goes to the device as N4_MASK_U8, any other dtype as float32).
        outputs: indices of the volumes whose arrays are downloaded and
        whose log bias is computed (None: all); the others report their
        iteration counts only.  Per-volume CV traces and level lattices:
        last_cv_trace(i) / last_level_lattice(i, level, ncpt) afterwards."""
        p = params if params is not None else default_params(**kw)
        nc = final_ncpt(p)
        sel = None if outputs is None else set(int(i) for i in outputs)
        keep, arr = self._batch_arrays(vols, p, want_log_bias, shrink, sel)
        reps = (Report * len(vols))()
        self._check(self.L.n4_correct_batch(self.h, arr, len(vols), C.byref(p), reps, None))
        res = []
        for i, ((shape, _, _, d_out, d_lat, d_lb), r) in enumerate(zip(keep, reps)):
            o = {"iterations": [int(r.iterations[k]) for k in range(p.number_of_fitting_levels)],
                 "last_cv": [float(r.last_convergence[k]) for k in range(p.number_of_fitting_levels)],
                 "n_mask_full": int(r.n_mask_full), "n_mask_work": int(r.n_mask_work),
                 "elapsed_ms": float(r.elapsed_ms)}
            if sel is None or i in sel:
                o["normalized"] = d_out.download(np.float32, shape)
                o["lattice"] = d_lat.download(np.float32, nc)
                if d_lb is not None:
                    o["log_bias"] = d_lb.download(np.float32, shape)
            res.append(o)
        return res

    def _batch_arrays(self, vols, p, want_log_bias, shrink, sel=None):
        nc = final_ncpt(p)
        keep, arr = [], (Volume * len(vols))()
        for i, (im, mask) in enumerate(vols):
            im = np.ascontiguousarray(im, dtype=np.float32)
            d_im = self.to_device(im)
            if mask is None:
                d_mask, mk = None, MASK_NONE
            elif np.asarray(mask).dtype == np.uint8:
                d_mask, mk = self.to_device(np.ascontiguousarray(mask)), MASK_U8
            else:
                d_mask, mk = self.to_device(np.ascontiguousarray(mask, dtype=np.float32)), MASK_F32
            d_out = self.empty(im.nbytes)
            d_lat = self.empty(4 * nc[0] * nc[1] * nc[2])
            d_lb = self.empty(im.nbytes) if want_log_bias and (sel is None or i in sel) else None
            keep.append((im.shape, d_im, d_mask, d_out, d_lat, d_lb))
            v = arr[i]
            v.d_im = d_im.ptr
            v.d_mask = d_mask.ptr if d_mask else None
            v.mask_kind = mk
            for d in range(3):
                v.dims[d] = im.shape[d]
            v.shrink_factor = int(shrink[i]) if shrink is not None else 0
            v.d_im_normalized = d_out.ptr
            v.d_lattice = d_lat.ptr
            v.d_bias_field = None
            v.d_log_bias_field = d_lb.ptr if d_lb else None
        return keep, arr

    # -------------------------------------------------------- stages ----
    def histogram(self, x, mask, nbins=200, with_index=True):
        x = np.ascontiguousarray(x, dtype=np.float32).ravel()
        m = np.ascontiguousarray(np.asarray(mask).ravel() != 0, dtype=np.uint8)
        dx, dm = self.to_device(x), self.to_device(m)
        dh = self.empty(8 * nbins)
        di = self.empty(4 * x.size) if with_index else None
        mms = (C.c_double * 3)()
        self._check(self.L.n4_histogram(self.h, dx.ptr, dm.ptr, x.size, nbins, mms, dh.ptr,
                                        di.ptr if di else None))
        H = dh.download(np.float64, nbins)
        idx = di.download(np.int32, x.size) if di else None
        return H, idx, (mms[0], mms[1], mms[2])

    def sharpen(self, x, mask, nbins=200, noise=0.01, fwhm=0.15):
        shape = np.shape(x)
        x = np.ascontiguousarray(x, dtype=np.float32).ravel()
        m = np.ascontiguousarray(np.asarray(mask).ravel() != 0, dtype=np.uint8)
        dx, dm = self.to_device(x), self.to_device(m)
        do = self.empty(4 * x.size)
        de = self.empty(8 * nbins)
        self._check(self.L.n4_sharpen_image(self.h, dx.ptr, dm.ptr, x.size, nbins, noise, fwhm, do.ptr, de.ptr))
        return do.download(np.float32, shape), de.download(np.float64, nbins)

    def mask_to_u8(self, mask, offset=0):
        """Float mask -> uint8 mask on the device (n4_mask_to_u8); offset
        (elements) shifts both device pointers, to exercise unaligned
        buffers."""
        m = np.ascontiguousarray(mask, dtype=np.float32).ravel()
        src = np.concatenate([np.zeros(offset, np.float32), m])
        dm = self.to_device(src)
        do = self.empty(m.size + offset + 4)
        self._check(self.L.n4_mask_to_u8(self.h, dm.ptr + 4 * offset, m.size, do.ptr + offset))
        return do.download(np.uint8, (m.size + offset + 4,))[offset:offset + m.size]

    def fit_bspline3d(self, field, mask, ncpt, eps=1e-3):
        field = np.ascontiguousarray(field, dtype=np.float32)
        m = np.ascontiguousarray(np.asarray(mask) != 0, dtype=np.uint8)
        nc = (C.c_uint32 * 3)(*[int(v) for v in ncpt])
        dims = (C.c_uint32 * 3)(*field.shape)
        df, dm = self.to_device(field), self.to_device(m)
        dl = self.empty(4 * int(np.prod(ncpt)))
        self._check(self.L.n4_fit_bspline3d(self.h, df.ptr, dm.ptr, nc, dims, eps, dl.ptr))
        return dl.download(np.float32, tuple(int(v) for v in ncpt))

    def evaluate_bspline3d(self, lattice, dims):
        lattice = np.ascontiguousarray(lattice, dtype=np.float32)
        nc = (C.c_uint32 * 3)(*lattice.shape)
        dd = (C.c_uint32 * 3)(*[int(v) for v in dims])
        dl = self.to_device(lattice)
        do = self.empty(4 * int(np.prod(dims)))
        self._check(self.L.n4_evaluate_bspline3d(self.h, dl.ptr, nc, dd, do.ptr))
        return do.download(np.float32, tuple(int(v) for v in dims))

    def convergence(self, f1, f2, mask):
        f1 = np.ascontiguousarray(f1, dtype=np.float32).ravel()
        f2 = np.ascontiguousarray(f2, dtype=np.float32).ravel()
        m = np.ascontiguousarray(np.asarray(mask).ravel() != 0, dtype=np.uint8)
        d1, d2, dm = self.to_device(f1), self.to_device(f2), self.to_device(m)
        out = C.c_double()
        self._check(self.L.n4_convergence_measurement(self.h, d1.ptr, d2.ptr, dm.ptr, f1.size, C.byref(out)))
        return out.value

    def upsample_lattice3d(self, lattice):
        lattice = np.ascontiguousarray(lattice, dtype=np.float32)
        nc = (C.c_uint32 * 3)(*lattice.shape)
        out_shape = tuple(2 * (s - 3) + 3 for s in lattice.shape)
        dl = self.to_device(lattice)
        do = self.empty(4 * int(np.prod(out_shape)))
        self._check(self.L.n4_upsample_lattice3d(self.h, dl.ptr, nc, do.ptr))
        return do.download(np.float32, out_shape)

    def reduce(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32).ravel()
        dx = self.to_device(x)
        s, mx, mn = C.c_double(), C.c_float(), C.c_float()
        self._check(self.L.n4_reduce_sum(self.h, x.size, dx.ptr, C.byref(s)))
        self._check(self.L.n4_reduce_max(self.h, x.size, dx.ptr, C.byref(mx)))
        self._check(self.L.n4_reduce_min(self.h, x.size, dx.ptr, C.byref(mn)))
        return s.value, mx.value, mn.value
