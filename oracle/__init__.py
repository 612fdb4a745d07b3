"""ctypes binding of the FP64 CPU oracle (oracle/n4_oracle.c).

TEST INFRASTRUCTURE ONLY.  The oracle restates the reference's N4.py in FP64
and is the parity reference for the HIP product path.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import it; the
product (the HIP library and its host wrappers) never does.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libn4oracle.so")
MAX_LEVELS = 16

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


class Params(C.Structure):
    _fields_ = [
        ("nbins", C.c_int32),
        ("wiener_noise", C.c_double),
        ("fwhm", C.c_double),
        ("ncpt", C.c_int32 * 3),
        ("levels", C.c_int32),
        ("max_iter", C.c_int32 * MAX_LEVELS),
        ("conv_threshold", C.c_double),
        ("low_value", C.c_double),
        ("shrink", C.c_int32),
        ("fit_eps", C.c_double),
        ("nthreads", C.c_int32),
        ("weighted", C.c_int32),
    ]


class Report(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32 * MAX_LEVELS),
        ("last_cv", C.c_double * MAX_LEVELS),
        ("n_mask_full", C.c_int64),
        ("n_mask_work", C.c_int64),
        ("work_dims", C.c_int32 * 3),
        ("final_ncpt", C.c_int32 * 3),
    ]


_lib = None


def build() -> None:
    """Compile the oracle with its Makefile (gcc, no GPU)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.n4o_default_params.argtypes = [C.POINTER(Params)]
        L.n4o_cubic_bspline.argtypes = [C.c_double, _f64p]
        L.n4o_axis_table.argtypes = [C.c_int32, C.c_double, C.c_int32, _i32p, _f64p]
        L.n4o_padded_size.argtypes = [C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.n4o_fft.argtypes = [_f64p, C.c_int32, C.c_int32]
        L.n4o_masked_minmax.argtypes = [_f64p, _u8p, C.c_int64, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.n4o_masked_minmax.restype = C.c_int64
        L.n4o_histogram.argtypes = [_f64p, _u8p, C.c_int64, C.c_int32, C.c_double, C.c_double, _f64p, C.c_void_p]
        L.n4o_wiener_etable.argtypes = [_f64p, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double, _f64p]
        L.n4o_sharpen.argtypes = [_f64p, _u8p, C.c_int64, C.c_int32, C.c_double, C.c_double, _f64p,
                                  _f64p, _f64p, _f64p]
        L.n4o_fit_bspline3d.argtypes = [_f64p, _u8p, _i32p, _i32p, C.c_double, C.c_int32, _f64p]
        L.n4o_fit_numden.argtypes = [_f64p, _u8p, _i32p, _i32p, C.c_int32, _f64p, _f64p]
        L.n4o_fit_numden_w.argtypes = [_f64p, _u8p, C.c_void_p, _i32p, _i32p, C.c_int32, _f64p, _f64p]
        L.n4o_evaluate_bspline3d.argtypes = [_f64p, _i32p, _i32p, _f64p, C.c_int32, _f64p]
        L.n4o_convergence.argtypes = [_f64p, _f64p, _u8p, C.c_int64]
        L.n4o_convergence.restype = C.c_double
        L.n4o_upsample_lattice3d.argtypes = [_f64p, _i32p, _f64p]
        L.n4o_threshold_log_shrink.argtypes = [_f32p, C.c_void_p, _i32p, C.c_double, C.c_int32, _i32p,
                                               C.c_void_p, C.c_void_p, C.POINTER(C.c_int64),
                                               C.POINTER(C.c_int64)]
        L.n4o_threshold_log_shrink_w.argtypes = [_f32p, C.c_void_p, C.c_int32, _i32p, C.c_double, C.c_int32,
                                                 _i32p, C.c_void_p, C.c_void_p, C.c_void_p,
                                                 C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.n4o_n4.argtypes = [_f32p, C.c_void_p, _i32p, C.POINTER(Params), C.c_void_p, C.c_void_p,
                             C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(Report)]
        L.n4o_n4.restype = C.c_int32
        _lib = L
    return _lib


def default_params(**kw) -> Params:
    p = Params()
    lib().n4o_default_params(C.byref(p))
    for k, v in kw.items():
        if k == "ncpt":
            v = (v, v, v) if np.isscalar(v) else v
            for i in range(3):
                p.ncpt[i] = int(v[i])
        elif k == "max_iter":
            v = [v] * MAX_LEVELS if np.isscalar(v) else list(v)
            for i in range(MAX_LEVELS):
                p.max_iter[i] = int(v[i]) if i < len(v) else int(v[-1])
        else:
            setattr(p, k, v)
    return p


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _mask_u8(m):
    return np.ascontiguousarray(np.asarray(m) != 0, dtype=np.uint8)


def cubic_bspline(t: float) -> np.ndarray:
    w = np.zeros(4)
    lib().n4o_cubic_bspline(float(t), w)
    return w


def axis_table(npos: int, domain: float, ncpt: int):
    b = np.zeros(npos, np.int32)
    w = np.zeros((npos, 4))
    lib().n4o_axis_table(npos, float(domain), ncpt, b, w)
    return b, w


def padded_size(nbins: int):
    P, off = C.c_int32(), C.c_int32()
    lib().n4o_padded_size(nbins, C.byref(P), C.byref(off))
    return P.value, off.value


def fft(x, inverse=False) -> np.ndarray:
    x = np.asarray(x, dtype=np.complex128)
    buf = np.ascontiguousarray(np.stack([x.real, x.imag], -1).reshape(-1))
    lib().n4o_fft(buf, len(x), 1 if inverse else 0)
    return buf[0::2] + 1j * buf[1::2]


def masked_minmax(x, mask):
    x = _c64(x).ravel()
    m = _mask_u8(mask).ravel()
    mn, mx = C.c_double(), C.c_double()
    n = lib().n4o_masked_minmax(x, m, x.size, C.byref(mn), C.byref(mx))
    return n, mn.value, mx.value


def histogram(x, mask, nbins, bin_min, slope, with_idx=False):
    x = _c64(x).ravel()
    m = _mask_u8(mask).ravel()
    H = np.zeros(nbins)
    idx = np.zeros(x.size, np.int32) if with_idx else None
    lib().n4o_histogram(x, m, x.size, nbins, bin_min, slope, H,
                        idx.ctypes.data_as(C.c_void_p) if with_idx else None)
    return (H, idx) if with_idx else H


def wiener_etable(H, bin_min, slope, noise=0.01, fwhm=0.15):
    H = _c64(H)
    E = np.zeros(H.size)
    lib().n4o_wiener_etable(H, H.size, bin_min, slope, noise, fwhm, E)
    return E


def sharpen(x, mask, nbins=200, noise=0.01, fwhm=0.15):
    """Returns (sharpened, H, E, (bin_min, bin_max, slope))."""
    shape = np.shape(x)
    x = _c64(x).ravel()
    m = _mask_u8(mask).ravel()
    out = np.zeros(x.size)
    H = np.zeros(nbins)
    E = np.zeros(nbins)
    mms = np.zeros(3)
    lib().n4o_sharpen(x, m, x.size, nbins, noise, fwhm, out, H, E, mms)
    return out.reshape(shape), H, E, tuple(mms)


def fit_bspline3d(field, mask, ncpt, eps=1e-3, nthreads=1):
    field = _c64(field)
    dims = np.array(field.shape, np.int32)
    nc = np.array(ncpt, np.int32)
    lat = np.zeros(tuple(nc))
    lib().n4o_fit_bspline3d(field.ravel(), _mask_u8(mask).ravel(), dims, nc, eps, nthreads, lat.ravel())
    return lat


def fit_numden(field, mask, ncpt, nthreads=1, conf=None):
    """Numerator and denominator of the fit; conf (optional) holds per-voxel
    confidence weights."""
    field = _c64(field)
    dims = np.array(field.shape, np.int32)
    nc = np.array(ncpt, np.int32)
    num = np.zeros(tuple(nc))
    den = np.zeros(tuple(nc))
    if conf is None:
        lib().n4o_fit_numden(field.ravel(), _mask_u8(mask).ravel(), dims, nc, nthreads, num.ravel(), den.ravel())
    else:
        cf = _c64(conf)
        lib().n4o_fit_numden_w(field.ravel(), _mask_u8(mask).ravel(), cf.ctypes.data_as(C.c_void_p), dims, nc,
                               nthreads, num.ravel(), den.ravel())
    return num, den


def evaluate_bspline3d(lattice, dims, domain=None, nthreads=1):
    lattice = _c64(lattice)
    nc = np.array(lattice.shape, np.int32)
    d = np.array(dims, np.int32)
    dom = np.array(dims if domain is None else domain, np.float64)
    out = np.zeros(tuple(d))
    lib().n4o_evaluate_bspline3d(lattice.ravel(), nc, d, dom, nthreads, out.ravel())
    return out


def convergence(f1, f2, mask):
    f1 = _c64(f1).ravel()
    f2 = _c64(f2).ravel()
    return lib().n4o_convergence(f1, f2, _mask_u8(mask).ravel(), f1.size)


def upsample_lattice3d(lattice):
    lattice = _c64(lattice)
    nc = np.array(lattice.shape, np.int32)
    out = np.zeros(tuple(2 * (nc - 3) + 3))
    lib().n4o_upsample_lattice3d(lattice.ravel(), nc, out.ravel())
    return out


def threshold_log_shrink(im, mask=None, low_value=10.0, shrink=1, weighted=False):
    """Working-grid log image and mask (and, with weighted=True, the working
    voxels' mean confidence weights as a fifth output)."""
    im = np.ascontiguousarray(im, dtype=np.float32)
    fd = np.array(im.shape, np.int32)
    wd = np.zeros(3, np.int32)
    W = [(int(f) + shrink - 1) // shrink for f in im.shape]
    imlog = np.zeros(W)
    mw = np.zeros(W, np.uint8)
    cw = np.zeros(W)
    mk = None if mask is None else np.ascontiguousarray(mask, dtype=np.float32)
    nf, nw = C.c_int64(), C.c_int64()
    lib().n4o_threshold_log_shrink_w(im.ravel(), None if mk is None else mk.ctypes.data_as(C.c_void_p),
                                     int(bool(weighted)), fd, low_value, shrink, wd,
                                     imlog.ctypes.data_as(C.c_void_p), mw.ctypes.data_as(C.c_void_p),
                                     cw.ctypes.data_as(C.c_void_p), C.byref(nf), C.byref(nw))
    if weighted:
        return imlog, mw, nf.value, nw.value, cw
    return imlog, mw, nf.value, nw.value


def n4(im, mask=None, params: Params | None = None, cv_cap=4096, **kw):
    """Whole N4 correction.  Returns a dict with normalized image, log bias
    field (full resolution), final lattice, per-level iterations, CV trace."""
    p = params if params is not None else default_params(**kw)
    im = np.ascontiguousarray(im, dtype=np.float32)
    fd = np.array(im.shape, np.int32)
    mk = None if mask is None else np.ascontiguousarray(mask, dtype=np.float32)
    norm = np.zeros(im.shape)
    logb = np.zeros(im.shape)
    nc = np.array(p.ncpt[:], np.int64)
    for _ in range(p.levels - 1):
        nc = 2 * (nc - 3) + 3
    lat = np.zeros(tuple(int(v) for v in nc))
    cv = np.zeros(cv_cap)
    rep = Report()
    rc = lib().n4o_n4(im.ravel(), None if mk is None else mk.ctypes.data_as(C.c_void_p), fd, C.byref(p),
                      norm.ctypes.data_as(C.c_void_p), logb.ctypes.data_as(C.c_void_p),
                      lat.ctypes.data_as(C.c_void_p), cv.ctypes.data_as(C.c_void_p), cv_cap, C.byref(rep))
    if rc != 0:
        raise ValueError("oracle n4: invalid arguments")
    its = [rep.iterations[i] for i in range(p.levels)]
    return {
        "normalized": norm,
        "log_bias": logb,
        "lattice": lat,
        "iterations": its,
        "last_cv": [rep.last_cv[i] for i in range(p.levels)],
        "cv_trace": cv[: min(sum(its), cv_cap)].copy(),
        "n_mask_full": rep.n_mask_full,
        "n_mask_work": rep.n_mask_work,
        "work_dims": tuple(rep.work_dims[:]),
    }
