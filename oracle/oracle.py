"""ctypes wrapper of the oracle (oracle/mvd_oracle.cpp).

TEST INFRASTRUCTURE ONLY.  The oracle is a CPU restatement of the reference's
float32 deconvolution path plus an FP64 twin.  Only tests/, the smoke check
and the benchmark's CPU-baseline leg import this module, and only as the
checker or the baseline; the product library never calls it.

Arrays are numpy, x-fastest, indexed [z, y, x].  Dimension tuples passed to
the C functions are (nx, ny, nz).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libmvd_oracle.so")

OOB_MIRROR, OOB_ONE, OOB_ZERO = 0, 1, 2
OPTIMIZATION_II, OPTIMIZATION_I, EFFICIENT_BAYESIAN, INDEPENDENT = 0, 1, 2, 3

_lib = None


def build() -> str:
    """Compile the oracle with its Makefile; returns the library path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_P = ctypes.c_void_p
_I64P = ctypes.POINTER(ctypes.c_int64)


def _declare(L):
    L.orc_last_error.restype = ctypes.c_char_p
    L.orc_divide_portions.argtypes = [ctypes.c_int64, ctypes.c_int, _P, _P]
    L.orc_divide_blocks.argtypes = [_P, _P, _P, _P, ctypes.c_int]
    L.orc_direct_convolve_f64.argtypes = [_P, _P, ctypes.c_int, _P, _P, _P, ctypes.c_int]
    for suf in ("f32", "f64"):
        getattr(L, f"orc_sum_img_{suf}").argtypes = [_P, ctypes.c_int64, ctypes.c_int]
        getattr(L, f"orc_sum_img_{suf}").restype = ctypes.c_double
        getattr(L, f"orc_norm_img_{suf}").argtypes = [_P, ctypes.c_int64, ctypes.c_int]
        getattr(L, f"orc_mirror_{suf}").argtypes = [_P, _P, ctypes.c_int]
        getattr(L, f"orc_fft_convolve_{suf}").argtypes = [_P, _P, ctypes.c_int, _P, _P, _P, ctypes.c_int]
        getattr(L, f"orc_convolve_{suf}").argtypes = [_P, _P, ctypes.c_int, _P, _P, _P, _P, ctypes.c_int]
        getattr(L, f"orc_init_kernels_{suf}").argtypes = [
            ctypes.c_int, _P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P]
        getattr(L, f"orc_first_iteration_{suf}").argtypes = [ctypes.c_int, _P, ctypes.c_int64, ctypes.c_int, _P]
        getattr(L, f"orc_first_iteration_{suf}").restype = ctypes.c_double
        getattr(L, f"orc_quotient_{suf}").argtypes = [_P, _P, ctypes.c_int64]
        getattr(L, f"orc_next_value_{suf}").argtypes = [_P, _P, _P, ctypes.c_int64, ctypes.c_double, _P]
        getattr(L, f"orc_deconvolve_{suf}").argtypes = [
            ctypes.c_int, _P, _P, _P, _P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_double, _P,
            ctypes.c_int, ctypes.c_int, _P, _P, _P, _P]


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().orc_last_error().decode())


def _dt(dtype):
    dtype = np.dtype(dtype)
    if dtype == np.float32:
        return "f32"
    if dtype == np.float64:
        return "f64"
    raise TypeError(f"oracle supports float32/float64, got {dtype}")


def _ptr(a):
    return a.ctypes.data_as(_P)


def _dims(a):
    """(nx, ny, nz) of a [z, y, x] array as an int64 array."""
    z, y, x = a.shape
    return np.array([x, y, z], dtype=np.int64)


def _ptr_array(arrays):
    return (_P * len(arrays))(*[a.ctypes.data for a in arrays])


def nproc() -> int:
    """CPUs this process may run on (what `nproc` prints)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def default_threads() -> int:
    """nproc, capped by OMP_NUM_THREADS when the host sets one (a share of a
    larger machine)."""
    n = nproc()
    try:
        n = min(n, int(os.environ.get("OMP_NUM_THREADS", n)))
    except ValueError:
        pass
    return max(1, n)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def divide_portions(n: int, k: int):
    s = np.zeros(k, np.int64)
    z = np.zeros(k, np.int64)
    _check(lib().orc_divide_portions(n, k, _ptr(s), _ptr(z)))
    return list(zip(s.tolist(), z.tolist()))


def divide_blocks(img_xyz, ker_xyz, block_xyz):
    """List of dicts with offset/eff_size/eff_offset/eff_local (x, y, z)."""
    i = np.asarray(img_xyz, np.int64)
    k = np.asarray(ker_xyz, np.int64)
    b = np.asarray(block_xyz, np.int64)
    cap = 4096
    out = np.zeros(cap * 12, np.int64)
    nb = lib().orc_divide_blocks(_ptr(i), _ptr(k), _ptr(b), _ptr(out), cap)
    if nb < 0:
        raise ValueError(lib().orc_last_error().decode())
    res = []
    for j in range(nb):
        r = out[j * 12:(j + 1) * 12]
        res.append(dict(offset=tuple(r[0:3]), eff_size=tuple(r[3:6]),
                        eff_offset=tuple(r[6:9]), eff_local=tuple(r[9:12])))
    return res


def sum_img(k, norm_threads: int) -> float:
    k = np.ascontiguousarray(k)
    return getattr(lib(), f"orc_sum_img_{_dt(k.dtype)}")(_ptr(k), k.size, norm_threads)


def norm_img(k, norm_threads: int):
    k = np.array(k, copy=True, order="C")
    _check(getattr(lib(), f"orc_norm_img_{_dt(k.dtype)}")(_ptr(k), k.size, norm_threads))
    return k


def mirror(img, axis_xyz: int):
    img = np.array(img, copy=True, order="C")
    _check(getattr(lib(), f"orc_mirror_{_dt(img.dtype)}")(_ptr(img), _ptr(_dims(img)), axis_xyz))
    return img


def fft_convolve(img, kernel, oob: int, threads: int | None = None):
    img = np.ascontiguousarray(img)
    kernel = np.ascontiguousarray(kernel, dtype=img.dtype)
    out = np.empty_like(img)
    _check(getattr(lib(), f"orc_fft_convolve_{_dt(img.dtype)}")(
        _ptr(img), _ptr(_dims(img)), oob, _ptr(kernel), _ptr(_dims(kernel)), _ptr(out),
        threads or default_threads()))
    return out


def convolve(img, kernel, oob: int, block_xyz=None, threads: int | None = None):
    img = np.ascontiguousarray(img)
    kernel = np.ascontiguousarray(kernel, dtype=img.dtype)
    out = np.empty_like(img)
    blk = None if block_xyz is None else np.asarray(block_xyz, np.int64)
    _check(getattr(lib(), f"orc_convolve_{_dt(img.dtype)}")(
        _ptr(img), _ptr(_dims(img)), oob, _ptr(kernel), _ptr(_dims(kernel)),
        None if blk is None else _ptr(blk), _ptr(out), threads or default_threads()))
    return out


def direct_convolve_f64(img, kernel, oob: int, threads: int | None = None):
    img = np.ascontiguousarray(img, dtype=np.float64)
    kernel = np.ascontiguousarray(kernel, dtype=np.float64)
    out = np.empty_like(img)
    _check(lib().orc_direct_convolve_f64(_ptr(img), _ptr(_dims(img)), oob, _ptr(kernel),
                                         _ptr(_dims(kernel)), _ptr(out), threads or default_threads()))
    return out


def init_kernels(kernels, psf_type: int, norm_threads: int, threads: int | None = None):
    """Returns (kernel1 normalised, kernel2) lists, as MVDeconFFT.init does."""
    dtype = kernels[0].dtype
    k1 = [np.array(k, copy=True, order="C", dtype=dtype) for k in kernels]
    k2 = [np.empty_like(k) for k in k1]
    kd = np.concatenate([_dims(k) for k in k1]).astype(np.int64)
    _check(getattr(lib(), f"orc_init_kernels_{_dt(dtype)}")(
        len(k1), _ptr_array(k1), _ptr(kd), psf_type, norm_threads, threads or default_threads(),
        _ptr_array(k2)))
    return k1, k2


def first_iteration(imgs, norm_threads: int):
    """Returns (average, counts image)."""
    imgs = [np.ascontiguousarray(i) for i in imgs]
    counts = np.empty_like(imgs[0])
    avg = getattr(lib(), f"orc_first_iteration_{_dt(imgs[0].dtype)}")(
        len(imgs), _ptr_array(imgs), imgs[0].size, norm_threads, _ptr(counts))
    return avg, counts


def quotient(blurred, img):
    b = np.array(blurred, copy=True, order="C")
    img = np.ascontiguousarray(img, dtype=b.dtype)
    _check(getattr(lib(), f"orc_quotient_{_dt(b.dtype)}")(_ptr(b), _ptr(img), b.size))
    return b


def next_value(last, integral, weight, lam: float):
    last = np.ascontiguousarray(last)
    integral = np.ascontiguousarray(integral, dtype=last.dtype)
    weight = np.ascontiguousarray(weight, dtype=last.dtype)
    out = np.empty_like(last)
    _check(getattr(lib(), f"orc_next_value_{_dt(last.dtype)}")(
        _ptr(last), _ptr(integral), _ptr(weight), last.size, lam, _ptr(out)))
    return out


# Weight types of the session API (the reference's WeightType,
# deconv/ProcessForDeconvolution.java:81; AS_GIVEN = weights handed to the
# operator as they are, LOAD_WEIGHTS' normalisation left to the caller).
W_NONE, W_AS_GIVEN, W_PRECOMPUTED, W_VIRTUAL = 0, 1, 2, 3


def overlap_counts(raw, norm_threads: int):
    """WeightNormalizer.process (deconv/WeightNormalizer.java:73-116): per
    portion of 2*numThreads the minimum number of views with weight > 0 and
    the average count; returns (min over portions, mean of portion averages),
    each then raised to at least 1 (ProcessForDeconvolution.java:349-350)."""
    cnt = np.zeros(raw[0].size, np.int64)
    for w in raw:
        cnt += (np.ravel(w) > 0)
    mins, avgs = [], []
    for start, size in divide_portions(cnt.size, 2 * norm_threads):
        part = cnt[start:start + size]
        mins.append(int(part.min()) if size else len(raw))
        avgs.append(part.sum() / size if size else float("nan"))
    return max(1, min(mins)), max(1.0, sum(avgs) / len(avgs))  # summed in portion order


def effective_weights(raw, mode: int, osem_index: int = 0, osem_value: float = 1.0, norm_threads: int = 8,
                      shape=None):
    """The weight each view's update reads (computeFinalValues,
    deconv/MVDeconvolution.java:603-680), float32 per view, and the OSEM factor.

    PRECOMPUTED: WeightNormalizer.ApplyDirectly (WeightNormalizer.java:138-185):
      w_v = (float)(w_v / sum_u w_u), the double sum in view order, no guard
      (0/0 gives NaN).
    VIRTUAL: ComputeSumImage (:187-253) keeps s = sum > 1 ? (float)sum : 1 and
      NormalizingRandomAccess.get (weights/NormalizingRandomAccess.java:56-66)
      reads (float)Math.min(1, (double)w / s * osem).
    OSEM factor (ProcessForDeconvolution.java:343-356): index 1 = minimum
      overlap, 2 = average overlap, 3 = the given value, 0 = 1; without weights
      the overlap counts stay 0.
    adjustForOSEM (:372-405), if the factor is not 1: PRECOMPUTED / AS_GIVEN
      w = Math.min(1, w * (float)osem) in float (NaN stays NaN); NO_WEIGHTS the
      constant 1 becomes Math.min(1, 1 * (float)osem).
    Returns (list of float32 arrays, osem)."""
    if mode == W_NONE:
        mn, avg = 0, 0.0
    elif mode in (W_PRECOMPUTED, W_VIRTUAL):
        mn, avg = overlap_counts(raw, norm_threads)
    else:
        mn, avg = 1, 1.0
    osem = {0: 1.0, 1: float(mn), 2: float(avg), 3: float(osem_value)}[osem_index]
    f1 = np.float32(1)
    if mode == W_NONE:
        c = f1 if osem == 1.0 else np.minimum(f1, f1 * np.float32(osem))
        return [np.full(shape, c, np.float32) for _ in range(len(raw) if raw else 1)], osem
    ws = [np.asarray(w, np.float32) for w in raw]
    total = np.zeros(ws[0].shape, np.float64)
    for w in ws:  # view order, as the cursors are read
        total += w.astype(np.float64)
    if mode == W_VIRTUAL:
        ssum = np.where(total > 1, total.astype(np.float32), f1).astype(np.float64)
        return [np.minimum(1.0, w.astype(np.float64) / ssum * osem).astype(np.float32) for w in ws], osem
    if mode == W_PRECOMPUTED:
        with np.errstate(invalid="ignore", divide="ignore"):
            ws = [(w.astype(np.float64) / total).astype(np.float32) for w in ws]
    if osem != 1.0:
        ws = [np.minimum(f1, w * np.float32(osem)) for w in ws]  # np.minimum keeps NaN, as Math.min
    return ws, osem


def deconvolve(views, kernels, psf_type: int, iterations: int, lam: float, weights=None,
               block_xyz=None, norm_threads: int = 8, threads: int | None = None, psi0=None,
               return_kernels: bool = False):
    """Full MVDeconvolution run.  Returns (psi, stats[iterations, views, 2]) and
    optionally kernel2 per view."""
    dtype = views[0].dtype
    views = [np.ascontiguousarray(v, dtype=dtype) for v in views]
    kernels = [np.ascontiguousarray(k, dtype=dtype) for k in kernels]
    nv = len(views)
    dims = _dims(views[0])
    kd = np.concatenate([_dims(k) for k in kernels]).astype(np.int64)
    w_arr = None
    if weights is not None:
        weights = [np.ascontiguousarray(w, dtype=dtype) for w in weights]
        w_arr = _ptr_array(weights)
    psi = np.empty_like(views[0])
    stats = np.zeros((iterations, nv, 2), np.float64)
    blk = None if block_xyz is None else np.asarray(block_xyz, np.int64)
    k2 = [np.empty_like(k) for k in kernels] if return_kernels else None
    p0 = None
    if psi0 is not None:
        p0 = np.ascontiguousarray(psi0, dtype=dtype)
    _check(getattr(lib(), f"orc_deconvolve_{_dt(dtype)}")(
        nv, _ptr(dims), _ptr_array(views), w_arr, _ptr_array(kernels), _ptr(kd), psf_type,
        iterations, lam, None if blk is None else _ptr(blk), norm_threads, threads or default_threads(),
        None if p0 is None else _ptr(p0), _ptr(psi), _ptr(stats),
        _ptr_array(k2) if k2 is not None else None))
    if return_kernels:
        return psi, stats, k2
    return psi, stats


def _declare_grad(L):
    L.orc_nll_view_fd_f64.argtypes = [_P, _P, _P, _P, _P, _P, _P, ctypes.c_int, ctypes.c_double, _P, ctypes.c_int]
    for suf in ("f32", "f64"):
        getattr(L, f"orc_nll_grad_{suf}").argtypes = [ctypes.c_int, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                                      ctypes.c_int, ctypes.c_int]


def nll_grad(views, psfs, psi, weights=None, threads=None, variant: int = 0):
    """Poisson NLL of the forward model b_v = E_mirror(psi) * P_v and its
    gradient.  Returns (nll, grad_psi, [grad_psf_v]).  variant != 0 plants a
    known mistake (1: the mirrored shell of the psi fold dropped); only the
    tests that show their checks catch it use it."""
    L = lib()
    _declare_grad(L)
    dtype = np.dtype(psi.dtype)
    views = [np.ascontiguousarray(v, dtype=dtype) for v in views]
    psfs = [np.ascontiguousarray(p, dtype=dtype) for p in psfs]
    psi = np.ascontiguousarray(psi)
    w = None
    if weights is not None:
        weights = [np.ascontiguousarray(x, dtype=dtype) for x in weights]
        w = _ptr_array(weights)
    gpsi = np.empty_like(psi)
    gpsf = [np.empty_like(p) for p in psfs]
    kd = np.concatenate([_dims(p) for p in psfs]).astype(np.int64)
    nll = ctypes.c_double()
    _check(getattr(L, f"orc_nll_grad_{_dt(dtype)}")(
        len(views), _ptr(_dims(psi)), _ptr_array(views), w, _ptr(psi), _ptr_array(psfs), _ptr(kd), _ptr(gpsi),
        _ptr_array(gpsf), ctypes.byref(nll), threads or default_threads(), variant))
    return nll.value, gpsi, gpsf


def nll_psf_fd(view, psf, psi, voxels, h=1e-3, weight=None, threads=None):
    """Central finite differences of one view's NLL w.r.t. single PSF voxels
    (flat indices), in FP64: each probe is a rank-one update of the blur."""
    L = lib()
    _declare_grad(L)
    v = np.ascontiguousarray(view, dtype=np.float64)
    p = np.ascontiguousarray(psf, dtype=np.float64)
    s = np.ascontiguousarray(psi, dtype=np.float64)
    blur = fft_convolve(s, p, OOB_MIRROR, threads)
    wt = None if weight is None else np.ascontiguousarray(weight, dtype=np.float64)
    vox = np.ascontiguousarray(voxels, dtype=np.int64)
    out = np.empty(len(vox), np.float64)
    _check(L.orc_nll_view_fd_f64(_ptr(v), None if wt is None else _ptr(wt), _ptr(blur), _ptr(s), _ptr(_dims(s)),
                                 _ptr(_dims(p)), _ptr(vox), len(vox), h, _ptr(out), threads or default_threads()))
    return out


def _declare_transform(L):
    L.orc_blend_lut.argtypes = [_P]
    L.orc_blend_lut.restype = None
    L.orc_blend_weight.argtypes = [_P, _P, _P, _P]
    L.orc_blend_weight.restype = ctypes.c_float
    L.orc_transform_view.argtypes = [_P, _P, _P, _P, _P, _P, _P, _P, _P, ctypes.c_int]


def blend_lut():
    """The blending lookup table (1001 doubles)."""
    L = lib()
    _declare_transform(L)
    out = np.empty(1001, np.float64)
    L.orc_blend_lut(_ptr(out))
    return out


def blend_weight(t_xyz, size_xyz, border_xyz, range_xyz) -> np.float32:
    """Blending weight of a view of size_xyz at one (float) position."""
    L = lib()
    _declare_transform(L)
    t = np.asarray(t_xyz, np.float32)
    return np.float32(L.orc_blend_weight(_ptr(t), _ptr(np.asarray(size_xyz, np.int64)),
                                         _ptr(np.asarray(border_xyz, np.float32)),
                                         _ptr(np.asarray(range_xyz, np.float32))))


def transform_view(raw, inverse, offset, size_xyz, border=None, rng=None, want_image=True, raw_size_xyz=None,
                   threads=None):
    """One view resampled into the bounding box (size_xyz, min = offset) by
    the row-packed 3x4 inverse map.  Returns (image or None, weight or None);
    the image is 0 where the position falls outside the raw view."""
    L = lib()
    _declare_transform(L)
    inv = np.ascontiguousarray(np.asarray(inverse, np.float64).reshape(12))
    off = np.ascontiguousarray(np.asarray(offset, np.int64).reshape(3))
    nx, ny, nz = (int(v) for v in size_xyz)
    if raw is not None:
        raw = np.ascontiguousarray(raw, np.float32)
        rs = np.array([raw.shape[2], raw.shape[1], raw.shape[0]], np.int64)
    else:
        rs = np.asarray(raw_size_xyz, np.int64)
    img = np.zeros((nz, ny, nx), np.float32) if want_image else None
    w = None
    b = r = None
    if border is not None:
        b = np.asarray(border, np.float32)
        r = np.asarray(rng, np.float32)
        w = np.empty((nz, ny, nx), np.float32)
    _check(L.orc_transform_view(None if raw is None else _ptr(raw), _ptr(rs), _ptr(inv), _ptr(off),
                                _ptr(np.array([nx, ny, nz], np.int64)), None if b is None else _ptr(b),
                                None if r is None else _ptr(r), None if img is None else _ptr(img),
                                None if w is None else _ptr(w), threads or default_threads()))
    return img, w


def _declare_psf(L):
    L.orc_extract_psf.argtypes = [_P, _P, _P, ctypes.c_int64, _P, _P, _P]
    L.orc_psf_transform_geometry.argtypes = [_P, _P, _P, _P]
    L.orc_psf_transform_geometry.restype = None
    L.orc_psf_transform.argtypes = [_P, _P, _P, _P, _P, _P]
    L.orc_psf_transform.restype = None


def extract_psf(img, locations, size_xyz):
    """extractPSFLocal + normalize: (raw sum, normalised PSF), float32 [z, y, x]."""
    L = lib()
    _declare_psf(L)
    img = np.ascontiguousarray(img, np.float32)
    loc = np.ascontiguousarray(np.asarray(locations, np.float64).reshape(-1, 3))
    sx, sy, sz = (int(v) for v in size_xyz)
    raw = np.empty((sz, sy, sx), np.float32)
    out = np.empty_like(raw)
    _check(L.orc_extract_psf(_ptr(img), _ptr(_dims(img)), _ptr(loc), len(loc),
                             _ptr(np.array([sx, sy, sz], np.int64)), _ptr(raw), _ptr(out)))
    return raw, out


def psf_transform_geometry(size_xyz, model):
    L = lib()
    _declare_psf(L)
    ns = np.empty(3, np.int64)
    off = np.empty(3, np.float64)
    L.orc_psf_transform_geometry(_ptr(np.asarray(size_xyz, np.int64)),
                                 _ptr(np.ascontiguousarray(np.asarray(model, np.float64).reshape(12))), _ptr(ns),
                                 _ptr(off))
    return tuple(int(v) for v in ns), off


def psf_transform(psf, model, inverse):
    """transformPSF: the PSF resampled by the forward model into its odd box."""
    L = lib()
    _declare_psf(L)
    psf = np.ascontiguousarray(psf, np.float32)
    size = (psf.shape[2], psf.shape[1], psf.shape[0])
    ns, off = psf_transform_geometry(size, model)
    out = np.empty((ns[2], ns[1], ns[0]), np.float32)
    L.orc_psf_transform(_ptr(psf), _ptr(np.asarray(size, np.int64)),
                        _ptr(np.ascontiguousarray(np.asarray(inverse, np.float64).reshape(12))),
                        _ptr(np.asarray(ns, np.int64)), _ptr(off), _ptr(out))
    return out
