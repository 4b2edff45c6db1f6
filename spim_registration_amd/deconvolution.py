"""Python mirror of the reference's deconvolution operator API, backed by the
GPU session of libmvdecon_hip (include/mvdecon.h).

Reference classes (src/main/java/spim/process/fusion/deconvolution/):
  MVDeconFFT      -- one view: image, weight, kernel, device list, blocks
                     (MVDeconFFT.java:79-172; getters :371-377; setKernel :364-369)
  MVDeconInput    -- the list of views (MVDeconInput.java:31-80)
  MVDeconvolution -- initial estimate, iterations, final mask
                     (MVDeconvolution.java:94-211; runIteration :349-465)

Volumes are numpy float32 arrays indexed [z, y, x] (x fastest, as the
reference's ArrayImg); block sizes are given (x, y, z) as in the reference.
Errors raise; there is no CPU fallback (device -1, the reference's CPU path,
is rejected).
"""
from __future__ import annotations

import ctypes
import enum
from typing import List, Optional, Sequence

import numpy as np

from . import _native


class PSFTYPE(enum.IntEnum):
    """MVDeconFFT.PSFTYPE (MVDeconFFT.java:49), same order."""
    OPTIMIZATION_II = 0
    OPTIMIZATION_I = 1
    EFFICIENT_BAYESIAN = 2
    INDEPENDENT = 3


class WeightType(enum.IntEnum):
    """How per-view weights are used (ProcessForDeconvolution.java:221-243)."""
    NO_WEIGHTS = 0
    AS_GIVEN = 1             # already normalised by the caller
    PRECOMPUTED_WEIGHTS = 2  # WeightNormalizer.ApplyDirectly on the GPU
    VIRTUAL_WEIGHTS = 3      # NormalizingRandomAccess, evaluated per voxel


def _f32(a) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim != 3:
        raise ValueError("expected a 3D [z, y, x] array")
    return a


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class MVDeconFFT:
    """One view of the deconvolution (MVDeconFFT.java:79-172)."""

    def __init__(self, image, weight, kernel, device_list: Sequence[int] = (0,), use_blocks: bool = False,
                 block_size: Optional[Sequence[int]] = None, save_memory: bool = False):
        self.image = _f32(image)
        self.weight = None if weight is None else _f32(weight)
        if self.weight is not None and self.weight.shape != self.image.shape:
            raise ValueError("weight and image differ in size")
        self.kernel1 = _f32(kernel)
        if any(s % 2 == 0 for s in self.kernel1.shape):
            raise ValueError("kernel sizes must be odd (BlockGeneratorFixedSizePrecise.java:43)")
        self.device_list = list(device_list)
        if not self.device_list or any(d < 0 for d in self.device_list):
            raise ValueError("only GPU devices (>= 0) are supported; the reference's CPU path (-1) is not")
        self.use_blocks = bool(use_blocks)
        if self.use_blocks:
            if block_size is None or len(block_size) != 3:
                raise ValueError("use_blocks needs a 3D block size (x, y, z)")
            self.block_size = tuple(int(b) for b in block_size)
        else:
            self.block_size = None
        self.save_memory = save_memory  # kernel spectra always stay resident in HBM
        self.kernel2: Optional[np.ndarray] = None
        self.num_views = 0
        self._session = None
        self._index = -1
        self._raw = None  # (raw view, inverse map, offset, blending) of a view resampled on the GPU

    @property
    def shape(self):
        return self.image.shape if self._raw is None else self._raw_shape

    @classmethod
    def from_raw(cls, raw, inverse, offset: Sequence[int], size_xyz: Sequence[int], kernel, blending=None,
                 device_list: Sequence[int] = (0,), use_blocks: bool = False,
                 block_size: Optional[Sequence[int]] = None) -> "MVDeconFFT":
        """A view given as acquired: resampled into the bounding box (size_xyz,
        min = offset) on the GPU when the deconvolution adds it, with its
        blending weight when `blending` (a transform.Blending) is given
        (ProcessForDeconvolution.java:187-277; TransformInputAndWeights /
        TransformInput).  `inverse`: row-packed 3x4 inverse registration."""
        from .transform import _inv12
        nx, ny, nz = (int(v) for v in size_xyz)
        placeholder = np.zeros((1, 1, 1), np.float32)
        v = cls(placeholder, None, kernel, device_list, use_blocks, block_size)
        v.image = None
        v._raw = (np.ascontiguousarray(raw, np.float32), _inv12(inverse),
                  np.ascontiguousarray(np.asarray(offset, np.int64).reshape(3)), blending)
        if v._raw[0].ndim != 3:
            raise ValueError("expected a 3D [z, y, x] raw view")
        v._raw_shape = (nz, ny, nx)
        v._has_blending = blending is not None
        return v

    def has_weight(self) -> bool:
        return self.weight is not None if self._raw is None else self._has_blending

    def _resample(self):
        from .transform import transform_view
        raw, inv, off, blending = self._raw
        nz, ny, nx = self._raw_shape
        dev = self.device_list[0]
        return transform_view(raw, inv, off, (nx, ny, nz), blending, device=dev)

    # getters of the reference (MVDeconFFT.java:371-377); a view given raw is
    # resampled on the GPU on first access
    def get_image(self):
        if self.image is None and self._raw is not None:
            self.image, w = self._resample()
            if self.weight is None:
                self.weight = w
        return self.image

    def get_weight(self):
        if self.weight is None and self._raw is not None and self._has_blending:
            self.get_image()
        return self.weight

    def get_kernel1(self):
        return self.kernel1

    def get_kernel2(self):
        return self.kernel2

    def set_num_views(self, n: int):
        self.num_views = n

    def set_kernel(self, kernel):
        """MVDeconFFT.setKernel (:364-369): new PSF, re-run init for this view."""
        k = _f32(kernel)
        if k.shape != self.kernel1.shape:
            raise ValueError("kernel size cannot change")
        if self._session is None:
            self.kernel1 = k
            return
        s = self._session
        _native.check(_native.lib().mvd_set_psf(s.handle, self._index, _ptr(k)), "mvd_set_psf")
        s._fetch_kernels()


class MVDeconInput:
    """List of views (MVDeconInput.java:31-80)."""

    min_value = np.float32(0.0001)

    def __init__(self):
        self.views: List[MVDeconFFT] = []

    def add(self, view: MVDeconFFT):
        self.views.append(view)
        for v in self.views:
            v.set_num_views(len(self.views))

    def get_views(self):
        return self.views

    def get_num_views(self):
        return len(self.views)


class _Session:
    """Owns one mvd_session handle."""

    def __init__(self, device: int, shape, nviews: int):
        z, y, x = shape
        h = ctypes.c_void_p()
        _native.check(_native.lib().mvd_create(device, x, y, z, nviews, ctypes.byref(h)), "mvd_create")
        self.handle = h
        self.shape = shape
        self.nviews = nviews
        self.views: List[MVDeconFFT] = []

    def add(self, view: MVDeconFFT):
        kz, ky, kx = view.kernel1.shape
        if view._raw is not None:
            raw, inv, off, blending = view._raw
            rz, ry, rx = raw.shape
            b = r = None
            if blending is not None:
                b, r = blending.arrays()
            _native.check(_native.lib().mvd_add_view_transformed(
                self.handle, _ptr(raw), rx, ry, rz, _ptr(inv), _ptr(off), _ptr(b), _ptr(r), _ptr(view.kernel1),
                kx, ky, kz), "mvd_add_view_transformed")
        else:
            _native.check(_native.lib().mvd_add_view(self.handle, _ptr(view.image), _ptr(view.weight),
                                                     _ptr(view.kernel1), kx, ky, kz), "mvd_add_view")
        view._session = self
        view._index = len(self.views)
        self.views.append(view)

    def _fetch_kernels(self):
        for i, v in enumerate(self.views):
            k1 = np.empty_like(v.kernel1)
            k2 = np.empty_like(v.kernel1)
            _native.check(_native.lib().mvd_get_kernels(self.handle, i, _ptr(k1), _ptr(k2)), "mvd_get_kernels")
            v.kernel1, v.kernel2 = k1, k2

    def close(self):
        if self.handle:
            _native.lib().mvd_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MVDeconvolution:
    """MVDeconvolution (MVDeconvolution.java:94-211): the constructor runs the
    whole deconvolution, like the reference; `get_psi()` returns the masked
    estimate.  `run=False` only prepares (kernels, initial estimate) so that
    `run_iteration()` can be driven step by step."""

    min_value = np.float32(0.0001)

    def __init__(self, views: MVDeconInput, iteration_type: PSFTYPE, num_iterations: int, lam: float,
                 osem_speedup: float = 1.0, osem_speedup_index: int = 0, name: str = "deconvolved",
                 norm_threads: int = 8, weight_type: Optional[WeightType] = None,
                 initial_psi=None, run: bool = True, fft_policy: int = 0):
        vs = views.get_views()
        if not vs:
            raise ValueError("no views")
        shape = vs[0].shape
        for v in vs:
            if v.shape != shape:
                raise ValueError("all views must have the same size")
        v0 = vs[0]
        self.name = name
        self.views = views
        self.norm_threads = int(norm_threads)
        self.num_iterations = int(num_iterations)
        self.lam = float(lam)
        self.i = 0
        self.stats: List[np.ndarray] = []
        self._s = _Session(v0.device_list[0], shape, len(vs))
        for v in vs:
            self._s.add(v)
        if weight_type is None:
            # blending weights resampled on the GPU are raw (ProcessForDeconvolution
            # normalises them: PRECOMPUTED); weights handed in are taken as given
            if all(v.has_weight() for v in vs):
                raw_blend = any(v._raw is not None for v in vs)
                weight_type = WeightType.PRECOMPUTED_WEIGHTS if raw_blend else WeightType.AS_GIVEN
            else:
                weight_type = WeightType.NO_WEIGHTS
        self.weight_type = WeightType(weight_type)
        mm = (ctypes.c_double * 2)()
        L = _native.lib()
        _native.check(L.mvd_set_weights(self._s.handle, int(self.weight_type), int(osem_speedup_index),
                                        float(osem_speedup), int(norm_threads), mm), "mvd_set_weights")
        self.min_overlapping_views, self.avg_overlapping_views = mm[0], mm[1]
        _native.check(L.mvd_set_fft_policy(self._s.handle, int(fft_policy)), "mvd_set_fft_policy")
        blk = None
        if v0.use_blocks:
            blk = (ctypes.c_int * 3)(*v0.block_size)
        _native.check(L.mvd_prepare(self._s.handle, int(iteration_type), self.lam,
                                    ctypes.cast(blk, ctypes.c_void_p) if blk is not None else None,
                                    int(norm_threads)), "mvd_prepare")
        self._s._fetch_kernels()
        avg = ctypes.c_double()
        p0 = None if initial_psi is None else _f32(initial_psi)
        _native.check(L.mvd_init_psi(self._s.handle, _ptr(p0), ctypes.byref(avg)), "mvd_init_psi")
        self.average = avg.value
        if run:
            self.iterate(self.num_iterations)

    # MVDeconvolution.runIteration (:349-352)
    def run_iteration(self):
        self.iterate(1)

    def iterate(self, n: int):
        if n <= 0:
            return
        nv = self._s.nviews
        st = np.zeros((n, nv, 2), np.float64)
        _native.check(_native.lib().mvd_iterate(self._s.handle, n, _ptr(st)), "mvd_iterate")
        self.stats.extend(list(st))
        self.i += n

    def get_current_iteration(self):
        return self.i

    def get_data(self):
        return self.views

    def get_name(self):
        return self.name

    def get_psi(self, masked: bool = True) -> np.ndarray:
        out = np.empty(self._s.shape, np.float32)
        _native.check(_native.lib().mvd_get_psi(self._s.handle, _ptr(out), 1 if masked else 0), "mvd_get_psi")
        return out

    def geometry(self, view: int = 0):
        g = (ctypes.c_int * 4)()
        _native.check(_native.lib().mvd_get_geometry(self._s.handle, view, g), "mvd_get_geometry")
        return dict(blocks=g[0], fft=(g[1], g[2], g[3]))

    def nll_grad(self, psi=None, psfs=None, want_psi: bool = True, want_psf: bool = True,
                 psi_on_device: bool = False):
        """Poisson NLL of the forward model (b_v = E_mirror(psi) * P_v) and its
        gradient w.r.t. the estimate and each view's PSF (mvd_nll_grad).
        psi / psfs default to the current estimate and the views' PSFs.
        psi_on_device keeps d/dpsi in device memory (fetch it with grad_psi()).
        Returns (nll, grad_psi or None, [grad_psf_v] or None)."""
        L = _native.lib()
        s = self._s
        p0 = None if psi is None else _f32(psi)
        kp = None
        keep = []
        if psfs is not None:
            ks = [_f32(k) for k in psfs]
            keep += ks
            kp = (ctypes.c_void_p * len(ks))(*[k.ctypes.data for k in ks])
        gpsf = [np.empty_like(v.kernel1) for v in s.views] if want_psf else None
        gp = None
        if gpsf is not None:
            gp = (ctypes.c_void_p * len(gpsf))(*[g.ctypes.data for g in gpsf])
        nll = ctypes.c_double()
        if psi_on_device:
            _native.check(L.mvd_nll_grad_device(s.handle, _ptr(p0), kp, ctypes.byref(nll), 1 if want_psi else 0, gp),
                          "mvd_nll_grad_device")
            return nll.value, None, gpsf
        gpsi = np.empty(s.shape, np.float32) if want_psi else None
        _native.check(L.mvd_nll_grad(s.handle, _ptr(p0), kp, ctypes.byref(nll), _ptr(gpsi), gp), "mvd_nll_grad")
        return nll.value, gpsi, gpsf

    def grad_psi(self) -> np.ndarray:
        """d/dpsi of the last nll_grad(..., psi_on_device=True) (mvd_get_grad_psi)."""
        out = np.empty(self._s.shape, np.float32)
        _native.check(_native.lib().mvd_get_grad_psi(self._s.handle, _ptr(out)), "mvd_get_grad_psi")
        return out

    def psi_sum(self) -> float:
        """Sum of the unmasked estimate in double precision, on the device
        (mvd_psi_sum)."""
        out = ctypes.c_double()
        _native.check(_native.lib().mvd_psi_sum(self._s.handle, ctypes.byref(out)), "mvd_psi_sum")
        return out.value

    def set_kernels(self, kernels):
        """setKernel for every view at once (mvd_set_psfs): the same kernels
        as calling MVDeconFFT.setKernel (:364-369) for all views twice, each
        view prepared once from the new PSFs of all views."""
        views = self._s.views
        if len(kernels) != len(views):
            raise ValueError("one kernel per view")
        ks = [_f32(k) for k in kernels]
        for k, v in zip(ks, views):
            if k.shape != v.kernel1.shape:
                raise ValueError("kernel size cannot change")
        arr = (ctypes.c_void_p * len(ks))(*[k.ctypes.data for k in ks])
        _native.check(_native.lib().mvd_set_psfs(self._s.handle, arr), "mvd_set_psfs")
        self._s._fetch_kernels()

    def export_tiff(self, path: str, bits: int = 32, min: float = 0.0, max: float = 1.0, calibration=None,
                    big_endian: bool = True, bigtiff: bool = False) -> str:
        """exporter.exportImage(getPsi(), bb, tp, vs, 0, 1)
        (EfficientBayesianBased.java:294-299) as a TIFF stack written from
        device memory (mvd_export_psi); NaN min/max -> minMax of the image."""
        import os
        cal = None if calibration is None else np.ascontiguousarray(calibration, np.float64).reshape(4)
        _native.check(_native.lib().mvd_export_psi(
            self._s.handle, os.fsencode(path), int(bits), float(min), float(max),
            None if cal is None else cal.ctypes.data_as(ctypes.c_void_p), 1 if big_endian else 0,
            1 if bigtiff else 0), "mvd_export_psi")
        return path

    def set_timing(self, on: bool):
        _native.check(_native.lib().mvd_set_timing(self._s.handle, 1 if on else 0), "mvd_set_timing")

    def timing(self):
        L = _native.lib()
        ms = (ctypes.c_double * 32)()
        cnt = (ctypes.c_longlong * 32)()
        n = L.mvd_get_timing(self._s.handle, ms, cnt, 32)
        return {L.mvd_timing_name(i).decode(): (ms[i], cnt[i]) for i in range(n)}

    def close(self):
        self._s.close()


class FFTPolicy(enum.IntEnum):
    """FFT shape of block windows (mvd_set_fft_policy)."""
    SMOOTH = 0    # smallest 5-smooth size per axis
    POW2 = 1      # smallest power of two per axis
    AUTOTUNE = 2  # both timed on the GPU, fastest kept


class MVDeconBatch:
    """Iterate several independent deconvolutions (stacks) together
    (mvd_batch_*): one launch per FFT pass and kernel class for all stacks,
    optionally replayed as a hipGraph.  Each stack's result equals running it
    alone.  The MVDeconvolution objects must be created with run=False and
    stay alive while the batch is used."""

    def __init__(self, deconvolutions: Sequence[MVDeconvolution], use_graph: bool = True, device: int = 0):
        L = _native.lib()
        h = ctypes.c_void_p()
        _native.check(L.mvd_batch_create(device, ctypes.byref(h)), "mvd_batch_create")
        self.handle = h
        self.stacks = list(deconvolutions)
        for d in self.stacks:
            _native.check(L.mvd_batch_add(h, d._s.handle), "mvd_batch_add")
        _native.check(L.mvd_batch_prepare(h, 1 if use_graph else 0), "mvd_batch_prepare")

    def iterate(self, n: int):
        """n batched iterations; returns per stack an array [n, views, 2] of statistics."""
        if n <= 0:
            return [np.zeros((0, d._s.nviews, 2)) for d in self.stacks]
        tot = sum(d._s.nviews for d in self.stacks)
        st = np.zeros((n, tot, 2), np.float64)
        _native.check(_native.lib().mvd_batch_iterate(self.handle, n, _ptr(st)), "mvd_batch_iterate")
        out, off = [], 0
        for d in self.stacks:
            nv = d._s.nviews
            out.append(st[:, off:off + nv, :].copy())
            d.stats.extend(list(out[-1]))
            d.i += n
            off += nv
        return out

    def close(self):
        if self.handle:
            _native.lib().mvd_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
