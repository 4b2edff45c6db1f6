"""Output path of the deconvolved image (the ImgExport side of
EfficientBayesianBased.java:294-299, which hands the result over with the
display range (0, 1)).

  min_max       -- FusionHelper.minMax (FusionHelper.java:108-172) on the GPU
  u16_range     -- ExportSpimData2HDF5.updateAndGetMinMax (:227-253)
  convert_u16   -- RealUnsignedShortConverter (imglib2 6.1.0) as the HDF5
                   export applies it (ExportSpimData2HDF5.java:256-262)
  Save3dTIFF    -- Save3dTIFF (Save3dTIFF.java:40-185): the image written as
                   a TIFF stack named after its title, calibrated from the
                   bounding box
  write_tiff    -- the writer behind it (mvd_write_tiff)

Inputs are numpy arrays (host memory) or contiguous float32 torch tensors on
the GPU; an MVDeconvolution is exported from device memory.  The HDF5 writer
of the reference (BigDataViewer over libhdf5) has no counterpart: libhdf5 is
not available to this package.
"""
from __future__ import annotations

import ctypes
import math
import os
from typing import Optional, Sequence

import numpy as np

from . import _native

_P = ctypes.c_void_p


def _is_tensor(img) -> bool:
    return type(img).__module__.startswith("torch") and hasattr(img, "data_ptr")


def _tensor_ptr(t):
    import torch
    if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
        raise ValueError("expected a contiguous float32 GPU tensor")
    return ctypes.c_void_p(t.data_ptr())


def min_max(img, device: int = 0, stream=None):
    """(min, max) as float32 with Math.min/max semantics: a NaN anywhere makes
    both NaN, -0.0 is below +0.0, start values +/-Float.MAX_VALUE."""
    out = np.empty(2, np.float32)
    L = _native.lib()
    if _is_tensor(img):
        import torch
        st = (stream or torch.cuda.current_stream()).cuda_stream
        _native.check(L.mvd_min_max_device(img.device.index or 0, _tensor_ptr(img), img.numel(),
                                           out.ctypes.data_as(_P), ctypes.c_void_p(st)), "mvd_min_max_device")
    else:
        a = np.ascontiguousarray(img, np.float32)
        _native.check(L.mvd_min_max(device, a.ctypes.data_as(_P), a.size, out.ctypes.data_as(_P)), "mvd_min_max")
    return out[0], out[1]


def u16_range(minmax) -> tuple:
    """The 16-bit conversion range widened from (min, max); the widening of
    max uses the widened min."""
    mm = np.ascontiguousarray(minmax, np.float32).reshape(2)
    out = np.empty(2, np.float64)
    _native.check(_native.lib().mvd_u16_range(mm.ctypes.data_as(_P), out.ctypes.data_as(_P)), "mvd_u16_range")
    return float(out[0]), float(out[1])


def convert_u16(img, lo: float, hi: float, big_endian: bool = False, out=None, stream=None):
    """RealUnsignedShortConverter(lo, hi) on a float32 GPU tensor; returns a
    uint16 tensor of the same shape (raw bytes byte-swapped when
    big_endian).  Queued on `stream` (torch's current stream by default)."""
    import torch
    if out is None:
        out = torch.empty(img.shape, dtype=torch.uint16, device=img.device)
    if out.dtype != torch.uint16 or not out.is_contiguous() or out.numel() != img.numel():
        raise ValueError("out must be a contiguous uint16 tensor of the image's size")
    st = (stream or torch.cuda.current_stream()).cuda_stream
    _native.check(_native.lib().mvd_convert_u16_device(
        img.device.index or 0, _tensor_ptr(img), img.numel(), float(lo), float(hi), 1 if big_endian else 0,
        ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(st)), "mvd_convert_u16_device")
    return out


def calibration(bb_min: Sequence[int], downsampling: int = 1) -> tuple:
    """Save3dTIFF.java:103-108: origin -(bb.min(d) / downsampling) with Java
    long division (toward zero), pixel size = downsampling."""
    ds = int(downsampling)
    if ds <= 0:
        raise ValueError("downsampling must be positive")

    def jdiv(a: int) -> int:
        q = abs(a) // ds
        return q if a >= 0 else -q

    return tuple(float(-jdiv(int(m))) for m in bb_min[:3]) + (float(ds),)


def write_tiff(path: str, img, bits: int = 32, min: float = math.nan, max: float = math.nan,
               calibration: Optional[Sequence[float]] = None, big_endian: bool = True, bigtiff: bool = False,
               device: int = 0) -> str:
    """Write a [z, y, x] float32 image (numpy, or a GPU tensor, which must be
    complete on the device) as a TIFF stack; see mvd_write_tiff."""
    cal = None if calibration is None else np.ascontiguousarray(calibration, np.float64).reshape(4)
    L = _native.lib()
    if _is_tensor(img):
        import torch
        if img.dim() != 3:
            raise ValueError("expected a [z, y, x] image")
        torch.cuda.synchronize(img.device)
        nz, ny, nx = img.shape
        ptr, on_dev, device = _tensor_ptr(img), 1, img.device.index or 0
        keep = img
    else:
        keep = np.ascontiguousarray(img, np.float32)
        if keep.ndim != 3:
            raise ValueError("expected a [z, y, x] image")
        nz, ny, nx = keep.shape
        ptr, on_dev = keep.ctypes.data_as(_P), 0
    _native.check(L.mvd_write_tiff(device, ptr, on_dev, nx, ny, nz, int(bits), float(min), float(max),
                                   None if cal is None else cal.ctypes.data_as(_P), 1 if big_endian else 0,
                                   1 if bigtiff else 0, os.fsencode(path)), "mvd_write_tiff")
    del keep
    return path


class Save3dTIFF:
    """Save3dTIFF (Save3dTIFF.java:40-185): exportImage writes `title`.tif
    (the title as is when it already ends in .tif) into the directory."""

    def __init__(self, path: str, compress: bool = False):
        if compress:
            raise NotImplementedError("ZIP-compressed TIFF (FileSaver.saveAsZip) is not provided")
        self.path = path

    def file_name(self, title: str) -> str:
        return os.path.abspath(os.path.join(self.path, title if title.endswith(".tif") else title + ".tif"))

    def export_image(self, img, title: str, bb_min: Optional[Sequence[int]] = None, downsampling: int = 1,
                     min: float = math.nan, max: float = math.nan, bits: int = 32, big_endian: bool = True) -> str:
        """exportImage(img, bb, tp, vs, min, max): NaN min/max -> minMax of the
        image.  img: numpy array, GPU tensor or MVDeconvolution (its masked
        estimate, written from device memory)."""
        cal = None if bb_min is None else calibration(bb_min, downsampling)
        fn = self.file_name(title)
        if hasattr(img, "export_tiff"):
            return img.export_tiff(fn, bits=bits, min=min, max=max, calibration=cal, big_endian=big_endian)
        return write_tiff(fn, img, bits=bits, min=min, max=max, calibration=cal, big_endian=big_endian)
