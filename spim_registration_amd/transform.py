"""Input preparation on the GPU: each view resampled into the fused bounding
box, with its blending weight at the same position.

Reference (src/main/java/spim/process/fusion/):
  ProcessForDeconvolution (deconvolution/ProcessForDeconvolution.java:187-277):
      per view, a new bounding-box image (and weight image), offset = the
      bounding box's min, then one of
  TransformInputAndWeights (deconvolution/TransformInputAndWeights.java:100-135):
      image + PRECOMPUTED blending weights,
  TransformInput (deconvolution/TransformInput.java:69-95): image only
      (NO_WEIGHTS, VIRTUAL_WEIGHTS, LOAD_WEIGHTS),
  TransformWeights (deconvolution/TransformWeights.java:53-91): weights only;
  Blending / BlendingRealRandomAccess (weights/Blending.java:50-56,
      weights/BlendingRealRandomAccess.java:44-121): cosine blending with a
      1001-entry lookup table.

The image becomes max(minValue, trilinear interpolation over the mirror-single
extension) where the mapped position lies inside the raw view, 0 elsewhere;
the weight is the blending function of the raw view at the mapped position.
Positions map from the bounding box to the raw view by the inverse of the
view's registration: pass it row-packed, 3 x 4 (a Java caller passes
AffineTransform3D.inverse().getRowPackedCopy(), so the arithmetic starts from
the same doubles as the reference); `inverse_affine()` computes it from a
forward map for Python callers.  Volumes are float32 [z, y, x]; sizes and
offsets are (x, y, z).
"""
from __future__ import annotations

import ctypes
import dataclasses
from typing import Optional, Sequence, Tuple

import numpy as np

from . import _native


@dataclasses.dataclass(frozen=True)
class Blending:
    """Blending border and range per axis (x, y, z); the reference's defaults
    (EfficientBayesianBased.java:96-97, :694-696): border -8, -8, round(-8/2.5),
    range 12."""
    border: Tuple[float, float, float] = (-8.0, -8.0, -3.0)
    range: Tuple[float, float, float] = (12.0, 12.0, 12.0)

    def arrays(self):
        b = np.asarray(self.border, np.float32)
        r = np.asarray(self.range, np.float32)
        if b.shape != (3,) or r.shape != (3,):
            raise ValueError("border and range need 3 entries (x, y, z)")
        if not np.all(r > 0):
            raise ValueError("blending range must be positive")
        return b, r


def inverse_affine(affine) -> np.ndarray:
    """Row-packed 3 x 4 inverse of a 3 x 4 (or 4 x 4) affine map, float64."""
    a = np.asarray(affine, np.float64)
    if a.shape == (12,):
        a = a.reshape(3, 4)
    if a.shape == (3, 4):
        a = np.vstack([a, [0.0, 0.0, 0.0, 1.0]])
    if a.shape != (4, 4):
        raise ValueError("expected a 3x4 or 4x4 affine map")
    return np.ascontiguousarray(np.linalg.inv(a)[:3].reshape(12))


def _inv12(inverse) -> np.ndarray:
    m = np.ascontiguousarray(np.asarray(inverse, np.float64).reshape(-1))
    if m.shape != (12,):
        raise ValueError("the inverse affine map needs 12 values (row-packed 3x4)")
    if not np.all(np.isfinite(m)):
        raise ValueError("the inverse affine map must be finite")
    return m


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def transform_view(raw: Optional[np.ndarray], inverse, offset: Sequence[int], size_xyz: Sequence[int],
                   blending: Optional[Blending] = None, want_image: bool = True, raw_size_xyz=None,
                   device: int = 0):
    """One view into the bounding box (size_xyz, min = offset) on the GPU.
    Returns (image or None, weight or None).  want_image=False with a blending
    is TransformWeights (then raw may be None if raw_size_xyz gives its size)."""
    inv = _inv12(inverse)
    off = np.ascontiguousarray(np.asarray(offset, np.int64).reshape(3))
    nx, ny, nz = (int(v) for v in size_xyz)
    if raw is not None:
        raw = np.ascontiguousarray(raw, np.float32)
        if raw.ndim != 3:
            raise ValueError("expected a 3D [z, y, x] raw view")
        rz, ry, rx = raw.shape
    elif raw_size_xyz is not None and not want_image:
        rx, ry, rz = (int(v) for v in raw_size_xyz)
    else:
        raise ValueError("the raw view is needed")
    img = np.empty((nz, ny, nx), np.float32) if want_image else None
    w = None
    b = r = None
    if blending is not None:
        b, r = blending.arrays()
        w = np.empty((nz, ny, nx), np.float32)
    if img is None and w is None:
        raise ValueError("nothing to compute")
    _native.check(_native.lib().mvd_transform_view(device, _p(raw) if want_image else None, rx, ry, rz, _p(inv),
                                                   _p(off), nx, ny, nz, _p(b), _p(r), _p(img), _p(w)),
                  "mvd_transform_view")
    return img, w


def bounding_box(raw_sizes_xyz, affines) -> Tuple[Tuple[int, int, int], Tuple[int, int, int]]:
    """Integer bounding box (min, size) of all views mapped by their forward
    affine maps: the corners of every raw view, floor / ceil per axis."""
    lo = np.full(3, np.inf)
    hi = np.full(3, -np.inf)
    for (rx, ry, rz), a in zip(raw_sizes_xyz, affines):
        m = np.asarray(a, np.float64).reshape(-1)[:12].reshape(3, 4)
        for cx in (0, rx - 1):
            for cy in (0, ry - 1):
                for cz in (0, rz - 1):
                    p = m[:, :3] @ np.array([cx, cy, cz], np.float64) + m[:, 3]
                    lo = np.minimum(lo, p)
                    hi = np.maximum(hi, p)
    mn = np.floor(lo).astype(np.int64)
    mx = np.ceil(hi).astype(np.int64)
    return tuple(int(v) for v in mn), tuple(int(v) for v in (mx - mn + 1))


def transform_view_device(raw, inverse, offset: Sequence[int], size_xyz: Sequence[int], out_img, out_weight=None,
                          blending: Optional[Blending] = None, raw_size_xyz=None, stream=None, device: int = 0):
    """transform_view on device memory (mvd_transform_view_device): raw,
    out_img and out_weight are contiguous float32 torch tensors on the GPU
    (or None), [z, y, x]; the launch is queued on `stream` (a torch stream;
    None = torch's current stream) and the call does not wait."""
    import torch
    inv = _inv12(inverse)
    off = np.ascontiguousarray(np.asarray(offset, np.int64).reshape(3))
    nx, ny, nz = (int(v) for v in size_xyz)

    def ptr(t, shape=None):
        if t is None:
            return None
        if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
            raise ValueError("expected a contiguous float32 GPU tensor")
        if shape is not None and tuple(t.shape) != shape:
            raise ValueError(f"expected shape {shape}, got {tuple(t.shape)}")
        return ctypes.c_void_p(t.data_ptr())

    if raw is not None:
        rz, ry, rx = raw.shape
    else:
        rx, ry, rz = (int(v) for v in raw_size_xyz)
    b = r = None
    if blending is not None:
        b, r = blending.arrays()
    if out_weight is not None and blending is None:
        raise ValueError("a weight output needs a blending")
    st = (stream or torch.cuda.current_stream()).cuda_stream
    _native.check(_native.lib().mvd_transform_view_device(
        device, ptr(raw), rx, ry, rz, _p(inv), _p(off), nx, ny, nz, _p(b), _p(r), ptr(out_img, (nz, ny, nx)),
        ptr(out_weight, (nz, ny, nx)), ctypes.c_void_p(st)), "mvd_transform_view_device")
