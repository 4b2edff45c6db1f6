"""PSF extraction from bead locations (ExtractPSF,
src/main/java/spim/process/fusion/deconvolution/ExtractPSF.java:60-608).

`extract_next_img` runs on the GPU (mvd_extract_psf): the image, extended
periodically, is interpolated trilinearly at every PSF voxel around every bead
and summed in bead order (extractPSFLocal, :374-415), normalised to [0, 1]
(normalize, :298-317) and resampled by the view's model into an odd box whose
centre is the transformed centre (transformPSF / transform, :325-458).  The
bookkeeping and the small averaging helpers (computeAveragePSF,
computeAverageTransformedPSF, computeMaxProjection, makeSameSize, commonSize)
are host code on PSF-sized arrays, with the reference's loop orders.
Volumes are float32 [z, y, x]; sizes, locations and models are (x, y, z).
"""
from __future__ import annotations

import ctypes
from typing import Dict, Hashable, List, Optional, Sequence

import numpy as np

from . import _native
from .transform import inverse_affine


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def transformed_psf_size(psf_size_xyz: Sequence[int], model) -> tuple:
    """Size (x, y, z) of the transformed PSF (transformPSF, :325-372)."""
    m = np.ascontiguousarray(np.asarray(model, np.float64).reshape(-1)[:12])
    size = np.asarray(psf_size_xyz, np.int32)
    out = np.empty(3, np.int32)
    _native.check(_native.lib().mvd_psf_transform_size(_p(size), _p(m), _p(out)), "mvd_psf_transform_size")
    return tuple(int(v) for v in out)


def extract_psf(img, locations, psf_size_xyz: Sequence[int], model=None, device: int = 0):
    """extractNextImg on the GPU.  Returns (normalised PSF in input calibration,
    transformed PSF or None when no model is given)."""
    img = np.ascontiguousarray(img, np.float32)
    if img.ndim != 3:
        raise ValueError("expected a 3D [z, y, x] image")
    loc = np.ascontiguousarray(np.asarray(locations, np.float64).reshape(-1, 3))
    sx, sy, sz = (int(v) for v in psf_size_xyz)
    size = np.array([sx, sy, sz], np.int32)
    psf = np.empty((sz, sy, sx), np.float32)
    out = m = inv = None
    if model is not None:
        m = np.ascontiguousarray(np.asarray(model, np.float64).reshape(-1)[:12])
        inv = inverse_affine(m)
        tx, ty, tz = transformed_psf_size((sx, sy, sz), m)
        out = np.empty((tz, ty, tx), np.float32)
    nz, ny, nx = img.shape
    _native.check(_native.lib().mvd_extract_psf(device, _p(img), nx, ny, nz, _p(loc) if len(loc) else None, len(loc),
                                                _p(size), _p(m), _p(inv), _p(psf), _p(out)), "mvd_extract_psf")
    return psf, out


class ExtractPSF:
    """The reference's PSF container (ExtractPSF.java:60-116)."""

    def __init__(self, device: int = 0):
        self.device = device
        self.view_ids: List[Hashable] = []
        self.point_spread_functions: Dict[Hashable, np.ndarray] = {}
        self.original_psfs: Dict[Hashable, np.ndarray] = {}
        self.map_view_ids: Dict[Hashable, Hashable] = {}

    # extractNextImg (:277-296)
    def extract_next_img(self, img, view_id: Hashable, model, locations, psf_size_xyz: Sequence[int]):
        original, transformed = extract_psf(img, locations, psf_size_xyz, model, self.device)
        self.view_ids.append(view_id)
        self.point_spread_functions[view_id] = transformed
        self.original_psfs[view_id] = original

    def get_view_id_mapping(self):
        return self.map_view_ids

    def get_psf_map(self):
        return self.point_spread_functions

    def get_transformed_psf(self, view_id: Hashable) -> np.ndarray:
        if view_id in self.point_spread_functions:
            return self.point_spread_functions[view_id]
        if view_id in self.map_view_ids:
            return self.point_spread_functions[self.map_view_ids[view_id]]
        raise KeyError(f"Cannot find PSF for {view_id!r}")

    def get_input_calibration_psfs(self):
        return self.original_psfs

    def get_view_ids_for_psfs(self):
        return self.view_ids

    # computeAveragePSF (:243-270): float sum of the input-calibration PSFs in view order
    def compute_average_psf(self) -> np.ndarray:
        first = self.original_psfs[self.view_ids[0]]
        avg = np.zeros_like(first)
        for v in self.view_ids:
            avg = (avg + self.original_psfs[v]).astype(np.float32)
        return avg

    # computeMaxDimTransformedPSF (:523-535)
    def compute_max_dim_transformed_psf(self):
        size = [0, 0, 0]
        for p in self.point_spread_functions.values():
            for d, n in enumerate((p.shape[2], p.shape[1], p.shape[0])):
                size[d] = max(size[d], n)
        return tuple(size)

    # computeAverageTransformedPSF (:194-241): every PSF added point-reflected
    # about the centres (loc -> psfCentre - loc + avgCentre), zero outside
    def compute_average_transformed_psf(self) -> np.ndarray:
        mx, my, mz = self.compute_max_dim_transformed_psf()
        avg = np.zeros((mz, my, mx), np.float32)
        ac = np.array([mx // 2, my // 2, mz // 2])
        for v in self.view_ids:
            p = self.point_spread_functions[v]
            pz, py, px = p.shape
            pc = np.array([px // 2, py // 2, pz // 2])
            for z in range(pz):
                tz = pc[2] - z + ac[2]
                if not 0 <= tz < mz:
                    continue
                for y in range(py):
                    ty = pc[1] - y + ac[1]
                    if not 0 <= ty < my:
                        continue
                    xs = np.arange(px)
                    tx = pc[0] - xs + ac[0]
                    ok = (tx >= 0) & (tx < mx)
                    avg[tz, ty, tx[ok]] = (avg[tz, ty, tx[ok]] + p[z, y, xs[ok]]).astype(np.float32)
        return avg

    # computeMaxProjection (:118-192): maximum along min_dim (x = 0, y = 1, z = 2;
    # < 0: the smallest dimension), as a 2D image of the remaining axes
    @staticmethod
    def compute_max_projection(img: np.ndarray, min_dim: int = -1) -> np.ndarray:
        dims = (img.shape[2], img.shape[1], img.shape[0])
        if min_dim < 0:
            min_dim = 0
            for d in range(3):
                if dims[d] < dims[min_dim]:
                    min_dim = d
        axis = 2 - min_dim
        return img.astype(np.float64).max(axis=axis).astype(img.dtype)


def common_size(images: Sequence[np.ndarray]) -> Optional[tuple]:
    """commonSize (:505-521): per-axis maximum of the sizes (x, y, z)."""
    if not images:
        return None
    size = [images[0].shape[2], images[0].shape[1], images[0].shape[0]]
    for im in images:
        for d, n in enumerate((im.shape[2], im.shape[1], im.shape[0])):
            size[d] = max(size[d], n)
    return tuple(size)


def make_same_size(img: np.ndarray, size_xyz: Sequence[int]) -> np.ndarray:
    """makeSameSize (:466-499): img centred in a box of size_xyz, padded with
    its minimum."""
    sx, sy, sz = (int(v) for v in size_xyz)
    mn = img.astype(np.float64).min()
    out = np.full((sz, sy, sx), np.float32(mn), img.dtype)
    nz, ny, nx = img.shape
    ox, oy, oz = sx // 2 - nx // 2, sy // 2 - ny // 2, sz // 2 - nz // 2
    for z in range(sz):
        iz = z - oz
        if not 0 <= iz < nz:
            continue
        for y in range(sy):
            iy = y - oy
            if not 0 <= iy < ny:
                continue
            xs = np.arange(sx)
            ix = xs - ox
            ok = (ix >= 0) & (ix < nx)
            out[z, y, xs[ok]] = img[iz, iy, ix[ok]]
    return out
