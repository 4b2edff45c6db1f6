"""Seeded synthetic multiview data for the benchmark workloads and the tests.

Backed by libmvd_synth.so (csrc/synth.cpp): Gaussian beads on a constant
background, per-view normalised anisotropic Gaussian PSFs rotated about y,
views = ground truth blurred by the view's PSF (continuous Gaussian
convolution sampled at voxel centres), Poisson-sampled.  Arrays are float32
indexed [z, y, x]; sizes are given (x, y, z).

The named workloads follow the deconvolution benchmark definitions:
  c1: 3 views of 128x128x64, PSF 19x19x31 (sxy 1.2, sz 4), 200 beads,
      10 iterations, lambda 0, whole image as one block.
  c2: 6 views of 512x512x256, PSF 31x31x51, 500 beads, blocks 256^3,
      lambda 0.006, 20 iterations.
  c4: 4 views of 256^3, PSF 31^3.
  c5: 8 views of 1024x1024x512, PSF 41x41x81 (sxy 1.5, sz 6), 2000 beads +
      low-frequency field, blocks 512^3, lambda 0.004, 25 iterations.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
import os
from typing import List, Optional, Sequence, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmvd_synth.so")
_lib = None
_P = ctypes.c_void_p


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from . import build
            build.build_synth()
        L = ctypes.CDLL(LIB_PATH)
        L.syn_beads.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                ctypes.c_double, ctypes.c_uint64, _P]
        L.syn_psf.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                              ctypes.c_double, _P]
        L.syn_render.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, ctypes.c_int, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, _P,
                                 ctypes.c_int, ctypes.c_uint64, ctypes.c_int, _P]
        L.syn_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, _P]
        L.syn_poisson.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_double, ctypes.c_int, _P]
        _lib = L
    return _lib


def _threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


def beads(n: int, size_xyz, seed: int, amp=(500.0, 2000.0)) -> np.ndarray:
    out = np.zeros((n, 4), np.float64)
    nx, ny, nz = size_xyz
    lib().syn_beads(n, nx, ny, nz, amp[0], amp[1], seed, out.ctypes.data_as(_P))
    return out


def psf(k_xyz, sxy: float, sz: float, angle_deg: float) -> np.ndarray:
    kx, ky, kz = k_xyz
    out = np.empty((kz, ky, kx), np.float32)
    if lib().syn_psf(kx, ky, kz, sxy, sz, angle_deg, out.ctypes.data_as(_P)) != 0:
        raise ValueError("degenerate PSF covariance")
    return out


def render(size_xyz, bead_table: np.ndarray, bead_sigma: float = 1.5, psf_sxy: float = 0.0, psf_sz: float = 0.0,
           angle_deg: float = 0.0, background: float = 100.0, field: Optional[np.ndarray] = None,
           poisson_seed: int = 0, threads: Optional[int] = None) -> np.ndarray:
    nx, ny, nz = size_xyz
    out = np.empty((nz, ny, nx), np.float32)
    f = None if field is None else np.ascontiguousarray(field, np.float64)
    rc = lib().syn_render(nx, ny, nz, bead_table.ctypes.data_as(_P), len(bead_table), bead_sigma, psf_sxy, psf_sz,
                          angle_deg, background, None if f is None else f.ctypes.data_as(_P),
                          0 if f is None else len(f), poisson_seed, threads or _threads(), out.ctypes.data_as(_P))
    if rc != 0:
        raise ValueError("degenerate covariance")
    return out


def uniform(seed: int, stream: int, n: int) -> np.ndarray:
    out = np.empty(n, np.float64)
    lib().syn_uniform(seed, stream, n, out.ctypes.data_as(_P))
    return out


def poisson_samples(seed: int, stream: int, lam: float, n: int) -> np.ndarray:
    out = np.empty(n, np.float64)
    lib().syn_poisson(seed, stream, lam, n, out.ctypes.data_as(_P))
    return out


# BlendingRealRandomAccess (spim/process/fusion/weights/BlendingRealRandomAccess.java:44-121):
# cosine blending with a 1001-entry lookup table; per axis the distance to the
# nearer face minus `border`, relative to `blending`.
def _blend_lut() -> np.ndarray:
    lut = np.zeros(1001, np.float64)
    d = 0.0
    while d <= 1.0001:
        lut[int(math.floor(d * 1000.0 + 0.5))] = (math.cos((1 - d) * math.pi) + 1) / 2
        d = d + 0.001
    return lut


def blending_weight(shape_zyx, border_xyz=(-8, -8, -3), blending_xyz=(12, 12, 12)) -> np.ndarray:
    """Blending weight of a view on its own (identity-transformed) grid."""
    lut = _blend_lut()
    f = []
    for d, ax in ((0, 2), (1, 1), (2, 0)):
        n = shape_zyx[ax]
        l = np.arange(n, dtype=np.float32)
        dist = np.maximum(np.float32(0), np.minimum(l - np.float32(border_xyz[d]),
                                                     np.float32(n - 1) - l - np.float32(border_xyz[d])))
        rel = dist / np.float32(blending_xyz[d])
        fa = np.where(rel < 1, lut[np.floor(rel.astype(np.float64) * 1000.0 + 0.5).clip(0, 1000).astype(int)], 1.0)
        f.append(np.where(dist == 0, 0.0, fa).astype(np.float64))
    fx, fy, fz = f
    # float products in the order x, y, z, each rounded to float32
    plane = (fx.astype(np.float32).astype(np.float64)[None, :] * fy[:, None]).astype(np.float32)
    w = np.empty(shape_zyx, np.float32)
    p64 = plane.astype(np.float64)
    for z in range(shape_zyx[0]):
        w[z] = (p64 * fz[z]).astype(np.float32)
    return w


@dataclasses.dataclass
class Dataset:
    size_xyz: Tuple[int, int, int]
    views: List[np.ndarray]
    psfs: List[np.ndarray]
    angles: List[float]
    ground_truth: Optional[np.ndarray] = None


def make_dataset(size_xyz, nviews: int, angles: Sequence[float], nbeads: int, k_xyz, sxy: float, sz: float,
                 seed: int, background: float = 100.0, field: Optional[np.ndarray] = None,
                 with_ground_truth: bool = False, poisson: bool = True, sz_per_view: Optional[Sequence[float]] = None,
                 threads: Optional[int] = None) -> Dataset:
    bt = beads(nbeads, size_xyz, seed)
    views, psfs = [], []
    for v in range(nviews):
        szv = sz if sz_per_view is None else sz_per_view[v]
        psfs.append(psf(k_xyz, sxy, szv, angles[v]))
        views.append(render(size_xyz, bt, 1.5, sxy, szv, angles[v], background, field,
                            poisson_seed=(seed * 1000003 + v + 1) if poisson else 0, threads=threads))
    gt = render(size_xyz, bt, 1.5, 0.0, 0.0, 0.0, background, field, 0, threads) if with_ground_truth else None
    return Dataset(tuple(size_xyz), views, psfs, list(angles), gt)


def low_frequency_field(seed: int, amplitude: float = 50.0, terms: int = 4) -> np.ndarray:
    u = uniform(seed, 0xF1E1D, terms * 4)
    rows = []
    for t in range(terms):
        fx, fy, fz, ph = u[4 * t:4 * t + 4]
        rows.append([amplitude / terms, 1 + int(fx * 3), 1 + int(fy * 3), 1 + int(fz * 2), 2 * math.pi * ph])
    return np.array(rows, np.float64)


WORKLOADS = {
    "c1": dict(size_xyz=(128, 128, 64), nviews=3, angles=[0, 120, 240], nbeads=200, k_xyz=(19, 19, 31),
               sxy=1.2, sz=4.0, seed=1, iterations=10, lam=0.0, block=None),
    "c2": dict(size_xyz=(512, 512, 256), nviews=6, angles=[0, 60, 120, 180, 240, 300], nbeads=500,
               k_xyz=(31, 31, 51), sxy=1.2, sz=4.0, seed=2, iterations=20, lam=0.006, block=(256, 256, 256)),
    "c4": dict(size_xyz=(256, 256, 256), nviews=4, angles=[0, 90, 180, 270], nbeads=3200, k_xyz=(31, 31, 31),
               sxy=1.2, sz=4.0, seed=4, iterations=50, lam=0.0, block=None),
    "c5": dict(size_xyz=(1024, 1024, 512), nviews=8, angles=[45 * k for k in range(8)], nbeads=2000,
               k_xyz=(41, 41, 81), sxy=1.5, sz=6.0, seed=5, iterations=25, lam=0.004, block=(512, 512, 512)),
}


def workload(name: str, with_ground_truth: bool = False, threads: Optional[int] = None) -> Tuple[Dataset, dict]:
    w = dict(WORKLOADS[name])
    field = low_frequency_field(w["seed"]) if name == "c5" else None
    ds = make_dataset(w["size_xyz"], w["nviews"], w["angles"], w["nbeads"], w["k_xyz"], w["sxy"], w["sz"],
                      w["seed"], field=field, with_ground_truth=with_ground_truth, threads=threads)
    return ds, w


def perturbed_psfs(k_xyz, sxy: float, sz: float, angles: Sequence[float], seed: int,
                   rel: float = 0.2) -> List[np.ndarray]:
    """Start PSFs of the refinement workload: every view's sigma_xy and sigma_z
    scaled by an independent seeded factor uniform in [1 - rel, 1 + rel]."""
    u = uniform(seed, 0xC4, 2 * len(angles))
    return [psf(k_xyz, sxy * (1 + rel * (2 * u[2 * i] - 1)), sz * (1 + rel * (2 * u[2 * i + 1] - 1)), a)
            for i, a in enumerate(angles)]


def c3_stack_specs(seed: int = 3, nstacks: int = 32):
    """Seeded stack parameters of the batched workload: per stack a size
    uniform in {96, 128, ..., 384} per axis, a view count uniform in {2..8},
    PSF sigma_z uniform in [3, 6] (sigma_xy 1.2, box 19x19x31) and beads at
    200 per million voxels."""
    u = uniform(seed, 0xC3, nstacks * 5)
    sizes = list(range(96, 385, 32))
    specs = []
    for i in range(nstacks):
        r = u[5 * i:5 * i + 5]
        size = tuple(sizes[min(len(sizes) - 1, int(r[j] * len(sizes)))] for j in range(3))
        nv = 2 + min(6, int(r[3] * 7))
        sz = 3.0 + 3.0 * r[4]
        nbeads = max(1, int(round(200 * size[0] * size[1] * size[2] / 1e6)))
        specs.append(dict(size_xyz=size, nviews=nv, sz=sz, sxy=1.2, k_xyz=(19, 19, 31), nbeads=nbeads,
                          angles=[360.0 * k / nv for k in range(nv)], seed=seed * 1000 + i))
    return specs


def c3_stack(spec, threads: Optional[int] = None) -> Dataset:
    return make_dataset(spec["size_xyz"], spec["nviews"], spec["angles"], spec["nbeads"], spec["k_xyz"], spec["sxy"],
                        spec["sz"], spec["seed"], threads=threads)
