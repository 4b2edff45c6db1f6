"""Oracle: the output path of the deconvolved image, in numpy.

TEST INFRASTRUCTURE ONLY: loaded by tests/ and bench.py's cpu_baseline leg as
the checker; the product library never calls it.

  * min_max: FusionHelper.minMax (spim/process/fusion/FusionHelper.java:108-172),
    java.lang.Math.min/max over the image from +/-Float.MAX_VALUE: a NaN
    anywhere makes both NaN, -0.0 counts as smaller than +0.0;
  * u16_range: ExportSpimData2HDF5.updateAndGetMinMax
    (spim/process/fusion/export/ExportSpimData2HDF5.java:227-253), the widening
    of max reads the already widened min;
  * to_u16: imglib2's RealUnsignedShortConverter (imglib2 6.1.0, pinned in the
    reference pom; the class is not in the reference tree), used at
    ExportSpimData2HDF5.java:256-262: min(65535, (int)(max(0, (a - min) /
    (max - min) * 65535) + 0.5)) in double, (int) of NaN = 0;
  * calibration: Save3dTIFF.exportImage (Save3dTIFF.java:103-108), origin
    -(bb.min(d) / downsampling) in Java long division, pixel size =
    downsampling.

The TIFF bytes themselves follow ImageJ's FileSaver, a library outside the
reference tree: files are checked by reading them back with an independent
reader (PIL), not against this module.
"""
from __future__ import annotations

import numpy as np

FLOAT_MAX = float(np.finfo(np.float32).max)


def min_max(img) -> tuple:
    a = np.asarray(img, np.float32).ravel()
    if a.size == 0:
        return (np.float32(FLOAT_MAX), np.float32(-FLOAT_MAX))
    if np.isnan(a).any():
        return (np.float32(np.nan), np.float32(np.nan))
    mn = np.float32(min(FLOAT_MAX, float(a.min())))
    mx = np.float32(max(-FLOAT_MAX, float(a.max())))
    zeros = a[a == 0]
    if mn == 0:
        mn = np.float32(-0.0) if np.signbit(zeros).any() else np.float32(0.0)
    if mx == 0:
        mx = np.float32(0.0) if (~np.signbit(zeros)).any() else np.float32(-0.0)
    return (mn, mx)


def u16_range(minmax) -> tuple:
    mn, mx = float(minmax[0]), float(minmax[1])
    t = mn - ((mn + mx) / 2.0) * 0.1
    mn = t if t != t else max(0.0, t)
    mx = mx + ((mn + mx) / 2.0) * 0.1
    return (mn, mx)


def to_u16(img, lo: float, hi: float) -> np.ndarray:
    a = np.asarray(img, np.float32).astype(np.float64)
    with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
        x = (a - lo) / (hi - lo) * 65535.0
        nan = np.isnan(x)
        x = np.where(nan, 0.0, np.maximum(x, 0.0))
        big = x >= 65535.0
        u = np.floor(np.where(big, 0.0, x) + 0.5)
    return np.where(big, 65535, u).astype(np.uint16)


def _jdiv(a: int, b: int) -> int:
    """Java long division: truncation toward zero."""
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b > 0) else -q


def calibration(bb_min, downsampling: int) -> tuple:
    ds = int(downsampling)
    return (float(-_jdiv(int(bb_min[0]), ds)), float(-_jdiv(int(bb_min[1]), ds)),
            float(-_jdiv(int(bb_min[2]), ds)), float(ds))
