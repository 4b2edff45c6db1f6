"""Oracle: separable 3D convolution of the reference's second native hook,
CUDASeparableConvolution.convolve_{7,15,31,63,127}
(spim/process/cuda/CUDASeparableConvolution.java:30-43), in float64 numpy.

TEST INFRASTRUCTURE ONLY: loaded by tests/ and bench.py's cpu_baseline leg as
the checker; the product library never calls it.

The native library behind that interface (SeparableConvolutionCUDALib) is not
in the reference tree; its contract is the interface's comments and its caller
(CUDASeparableConvolutionFunctions.gauss, :123-217): in-place convolution of a
w x h x d float image (x fastest) with one kernel of odd size N per axis,
centred at N/2, each axis optional, out of bounds 0 = zero, 1 = a given value,
2 = the nearest border pixel.  Restated here as the textbook separable
convolution out(x) = sum_j in(x - j) k(N/2 + j), j in [-N/2, N/2], applied
along x, then y, then z; the caller only passes symmetric Gaussian kernels
(Util.createGaussianKernel1DDouble), for which convolution and correlation
agree.  The arithmetic order of the external library is unknown (parity
unpinned at the float rounding level; compared within a tolerance).
"""
from __future__ import annotations

import numpy as np

ZERO, VALUE, EXTEND_BORDER_PIXELS = 0, 1, 2


def _axis(a: np.ndarray, k: np.ndarray, axis: int, oob: int, value: float) -> np.ndarray:
    n = len(k)
    r = n // 2
    pad = [(0, 0)] * a.ndim
    pad[axis] = (r, r)
    if oob == EXTEND_BORDER_PIXELS:
        p = np.pad(a, pad, mode="edge")
    else:
        p = np.pad(a, pad, mode="constant", constant_values=0.0 if oob == ZERO else float(value))
    out = np.zeros_like(a)
    length = a.shape[axis]
    for j in range(-r, r + 1):
        # out(x) += in(x - j) * k(r + j); in(x - j) sits at padded index x - j + r
        sl = [slice(None)] * a.ndim
        sl[axis] = slice(r - j, r - j + length)
        out += p[tuple(sl)] * float(k[r + j])
    return out


def separable_convolve(img, kx=None, ky=None, kz=None, oob: int = EXTEND_BORDER_PIXELS, value: float = 0.0):
    """img: [z, y, x] array; kernels of odd size or None (axis skipped).
    Returns the float64 result."""
    a = np.asarray(img, np.float64)
    if kx is not None:
        a = _axis(a, np.asarray(kx, np.float64), 2, oob, value)
    if ky is not None:
        a = _axis(a, np.asarray(ky, np.float64), 1, oob, value)
    if kz is not None:
        a = _axis(a, np.asarray(kz, np.float64), 0, oob, value)
    return a


def gaussian_kernel(sigma: float) -> np.ndarray:
    """Util.createGaussianKernel1DDouble(sigma, true) as mpicbg defines it:
    size = max(3, 2 * round(3 sigma) + 1), exp(-x^2 / (2 sigma^2)) normalised
    to unit sum (used to build realistic inputs; the convolution itself takes
    any kernel)."""
    size = max(3, 2 * int(np.floor(3 * sigma + 0.5)) + 1)
    x = np.arange(size) - size // 2
    k = np.exp(-(x.astype(np.float64) ** 2) / (2 * sigma * sigma))
    return k / k.sum()


def padded_kernel(k, size: int) -> np.ndarray:
    """CUDASeparableConvolutionFunctions.getFloatKernelPadded (:246-258):
    the kernel centred in `size` float zeros."""
    out = np.zeros(size, np.float32)
    s = (size - len(k)) // 2
    out[s:s + len(k)] = np.asarray(k, np.float32)
    return out
