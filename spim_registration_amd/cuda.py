"""Python mirror of the reference's native-library glue (src/main/java/spim/process/cuda/).

  CUDAFourierConvolution       -- the JNA interface (CUDAFourierConvolution.java:25-32,
                                  CUDAStandardFunctions.java:26-45), calling the same
                                  C symbols JNA binds, with Java-style arrays
                                  (flat float32 / int32 numpy arrays, dims reversed).
  Block, BlockGeneratorFixedSizePrecise
                               -- block split, copy and paste (Block.java:47-423,
                                  BlockGeneratorFixedSizePrecise.java:30-124) on host arrays.
  convolve_blocks_jna          -- MVDeconFFTThreads.convolve{1,2}BlockCUDA
                                  (MVDeconFFTThreads.java:73-115): the reference's own
                                  per-block GPU loop, driven through the JNA symbols.
"""
from __future__ import annotations

import ctypes
from typing import List, Sequence

import numpy as np

from . import _native

OOB_MIRROR_SINGLE, OOB_VALUE_ONE, OOB_ZERO = 0, 1, 2


class CUDAFourierConvolution:
    """What `Native.load(path, CUDAFourierConvolution.class)` returns in the reference."""

    def __init__(self):
        self._L = _native.lib()

    def getNumDevicesCUDA(self) -> int:
        return self._L.getNumDevicesCUDA()

    def getNameDeviceCUDA(self, devCUDA: int, name: bytearray) -> None:
        buf = ctypes.create_string_buffer(256)
        self._L.getNameDeviceCUDA(devCUDA, buf)
        name[:256] = buf.raw[:256]

    def getMemDeviceCUDA(self, devCUDA: int) -> int:
        return self._L.getMemDeviceCUDA(devCUDA)

    def getFreeMemDeviceCUDA(self, devCUDA: int) -> int:
        return self._L.getFreeMemDeviceCUDA(devCUDA)

    def getCUDAcomputeCapabilityMajorVersion(self, devCUDA: int) -> int:
        return self._L.getCUDAcomputeCapabilityMajorVersion(devCUDA)

    def getCUDAcomputeCapabilityMinorVersion(self, devCUDA: int) -> int:
        return self._L.getCUDAcomputeCapabilityMinorVersion(devCUDA)

    @staticmethod
    def _arrays(im, imDim, kernel, kernelDim):
        if im.dtype != np.float32 or not im.flags["C_CONTIGUOUS"]:
            raise TypeError("im must be a contiguous float32 array (Java float[])")
        kernel = np.ascontiguousarray(kernel, dtype=np.float32)
        imDim = np.ascontiguousarray(imDim, dtype=np.int32)
        kernelDim = np.ascontiguousarray(kernelDim, dtype=np.int32)
        if im.size != int(np.prod(imDim, dtype=np.int64)) or kernel.size != int(np.prod(kernelDim, dtype=np.int64)):
            raise ValueError("array length does not match its dimensions")
        return im, imDim, kernel, kernelDim

    def convolution3DfftCUDAInPlace(self, im, imDim, kernel, kernelDim, devCUDA: int) -> None:
        im, imDim, kernel, kernelDim = self._arrays(im, imDim, kernel, kernelDim)
        p = ctypes.c_void_p
        self._L.convolution3DfftCUDAInPlace(im.ctypes.data_as(p), imDim.ctypes.data_as(p),
                                            kernel.ctypes.data_as(p), kernelDim.ctypes.data_as(p), devCUDA)

    def convolution3DfftCUDA(self, im, imDim, kernel, kernelDim, devCUDA: int):
        im, imDim, kernel, kernelDim = self._arrays(im, imDim, kernel, kernelDim)
        p = ctypes.c_void_p
        r = self._L.convolution3DfftCUDA(im.ctypes.data_as(p), imDim.ctypes.data_as(p),
                                         kernel.ctypes.data_as(p), kernelDim.ctypes.data_as(p), devCUDA)
        if not r:
            return None
        try:
            return np.ctypeslib.as_array(r, shape=(im.size,)).copy()
        finally:
            self._L.mvd_free(ctypes.cast(r, ctypes.c_void_p))


def _mirror_single(p: np.ndarray, n: int) -> np.ndarray:
    """imglib2 mirror-single out-of-bounds index map (period 2n-2)."""
    if n == 1:
        return np.zeros_like(p)
    period = 2 * n - 2
    q = np.mod(p, period)
    return np.where(q < n, q, period - q)


class Block:
    """cuda/Block.java:47-423 (3D, x-fastest arrays indexed [z, y, x])."""

    def __init__(self, block_size, offset, effective_size, effective_offset, effective_local_offset):
        self.block_size = tuple(block_size)
        self.offset = tuple(offset)
        self.effective_size = tuple(effective_size)
        self.effective_offset = tuple(effective_offset)
        self.effective_local_offset = tuple(effective_local_offset)

    def copy_block(self, source: np.ndarray, oob: int) -> np.ndarray:
        """Block.copyBlock (:134-173) from source extended by `oob`."""
        idx = []
        inside = []
        for d, ax in ((0, 2), (1, 1), (2, 0)):
            n = source.shape[ax]
            p = np.arange(self.offset[d], self.offset[d] + self.block_size[d])
            inside.append((p >= 0) & (p < n))
            idx.append(_mirror_single(p, n) if oob == OOB_MIRROR_SINGLE else np.clip(p, 0, n - 1))
        blk = source[np.ix_(idx[2], idx[1], idx[0])].astype(np.float32)
        if oob != OOB_MIRROR_SINGLE:
            mask = inside[2][:, None, None] & inside[1][None, :, None] & inside[0][None, None, :]
            blk = np.where(mask, blk, np.float32(1.0 if oob == OOB_VALUE_ONE else 0.0)).astype(np.float32)
        return np.ascontiguousarray(blk)

    def paste_block(self, target: np.ndarray, block: np.ndarray) -> None:
        """Block.pasteBlock (:175-215): effective region only."""
        lo, eo, es = self.effective_local_offset, self.effective_offset, self.effective_size
        target[eo[2]:eo[2] + es[2], eo[1]:eo[1] + es[1], eo[0]:eo[0] + es[0]] = \
            block[lo[2]:lo[2] + es[2], lo[1]:lo[1] + es[1], lo[0]:lo[0] + es[0]]


class BlockGeneratorFixedSizePrecise:
    """cuda/BlockGeneratorFixedSizePrecise.java:30-124."""

    def __init__(self, block_size: Sequence[int]):
        self.block_size = tuple(int(b) for b in block_size)

    def divide_into_blocks(self, img_size: Sequence[int], kernel_size: Sequence[int]) -> List[Block]:
        eff = [b - k + 1 for b, k in zip(self.block_size, kernel_size)]
        if any(e <= 0 for e in eff):
            raise ValueError("block size smaller than the kernel")
        local = [k // 2 for k in kernel_size]
        nb = [i // e + (1 if i % e else 0) for i, e in zip(img_size, eff)]
        blocks = []
        for k in range(nb[2]):
            for j in range(nb[1]):
                for i in range(nb[0]):
                    cur = (i, j, k)
                    eoff = [c * e for c, e in zip(cur, eff)]
                    off = [eo - ks // 2 for eo, ks in zip(eoff, kernel_size)]
                    esz = [e if eo + e <= n else n - eo for e, eo, n in zip(eff, eoff, img_size)]
                    blocks.append(Block(self.block_size, off, esz, eoff, local))
        return blocks


def convolve_blocks_jna(cuda: CUDAFourierConvolution, image: np.ndarray, kernel: np.ndarray, oob: int,
                        blocks: Sequence[Block], device: int = 0) -> np.ndarray:
    """convolve1BlockCUDA / convolve2BlockCUDA (MVDeconFFTThreads.java:73-115) for all blocks."""
    result = np.empty_like(image, dtype=np.float32)
    kz, ky, kx = kernel.shape
    kdim = np.array([kz, ky, kx], np.int32)  # getCUDACoordinates: reversed (:157-165)
    kflat = np.ascontiguousarray(kernel, dtype=np.float32).ravel()
    for b in blocks:
        blk = b.copy_block(image, oob)
        flat = blk.ravel()
        bx, by, bz = b.block_size
        cuda.convolution3DfftCUDAInPlace(flat, np.array([bz, by, bx], np.int32), kflat, kdim, device)
        b.paste_block(result, flat.reshape(bz, by, bx))
    return result


class CUDASeparableConvolution(CUDAFourierConvolution):
    """What `Native.load(path, CUDASeparableConvolution.class)` returns in the
    reference (CUDASeparableConvolution.java:30-43): in-place separable
    convolution with kernels of N in {7, 15, 31, 63, 127} taps."""

    def _convolve(self, n, image, kernelX, kernelY, kernelZ, imageW, imageH, imageD, convolveX, convolveY,
                  convolveZ, outofbounds, outofboundsvalue, devCUDA) -> bool:
        if image.dtype != np.float32 or not image.flags["C_CONTIGUOUS"]:
            raise TypeError("image must be a contiguous float32 array (Java float[])")
        if image.size != imageW * imageH * imageD:
            raise ValueError("array length does not match its dimensions")
        p = ctypes.c_void_p
        ks = []
        for k, on in ((kernelX, convolveX), (kernelY, convolveY), (kernelZ, convolveZ)):
            if k is None:
                ks.append(None)
                continue
            k = np.ascontiguousarray(k, dtype=np.float32)
            if on and k.size != n:
                raise ValueError(f"kernels of convolve_{n} have {n} taps")
            ks.append(k)
        f = getattr(self._L, f"convolve_{n}")
        return bool(f(image.ctypes.data_as(p), *[None if k is None else k.ctypes.data_as(p) for k in ks],
                      imageW, imageH, imageD, int(bool(convolveX)), int(bool(convolveY)), int(bool(convolveZ)),
                      int(outofbounds), float(outofboundsvalue), devCUDA))

    def convolve_7(self, *a) -> bool:
        return self._convolve(7, *a)

    def convolve_15(self, *a) -> bool:
        return self._convolve(15, *a)

    def convolve_31(self, *a) -> bool:
        return self._convolve(31, *a)

    def convolve_63(self, *a) -> bool:
        return self._convolve(63, *a)

    def convolve_127(self, *a) -> bool:
        return self._convolve(127, *a)


SUPPORTED_KERNEL_SIZES = (7, 15, 31, 63, 127)
OOB_ZERO_SEP, OOB_VALUE_SEP, OOB_EXTEND_BORDER_PIXELS = 0, 1, 2


def gauss(img: np.ndarray, dim: Sequence[int], kernels: Sequence[np.ndarray], oobs: int = OOB_EXTEND_BORDER_PIXELS,
          oobs_value: float = 0.0, cuda: CUDASeparableConvolution = None, device: int = 0) -> bool:
    """CUDASeparableConvolutionFunctions.gauss (:123-217) given the per-axis
    kernels (1 to 3, x first): every kernel is centred in zeros of the
    smallest supported size that holds the longest one
    (getCUDAKernels / getFloatKernelPadded, :219-258), then convolve_N."""
    n = len(dim)
    if n == 0 or n > 3 or len(kernels) != n:
        return False
    longest = max(len(k) for k in kernels)
    size = next((s for s in SUPPORTED_KERNEL_SIZES if longest <= s), None)
    if size is None:
        return False
    padded = []
    for k in kernels:
        out = np.zeros(size, np.float32)
        s0 = (size - len(k)) // 2
        out[s0:s0 + len(k)] = np.asarray(k, np.float32)
        padded.append(out)
    w = int(dim[0])
    h = int(dim[1]) if n > 1 else 1
    d = int(dim[2]) if n > 2 else 1
    kx = padded[0]
    ky = padded[1] if n > 1 else None
    kz = padded[2] if n > 2 else None
    cuda = cuda or CUDASeparableConvolution()
    return getattr(cuda, f"convolve_{size}")(img, kx, ky, kz, w, h, d, kx is not None, ky is not None,
                                            kz is not None, oobs, oobs_value, device)
