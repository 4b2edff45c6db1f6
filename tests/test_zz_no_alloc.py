"""The no-allocation block convolution (mvd_convolution3d_inplace_noalloc)
against the JNA in-place convolution it is a variant of
(convolution3DfftCUDAInPlace), both through the C ABI: two calls in a row,
the second on a smaller block, each equal to the JNA result within float
rounding (2e-6 of the maximum; the two share kernels, so the results are
expected to be identical) and, on the interior, to the oracle's linear
convolution.  Sizes past the scratch limit are rejected with an error.

Device arrays are made with torch (test infrastructure only)."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O
from spim_registration_amd import _native
from spim_registration_amd import cuda as C

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.abs(a.astype(np.float64) - b.astype(np.float64)).max() / np.abs(b.astype(np.float64)).max())


def _noalloc(blk, k):
    import torch
    d_im = torch.from_numpy(blk.copy()).to("cuda")
    d_k = torch.from_numpy(k.copy()).to("cuda")
    dims = np.array(blk.shape, np.int32)
    kd = np.array(k.shape, np.int32)
    stream = torch.cuda.current_stream().cuda_stream
    torch.cuda.synchronize()
    rc = _native.lib().mvd_convolution3d_inplace_noalloc(
        ctypes.c_void_p(d_im.data_ptr()), dims.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(d_k.data_ptr()),
        kd.ctypes.data_as(ctypes.c_void_p), 0, ctypes.c_void_p(stream))
    _native.check(rc, "mvd_convolution3d_inplace_noalloc")
    torch.cuda.synchronize()
    return d_im.cpu().numpy()


def _jna(blk, k):
    flat = blk.ravel().copy()
    C.CUDAFourierConvolution().convolution3DfftCUDAInPlace(flat, np.array(blk.shape, np.int32), k.ravel(),
                                                           np.array(k.shape, np.int32), 0)
    return flat.reshape(blk.shape)


def test_noalloc_equals_jna_entry_two_calls():
    rng = np.random.default_rng(12)
    for bshape, kshape in (((40, 48, 56), (5, 7, 9)), ((20, 24, 30), (3, 5, 5))):
        blk = rng.random(bshape).astype(np.float32)
        k = rng.random(kshape).astype(np.float32)
        k /= k.sum()
        print("no-alloc block", bshape, "kernel", kshape, flush=True)
        out = _noalloc(blk, k)
        ref = _jna(blk, k)
        err = _rel(out, ref)
        print(f"  vs JNA entry: max rel diff {err:.2e}, identical: {np.array_equal(out, ref)}", flush=True)
        assert err < 2e-6
        lin = O.fft_convolve(blk.astype(np.float64), k.astype(np.float64), O.OOB_ZERO)
        inner = tuple(slice(s // 2, b - s // 2) for s, b in zip(kshape, bshape))
        assert _rel(out[inner], lin[inner]) < 2e-6


def test_noalloc_rejects_blocks_past_the_limit():
    import torch
    d_im = torch.zeros(8, device="cuda")
    L = _native.lib()
    dims = np.array([8, 8, 130], np.int32)  # nx = 130 > 128
    kd = np.array([3, 3, 3], np.int32)
    rc = L.mvd_convolution3d_inplace_noalloc(ctypes.c_void_p(d_im.data_ptr()), dims.ctypes.data_as(ctypes.c_void_p),
                                             ctypes.c_void_p(d_im.data_ptr()), kd.ctypes.data_as(ctypes.c_void_p), 0,
                                             None)
    assert rc == -1
    assert "128" in L.mvd_last_error().decode()
    rc = L.mvd_convolution3d_inplace_noalloc(None, dims.ctypes.data_as(ctypes.c_void_p), None,
                                             kd.ctypes.data_as(ctypes.c_void_p), 0, None)
    assert rc == -1
