"""ctypes bindings of libmvdecon_hip.so (declared in include/mvdecon.h).

The library is the product path: there is no CPU fallback.  If it has not
been built, or cannot be loaded, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmvdecon_hip.so")

_P = ctypes.c_void_p
_lib = None

# Every symbol declared in include/mvdecon.h.
EXPORTS = (
    "getNumDevicesCUDA", "getNameDeviceCUDA", "getMemDeviceCUDA", "getFreeMemDeviceCUDA",
    "getCUDAcomputeCapabilityMajorVersion", "getCUDAcomputeCapabilityMinorVersion",
    "convolution3DfftCUDAInPlace", "convolution3DfftCUDA", "mvd_free", "mvd_convolution3d_inplace_noalloc",
    "convolve_7", "convolve_15", "convolve_31", "convolve_63", "convolve_127", "convolutionCPU",
    "mvd_create", "mvd_destroy", "mvd_add_view", "mvd_add_view_transformed", "mvd_transform_view",
    "mvd_transform_view_device", "mvd_extract_psf", "mvd_psf_transform_size",
    "mvd_set_weights", "mvd_prepare", "mvd_init_psi",
    "mvd_iterate", "mvd_get_psi", "mvd_get_kernels", "mvd_set_psf", "mvd_set_psfs", "mvd_psi_sum", "mvd_set_timing", "mvd_get_timing",
    "mvd_timing_name", "mvd_get_geometry", "mvd_last_error", "mvd_nll_grad", "mvd_nll_grad_device",
    "mvd_get_grad_psi", "mvd_set_fft_policy",
    "mvd_min_max", "mvd_min_max_device", "mvd_u16_range", "mvd_convert_u16_device", "mvd_write_tiff",
    "mvd_export_psi",
    "mvd_batch_create", "mvd_batch_add", "mvd_batch_prepare", "mvd_batch_iterate", "mvd_batch_destroy",
)


class NativeError(RuntimeError):
    pass


def _declare(L):
    L.getNumDevicesCUDA.restype = ctypes.c_int
    L.getNameDeviceCUDA.argtypes = [ctypes.c_int, ctypes.c_char_p]
    L.getNameDeviceCUDA.restype = None
    L.getMemDeviceCUDA.argtypes = [ctypes.c_int]
    L.getMemDeviceCUDA.restype = ctypes.c_longlong
    L.getFreeMemDeviceCUDA.argtypes = [ctypes.c_int]
    L.getFreeMemDeviceCUDA.restype = ctypes.c_longlong
    L.getCUDAcomputeCapabilityMajorVersion.argtypes = [ctypes.c_int]
    L.getCUDAcomputeCapabilityMinorVersion.argtypes = [ctypes.c_int]
    L.convolution3DfftCUDAInPlace.argtypes = [_P, _P, _P, _P, ctypes.c_int]
    L.convolution3DfftCUDAInPlace.restype = None
    L.convolution3DfftCUDA.argtypes = [_P, _P, _P, _P, ctypes.c_int]
    L.convolution3DfftCUDA.restype = ctypes.POINTER(ctypes.c_float)
    for n in (7, 15, 31, 63, 127):
        f = getattr(L, f"convolve_{n}")
        f.argtypes = [_P, _P, _P, _P] + [ctypes.c_int] * 7 + [ctypes.c_float, ctypes.c_int]
        f.restype = ctypes.c_int
    L.convolutionCPU.argtypes = [_P, _P, _P, _P] + [ctypes.c_int] * 7 + [ctypes.c_float]
    L.convolutionCPU.restype = None
    L.mvd_free.argtypes = [_P]
    L.mvd_free.restype = None
    L.mvd_convolution3d_inplace_noalloc.argtypes = [_P, _P, _P, _P, ctypes.c_int, _P]
    L.mvd_convolution3d_inplace_noalloc.restype = ctypes.c_int
    L.mvd_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P]
    L.mvd_destroy.argtypes = [_P]
    L.mvd_add_view.argtypes = [_P, _P, _P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.mvd_add_view_transformed.argtypes = [_P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P, _P, _P, _P,
                                           ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.mvd_transform_view.argtypes = [ctypes.c_int, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P, _P, _P]
    L.mvd_transform_view_device.argtypes = [ctypes.c_int, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P,
                                            ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P, _P, _P, _P]
    L.mvd_extract_psf.argtypes = [ctypes.c_int, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, ctypes.c_int, _P,
                                  _P, _P, _P, _P]
    L.mvd_psf_transform_size.argtypes = [_P, _P, _P]
    L.mvd_set_weights.argtypes = [_P, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, _P]
    L.mvd_prepare.argtypes = [_P, ctypes.c_int, ctypes.c_double, _P, ctypes.c_int]
    L.mvd_init_psi.argtypes = [_P, _P, _P]
    L.mvd_iterate.argtypes = [_P, ctypes.c_int, _P]
    L.mvd_get_psi.argtypes = [_P, _P, ctypes.c_int]
    L.mvd_get_kernels.argtypes = [_P, ctypes.c_int, _P, _P]
    L.mvd_set_psf.argtypes = [_P, ctypes.c_int, _P]
    L.mvd_set_psfs.argtypes = [_P, _P]
    L.mvd_psi_sum.argtypes = [_P, _P]
    L.mvd_min_max.argtypes = [ctypes.c_int, _P, ctypes.c_longlong, _P]
    L.mvd_min_max_device.argtypes = [ctypes.c_int, _P, ctypes.c_longlong, _P, _P]
    L.mvd_u16_range.argtypes = [_P, _P]
    L.mvd_convert_u16_device.argtypes = [ctypes.c_int, _P, ctypes.c_longlong, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_int, _P, _P]
    L.mvd_write_tiff.argtypes = [ctypes.c_int, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_double, ctypes.c_double, _P, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_char_p]
    L.mvd_export_psi.argtypes = [_P, ctypes.c_char_p, ctypes.c_int, ctypes.c_double, ctypes.c_double, _P,
                                 ctypes.c_int, ctypes.c_int]
    L.mvd_set_timing.argtypes = [_P, ctypes.c_int]
    L.mvd_get_timing.argtypes = [_P, _P, _P, ctypes.c_int]
    L.mvd_timing_name.argtypes = [ctypes.c_int]
    L.mvd_timing_name.restype = ctypes.c_char_p
    L.mvd_get_geometry.argtypes = [_P, ctypes.c_int, _P]
    L.mvd_last_error.restype = ctypes.c_char_p
    L.mvd_nll_grad.argtypes = [_P, _P, _P, _P, _P, _P]
    L.mvd_nll_grad_device.argtypes = [_P, _P, _P, _P, ctypes.c_int, _P]
    L.mvd_get_grad_psi.argtypes = [_P, _P]
    L.mvd_set_fft_policy.argtypes = [_P, ctypes.c_int]
    L.mvd_batch_create.argtypes = [ctypes.c_int, _P]
    L.mvd_batch_add.argtypes = [_P, _P]
    L.mvd_batch_prepare.argtypes = [_P, ctypes.c_int]
    L.mvd_batch_iterate.argtypes = [_P, ctypes.c_int, _P]
    L.mvd_batch_destroy.argtypes = [_P]


def lib():
    """The loaded library; raises NativeError if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeError(
                f"{LIB_PATH} is missing: build it with `python -m spim_registration_amd.build` "
                "(there is no CPU fallback)")
        # torch ships its own HIP runtime under the soname this library links
        # (libamdhip64.so.7); loaded first, it is the one copy both use.  Loaded
        # after this library, torch would bring a second runtime that cannot
        # open the device, so torch is imported first where it is installed.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        try:
            L = ctypes.CDLL(LIB_PATH)
        except OSError as e:
            raise NativeError(f"cannot load {LIB_PATH}: {e}") from e
        _declare(L)
        _lib = L
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().mvd_last_error().decode(errors="replace")
        raise NativeError(f"{what}: {msg} (code {rc})")
