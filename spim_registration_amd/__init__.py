"""MI355X-native efficient-Bayesian multiview deconvolution.

A drop-in for the MVDeconvolution / MVDeconFFT path of fiji/SPIM_Registration:
HIP kernels for gfx950 behind a plain C ABI (include/mvdecon.h), with a Python
mirror of the reference's operator API on top.
"""
from .deconvolution import (FFTPolicy, MVDeconBatch, MVDeconFFT, MVDeconInput,  # noqa: F401
                            MVDeconvolution, PSFTYPE, WeightType)
from .cuda import CUDAFourierConvolution  # noqa: F401

__all__ = ["MVDeconFFT", "MVDeconInput", "MVDeconvolution", "MVDeconBatch", "PSFTYPE", "WeightType",
           "FFTPolicy", "CUDAFourierConvolution"]
