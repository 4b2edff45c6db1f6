"""Writes convolution_fixtures.npz: linear 3D convolutions of one seeded image
with one seeded odd-sized kernel under the three out-of-bounds rules of the
deconvolution (mirror-single for the estimate, MVDeconFFTThreads.java:44,78;
the constant 1 for the ratio, MVDeconFFT.java:515; zero for the compound
kernels, MVDeconFFT.java:222-228), computed in float64 by scipy.ndimage.convolve
-- a library routine independent of the oracle and of the HIP path
(mode 'mirror' is d c b | a b c d | c b a, i.e. mirror-single; the kernel is
flipped, so this is a convolution centred at K // 2).

Run from the repository root: python tests/golden/make_golden.py
"""
import os

import numpy as np
import scipy.ndimage as ndi

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rng = np.random.default_rng(20240531)
    img = (50.0 + 100.0 * rng.random((14, 12, 10))).astype(np.float32)  # [z, y, x]
    ker = rng.random((5, 3, 7)).astype(np.float32)
    ker /= np.float32(ker.sum())
    a, k = img.astype(np.float64), ker.astype(np.float64)
    out = {
        "image": img,
        "kernel": ker,
        "mirror_single": ndi.convolve(a, k, mode="mirror"),
        "constant_one": ndi.convolve(a, k, mode="constant", cval=1.0),
        "zero": ndi.convolve(a, k, mode="constant", cval=0.0),
    }
    np.savez(os.path.join(HERE, "convolution_fixtures.npz"), **out)


if __name__ == "__main__":
    main()
