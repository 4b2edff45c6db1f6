"""The separable-convolution oracle (oracle/sepconv.py) pinned against
scipy.ndimage (a library routine, not the reference) for the three
out-of-bounds strategies, plus closed forms: a delta kernel is the identity, a
shifted delta a shift, and a constant image stays constant under a unit-sum
kernel with border extension."""
import numpy as np
import pytest
from scipy import ndimage

from oracle import sepconv as SC


def _rand(shape, seed):
    return np.random.default_rng(seed).uniform(-1.0, 2.0, shape)


@pytest.mark.parametrize("oob,mode,cval", [(SC.ZERO, "constant", 0.0), (SC.VALUE, "constant", 3.5),
                                           (SC.EXTEND_BORDER_PIXELS, "nearest", 0.0)])
def test_matches_scipy(oob, mode, cval):
    img = _rand((9, 13, 17), 1)
    kx, ky, kz = _rand(7, 2), _rand(15, 3), _rand(7, 4)
    got = SC.separable_convolve(img, kx, ky, kz, oob, cval)
    exp = img
    for k, ax in ((kx, 2), (ky, 1), (kz, 0)):
        exp = ndimage.convolve1d(exp, k, axis=ax, mode=mode, cval=cval)
    np.testing.assert_allclose(got, exp, rtol=1e-12, atol=1e-12)


def test_closed_forms():
    img = _rand((6, 7, 8), 5)
    delta = np.zeros(7)
    delta[3] = 1
    np.testing.assert_array_equal(SC.separable_convolve(img, delta, delta, delta, SC.ZERO), img)
    shift = np.zeros(7)
    shift[4] = 1  # out(x) = in(x - 1)
    got = SC.separable_convolve(img, shift, None, None, SC.ZERO)
    np.testing.assert_array_equal(got[:, :, 1:], img[:, :, :-1])
    assert np.all(got[:, :, 0] == 0)
    const = np.full((5, 6, 7), 2.5)
    g = SC.gaussian_kernel(2.0)
    np.testing.assert_allclose(SC.separable_convolve(const, g, g, g, SC.EXTEND_BORDER_PIXELS), const, rtol=1e-12)


def test_padded_kernel():
    g = SC.gaussian_kernel(1.5)
    p = SC.padded_kernel(g, 31)
    assert len(p) == 31 and p[15] == np.float32(g[len(g) // 2]) and p[0] == 0
