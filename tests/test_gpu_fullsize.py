"""Properties checked at the full 512x512x256 workload (6 views, PSF 31x31x51,
blocks of 256^3), where the oracle would take minutes:

  * scale equivariance, bit for bit: with lambda = 0 every step of an
    iteration commutes with multiplying the views by 2 (the first-iteration
    average, the FFT convolutions, the quotient img / blurred and the
    multiplicative update are exact under scaling by a power of two), so the
    estimate of the doubled views is exactly twice the estimate, and the
    per-view statistics are doubled;
  * blocked equals whole-image: the 18 blocks of 256^3 and one window of
    img + K - 1 compute the same operator, up to float rounding.
"""
import numpy as np
import pytest

from spim_registration_amd import MVDeconFFT, MVDeconInput, MVDeconvolution, PSFTYPE
from spim_registration_amd import synthetic as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c2():
    print("generating the 512x512x256 workload", flush=True)
    ds, w = S.workload("c2")
    return ds, w


def _run(views, psfs, iters, block):
    inp = MVDeconInput()
    for v, p in zip(views, psfs):
        inp.add(MVDeconFFT(v, None, p, [0], block is not None, block))
    d = MVDeconvolution(inp, PSFTYPE.EFFICIENT_BAYESIAN, iters, 0.0)
    psi = d.get_psi(masked=False)
    stats = np.array(d.stats)
    d.close()
    return psi, stats


def test_scale_equivariance_full_size(c2):
    ds, w = c2
    print("c2 x1 and x2, 2 iterations, blocks", w["block"], flush=True)
    psi1, st1 = _run(ds.views, ds.psfs, 2, w["block"])
    psi2, st2 = _run([2.0 * v for v in ds.views], ds.psfs, 2, w["block"])
    np.testing.assert_array_equal(psi2, 2.0 * psi1)
    np.testing.assert_array_equal(st2, 2.0 * st1)
    assert psi1.min() > 1e-3  # far from the minValue clamp, which is not scale-free


def test_blocked_equals_whole_image_full_size(c2):
    ds, w = c2
    print("c2 blocked vs whole image, 1 iteration", flush=True)
    blocked, _ = _run(ds.views, ds.psfs, 1, w["block"])
    whole, _ = _run(ds.views, ds.psfs, 1, None)
    err = float(np.abs(blocked.astype(np.float64) - whole).max() / np.abs(whole).max())
    print("max rel diff blocked vs whole", err, flush=True)
    assert err < 2e-5
