"""GPU likelihood gradient (mvd_nll_grad) against the FP64 oracle gradient and
FP64 central finite differences of the PSF.

Tolerances: NLL relative 1e-6; PSF gradient vs FP64 central differences
(h = 1e-3) max|g - fd| / max|fd| <= 1e-3 over the probed voxels; psi gradient
vs the FP64 oracle max-relative 1e-4."""
import numpy as np
import pytest

from oracle import oracle as O
from spim_registration_amd import MVDeconFFT, MVDeconInput, MVDeconvolution, PSFTYPE
from spim_registration_amd import synthetic as S

pytestmark = pytest.mark.gpu


def _setup(size=(40, 36, 28), k=(9, 9, 11), nv=3, seed=21, perturb=0.2):
    ds = S.make_dataset(size, nv, [0, 120, 240, 60][:nv], 40, k, 1.2, 3.0, seed=seed)
    rng = np.random.default_rng(seed)
    start = [S.psf(k, 1.2 * (1 + perturb * rng.uniform(-1, 1)), 3.0 * (1 + perturb * rng.uniform(-1, 1)), a)
             for a in ds.angles]
    inp = MVDeconInput()
    for v, p in zip(ds.views, start):
        inp.add(MVDeconFFT(v, None, p, [0]))
    d = MVDeconvolution(inp, PSFTYPE.EFFICIENT_BAYESIAN, 2, 0.0)
    psi = d.get_psi(masked=False)
    return ds, start, d, psi


def test_gradient_matches_fp64_oracle():
    ds, start, d, psi = _setup()
    print("nll_grad gpu", flush=True)
    nll, gpsi, gpsf = d.nll_grad()
    ref_nll, ref_gpsi, ref_gpsf = O.nll_grad([v.astype(np.float64) for v in ds.views],
                                             [p.astype(np.float64) for p in start], psi.astype(np.float64))
    assert abs(nll - ref_nll) / abs(ref_nll) < 1e-6
    for g, r in zip(gpsf, ref_gpsf):
        assert np.abs(g - r).max() / np.abs(r).max() < 1e-4
    assert np.abs(gpsi - ref_gpsi).max() / np.abs(ref_gpsi).max() < 1e-4
    d.close()


def test_psf_gradient_matches_central_differences():
    ds, start, d, psi = _setup(seed=22)
    _, _, gpsf = d.nll_grad(want_psi=False)
    rng = np.random.default_rng(4)
    for v in range(len(start)):
        vox = rng.choice(start[v].size, 64, replace=False)
        fd = O.nll_psf_fd(ds.views[v], start[v], psi, vox, h=1e-3)
        g = gpsf[v].ravel()[vox].astype(np.float64)
        err = np.abs(g - fd).max() / np.abs(fd).max()
        print(f"view {v}: PSF gradient vs FP64 central differences, max rel err {err:.2e}", flush=True)
        assert err < 1e-3
    d.close()


def test_explicit_psi_and_psfs_arguments():
    ds, start, d, psi = _setup(seed=23)
    other = psi * np.float32(1.1)
    nll1, g1, _ = d.nll_grad(psi=other, psfs=ds.psfs, want_psf=False)
    ref_nll, ref_g, _ = O.nll_grad([v.astype(np.float64) for v in ds.views], [p.astype(np.float64) for p in ds.psfs],
                                   other.astype(np.float64))
    assert abs(nll1 - ref_nll) / abs(ref_nll) < 1e-6
    assert np.abs(g1 - ref_g).max() / np.abs(ref_g).max() < 1e-4
    d.close()


@pytest.mark.parametrize("block", [(32, 30, 24), (256, 256, 24), (64, 64, 64)])
def test_blocked_gradient_matches_oracle(block):
    """With blocks the gradient runs block by block (residual everywhere first,
    then per block a recomputed window spectrum); it must equal the whole-image
    FP64 oracle.  (256, 256, 24) takes the fused plane kernels; the 24-point
    z passes the gradient z kernel, which sums the lag-shifted correlation
    spectra over the blocks; (64, 64, 64) the register z kernels and the
    separate spectrum products."""
    k = (9, 7, 11)
    ds = S.make_dataset((60, 50, 40), 3, [0, 120, 240], 40, k, 1.2, 3.0, seed=24)
    inp = MVDeconInput()
    for v, p in zip(ds.views, ds.psfs):
        inp.add(MVDeconFFT(v, None, p, [0], True, block))
    d = MVDeconvolution(inp, PSFTYPE.EFFICIENT_BAYESIAN, 2, 0.0)
    print("blocked gradient", block, d.geometry(0), flush=True)
    psi = d.get_psi(masked=False)
    kernels = [v.get_kernel1() for v in d.get_data().get_views()]
    nll, gpsi, gpsf = d.nll_grad(psfs=kernels)
    ref_nll, ref_gpsi, ref_gpsf = O.nll_grad([v.astype(np.float64) for v in ds.views],
                                             [p.astype(np.float64) for p in kernels], psi.astype(np.float64))
    assert abs(nll - ref_nll) / abs(ref_nll) < 1e-6
    for g, r in zip(gpsf, ref_gpsf):
        assert np.abs(g - r).max() / np.abs(r).max() < 1e-4
    assert np.abs(gpsi - ref_gpsi).max() / np.abs(ref_gpsi).max() < 1e-4
    # the device-resident variant gives the same values
    nll2, none, gpsf2 = d.nll_grad(psfs=kernels, psi_on_device=True)
    assert none is None and nll2 == pytest.approx(nll, rel=1e-12)
    np.testing.assert_allclose(d.grad_psi(), gpsi, rtol=0, atol=1e-6 * np.abs(gpsi).max())
    for a, b in zip(gpsf, gpsf2):
        np.testing.assert_array_equal(a, b)
    d.close()
