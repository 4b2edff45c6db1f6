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
    # dL/dP(j) = sum_x E(psi)(x + c - j) (1 - y / b): scaling the kernel by
    # 1 + e moves every PSF voxel's gradient by about e sum(y), here ~1e-3 of
    # the gradient per ulp of e.  The FP32 kernel normalisation is exact only
    # to its last bit, so the gradient is compared in two parts: its shape
    # (the constant offset removed) at 1e-4, and the offset within what two
    # ulps of the normalisation give.
    for v, (g, r) in enumerate(zip(gpsf, ref_gpsf)):
        diff = g.astype(np.float64) - r
        off = float(diff.mean())
        shape_err = float(np.abs(diff - off).max() / np.abs(r).max())
        bound = 2 * 2.0 ** -24 * float(ds.views[v].sum(dtype=np.float64))
        print(f"view {v}: dP shape err {shape_err:.2e}, offset {off:.3f} (2-ulp bound {bound:.3f})", flush=True)
        assert shape_err < 1e-4
        assert abs(off) < bound
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


@pytest.mark.parametrize("size,k", [((260, 20, 30), (29, 5, 5)), ((20, 260, 250), (5, 29, 39)),
                                    ((100, 100, 20), (29, 29, 9)), ((260, 270, 24), (29, 19, 5)),
                                    ((20, 472, 472), (5, 41, 41)), ((20, 162, 344), (5, 31, 41))])
def test_gradient_288_point_windows(size, k):
    """Windows of 288 points (image + kernel - 1) along x (a 144-point complex
    transform of the packed row) or along y and z take the 288 = 24 x 12
    register kernels, including the gradient z pass.  With register x and y
    passes (the last two 288 shapes: 128 x 128 and 288 x 288 windows) the
    pass-1 x inverse transforms the residual rows itself (no residual image).
    The shape with 512-point y and z windows takes the line kernels whose
    workgroups share 128-byte row lines in pairs, the gradient z pass too; the
    last one 192- and 384-point y and z windows."""
    # start PSFs off by up to 20 % in width, as in a refinement
    ds = S.make_dataset(size, 2, [0, 90], 12, k, 1.2, 2.0, seed=sum(size))
    start = S.perturbed_psfs(k, 1.2, 2.0, [0, 90], seed=sum(size))
    inp = MVDeconInput()
    for v, p in zip(ds.views, start):
        inp.add(MVDeconFFT(v, None, p, [0]))
    d = MVDeconvolution(inp, PSFTYPE.EFFICIENT_BAYESIAN, 2, 0.0)
    print("gradient 288", size, k, d.geometry(0), flush=True)
    psi = d.get_psi(masked=False)
    ref_psi, _ = O.deconvolve(ds.views, start, O.EFFICIENT_BAYESIAN, 2, 0.0, norm_threads=8)
    assert np.abs(psi - ref_psi).max() / np.abs(ref_psi).max() < 1e-4
    kernels = [v.get_kernel1() for v in d.get_data().get_views()]
    nll, gpsi, gpsf = d.nll_grad(psfs=kernels)
    ref_nll, ref_gpsi, ref_gpsf = O.nll_grad([v.astype(np.float64) for v in ds.views],
                                             [p.astype(np.float64) for p in kernels], psi.astype(np.float64))
    assert abs(nll - ref_nll) / abs(ref_nll) < 1e-6
    # dL/dP(j) = sum_x E(psi)(x + c - j) (1 - y / b): scaling the kernel by
    # 1 + e moves every PSF voxel's gradient by about e sum(y), here ~1e-3 of
    # the gradient per ulp of e.  The FP32 kernel normalisation is exact only
    # to its last bit, so the gradient is compared in two parts: its shape
    # (the constant offset removed) at 1e-4, and the offset within what two
    # ulps of the normalisation give.
    for v, (g, r) in enumerate(zip(gpsf, ref_gpsf)):
        diff = g.astype(np.float64) - r
        off = float(diff.mean())
        shape_err = float(np.abs(diff - off).max() / np.abs(r).max())
        bound = 2 * 2.0 ** -24 * float(ds.views[v].sum(dtype=np.float64))
        print(f"view {v}: dP shape err {shape_err:.2e}, offset {off:.3f} (2-ulp bound {bound:.3f})", flush=True)
        assert shape_err < 1e-4
        assert abs(off) < bound
    assert np.abs(gpsi - ref_gpsi).max() / np.abs(ref_gpsi).max() < 1e-4
    d.close()
