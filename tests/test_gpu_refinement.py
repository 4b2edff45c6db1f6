"""PSF refinement on the GPU (refine_psfs: deconvolution iterations, GPU PSF
gradient, EM update, setKernel) against the same trajectory computed with the
oracle (float32 deconvolution, FP64 gradient).

Tolerances: PSFs after every step max|P_gpu - P_ref| / max|P_ref| <= 1e-4;
NLL per step relative 1e-6."""
import numpy as np
import pytest

from oracle import oracle as O
from spim_registration_amd import MVDeconFFT, MVDeconInput, MVDeconvolution, PSFTYPE
from spim_registration_amd import synthetic as S
from spim_registration_amd.refinement import em_psf_update, refine_psfs

pytestmark = pytest.mark.gpu


def test_refinement_trajectory_matches_oracle():
    k = (7, 7, 9)
    ds = S.make_dataset((28, 24, 20), 2, [0, 90], 20, k, 1.2, 3.0, seed=21)
    rng = np.random.default_rng(21)
    start = [S.psf(k, 1.2 * (1 + 0.2 * rng.uniform(-1, 1)), 3.0 * (1 + 0.2 * rng.uniform(-1, 1)), a)
             for a in ds.angles]
    inp = MVDeconInput()
    for v, p in zip(ds.views, start):
        inp.add(MVDeconFFT(v, None, p, [0]))
    d = MVDeconvolution(inp, PSFTYPE.EFFICIENT_BAYESIAN, 3, 0.0)
    print("refinement: 3 steps", flush=True)
    hist = refine_psfs(d, 3)
    gpu_psfs = [v.get_kernel1() for v in d.get_data().get_views()]

    psfs = [O.norm_img(p, 8) for p in start]
    psi, _ = O.deconvolve(ds.views, psfs, O.EFFICIENT_BAYESIAN, 3, 0.0)
    for h in hist:
        psi, _ = O.deconvolve(ds.views, psfs, O.EFFICIENT_BAYESIAN, 1, 0.0, psi0=psi)
        nll, _, g = O.nll_grad([v.astype(np.float64) for v in ds.views], [p.astype(np.float64) for p in psfs],
                               psi.astype(np.float64))
        assert abs(h.nll - nll) / abs(nll) < 1e-6
        tot = psi.astype(np.float64).sum()
        psfs = [O.norm_img(em_psf_update(p, gg, tot), 8) for p, gg in zip(psfs, g)]
    for a, b in zip(gpu_psfs, psfs):
        err = np.abs(a.astype(np.float64) - b).max() / np.abs(b).max()
        print(f"PSF after 3 steps: max rel err {err:.2e}", flush=True)
        assert err < 1e-4
    d.close()


def test_psi_sum_and_weighted_totals():
    """mvd_psi_sum equals the double sum of the fetched estimate; unit weights
    in refine_psfs (host totals) give the trajectory of no weights (device
    total)."""
    k = (7, 7, 9)
    ds = S.make_dataset((28, 24, 20), 2, [0, 90], 20, k, 1.2, 3.0, seed=22)
    runs = []
    for weights in (None, [np.ones((20, 24, 28), np.float32), None]):
        inp = MVDeconInput()
        for v, p in zip(ds.views, ds.psfs):
            inp.add(MVDeconFFT(v, None, p, [0]))
        d = MVDeconvolution(inp, PSFTYPE.EFFICIENT_BAYESIAN, 2, 0.0)
        tot = d.psi_sum()
        ref = d.get_psi(masked=False).astype(np.float64).sum()
        assert abs(tot - ref) <= 1e-12 * abs(ref)
        hist = refine_psfs(d, 2, weights=weights)
        runs.append(([v.get_kernel1() for v in d.get_data().get_views()], [h.nll for h in hist]))
        d.close()
    for a, b in zip(runs[0][0], runs[1][0]):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-6 * np.abs(b).max())
    np.testing.assert_allclose(runs[0][1], runs[1][1], rtol=1e-12)


def test_parametric_refinement_corrects_view_differences():
    """Gaussian model per view (sigma_xy, sigma_z), views at 0/90/180/270
    degrees with independently perturbed widths: every step keeps the NLL at
    the iteration's kernels from rising, the geometric mean of each width over
    the views is kept (the common scale the likelihood cannot identify), the
    spread of the widths over the views shrinks and the PSFs move towards the
    ones the data were made with."""
    from spim_registration_amd.refinement import gaussian_psf_model, refine_psfs_parametric
    k, angles = (15, 15, 15), [0, 90, 180, 270]
    ds = S.make_dataset((48, 48, 48), 4, angles, 40, k, 1.2, 3.0, seed=31)
    truth = [p.astype(np.float64) / p.sum(dtype=np.float64) for p in ds.psfs]
    start = np.array([[1.2 * f, 3.0 * g] for f, g in ((0.85, 1.15), (1.15, 0.9), (1.0, 1.1), (1.1, 0.85))])
    model = gaussian_psf_model(k, angles)
    inp = MVDeconInput()
    for v, a in enumerate(angles):
        inp.add(MVDeconFFT(ds.views[v], None, model(v, start[v]), [0]))
    d = MVDeconvolution(inp, PSFTYPE.EFFICIENT_BAYESIAN, 2, 0.0)

    def err():
        ks = [v.get_kernel1().astype(np.float64) for v in d.get_data().get_views()]
        return float(np.mean([np.linalg.norm(a / a.sum() - b) / np.linalg.norm(b) for a, b in zip(ks, truth)]))

    e0 = err()
    print("parametric refinement: 20 steps", flush=True)
    hist = refine_psfs_parametric(d, start, model, 20, iterations=1)
    e1 = err()
    nll = [h.nll for h in hist]
    theta = hist[-1].params
    print(f"PSF error {e0:.4f} -> {e1:.4f}; widths {np.round(theta, 3).tolist()}", flush=True)
    assert all(b <= a + 1e-9 * abs(a) for a, b in zip(nll, nll[1:]))
    np.testing.assert_allclose(np.exp(np.log(theta).mean(axis=0)), np.exp(np.log(start).mean(axis=0)), rtol=1e-9)
    assert np.all(np.log(theta).std(axis=0) < 0.6 * np.log(start).std(axis=0))
    assert e1 < 0.8 * e0
    d.close()
