"""PSF refinement driven by the likelihood gradient.

The reference path has the hook (MVDeconFFT.setKernel re-runs the kernel
preparation, deconv/MVDeconFFT.java:364-369) but no refinement; this module
adds it on top of the GPU gradient (mvd_nll_grad).

Forward model of view v: b_v = E_mirror(psi) * P_v, Poisson NLL
L = sum_v sum_x w_v(x) (b_v(x) - y_v(x) log b_v(x)).  Its PSF gradient is
dL/dP_v(j) = S_v(j) - C_v(j) with S_v(j) = sum_x w_v(x) E(psi)(x - j) and
C_v(j) = sum_x w_v(x) (y_v / b_v)(x) E(psi)(x - j).  A refinement step runs
Richardson-Lucy iterations with the current PSFs, then applies the
expectation-maximisation (blind Richardson-Lucy) update of every PSF,

    P_v <- P_v * C_v / S_v = P_v * (1 - eta * g_v / S_v),

with S_v(j) taken as the total weighted estimate (exact away from the image
border, where the mirror extension repeats the estimate), eta = 1 the plain
EM step, clipped at zero and renormalised to unit sum like every kernel the
reference accepts (normImg in MVDeconFFT.init).

The likelihood alone does not fix the PSFs: the estimate blurred by a common
kernel G with every PSF sharpened by G fits the data about as well, so
descent on the free PSF values (or on a PSF model's width) drifts towards
sharper PSFs and a blurrier estimate.  What the data of several views do fix
is how the views' PSFs differ from one another, since every view sees the
same estimate.  `refine_psfs_parametric` uses that: the PSFs come from a model
with positive parameters per view (e.g. the widths of a rotated Gaussian), the
gradient in the log-parameters is the chain rule through the model of the GPU
PSF gradient, its mean over the views -- the common scale the likelihood
cannot identify -- is removed, and each step is backtracked until the NLL at
the current estimate does not rise.
"""
from __future__ import annotations

import dataclasses
from typing import List, Optional, Sequence

import numpy as np

from .deconvolution import MVDeconvolution
from . import synthetic


@dataclasses.dataclass
class RefinementStep:
    step: int
    nll: float          # NLL at the PSFs the step started from (after its RL iterations)
    max_change: float   # max |P_new - P_old| / max P_old over the views


def em_psf_update(psf: np.ndarray, grad: np.ndarray, total, eta: float = 1.0) -> np.ndarray:
    """One multiplicative step P * (1 - eta * g / S), clipped at zero and
    normalised to unit sum (float64 arithmetic, float32 result).  S is the
    scalar total of the (weighted) estimate or the exact per-voxel S(j)."""
    t = np.asarray(total, np.float64)
    if np.any(t <= 0):
        raise ValueError("the estimate must have a positive sum")
    p = psf.astype(np.float64) * (1.0 - eta * grad.astype(np.float64) / t)
    np.maximum(p, 0.0, out=p)
    s = p.sum()
    if not np.isfinite(s) or s <= 0:
        raise FloatingPointError("PSF update produced no positive mass")
    return (p / s).astype(np.float32)


def refine_psfs(decon: MVDeconvolution, steps: int, iterations: int = 1, eta: float = 1.0,
                weights: Optional[Sequence[Optional[np.ndarray]]] = None) -> List[RefinementStep]:
    """Interleave `iterations` deconvolution iterations and one PSF update per
    step, `steps` times.  Weights (optional, per view) enter S_v as the sum of
    w_v * psi; None means unit weights, the model mvd_nll_grad uses without
    weights."""
    views = decon.get_data().get_views()
    history: List[RefinementStep] = []
    for step in range(steps):
        decon.iterate(iterations)
        if weights is None:
            # S_v: the total estimate, the same for every view
            totals = [decon.psi_sum()] * len(views)
        else:
            psi = decon.get_psi(masked=False).astype(np.float64)
            totals = [float(psi.sum() if w is None else (psi * w).sum()) for w in weights]
        # the gradient at the kernels the iteration uses (kernel1: the given PSF
        # after normImg), not at the PSFs as handed in
        kernels = [view.get_kernel1() for view in views]
        nll, _, gpsf = decon.nll_grad(psfs=kernels, want_psi=False, want_psf=True)
        change = 0.0
        new_psfs = []
        for v in range(len(views)):
            old = kernels[v]
            new = em_psf_update(old, gpsf[v], totals[v], eta)
            change = max(change, float(np.abs(new - old).max() / max(np.abs(old).max(), 1e-30)))
            new_psfs.append(new)
        # every view prepared once from the new PSFs of all views (the compound
        # kernels couple the views)
        decon.set_kernels(new_psfs)
        history.append(RefinementStep(step, nll, change))
    return history


@dataclasses.dataclass
class ParametricStep:
    step: int
    nll: float              # NLL after the step's deconvolution iterations, at the PSFs it started from
    params: np.ndarray      # per view, after the step
    accepted: bool          # the backtracking found a step that does not raise the NLL


def gaussian_psf_model(k_xyz, angles: Sequence[float]):
    """Model of the synthetic PSFs: view v's PSF is a normalised anisotropic
    Gaussian (parameters sigma_xy, sigma_z) rotated by angles[v] about y, in
    the box k_xyz."""
    def model(view: int, params) -> np.ndarray:
        return synthetic.psf(k_xyz, float(params[0]), float(params[1]), angles[view])
    return model


def _norm_img(p: np.ndarray, threads: int) -> np.ndarray:
    """The kernel normalisation of the deconvolution (AdjustInput.normImg,
    deconv/AdjustInput.java:50-121): divided by its sum plus the sum of its
    first portion of 2 * threads (float64 sums; the reference's pairwise
    summation order moves only the last bits)."""
    q = p.astype(np.float64).ravel()
    first = q[: q.size // (2 * threads)].sum()
    return (q / (q.sum() + first)).astype(np.float32).reshape(p.shape)


def log_param_gradient(gpsf: Sequence[np.ndarray], theta, model, threads: int, rel_h: float = 1e-3) -> np.ndarray:
    """dL/dlog(theta_v,j) = theta_v,j sum_k dL/dP_v(k) dP_v(k)/dtheta_v,j for
    PSFs P_v = normImg(model(v, theta_v)), with dP/dtheta by central
    differences of the normalised model (step rel_h * theta_v,j).  gpsf: the
    PSF gradient per view at those PSFs (float64 arithmetic)."""
    theta = np.asarray(theta, np.float64)
    grad = np.zeros_like(theta)
    for v in range(theta.shape[0]):
        gv = np.asarray(gpsf[v], np.float64)
        for j in range(theta.shape[1]):
            h = rel_h * theta[v, j]
            up, dn = theta[v].copy(), theta[v].copy()
            up[j] += h
            dn[j] -= h
            dp = (_norm_img(model(v, up), threads).astype(np.float64) -
                  _norm_img(model(v, dn), threads).astype(np.float64)) / (2 * h)
            grad[v, j] = theta[v, j] * float((gv * dp).sum())
    return grad


def refine_psfs_parametric(decon: MVDeconvolution, params, model, steps: int, iterations: int = 1,
                           rate: float = 0.05, rel_h: float = 1e-3, max_halvings: int = 12) -> List[ParametricStep]:
    """Interleave `iterations` deconvolution iterations and one update of the
    PSF model parameters, `steps` times.  params: [views][n] positive
    parameters; model(view, params) -> PSF (the kernel box, float32).

    Per step: dL/dP_v from the GPU, dL/dlog(theta_v,j) = theta_v,j sum_k
    dL/dP_v(k) dP_v(k)/dtheta_v,j (central differences of the unit-sum model,
    step rel_h, through normImg), the mean over the views removed per
    parameter, and log theta <- log theta - t rate g / max_v |g| with t halved
    until the NLL at the current estimate does not rise.  The views' PSFs are set through set_kernels (normImg, kernel
    preparation of all views)."""
    views = decon.get_data().get_views()
    nt = decon.norm_threads
    theta = np.array(params, np.float64)
    if theta.ndim != 2 or theta.shape[0] != len(views) or np.any(theta <= 0):
        raise ValueError("params: one row of positive parameters per view")
    decon.set_kernels([model(v, theta[v]) for v in range(len(views))])
    history: List[ParametricStep] = []
    for step in range(steps):
        decon.iterate(iterations)
        # the kernels the iteration uses (kernel1 after normImg)
        cur = [view.get_kernel1() for view in views]
        nll, _, g = decon.nll_grad(psfs=cur, want_psi=False, want_psf=True)
        grad = log_param_gradient(g, theta, model, nt, rel_h)
        grad -= grad.mean(axis=0, keepdims=True)
        scale = np.abs(grad).max(axis=0)
        scale[scale == 0] = 1.0
        direction = -grad / scale
        accepted = False
        t = 1.0
        for _ in range(max_halvings):
            cand = theta * np.exp(t * rate * direction)
            new = [_norm_img(model(v, cand[v]), nt) for v in range(len(views))]
            n2 = decon.nll_grad(psfs=new, want_psi=False, want_psf=False)[0]
            if n2 <= nll:
                theta = cand
                accepted = True
                break
            t *= 0.5
        if accepted:
            decon.set_kernels([model(v, theta[v]) for v in range(len(views))])
        history.append(ParametricStep(step, nll, theta.copy(), accepted))
    return history
