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
"""
from __future__ import annotations

import dataclasses
from typing import List, Optional, Sequence

import numpy as np

from .deconvolution import MVDeconvolution


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
