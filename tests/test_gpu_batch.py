"""Batched iteration of several stacks (mvd_batch_*) and FFT shape policies.

The batch runs the same kernels on the same data as the per-stack path, so
per stack the estimate must be bit-identical to mvd_iterate; the FFT shape
policies change the FFT size, so they are checked against the oracle gate."""
import numpy as np
import pytest

from oracle import oracle as O
from spim_registration_amd import FFTPolicy, MVDeconBatch, MVDeconFFT, MVDeconInput, MVDeconvolution, PSFTYPE
from spim_registration_amd import synthetic as S

pytestmark = pytest.mark.gpu

GATE = 1e-4


def _stack(size, nv, seed, k=(7, 7, 9), block=None, policy=0, run=False, iters=0):
    ds = S.make_dataset(size, nv, [0, 60, 120, 180, 240, 300, 30, 90][:nv], 20, k, 1.2, 2.5, seed=seed)
    inp = MVDeconInput()
    for v, p in zip(ds.views, ds.psfs):
        inp.add(MVDeconFFT(v, None, p, [0], block is not None, block))
    d = MVDeconvolution(inp, PSFTYPE.EFFICIENT_BAYESIAN, iters, 0.006, run=run, fft_policy=policy)
    return ds, d


SPECS = [((40, 36, 24), 3, None), ((57, 30, 44), 2, None), ((64, 64, 48), 4, (64, 64, 64)),
         ((100, 90, 70), 2, (64, 64, 64)), ((260, 250, 20), 2, (256, 256, 32))]


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_equals_per_stack(use_graph):
    iters = 3
    plain, batched = [], []
    for i, (size, nv, blk) in enumerate(SPECS):
        print("stack", size, nv, blk, flush=True)
        _, d = _stack(size, nv, 30 + i, block=blk, run=True, iters=iters)
        plain.append((d.get_psi(), np.array(d.stats)))
        d.close()
        _, d = _stack(size, nv, 30 + i, block=blk)
        batched.append(d)
    b = MVDeconBatch(batched, use_graph=use_graph)
    st = b.iterate(iters)
    for (psi_p, st_p), d, s in zip(plain, batched, st):
        np.testing.assert_array_equal(d.get_psi(), psi_p)
        np.testing.assert_array_equal(s, st_p)
    b.close()
    for d in batched:
        d.close()


def test_batch_graph_replays_match_eager():
    """Graph replays with eager work between them give the same as eager runs."""
    _, d1 = _stack((48, 40, 32), 3, 41)
    _, d2 = _stack((48, 40, 32), 3, 41)
    g = MVDeconBatch([d1], use_graph=True)
    e = MVDeconBatch([d2], use_graph=False)
    for _ in range(3):
        g.iterate(1)
        e.iterate(1)
        np.testing.assert_array_equal(d1.get_psi(masked=False), d2.get_psi(masked=False))
    g.close()
    e.close()
    d1.close()
    d2.close()


@pytest.mark.parametrize("policy,size", [(FFTPolicy.POW2, (50, 45, 30)), (FFTPolicy.AUTOTUNE, (50, 45, 30)),
                                         (FFTPolicy.AUTOTUNE, (300, 280, 40))])
def test_fft_policies_match_oracle(policy, size):
    """The autotuner may also split an unblocked stack into 256 x 256 x bz
    blocks (fused plane kernels); any choice must stay within the gate of the
    whole-image oracle."""
    ds, d = _stack(size, 3, 51, policy=int(policy), run=True, iters=4)
    psi = d.get_psi()
    geo = d.geometry(0)
    print("policy", policy.name, size, "fft", geo["fft"], "blocks", geo["blocks"], flush=True)
    if policy == FFTPolicy.POW2:
        assert geo["fft"] == (64, 64, 64)
    d.close()
    ref, _ = O.deconvolve(ds.views, ds.psfs, O.EFFICIENT_BAYESIAN, 4, 0.006, norm_threads=8)
    assert float(np.abs(psi.astype(np.float64) - ref).max() / np.abs(ref).max()) < GATE
