"""Pins of the oracle's likelihood gradient (oracle/mvd_oracle.cpp nll_grad) on
the CPU, independent of its FFTs.

Forward operator of view v (the reference's convolve1, deconv/MVDeconFFT.java:384-470,
with the block's mirror-single extension, deconv/MVDeconFFTThreads.java:44,78):
    (A_v psi)(x) = sum_j P_v(j) E_mirror(psi)(x + c - j)
NLL L = sum_v sum_x w_v (b_v - y_v ln b_v), b_v = A_v psi.
    dL/dpsi  = sum_v A_v^T r_v,  r_v = w_v (1 - y_v / b_v)
    dL/dP_v(j) = sum_x r_v(x) E_mirror(psi)(x + c - j)

Checks (all FP64):
  * A built column by column from the brute-force direct convolution
    (itself pinned against scipy): dL/dpsi equals A^T r exactly to rounding,
    on shapes whose kernel is longer than the image (several reflections);
  * the adjoint identity <A x, w> = <x, A^T w> for random signed w (zero data
    makes r = w);
  * central differences in psi (h = 1e-3) at 64 voxels, more than half of them
    within K/2 of a face, edge or corner, where the mirror fold acts;
  * central differences in the PSF at 64 voxels per view (h = 1e-5: with
    h = 1e-3 the O(h^2) truncation of the difference quotient alone is about
    1e-4 of the gradient here);
  * the same checks fail on the oracle's planted variant whose fold drops the
    mirrored shell.
"""
import numpy as np
import pytest

from oracle import oracle as O

SHAPES = [((6, 5, 4), (5, 3, 7)), ((14, 12, 10), (5, 7, 9)), ((9, 11, 8), (3, 5, 5))]


def _problem(shape_xyz, k_xyz, seed, nviews=2, zero_data=False):
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape_xyz
    kx, ky, kz = k_xyz
    psi = rng.uniform(50, 150, (nz, ny, nx))
    psfs = []
    for _ in range(nviews):
        p = rng.uniform(0.1, 1.0, (kz, ky, kx))
        psfs.append(p / p.sum())
    blur = [O.direct_convolve_f64(psi, p, O.OOB_MIRROR) for p in psfs]
    if zero_data:
        views = [np.zeros_like(psi) for _ in psfs]
    else:
        views = [rng.poisson(b).astype(np.float64) for b in blur]
    weights = [rng.uniform(0.2, 1.0, psi.shape) for _ in psfs]
    return psi, psfs, views, weights, blur


def _matrix(shape_zyx, psf):
    """Dense A: column i = A e_i by the brute-force direct convolution."""
    n = int(np.prod(shape_zyx))
    a = np.empty((n, n))
    e = np.zeros(n)
    for i in range(n):
        e[i] = 1.0
        a[:, i] = O.direct_convolve_f64(e.reshape(shape_zyx), psf, O.OOB_MIRROR, threads=1).ravel()
        e[i] = 0.0
    return a


def _near_face_voxels(shape_zyx, k_zyx, rng, count):
    """Flat indices, more than half within k//2 of some face."""
    nz, ny, nx = shape_zyx
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    edge = np.zeros(shape_zyx, bool)
    for c, n, k in ((z, nz, k_zyx[0]), (y, ny, k_zyx[1]), (x, nx, k_zyx[2])):
        edge |= (c < k // 2) | (c >= n - k // 2)
    e_idx, i_idx = np.flatnonzero(edge), np.flatnonzero(~edge)
    ne = min(len(e_idx), count if len(i_idx) == 0 else (2 * count) // 3)
    pick = list(rng.choice(e_idx, ne, replace=False))
    if len(i_idx):
        pick += list(rng.choice(i_idx, min(len(i_idx), count - ne), replace=False))
    pick = np.array(pick)
    assert edge.ravel()[pick].mean() > 0.5
    return pick


def _fd_psi(psi, psfs, views, weights, blur, voxels, h=1e-3):
    """Central differences of L in psi at single voxels, FP64.  The blur of
    psi +- h e_p is b +- h A e_p; L(+) - L(-) is summed over the support of
    A e_p only, as (b+ - b-) - y (log1p(h d / b) - log1p(-h d / b)), which is
    the same difference without the cancellation of two full sums."""
    out = np.zeros(len(voxels))
    shape = psi.shape
    for t, p in enumerate(voxels):
        e = np.zeros(psi.size)
        e[p] = 1.0
        e = e.reshape(shape)
        acc = 0.0
        for pv, y, w, b in zip(psfs, views, weights, blur):
            d = O.direct_convolve_f64(e, pv, O.OOB_MIRROR, threads=1)
            m = d != 0
            hd, bb = h * d[m], b[m]
            acc += float(np.sum(w[m] * (2 * hd - y[m] * (np.log1p(hd / bb) - np.log1p(-hd / bb)))))
        out[t] = acc / (2 * h)
    return out


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("shape,k", SHAPES)
def test_psi_gradient_equals_dense_adjoint(shape, k):
    psi, psfs, views, weights, blur = _problem(shape, k, seed=sum(shape))
    _, gpsi, _ = O.nll_grad(views, psfs, psi, weights=weights, threads=2)
    expect = np.zeros(psi.size)
    for p, y, w, b in zip(psfs, views, weights, blur):
        a = _matrix(psi.shape, p)
        np.testing.assert_allclose(a @ psi.ravel(), b.ravel(), rtol=1e-12)  # A is the forward operator
        r = (w * (1 - y / b)).ravel()
        expect += a.T @ r
    assert _rel(gpsi.ravel(), expect) < 1e-10
    # the planted fold without the mirrored shell is caught
    _, bad, _ = O.nll_grad(views, psfs, psi, weights=weights, threads=2, variant=1)
    assert _rel(bad.ravel(), expect) > 1e-2


@pytest.mark.parametrize("shape,k", SHAPES)
def test_adjoint_identity(shape, k):
    """Zero data gives r = w, so dL/dpsi = sum_v A_v^T w_v for any signed w."""
    psi, psfs, views, _, _ = _problem(shape, k, seed=7 + sum(k), nviews=1, zero_data=True)
    rng = np.random.default_rng(11)
    for _ in range(3):
        w = rng.normal(size=psi.shape)
        x = rng.normal(size=psi.shape)
        _, atw, _ = O.nll_grad(views, psfs, psi, weights=[w], threads=2)
        lhs = float(np.sum(O.direct_convolve_f64(x, psfs[0], O.OOB_MIRROR) * w))
        rhs = float(np.sum(x * atw))
        assert abs(lhs - rhs) <= 1e-10 * (np.abs(x).sum() * np.abs(atw).max() + abs(lhs))
        _, bad, _ = O.nll_grad(views, psfs, psi, weights=[w], threads=2, variant=1)
        assert abs(lhs - float(np.sum(x * bad))) > 1e-3 * abs(lhs)


@pytest.mark.parametrize("shape,k", SHAPES[1:])
def test_psi_gradient_matches_central_differences(shape, k):
    psi, psfs, views, weights, blur = _problem(shape, k, seed=3 + sum(shape))
    rng = np.random.default_rng(5)
    vox = _near_face_voxels(psi.shape, (k[2], k[1], k[0]), rng, 64)
    fd = _fd_psi(psi, psfs, views, weights, blur, vox)
    _, gpsi, _ = O.nll_grad(views, psfs, psi, weights=weights, threads=2)
    err = _rel(gpsi.ravel()[vox], fd)
    print(f"dL/dpsi vs central differences: {err:.2e}")
    assert err < 1e-6
    _, bad, _ = O.nll_grad(views, psfs, psi, weights=weights, threads=2, variant=1)
    assert _rel(bad.ravel()[vox], fd) > 1e-2


def test_psf_gradient_matches_central_differences():
    psi, psfs, views, _, _ = _problem((14, 12, 10), (5, 7, 9), seed=31, nviews=2)
    _, _, gpsf = O.nll_grad(views, psfs, psi, threads=2)
    rng = np.random.default_rng(8)
    for v in range(2):
        vox = rng.choice(psfs[v].size, 64, replace=False)
        fd = O.nll_psf_fd(views[v], psfs[v], psi, vox, h=1e-5, threads=2)
        err = _rel(gpsf[v].ravel()[vox], fd)
        print(f"view {v}: dL/dP vs central differences: {err:.2e}")
        assert err < 1e-6
        # the difference quotient converges at second order: h = 1e-3 is off by ~1e-4
        coarse = _rel(O.nll_psf_fd(views[v], psfs[v], psi, vox, h=1e-3, threads=2), fd)
        assert 1e-6 < coarse < 1e-3
