"""Pins the oracle (CPU restatement of the reference path) against closed
forms, brute force, an independent library routine (scipy.ndimage) and
invariants of the algorithm.  The reference ships no tests or golden vectors
(it is Java without a test tree), so these checks are the anchor."""
import numpy as np
import pytest
import scipy.ndimage as ndi

from oracle import oracle as O

RNG = np.random.default_rng(1234)


def _scipy_conv(img, k, oob):
    """Same definition via scipy: out(x) = sum_j E(img)(x + k//2 - j) k(j)."""
    if oob == O.OOB_MIRROR:
        return ndi.convolve(img, k, mode="mirror")
    return ndi.convolve(img, k, mode="constant", cval=1.0 if oob == O.OOB_ONE else 0.0)


@pytest.mark.parametrize("oob", [O.OOB_MIRROR, O.OOB_ONE, O.OOB_ZERO])
@pytest.mark.parametrize("shape,kshape", [((9, 10, 12), (7, 3, 5)), ((5, 6, 4), (3, 5, 3)), ((3, 4, 5), (7, 9, 11))])
def test_fft_convolve_matches_direct_and_scipy(oob, shape, kshape):
    img = RNG.random(shape)
    k = RNG.random(kshape)
    d = O.direct_convolve_f64(img, k, oob)
    np.testing.assert_allclose(d, _scipy_conv(img, k, oob), rtol=1e-12, atol=1e-12)
    f64 = O.fft_convolve(img, k, oob)
    np.testing.assert_allclose(f64, d, rtol=1e-11, atol=1e-11)
    f32 = O.fft_convolve(img.astype(np.float32), k.astype(np.float32), oob)
    assert np.abs(f32 - d).max() / np.abs(d).max() < 2e-6


def test_mirror_single_period_beyond_one_reflection():
    # kernels far larger than the image reflect more than once (period 2n-2)
    img = RNG.random((4, 3, 5))
    k = RNG.random((9, 11, 13))
    np.testing.assert_allclose(O.direct_convolve_f64(img, k, O.OOB_MIRROR), ndi.convolve(img, k, mode="mirror"),
                               rtol=1e-12)


def test_delta_is_identity_and_shift():
    img = RNG.random((6, 7, 8))
    k = np.zeros((3, 5, 7))
    k[1, 2, 3] = 1
    np.testing.assert_allclose(O.fft_convolve(img, k, O.OOB_MIRROR), img, atol=1e-14)
    k = np.zeros((3, 5, 7))
    k[1, 2, 4] = 1  # j = c + 1 along x: out(x) = in(x - 1)
    out = O.fft_convolve(img, k, O.OOB_ZERO)
    np.testing.assert_allclose(out[:, :, 1:], img[:, :, :-1], atol=1e-14)
    np.testing.assert_allclose(out[:, :, 0], 0, atol=1e-14)
    out = O.fft_convolve(img, k, O.OOB_MIRROR)  # x = 0 reads in(-1) = in(1)
    np.testing.assert_allclose(out[:, :, 0], img[:, :, 1], atol=1e-14)
    out = O.fft_convolve(img, k, O.OOB_ONE)
    np.testing.assert_allclose(out[:, :, 0], 1, atol=1e-14)


def test_convolution_is_linear():
    a, b = RNG.random((2, 5, 6, 7))
    k = RNG.random((3, 3, 5))
    lhs = O.fft_convolve(2 * a - 3 * b, k, O.OOB_ZERO)
    rhs = 2 * O.fft_convolve(a, k, O.OOB_ZERO) - 3 * O.fft_convolve(b, k, O.OOB_ZERO)
    np.testing.assert_allclose(lhs, rhs, atol=1e-12)


@pytest.mark.parametrize("threads", [1, 2, 3, 8])
@pytest.mark.parametrize("n", [1, 7, 64, 19 * 19 * 31])
def test_norm_img_double_counts_first_portion(threads, n):
    # a kernel of ones: portion 0 holds floor(n / 2T) elements, so k = 1 / (n + floor(n / 2T))
    k = np.ones(n, np.float64)
    expect = 1.0 / (n + n // (2 * threads))
    np.testing.assert_allclose(O.norm_img(k, threads), expect, rtol=1e-15)
    k32 = np.ones(n, np.float32)
    assert np.all(O.norm_img(k32, threads) == np.float32(expect))


def test_norm_img_sum_formula_random():
    k = RNG.random(1000)
    for t in (1, 4, 8):
        s0 = k[: 1000 // (2 * t)].sum()
        assert abs(O.sum_img(k, t) - (k.sum() + s0)) < 1e-11


def test_divide_portions():
    assert O.divide_portions(10, 4) == [(0, 2), (2, 2), (4, 2), (6, 4)]
    assert O.divide_portions(3, 4) == [(0, 0), (0, 0), (0, 0), (0, 3)]


@pytest.mark.parametrize("shape", [(5, 3, 7), (4, 6, 2), (1, 1, 1), (3, 8, 5)])
def test_mirror_axis(shape):
    a = RNG.random(shape)
    for axis_xyz, ax in ((0, 2), (1, 1), (2, 0)):
        m = O.mirror(a, axis_xyz)
        n = shape[ax]
        expect = np.flip(a, axis=ax).copy()
        if n % 2 == 0:
            # the middle pair is swapped twice and stays (Mirror.java:88-107)
            sl = [slice(None)] * 3
            for p in (n // 2 - 1, n // 2):
                sl[ax] = p
                expect[tuple(sl)] = a[tuple(sl)]
        np.testing.assert_array_equal(m, expect)


def test_block_geometry_tiles_image():
    for img, ker, bs in [((512, 512, 256), (31, 31, 51), (256, 256, 256)),
                         ((1024, 1024, 512), (41, 41, 81), (512, 512, 512)),
                         ((128, 128, 64), (19, 19, 31), (146, 146, 94)),
                         ((100, 37, 20), (5, 7, 3), (24, 16, 9))]:
        blocks = O.divide_blocks(img, ker, bs)
        cover = np.zeros(img[::-1], np.int32)
        for b in blocks:
            eo, es = b["eff_offset"], b["eff_size"]
            assert b["eff_local"] == tuple(k // 2 for k in ker)
            assert b["offset"] == tuple(e - k // 2 for e, k in zip(eo, ker))
            cover[eo[2]:eo[2] + es[2], eo[1]:eo[1] + es[1], eo[0]:eo[0] + es[0]] += 1
        assert np.all(cover == 1)
    assert len(O.divide_blocks((512, 512, 256), (31, 31, 51), (256,) * 3)) == 18
    with pytest.raises(ValueError):
        O.divide_blocks((10, 10, 10), (5, 5, 5), (4, 4, 4))


@pytest.mark.parametrize("oob", [O.OOB_MIRROR, O.OOB_ONE])
def test_blocked_equals_whole(oob):
    img = RNG.random((20, 30, 25))
    k = RNG.random((5, 7, 3))
    whole = O.convolve(img, k, oob)
    blocked = O.convolve(img, k, oob, block_xyz=(12, 16, 9))
    np.testing.assert_allclose(blocked, whole, atol=1e-11)
    np.testing.assert_allclose(whole, _scipy_conv(img, k, oob), atol=1e-11)


def test_quotient_cases():
    blurred = np.array([2, 0.5, 3, 7, 0, 1], np.float32)
    img = np.array([4, 1, 0, -1, 5, np.nan], np.float32)
    q = O.quotient(blurred, img)
    assert q[0] == 2 and q[1] == 2 and q[2] == 1 and q[3] == 1  # img <= 0 -> 1
    assert np.isinf(q[4])
    assert q[5] == 1  # NaN > 0 is false


def test_next_value_properties():
    minv = np.float32(1e-4)
    last = np.array([1, 2, 0.5, 3, 1, 1e-6, np.nan], np.float32)
    integ = np.array([1.5, 0.25, 4, -1, 0, 1, 1], np.float32)
    w0 = np.zeros_like(last)
    # weight 0 leaves psi unchanged (NaN stays NaN)
    np.testing.assert_array_equal(O.next_value(last, integ, w0, 0.0)[:-1], last[:-1])
    assert np.isnan(O.next_value(last, integ, w0, 0.0)[-1])
    # weight 1, no Tikhonov, cases where the float arithmetic is exact:
    # psi * I when positive, minValue when psi * I <= 0
    last = np.array([1, 2, 0.5, 2 * minv, 2 * minv, 2 * minv], np.float32)
    integ = np.array([1.5, 0.25, 4, -1, 0, 0.25], np.float32)
    nv = O.next_value(last, integ, np.ones_like(last), 0.0)
    np.testing.assert_array_equal(nv, np.array([1.5, 0.5, 2, minv, minv, minv], np.float32))
    # NaN estimate stays NaN (value > 0 is false -> minValue, then psi + (min - psi) * w)
    assert np.isnan(O.next_value(np.float32([np.nan]), np.float32([1]), np.float32([1]), 0.0)[0])
    # Tikhonov: a = (sqrt(1 + 2 lam v) - 1) / lam is the positive root of lam/2 a^2 + a - v = 0
    lam = 0.006
    v = (np.float32(1) * RNG.uniform(1, 5000, 1000).astype(np.float32))
    a = O.next_value(np.ones_like(v), v, np.ones_like(v), lam).astype(np.float64)
    resid = lam / 2 * a * a + a - v.astype(np.float64)
    assert np.all(np.abs(resid) / v < 1e-6)
    assert np.all(a < v)
    # partial weight interpolates: last + (next - last) * w
    w = np.float32(0.25)
    nv = O.next_value(np.float32([2.0]), np.float32([3.0]), np.float32([w]), 0.0)
    assert nv[0] == np.float32(2.0) + (np.float32(6.0) - np.float32(2.0)) * w


def test_first_iteration_closed_form():
    a = np.array([[[1, 0, 3, 0]]], np.float32)
    b = np.array([[[3, 0, -1, 0]]], np.float32)
    c = np.array([[[2, 4, 0, 0]]], np.float32)
    avg, counts = O.first_iteration([a, b, c], 2)
    np.testing.assert_array_equal(counts.ravel(), [3, 1, 1, 0])
    # means: 2, 4, 3 over 3 voxels with data
    assert abs(avg - 3.0) < 1e-15


def _norm_sum(k, t):
    return O.norm_img(k, t).sum()


@pytest.mark.parametrize("ptype", [O.INDEPENDENT, O.EFFICIENT_BAYESIAN, O.OPTIMIZATION_I, O.OPTIMIZATION_II])
def test_kernel2_construction(ptype):
    from spim_registration_amd import synthetic as S
    ks = [S.psf((7, 7, 9), 1.0, 2.0, a).astype(np.float64) for a in (0, 90, 45)]
    k1, k2 = O.init_kernels(ks, ptype, 2)
    for k, kn in zip(ks, k1):
        np.testing.assert_allclose(kn, O.norm_img(k, 2), rtol=1e-15)
    for v in range(3):
        if ptype == O.INDEPENDENT:
            np.testing.assert_allclose(k2[v], k1[v][::-1, ::-1, ::-1], rtol=1e-15)
        elif ptype == O.OPTIMIZATION_II:
            np.testing.assert_allclose(k2[v], O.norm_img(k1[v] ** 3, 2)[::-1, ::-1, ::-1], rtol=1e-13)
        elif ptype == O.OPTIMIZATION_I:
            # K1_v * prod_w (K1_v conv mirror(K1_w)), normalised, mirrored: brute force
            tmp = k1[v].copy()
            for w in range(3):
                if w != v:
                    tmp = O.direct_convolve_f64(k1[v], k1[w][::-1, ::-1, ::-1].copy(), O.OOB_ZERO) * tmp
            np.testing.assert_allclose(k2[v], O.norm_img(tmp, 2)[::-1, ::-1, ::-1], rtol=1e-9, atol=1e-14)
        else:
            tmp = k1[v][::-1, ::-1, ::-1].copy()
            for w in range(3):
                if w == v:
                    continue
                # view v reads K1_w as it is at that time: normalised only if w < v
                kw = k1[w] if w < v else ks[w]
                a = O.direct_convolve_f64(k1[v][::-1, ::-1, ::-1].copy(), kw, O.OOB_ZERO)
                b = O.direct_convolve_f64(a, kw[::-1, ::-1, ::-1].copy(), O.OOB_ZERO)
                tmp = b * tmp
            np.testing.assert_allclose(k2[v], O.norm_img(tmp, 2), rtol=1e-9, atol=1e-14)


def test_fixed_point_invariant():
    """Noise-free data img = psi* conv K1 (mirror OOB), lambda = 0: psi = c2 psi*
    (c2 = sum of kernel2, != 1 because of the double count) is a fixed point of
    the update at voxels at least k/2 from the border.  Pins kernel centring,
    convolution sense, OOB rules and the normalisation quirk at once."""
    from spim_registration_amd import synthetic as S
    k = S.psf((5, 7, 9), 1.0, 1.7, 30).astype(np.float64)
    k1, k2 = O.init_kernels([k], O.INDEPENDENT, 3)
    c2 = k2[0].sum()
    assert abs(c2 - 1) > 1e-3  # the quirk is visible
    psi_star = 100 + 50 * RNG.random((16, 18, 20))
    img = O.direct_convolve_f64(psi_star, k1[0], O.OOB_MIRROR)
    psi, _ = O.deconvolve([img], [k], O.INDEPENDENT, 1, 0.0, norm_threads=3, psi0=c2 * psi_star)
    inner = (slice(4, -4), slice(3, -3), slice(2, -2))
    np.testing.assert_allclose(psi[inner], (c2 * psi_star)[inner], rtol=1e-12)
    # a wrong centre or sense breaks it
    shifted = np.roll(k, 1, axis=2)
    psi2, _ = O.deconvolve([img], [shifted], O.INDEPENDENT, 1, 0.0, norm_threads=3, psi0=c2 * psi_star)
    assert np.abs(psi2[inner] / (c2 * psi_star)[inner] - 1).max() > 1e-4


def test_fp32_and_fp64_oracles_agree_on_small_run():
    from spim_registration_amd import synthetic as S
    ds = S.make_dataset((24, 20, 16), 3, [0, 120, 240], 20, (5, 5, 7), 1.0, 2.0, seed=7)
    psi32, st32 = O.deconvolve(ds.views, ds.psfs, O.EFFICIENT_BAYESIAN, 4, 0.006, norm_threads=4)
    psi64, st64 = O.deconvolve([v.astype(np.float64) for v in ds.views], [p.astype(np.float64) for p in ds.psfs],
                               O.EFFICIENT_BAYESIAN, 4, 0.006, norm_threads=4)
    assert np.abs(psi32 - psi64).max() / np.abs(psi64).max() < 1e-5
    assert np.all(st32[:, :, 1] >= 0) and np.all(st64[:, :, 0] > 0)


def test_mask_zero_where_no_data():
    from spim_registration_amd import synthetic as S
    ds = S.make_dataset((12, 10, 8), 2, [0, 90], 5, (3, 3, 5), 1.0, 1.5, seed=3)
    v0, v1 = ds.views[0].copy(), ds.views[1].copy()
    v0[:, :, :3] = 0
    v1[:, :, :3] = 0
    psi, _ = O.deconvolve([v0, v1], ds.psfs, O.INDEPENDENT, 2, 0.0, norm_threads=2)
    assert np.all(psi[:, :, :3] == 0) and np.all(psi[:, :, 3:] > 0)


def test_overlap_counts_closed_form():
    """4 voxels with 0, 1, 2, 2 weighted views, numThreads 1: portions of 2
    voxels give minima 0 and 2 and averages 0.5 and 2."""
    a = np.array([0, 1, 1, 1], np.float32).reshape(1, 1, 4)
    b = np.array([0, 0, 2, 0.5], np.float32).reshape(1, 1, 4)
    mn, avg = O.overlap_counts([a, b], 1)
    assert mn == 1  # max(1, 0)
    assert avg == 1.25


def test_effective_weights_closed_forms():
    rng = np.random.default_rng(2)
    raw = [rng.random((3, 4, 5)).astype(np.float32) for _ in range(3)]
    for r in raw:
        r[0, 0, :] = 0  # a row no view weights
    raw[1][1] = 0
    with np.errstate(invalid="ignore"):
        ws, osem = O.effective_weights(raw, O.W_PRECOMPUTED, 0)
    assert osem == 1.0
    total = sum(w.astype(np.float64) for w in ws)
    assert np.all(np.isnan(total[0, 0, :]))  # 0/0, no guard
    np.testing.assert_allclose(total[1:], 1.0, rtol=1e-6)
    # OSEM keeps NaN (Math.min) and clamps at 1
    ws2, osem2 = O.effective_weights(raw, O.W_PRECOMPUTED, 3, 2.5)
    assert osem2 == 2.5
    assert np.all(np.isnan(ws2[0][0, 0, :]))
    fin = ~np.isnan(ws[0])
    np.testing.assert_array_equal(ws2[0][fin], np.minimum(np.float32(1), ws[0][fin] * np.float32(2.5)))
    # virtual weights: w / max(sum, 1) * osem, clamped at 1, never NaN
    wv, _ = O.effective_weights(raw, O.W_VIRTUAL, 3, 1.5)
    s = sum(r.astype(np.float64) for r in raw)
    expect = np.minimum(1.0, raw[2].astype(np.float64) / np.where(s > 1, s.astype(np.float32), 1) * 1.5)
    np.testing.assert_array_equal(wv[2], expect.astype(np.float32))
    assert not np.isnan(wv[2]).any() and wv[2].max() <= 1
    # no weights: the overlap counts stay 0, so OSEM index 1 and 2 give weight 0
    for idx, val, c in ((0, 7.0, 1.0), (1, 7.0, 0.0), (2, 7.0, 0.0), (3, 0.25, 0.25), (3, 3.0, 1.0)):
        wn, _ = O.effective_weights(raw, O.W_NONE, idx, val, shape=(2, 2, 2))
        assert np.all(wn[0] == np.float32(c))
