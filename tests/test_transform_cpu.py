"""The oracle's input preparation (view resampling + blending weights) pinned
by closed forms, on the CPU.

What pins it (the reference has no tests; the interpolator and the affine
arithmetic are imglib2's, not in the reference tree):
  * identity and integer shifts reproduce the samples exactly (clamped at
    minValue), outside the view the image is 0 (FusionHelper.intersects);
  * trilinear interpolation reproduces a linear function of position (a wrong
    corner weight, a transposed axis or a wrong fraction breaks this);
  * at the far face the upper corner is the mirror-single sample n - 2;
  * the blending table equals the cosine formula at its nodes, and the weight
    equals an independent numpy evaluation of BlendingRealRandomAccess and the
    closed form (cos((1 - d) pi) + 1) / 2 within the table's resolution;
  * the weight does not depend on the image (TransformWeights == the weight of
    TransformInputAndWeights).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from spim_registration_amd import synthetic as S
from spim_registration_amd import transform as T

MINV = np.float32(1e-4)
IDENT = np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], np.float64)


def _raw(shape, seed=0, lo=1.0, hi=100.0):
    return np.random.default_rng(seed).uniform(lo, hi, shape).astype(np.float32)


def test_identity_reproduces_samples():
    raw = _raw((7, 9, 11))
    raw[2, 3, 4] = 0.0  # clamped to minValue
    raw[1, 1, 1] = -5.0
    img, _ = O.transform_view(raw, IDENT, (0, 0, 0), (11, 9, 7))
    np.testing.assert_array_equal(img, np.maximum(raw, MINV))


def test_integer_shift_and_outside_zero():
    raw = _raw((6, 8, 10), 1)
    # bounding box larger than the view, min at (-2, 1, -3): voxel s maps to t = s + offset
    img, _ = O.transform_view(raw, IDENT, (-2, 1, -3), (14, 9, 12))
    exp = np.zeros((12, 9, 14), np.float32)
    # s + off in [0, n): x in [2, 12), y in [0, 7), z in [3, 9)
    exp[3:9, 0:7, 2:12] = raw[:, 1:8, :]
    np.testing.assert_array_equal(img, exp)
    # the same shift written into the matrix's translation instead of the offset
    inv = IDENT.copy()
    inv[3], inv[7], inv[11] = -2.0, 1.0, -3.0
    img2, _ = O.transform_view(raw, inv, (0, 0, 0), (14, 9, 12))
    np.testing.assert_array_equal(img2, exp)


def _rot_scale(angle_deg, scale_z, centre):
    a = math.radians(angle_deg)
    r = np.array([[math.cos(a), 0, math.sin(a)], [0, 1, 0], [-math.sin(a), 0, math.cos(a)]])
    m = r @ np.diag([1.0, 1.0, scale_z])
    c = np.asarray(centre, np.float64)
    t = c - m @ c
    return np.hstack([m, t[:, None]])


@pytest.mark.parametrize("angle,scale", [(0.0, 1.0), (30.0, 2.0), (-71.0, 1.5), (90.0, 0.8)])
def test_linear_function_is_reproduced(angle, scale):
    rz, ry, rx = 12, 13, 15
    z, y, x = np.meshgrid(np.arange(rz), np.arange(ry), np.arange(rx), indexing="ij")
    coef = (3.0, 0.75, -0.5, 1.25)
    raw = (coef[0] + 50 + coef[1] * x + coef[2] * y + coef[3] * z).astype(np.float32)
    fwd = _rot_scale(angle, scale, (7, 6, 5))
    inv = T.inverse_affine(fwd)
    (mn, size) = T.bounding_box([(rx, ry, rz)], [fwd])
    img, _ = O.transform_view(raw, inv, mn, size)
    nz, ny, nx = img.shape
    sz, sy, sx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    s = np.stack([sx + mn[0], sy + mn[1], sz + mn[2]], -1).astype(np.float32).astype(np.float64)
    m = inv.reshape(3, 4)
    t = (s @ m[:, :3].T + m[:, 3]).astype(np.float32).astype(np.float64)
    inside = np.all((t >= 0) & (t <= np.array([rx - 1, ry - 1, rz - 1])), -1)
    expect = coef[0] + 50 + coef[1] * t[..., 0] + coef[2] * t[..., 1] + coef[3] * t[..., 2]
    assert inside.sum() > 200
    err = np.abs(img[inside] - expect[inside]) / np.abs(expect[inside])
    assert err.max() < 4e-6
    outside = np.any((t < 0) | (t >= np.array([rx, ry, rz])), -1)
    assert np.all(img[outside] == 0)


def test_far_face_uses_mirror_single():
    raw = _raw((3, 4, 5), 2)
    inv = IDENT.copy()
    inv[3] = 0.5  # t.x = s.x + 0.5: the last column interpolates with column n - 2
    img, _ = O.transform_view(raw, inv, (0, 0, 0), (5, 4, 3))
    got = img[:, :, 4].astype(np.float64)
    exp = 0.5 * raw[:, :, 4].astype(np.float64) + 0.5 * raw[:, :, 3]
    np.testing.assert_allclose(got, exp, rtol=2e-7)
    inner = 0.5 * raw[:, :, 1].astype(np.float64) + 0.5 * raw[:, :, 2]
    np.testing.assert_allclose(img[:, :, 1], inner, rtol=2e-7)


def test_single_slice_axis():
    raw = _raw((1, 5, 6), 3)
    inv = IDENT.copy()
    inv[11] = 0.25  # t.z in [0.25, 1.25): only s.z = 0 maps inside; the upper corner mirrors to 0
    img, _ = O.transform_view(raw, inv, (0, 0, 0), (6, 5, 2))
    np.testing.assert_allclose(img[0], raw[0], rtol=2e-7)
    assert np.all(img[1] == 0)


def test_blend_lut_nodes():
    lut = O.blend_lut()
    i = np.arange(1001)
    np.testing.assert_allclose(lut, (np.cos((1 - i / 1000.0) * np.pi) + 1) / 2, atol=1e-12)
    assert lut[0] == 0.0 and abs(lut[1000] - 1.0) < 1e-15
    assert np.all(np.diff(lut) >= 0)


def test_blend_weight_matches_numpy_and_closed_form():
    shape = (20, 33, 41)
    border = (-8.0, -8.0, -3.0)
    rng = (12.0, 12.0, 12.0)
    _, w = O.transform_view(None, IDENT, (0, 0, 0), (41, 33, 20), border, rng, want_image=False,
                            raw_size_xyz=(41, 33, 20))
    np.testing.assert_array_equal(w, S.blending_weight(shape, (-8, -8, -3), (12, 12, 12)))
    # closed form per axis, product over axes (table resolution 1e-3 in d)
    def f(l, n, b, r):
        d = max(0.0, min(l - b, n - 1 - l - b))
        if d == 0:
            return 0.0
        d /= r
        return 1.0 if d >= 1 else (math.cos((1 - d) * math.pi) + 1) / 2
    for (z, y, x) in [(0, 0, 0), (5, 7, 9), (19, 32, 40), (10, 16, 20), (2, 30, 3)]:
        ref = f(x, 41, -8, 12) * f(y, 33, -8, 12) * f(z, 20, -3, 12)
        assert abs(float(w[z, y, x]) - ref) < 2e-3


def test_weights_do_not_depend_on_image():
    raw = _raw((9, 10, 12), 4)
    fwd = _rot_scale(40.0, 1.7, (6, 5, 4))
    inv = T.inverse_affine(fwd)
    mn, size = T.bounding_box([(12, 10, 9)], [fwd])
    img, w = O.transform_view(raw, inv, mn, size, (-2, -2, -1), (5, 5, 3))
    _, w2 = O.transform_view(None, inv, mn, size, (-2, -2, -1), (5, 5, 3), want_image=False,
                             raw_size_xyz=(12, 10, 9))
    np.testing.assert_array_equal(w, w2)
    assert np.all(w >= 0) and np.all(w <= 1) and w.max() > 0.9


def test_blend_weight_single_position():
    # x distance 5 - (-8) = 13 >= 12: factor 1; y: l - b = 2 + 8 = 10 -> 10/12; z: min(4+3, 9-4+3) = 7 -> 7/12
    lut = O.blend_lut()
    got = O.blend_weight((5.0, 2.0, 4.0), (30, 30, 10), (-8, -8, -3), (12, 12, 12))
    ey = lut[int(math.floor(np.float32(10 / 12) * 1000.0 + 0.5))]
    ez = lut[int(math.floor(np.float32(7 / 12) * 1000.0 + 0.5))]
    exp = np.float32(np.float32(1.0 * ey) * ez)
    assert got == exp


def test_inverse_affine_and_bounding_box():
    fwd = _rot_scale(25.0, 2.0, (3, 4, 5))
    inv = T.inverse_affine(fwd).reshape(3, 4)
    p = np.array([1.0, 2.0, 3.0])
    q = fwd[:, :3] @ p + fwd[:, 3]
    np.testing.assert_allclose(inv[:, :3] @ q + inv[:, 3], p, atol=1e-12)
    mn, size = T.bounding_box([(10, 10, 10)], [np.hstack([np.eye(3), [[2], [-3], [0.5]]])])
    assert mn == (2, -3, 0) and size == (10, 10, 11)


def test_raw_view_host_logic():
    """MVDeconFFT.from_raw keeps the raw view and reports the box shape; the
    deconvolution defaults to PRECOMPUTED weights for blended raw views."""
    from spim_registration_amd import MVDeconFFT
    raw = _raw((6, 7, 8), 9)
    psf = np.ones((3, 3, 3), np.float32)
    v = MVDeconFFT.from_raw(raw, IDENT, (0, 0, 0), (10, 9, 8), psf, T.Blending())
    assert v.shape == (8, 9, 10) and v.has_weight()
    w = MVDeconFFT.from_raw(raw, IDENT, (0, 0, 0), (10, 9, 8), psf)
    assert not w.has_weight()
    with pytest.raises(ValueError):
        MVDeconFFT.from_raw(raw, IDENT[:6], (0, 0, 0), (10, 9, 8), psf)
    with pytest.raises(ValueError):
        T.Blending((1.0, 2.0, 3.0), (0.0, 1.0, 1.0)).arrays()
