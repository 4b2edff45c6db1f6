"""The oracle's PSF extraction (ExtractPSF) pinned by closed forms, on the CPU:
integer bead positions give exact sums of periodic crops; the interpolation
reproduces linear functions; normalisation maps to [0, 1]; the transform
geometry gives odd sizes around the mapped centre, the identity transform is a
copy and a 90-degree rotation a permutation of the voxels."""
import math

import numpy as np

from oracle import oracle as O
from spim_registration_amd import psf as P


def _img(shape, seed=0):
    return np.random.default_rng(seed).uniform(0.0, 100.0, shape).astype(np.float32)


def _crop_periodic(img, centre_xyz, size_xyz):
    sx, sy, sz = size_xyz
    cx, cy, cz = centre_xyz
    zi = (np.arange(sz) - sz // 2 + cz) % img.shape[0]
    yi = (np.arange(sy) - sy // 2 + cy) % img.shape[1]
    xi = (np.arange(sx) - sx // 2 + cx) % img.shape[2]
    return img[np.ix_(zi, yi, xi)]


def test_integer_locations_sum_periodic_crops():
    img = _img((20, 24, 28))
    locs = [(5, 6, 7), (0, 0, 0), (27, 23, 19), (14, 3, 18)]  # the last ones wrap around the faces
    size = (9, 7, 5)
    raw, norm = O.extract_psf(img, locs, size)
    exp = np.zeros((5, 7, 9), np.float32)
    for l in locs:
        exp = (exp + _crop_periodic(img, l, size)).astype(np.float32)
    np.testing.assert_array_equal(raw, exp)
    mn, mx = float(raw.min()), float(raw.max())
    np.testing.assert_array_equal(norm, ((raw.astype(np.float64) - mn) / (mx - mn)).astype(np.float32))
    assert norm.min() == 0.0 and norm.max() == 1.0


def test_linear_image_is_reproduced():
    nz, ny, nx = 16, 18, 20
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    img = (10 + 0.5 * x + 0.25 * y + 0.75 * z).astype(np.float32)
    locs = np.array([[8.3, 9.6, 7.45], [10.9, 5.2, 8.01]])
    size = (5, 5, 3)
    raw, _ = O.extract_psf(img, locs, size)
    pz, py, px = np.meshgrid(np.arange(3) - 1, np.arange(5) - 2, np.arange(5) - 2, indexing="ij")
    exp = np.zeros(raw.shape)
    for l in locs:
        exp += 10 + 0.5 * (px + l[0]) + 0.25 * (py + l[1]) + 0.75 * (pz + l[2])
    np.testing.assert_allclose(raw, exp, rtol=3e-6)


def test_transform_geometry_and_identity():
    ident = np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], np.float64)
    ns, off = O.psf_transform_geometry((31, 30, 51), ident)
    assert ns == (31, 31, 51)  # extent 30 -> 31; 29 -> 30 -> made odd
    np.testing.assert_array_equal(off, [0.0, 0.0, 0.0])
    psf = _img((51, 30, 31), 1)
    out = O.psf_transform(psf, ident, ident)
    np.testing.assert_array_equal(out[:, :30, :], psf)
    assert np.all(out[:, 30, :] == 0)


def test_rotation_by_90_is_a_permutation():
    a = math.pi / 2
    rot = np.array([[math.cos(a), 0, math.sin(a), 0], [0, 1, 0, 0], [-math.sin(a), 0, math.cos(a), 0]], np.float64)
    from spim_registration_amd.transform import inverse_affine
    inv = inverse_affine(rot)
    psf = _img((11, 9, 7), 2)  # z 11, y 9, x 7
    ns, off = O.psf_transform_geometry((7, 9, 11), rot)
    assert ns == (11, 9, 7)
    out = O.psf_transform(psf, rot.reshape(12), inv)
    # x' = z, z' = -x (about the centres): out[z', y, x'] = psf[x = ..., y, z = ...]
    exp = np.empty((7, 9, 11), np.float32)
    for zp in range(7):
        for xp in range(11):
            # inverse: source x = -(z' + off_z) and z = x' + off_x, in voxel units
            sx = -(zp + off[2])
            sz = xp + off[0]
            exp[zp, :, xp] = psf[int(round(sz)), :, int(round(sx))]
    np.testing.assert_allclose(out, exp, rtol=1e-6, atol=1e-4)


def test_gaussian_beads_give_the_gaussian():
    rng = np.random.default_rng(3)
    nz, ny, nx = 48, 60, 60
    sig = (2.5, 2.5, 3.5)
    # beads on a coarse grid with sub-voxel jitter, far enough apart not to overlap
    gx, gy, gz = np.meshgrid([15.0, 45.0], [15.0, 45.0], [14.0, 34.0], indexing="ij")
    locs = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], 1) + rng.uniform(-0.5, 0.5, (8, 3))
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    img = np.zeros((nz, ny, nx))
    for l in locs:
        img += np.exp(-0.5 * (((x - l[0]) / sig[0]) ** 2 + ((y - l[1]) / sig[1]) ** 2 + ((z - l[2]) / sig[2]) ** 2))
    _, norm = O.extract_psf(img.astype(np.float32), locs, (9, 9, 13))
    pz, py, px = np.meshgrid(np.arange(13) - 6, np.arange(9) - 4, np.arange(9) - 4, indexing="ij")
    g = np.exp(-0.5 * ((px / sig[0]) ** 2 + (py / sig[1]) ** 2 + (pz / sig[2]) ** 2))
    g = (g - g.min()) / (g.max() - g.min())
    assert np.abs(norm - g).max() < 0.03


def test_host_helpers():
    e = P.ExtractPSF()
    a = _img((5, 7, 9), 4)
    b = _img((7, 5, 3), 5)
    e.view_ids = ["a", "b"]
    e.point_spread_functions = {"a": a, "b": b}
    assert e.compute_max_dim_transformed_psf() == (9, 7, 7)
    avg = e.compute_average_transformed_psf()
    assert avg.shape == (7, 7, 9)
    # point reflection about the centres: a[z, y, x] lands at (5 - z, 6 - y, 8 - x) and
    # b[z, y, x] at (6 - z, 5 - y, 5 - x); row y = 6 only receives a
    np.testing.assert_array_equal(avg[5:0:-1, 6, 8::-1], a[:, 0, :])
    assert abs(float(avg.sum(dtype=np.float64)) - float(a.sum(dtype=np.float64) + b.sum(dtype=np.float64))) < 1e-2
    np.testing.assert_array_equal(P.common_size([a, b]), (9, 7, 7))
    s = P.make_same_size(a, (11, 9, 7))
    np.testing.assert_array_equal(s[1:6, 1:8, 1:10], a)
    assert s[0, 0, 0] == a.min()
    proj = P.ExtractPSF.compute_max_projection(a)  # smallest dimension: z (5)
    np.testing.assert_array_equal(proj, a.max(axis=0))
