"""GPU input preparation (mvd_transform_view, mvd_add_view_transformed)
against the oracle: bit-identical images and weights (the kernel keeps the
oracle's types and operation order), and a deconvolution of resampled views
identical to one of views resampled by the oracle."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from spim_registration_amd import MVDeconFFT, MVDeconInput, MVDeconvolution, PSFTYPE, WeightType
from spim_registration_amd import synthetic as S
from spim_registration_amd import transform as T

pytestmark = pytest.mark.gpu

IDENT = np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], np.float64)
BLEND = T.Blending((-8.0, -8.0, -3.0), (12.0, 12.0, 12.0))


def _fwd(angle_deg, scale_z, centre, shift=(0.0, 0.0, 0.0)):
    a = math.radians(angle_deg)
    r = np.array([[math.cos(a), 0, math.sin(a)], [0, 1, 0], [-math.sin(a), 0, math.cos(a)]])
    m = r @ np.diag([1.0, 1.0, scale_z])
    c = np.asarray(centre, np.float64)
    return np.hstack([m, (c - m @ c + np.asarray(shift))[:, None]])


CASES = [
    ("identity", (23, 17, 11), None),
    ("rot30_z2", (40, 36, 18), (30.0, 2.0, (0.3, -0.7, 0.1))),
    ("rot-100_z1.3", (37, 29, 17), (-100.0, 1.3, (0.0, 0.0, 0.0))),
    ("rot45_z3", (64, 48, 20), (45.0, 3.0, (1.5, 0.25, -2.0))),
    ("flat", (31, 27, 1), (12.0, 1.0, (0.0, 0.0, 0.0))),
]


@pytest.mark.parametrize("name,raw_xyz,tf", CASES, ids=[c[0] for c in CASES])
def test_transform_view_matches_oracle(name, raw_xyz, tf):
    rx, ry, rz = raw_xyz
    raw = np.random.default_rng(rx * 7 + rz).uniform(0.0, 500.0, (rz, ry, rx)).astype(np.float32)
    raw[0, 0, :3] = 0.0
    fwd = np.hstack([np.eye(3), np.zeros((3, 1))]) if tf is None else _fwd(tf[0], tf[1], (rx / 2, ry / 2, rz / 2),
                                                                           tf[2])
    inv = T.inverse_affine(fwd)
    mn, size = T.bounding_box([raw_xyz], [fwd])
    # a box a little larger than the view on every side
    mn = tuple(m - 2 for m in mn)
    size = tuple(s + 4 for s in size)
    print("case", name, "raw", raw_xyz, "box", mn, size, flush=True)
    img, w = T.transform_view(raw, inv, mn, size, BLEND)
    rimg, rw = O.transform_view(raw, inv, mn, size, BLEND.border, BLEND.range)
    np.testing.assert_array_equal(img, rimg)
    np.testing.assert_array_equal(w, rw)
    assert (img > 0).sum() > (0.2 * img.size if raw_xyz[2] > 1 else 0)
    # image only (TransformInput) and weights only (TransformWeights)
    img2, none = T.transform_view(raw, inv, mn, size, None)
    assert none is None
    np.testing.assert_array_equal(img2, rimg)
    none, w2 = T.transform_view(None, inv, mn, size, BLEND, want_image=False, raw_size_xyz=raw_xyz)
    np.testing.assert_array_equal(w2, rw)


def test_transform_full_size_properties():
    """A 512 x 512 x 128 view (z spacing 2) into a 512 x 512 x 256 box: equals
    the oracle on sampled planes; identity part of the map checked by a second
    run with a pure z scaling, whose even planes reproduce the raw planes."""
    rx, ry, rz = 512, 512, 128
    raw = np.random.default_rng(5).uniform(50.0, 2000.0, (rz, ry, rx)).astype(np.float32)
    fwd = np.hstack([np.diag([1.0, 1.0, 2.0]), np.zeros((3, 1))])
    img, w = T.transform_view(raw, T.inverse_affine(fwd), (0, 0, 0), (512, 512, 256), BLEND)
    np.testing.assert_array_equal(img[0:256:2][:rz], raw)
    mid = 0.5 * raw[:-1].astype(np.float64) + 0.5 * raw[1:]
    np.testing.assert_allclose(img[1:255:2], mid, rtol=3e-7)
    rot = _fwd(60.0, 2.0, (256, 256, 64))
    inv = T.inverse_affine(rot)
    img, w = T.transform_view(raw, inv, (-60, 0, -60), (640, 512, 380), BLEND)
    planes = [0, 77, 190, 379]
    rimg, rw = O.transform_view(raw, inv, (-60, 0, -60), (640, 512, 380), BLEND.border, BLEND.range)
    for p in planes:
        np.testing.assert_array_equal(img[p], rimg[p])
        np.testing.assert_array_equal(w[p], rw[p])
    np.testing.assert_array_equal(img, rimg)


def _views(nv, raw_xyz, seed):
    rx, ry, rz = raw_xyz
    rng = np.random.default_rng(seed)
    raws, fwds = [], []
    for v in range(nv):
        raws.append(rng.uniform(80.0, 400.0, (rz, ry, rx)).astype(np.float32))
        fwds.append(_fwd(360.0 * v / nv, 1.5, (rx / 2, ry / 2, rz / 2)))
    mn, size = T.bounding_box([raw_xyz] * nv, fwds)
    return raws, [T.inverse_affine(f) for f in fwds], mn, size


@pytest.mark.parametrize("blending", [None, BLEND], ids=["no_weights", "precomputed"])
def test_session_resampling_equals_host_resampling(blending):
    """Views resampled inside the session (mvd_add_view_transformed) give the
    same deconvolution, bit for bit, as the oracle's resampled views added as
    plain images."""
    raws, invs, mn, size = _views(3, (30, 26, 14), 11)
    psfs = [S.psf((7, 7, 9), 1.2, 2.0, a) for a in (0, 120, 240)]
    wt = WeightType.PRECOMPUTED_WEIGHTS if blending is not None else WeightType.NO_WEIGHTS
    a = MVDeconInput()
    for r, inv, p in zip(raws, invs, psfs):
        a.add(MVDeconFFT.from_raw(r, inv, mn, size, p, blending))
    da = MVDeconvolution(a, PSFTYPE.EFFICIENT_BAYESIAN, 3, 0.006, weight_type=wt)
    b = MVDeconInput()
    for r, inv, p in zip(raws, invs, psfs):
        if blending is None:
            img, w = O.transform_view(r, inv, mn, size)
        else:
            img, w = O.transform_view(r, inv, mn, size, blending.border, blending.range)
        b.add(MVDeconFFT(img, w, p))
    db = MVDeconvolution(b, PSFTYPE.EFFICIENT_BAYESIAN, 3, 0.006, weight_type=wt)
    np.testing.assert_array_equal(da.get_psi(), db.get_psi())
    np.testing.assert_array_equal(np.array(da.stats), np.array(db.stats))
    # the view's image on request equals the oracle's
    v0 = a.get_views()[0]
    rimg, _ = O.transform_view(raws[0], invs[0], mn, size)
    np.testing.assert_array_equal(v0.get_image(), rimg)
    da.close()
    db.close()


def test_device_entry_matches_host_entry():
    import torch
    rx, ry, rz = 45, 38, 21
    raw = np.random.default_rng(9).uniform(1.0, 900.0, (rz, ry, rx)).astype(np.float32)
    fwd = _fwd(-35.0, 2.2, (rx / 2, ry / 2, rz / 2), (0.4, 0.0, -0.3))
    inv = T.inverse_affine(fwd)
    mn, size = T.bounding_box([(rx, ry, rz)], [fwd])
    img, w = T.transform_view(raw, inv, mn, size, BLEND)
    dev = torch.device("cuda:0")
    raw_t = torch.from_numpy(raw).to(dev)
    oi = torch.empty((size[2], size[1], size[0]), dtype=torch.float32, device=dev)
    ow = torch.empty_like(oi)
    for _ in range(2):  # the second call reuses the buffers
        oi.fill_(-1.0)
        ow.fill_(-1.0)
        T.transform_view_device(raw_t, inv, mn, size, oi, ow, BLEND)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(oi.cpu().numpy(), img)
        np.testing.assert_array_equal(ow.cpu().numpy(), w)


def test_invalid_arguments_are_rejected():
    """Argument errors come back as NativeError before any launch."""
    from spim_registration_amd import _native
    raw = np.ones((4, 5, 6), np.float32)
    with pytest.raises(_native.NativeError):
        T.transform_view(raw, IDENT, (0, 0, 0), (0, 5, 4))  # empty box
    bad = IDENT.copy()
    bad[5] = np.nan
    with pytest.raises(ValueError):
        T.transform_view(raw, bad, (0, 0, 0), (6, 5, 4))  # caught by the wrapper
    import ctypes
    L = _native.lib()
    inv = np.ascontiguousarray(bad)
    off = np.zeros(3, np.int64)
    out = np.empty((4, 5, 6), np.float32)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    assert L.mvd_transform_view(0, p(raw), 6, 5, 4, p(inv), p(off), 6, 5, 4, None, None, p(out), None) != 0
    b = np.zeros(3, np.float32)
    r = np.zeros(3, np.float32)  # range 0 is invalid
    w = np.empty_like(out)
    assert L.mvd_transform_view(0, p(raw), 6, 5, 4, p(IDENT), p(off), 6, 5, 4, p(b), p(r), None, p(w)) != 0
    assert L.mvd_transform_view(99, p(raw), 6, 5, 4, p(IDENT), p(off), 6, 5, 4, None, None, p(out), None) != 0
