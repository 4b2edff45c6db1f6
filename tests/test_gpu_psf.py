"""GPU PSF extraction (mvd_extract_psf) against the oracle: bit-identical
normalised and transformed PSFs (the kernels keep the oracle's arithmetic and
the bead order of the float sums)."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from spim_registration_amd import psf as P
from spim_registration_amd.transform import inverse_affine

pytestmark = pytest.mark.gpu


def _model(angle_deg, scale_z):
    a = math.radians(angle_deg)
    r = np.array([[math.cos(a), 0, math.sin(a)], [0, 1, 0], [-math.sin(a), 0, math.cos(a)]])
    m = r @ np.diag([1.0, 1.0, scale_z])
    return np.hstack([m, np.array([[3.5], [-2.0], [10.25]])]).reshape(12)


CASES = [
    ("small_odd", (40, 36, 24), 7, (9, 9, 13), 0.0, 1.0),
    ("rot60_z2", (64, 60, 32), 25, (15, 13, 21), 60.0, 2.0),
    ("even_sizes", (33, 47, 20), 11, (10, 12, 8), -35.0, 1.7),
    ("edge_beads", (50, 40, 30), 40, (21, 21, 31), 90.0, 3.0),
]


@pytest.mark.parametrize("name,img_xyz,nbeads,psf_xyz,angle,sz", CASES, ids=[c[0] for c in CASES])
def test_extract_psf_matches_oracle(name, img_xyz, nbeads, psf_xyz, angle, sz):
    nx, ny, nz = img_xyz
    rng = np.random.default_rng(nbeads)
    img = rng.uniform(0.0, 1000.0, (nz, ny, nx)).astype(np.float32)
    # bead positions anywhere in the image (near faces too: periodic extension)
    locs = np.stack([rng.uniform(0, nx - 1, nbeads), rng.uniform(0, ny - 1, nbeads),
                     rng.uniform(0, nz - 1, nbeads)], 1)
    model = _model(angle, sz)
    print("case", name, img_xyz, nbeads, psf_xyz, flush=True)
    psf, tpsf = P.extract_psf(img, locs, psf_xyz, model)
    _, rpsf = O.extract_psf(img, locs, psf_xyz)
    np.testing.assert_array_equal(psf, rpsf)
    rt = O.psf_transform(rpsf, model, inverse_affine(model))
    assert tpsf.shape == rt.shape
    np.testing.assert_array_equal(tpsf, rt)
    assert psf.min() == 0.0 and psf.max() == 1.0


def test_extract_psf_container():
    rng = np.random.default_rng(8)
    img = rng.uniform(0.0, 10.0, (20, 30, 30)).astype(np.float32)
    e = P.ExtractPSF()
    for v, ang in enumerate((0.0, 90.0)):
        locs = np.stack([rng.uniform(5, 25, 6), rng.uniform(5, 25, 6), rng.uniform(3, 17, 6)], 1)
        e.extract_next_img(img, v, _model(ang, 1.5), locs, (7, 7, 9))
    assert e.get_view_ids_for_psfs() == [0, 1]
    assert e.get_transformed_psf(1).shape == tuple(reversed(P.transformed_psf_size((7, 7, 9), _model(90.0, 1.5))))
    avg = e.compute_average_psf()
    np.testing.assert_array_equal(avg, (e.original_psfs[0] + e.original_psfs[1]).astype(np.float32))


def test_invalid_arguments_are_rejected():
    from spim_registration_amd import _native
    img = np.ones((10, 10, 10), np.float32)
    with pytest.raises(_native.NativeError):
        P.extract_psf(img, [(5.0, 5.0, 5.0)], (0, 5, 5))
    m = _model(0.0, 1.0).copy()
    m[0] = np.inf
    with pytest.raises(_native.NativeError):
        P.transformed_psf_size((5, 5, 5), m)
