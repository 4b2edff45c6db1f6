"""The oracle's convolution against the committed golden fixtures
(tests/golden/convolution_fixtures.npz, written by scipy.ndimage.convolve in
float64 with tests/golden/make_golden.py): the three out-of-bounds rules of
the path, whole image and split into blocks."""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = np.load(os.path.join(os.path.dirname(__file__), "golden", "convolution_fixtures.npz"))
MODES = [("mirror_single", O.OOB_MIRROR), ("constant_one", O.OOB_ONE), ("zero", O.OOB_ZERO)]


def _rel(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).max() / np.abs(b).max())


@pytest.mark.parametrize("name,oob", MODES, ids=[m[0] for m in MODES])
@pytest.mark.parametrize("block", [None, (8, 6, 9)])
def test_oracle_matches_golden(name, oob, block):
    img, ker, exp = GOLDEN["image"], GOLDEN["kernel"], GOLDEN[name]
    out64 = O.convolve(img.astype(np.float64), ker.astype(np.float64), oob, block_xyz=block)
    assert _rel(out64, exp) < 1e-12
    out32 = O.convolve(img, ker, oob, block_xyz=block)
    assert _rel(out32, exp) < 2e-6
