"""CPU-side checks of the native libraries and the host logic (no GPU calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import oracle as O
from spim_registration_amd import _native, cuda, synthetic as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    txt = open(os.path.join(ROOT, "include", "mvdecon.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", txt, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while")))


def test_library_exports_every_header_symbol():
    L = _native.lib()
    declared = _header_functions()
    assert len(declared) >= 20
    missing = [n for n in declared if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(declared) == sorted(_native.EXPORTS)


def test_bindings_match_header_arity():
    """Every ctypes declaration takes as many arguments as the header's prototype."""
    txt = open(os.path.join(ROOT, "include", "mvdecon.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    L = _native.lib()
    checked = 0
    for name, params in re.findall(r"\b([A-Za-z_]\w*)\s*\(([^;{]*?)\)\s*;", txt, flags=re.S):
        if name not in _native.EXPORTS:
            continue
        p = params.strip()
        n = 0 if p in ("", "void") else p.count(",") + 1
        argtypes = getattr(L, name).argtypes
        if argtypes is not None or n == 0:
            assert len(argtypes or []) == n, (name, n, argtypes)
            checked += 1
    assert checked >= 30


def test_library_is_built_for_gfx950():
    data = open(_native.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(_native, "_lib", None)
    monkeypatch.setattr(_native, "LIB_PATH", "/nonexistent/libmvdecon_hip.so")
    with pytest.raises(_native.NativeError):
        _native.lib()


def test_generator_is_deterministic():
    a = S.make_dataset((20, 16, 12), 2, [0, 90], 10, (5, 5, 7), 1.2, 3.0, seed=11)
    b = S.make_dataset((20, 16, 12), 2, [0, 90], 10, (5, 5, 7), 1.2, 3.0, seed=11, threads=1)
    for x, y in zip(a.views + a.psfs, b.views + b.psfs):
        np.testing.assert_array_equal(x, y)
    c = S.make_dataset((20, 16, 12), 2, [0, 90], 10, (5, 5, 7), 1.2, 3.0, seed=12)
    assert not np.array_equal(a.views[0], c.views[0])


@pytest.mark.parametrize("lam", [3.5, 25.0, 150.0, 1800.0])
def test_poisson_sampler_moments(lam):
    x = S.poisson_samples(5, 1, lam, 200000)
    assert np.all(x >= 0) and np.all(x == np.floor(x))
    assert abs(x.mean() - lam) < 5 * np.sqrt(lam / x.size)
    assert abs(x.var() / lam - 1) < 0.03


def test_uniform_range():
    u = S.uniform(1, 2, 100000)
    assert u.min() >= 0 and u.max() < 1 and abs(u.mean() - 0.5) < 0.01


def test_psf_normalised_centred_and_rotated():
    k = S.psf((19, 19, 31), 1.2, 4.0, 0.0)
    assert abs(k.astype(np.float64).sum() - 1) < 1e-5
    assert np.unravel_index(np.argmax(k), k.shape) == (15, 9, 9)
    np.testing.assert_array_equal(k, k[::-1, ::-1, ::-1])  # centro-symmetric
    k90 = S.psf((31, 31, 31), 1.2, 4.0, 90.0)
    # rotated by 90 deg about y the long axis lies along x
    prof_x = k90[15, 15, :]
    prof_z = k90[:, 15, 15]
    assert prof_x[20] > prof_z[20]


def test_blending_weight_profile():
    w = S.blending_weight((20, 30, 40))
    assert w.shape == (20, 30, 40) and np.all(w > 0) and np.all(w <= 1)
    assert w[10, 15, 20] == 1.0
    # border -8, range 12: at the face the relative distance is 8/12
    lut = S._blend_lut()
    assert abs(w[10, 15, 0] - lut[667]) < 1e-6


def test_python_block_generator_matches_oracle():
    for img, ker, bs in [((100, 37, 20), (5, 7, 3), (24, 16, 9)), ((512, 512, 256), (31, 31, 51), (256,) * 3)]:
        py = cuda.BlockGeneratorFixedSizePrecise(bs).divide_into_blocks(img, ker)
        ref = O.divide_blocks(img, ker, bs)
        assert len(py) == len(ref)
        for b, r in zip(py, ref):
            assert tuple(b.offset) == r["offset"]
            assert tuple(b.effective_size) == r["eff_size"]
            assert tuple(b.effective_offset) == r["eff_offset"]
            assert tuple(b.effective_local_offset) == r["eff_local"]


def test_block_copy_paste_round_trip():
    rng = np.random.default_rng(0)
    img = rng.random((20, 30, 25)).astype(np.float32)
    blocks = cuda.BlockGeneratorFixedSizePrecise((12, 16, 9)).divide_into_blocks((25, 30, 20), (5, 7, 3))
    out = np.zeros_like(img)
    for b in blocks:
        blk = b.copy_block(img, cuda.OOB_MIRROR_SINGLE)
        b.paste_block(out, blk)
    np.testing.assert_array_equal(out, img)
    # constant-one extension outside the image
    b = blocks[0]
    blk = b.copy_block(img, cuda.OOB_VALUE_ONE)
    assert np.all(blk[:, :, 0] == 1)
