"""GPU output path against the oracle: min/max and the 16-bit conversion
bit-exact (host and device inputs, odd sizes, misaligned pointers, NaN, signed
zeros, infinities), and TIFF stacks read back with an independent reader
(PIL) equal to the image or its oracle conversion."""
import math
import os

import numpy as np
import pytest
import torch

from oracle import export as OE
from spim_registration_amd import export as E
from spim_registration_amd import MVDeconFFT, MVDeconInput, MVDeconvolution, PSFTYPE
from spim_registration_amd import synthetic as S

pytestmark = pytest.mark.gpu


def _read_tiff(path):
    from PIL import Image
    im = Image.open(path)
    tags = dict(im.tag_v2)  # the first page carries the description
    pages = []
    for i in range(im.n_frames):
        im.seek(i)
        pages.append(np.array(im))
    return np.stack(pages), tags


def _img(shape, seed=0):
    return np.random.default_rng(seed).uniform(-3.0, 250.0, shape).astype(np.float32)


@pytest.mark.parametrize("n", [1, 3, 4, 1023, 1 << 20, (1 << 22) + 5])
def test_min_max_matches_oracle(n):
    a = _img(n, n)
    assert E.min_max(a) == OE.min_max(a)
    t = torch.from_numpy(a).cuda()
    assert E.min_max(t) == OE.min_max(a)
    # misaligned device view (offset of one float)
    if n > 1:
        assert E.min_max(t[1:]) == OE.min_max(a[1:])


def test_min_max_special_values():
    a = np.zeros(5000, np.float32)
    a[17] = -0.0
    mn, mx = E.min_max(torch.from_numpy(a).cuda())
    assert mn == 0 and np.signbit(mn) and not np.signbit(mx)
    a[4000] = np.nan
    mn, mx = E.min_max(a)
    assert np.isnan(mn) and np.isnan(mx)
    b = np.full(777, np.inf, np.float32)
    assert E.min_max(torch.from_numpy(b).cuda()) == OE.min_max(b)
    assert E.min_max(np.zeros(0, np.float32)) == OE.min_max([])


@pytest.mark.parametrize("big_endian", [False, True])
@pytest.mark.parametrize("n,off", [(4096, 0), (100003, 0), (100003, 1), (7, 3)])
def test_convert_u16_matches_oracle(n, off, big_endian):
    a = _img(n + off, 7 + n)
    a[::97] = np.nan
    a[::89] = np.inf
    t = torch.from_numpy(a).cuda()[off:]
    lo, hi = E.u16_range(E.min_max(a[off:][np.isfinite(a[off:])]))
    u = E.convert_u16(t, lo, hi, big_endian=big_endian)
    torch.cuda.synchronize()
    got = u.cpu().numpy().view(np.uint16)
    exp = OE.to_u16(a[off:], lo, hi)
    if big_endian:
        exp = exp.byteswap()
    np.testing.assert_array_equal(got, exp)


@pytest.mark.parametrize("big_endian", [True, False])
@pytest.mark.parametrize("on_device", [False, True])
def test_tiff_float_stack_roundtrip(tmp_path, big_endian, on_device):
    img = _img((5, 33, 47), 3)
    src = torch.from_numpy(img).cuda() if on_device else img
    p = E.write_tiff(str(tmp_path / "f.tif"), src, bits=32, min=0.0, max=1.0, big_endian=big_endian)
    pages, tags = _read_tiff(p)
    np.testing.assert_array_equal(pages.astype(np.float32), img)
    desc = tags[270]
    assert "images=5" in desc and "slices=5" in desc and "min=0.0" in desc and "max=1.0" in desc
    with open(p, "rb") as f:
        assert f.read(2) == (b"MM" if big_endian else b"II")


def test_tiff_u16_stack_and_calibration(tmp_path):
    img = _img((4, 20, 30), 4)
    cal = E.calibration((-7, 3, -9), 2)
    s = E.Save3dTIFF(str(tmp_path))
    p = s.export_image(torch.from_numpy(img).cuda(), "TP0_Ch0", bb_min=(-7, 3, -9), downsampling=2, bits=16)
    assert p == str(tmp_path / "TP0_Ch0.tif")
    pages, tags = _read_tiff(p)
    lo, hi = OE.u16_range(OE.min_max(img))
    np.testing.assert_array_equal(pages.astype(np.uint16), OE.to_u16(img, lo, hi))
    desc = tags[270]
    assert "spacing=2.0" in desc and f"xorigin={cal[0]}" in desc and f"zorigin={cal[2]}" in desc
    assert "yorigin" not in desc or cal[1] != 0
    assert float(tags[282]) == pytest.approx(0.5)


def _read_bigtiff(path):
    """Minimal BigTIFF reader (both byte orders): one strip per page.  PIL
    recognises BigTIFF only in little-endian files (it tests byte 2 of the
    header for 43), so big-endian files are read here."""
    import struct
    b = open(path, "rb").read()
    e = {b"MM": ">", b"II": "<"}[b[:2]]
    assert struct.unpack(e + "HHH", b[2:8]) == (43, 8, 0)
    off = struct.unpack(e + "Q", b[8:16])[0]
    pages, desc = [], None
    while off:
        n = struct.unpack(e + "Q", b[off:off + 8])[0]
        tags = {}
        for k in range(n):
            ent = b[off + 8 + 20 * k: off + 28 + 20 * k]
            tag, typ, cnt = struct.unpack(e + "HHQ", ent[:12])
            fmt = {3: "H", 4: "L", 16: "Q"}.get(typ)
            tags[tag] = (typ, cnt, struct.unpack(e + fmt, ent[12:12 + struct.calcsize(e + fmt)])[0] if fmt else
                         struct.unpack(e + "Q", ent[12:])[0])
        w, h, bits = tags[256][2], tags[257][2], tags[258][2]
        dt = np.dtype(e + ("f4" if tags[339][2] == 3 else "u2"))
        so, bc = tags[273][2], tags[279][2]
        assert bc == w * h * bits // 8
        pages.append(np.frombuffer(b[so:so + bc], dt).reshape(h, w))
        if 270 in tags:
            desc = b[tags[270][2]:tags[270][2] + tags[270][1] - 1].decode()
        off = struct.unpack(e + "Q", b[off + 8 + 20 * n: off + 16 + 20 * n])[0]
    return np.stack(pages), desc


@pytest.mark.parametrize("big_endian", [False, True])
def test_tiff_range_from_image_and_bigtiff(tmp_path, big_endian):
    img = _img((3, 16, 18), 5)
    p = E.write_tiff(str(tmp_path / "b.tif"), img, bits=32, bigtiff=True, big_endian=big_endian)
    with open(p, "rb") as f:
        assert f.read(4) == (b"MM\x00\x2b" if big_endian else b"II\x2b\x00")
    pages, desc = _read_bigtiff(p)
    np.testing.assert_array_equal(pages.astype(np.float32), img)
    if not big_endian:
        pil_pages, tags = _read_tiff(p)
        np.testing.assert_array_equal(pil_pages.astype(np.float32), img)
        assert tags[270] == desc
    mn, mx = OE.min_max(img)
    assert f"min={float(mn)!r}" in desc and f"max={float(mx)!r}" in desc
    p16 = E.write_tiff(str(tmp_path / "b16.tif"), torch.from_numpy(img).cuda(), bits=16, bigtiff=True,
                       big_endian=big_endian)
    pages16, desc16 = _read_bigtiff(p16)
    np.testing.assert_array_equal(pages16.astype(np.uint16), OE.to_u16(img, *OE.u16_range((mn, mx))))
    assert "min=0.0" in desc16 and "max=65535.0" in desc16


def test_invalid_arguments_are_rejected(tmp_path):
    from spim_registration_amd import _native
    img = _img((2, 4, 4), 6)
    with pytest.raises(_native.NativeError):
        E.write_tiff(str(tmp_path / "x.tif"), img, bits=8)
    with pytest.raises(_native.NativeError):
        E.write_tiff(str(tmp_path / "no_such_dir" / "x.tif"), img)
    with pytest.raises(_native.NativeError):
        E.write_tiff(str(tmp_path / "x.tif"), img, calibration=(0, 0, 0, -1.0))


def test_export_psi_equals_masked_estimate(tmp_path):
    ds = S.make_dataset((40, 36, 24), 3, [0, 120, 240], 30, (7, 7, 9), 1.2, 2.5, seed=8)
    inp = MVDeconInput()
    for v, p in zip(ds.views, ds.psfs):
        inp.add(MVDeconFFT(v, None, p, [0]))
    d = MVDeconvolution(inp, PSFTYPE.EFFICIENT_BAYESIAN, 2, 0.0)
    psi = d.get_psi(masked=True)
    # EfficientBayesianBased.java:294-299 exports with the range (0, 1)
    p = E.Save3dTIFF(str(tmp_path)).export_image(d, "deconvolved", min=0.0, max=1.0)
    pages, tags = _read_tiff(p)
    np.testing.assert_array_equal(pages.astype(np.float32), psi)
    assert "min=0.0" in tags[270] and "max=1.0" in tags[270]
    p16 = d.export_tiff(str(tmp_path / "d16.tif"), bits=16, min=math.nan, max=math.nan)
    pages16, _ = _read_tiff(p16)
    lo, hi = OE.u16_range(OE.min_max(psi))
    np.testing.assert_array_equal(pages16.astype(np.uint16), OE.to_u16(psi, lo, hi))
    d.close()
