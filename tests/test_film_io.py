"""CPU checks of the film-file oracle (oracle/film_oracle.py), after the
reference's ExrImageTests.cs / PfmImageTests.cs expectations, plus
independent pins: the OpenEXR ZIP transform restated byte by byte from the
file-layout description, known half encodings (round to nearest even,
subnormals, clamping), Pillow's PNG and PFM decoders, and the library
exporting the encoders (no GPU needed)."""
import io
import struct
import zlib

import numpy as np
import pytest

import film_oracle as F


def _random_plane(n, seed, scale=100.0):
    """ExrImageTests.RandomPlane: HDR, tiny and negative values, all finite."""
    rng = np.random.default_rng(seed)
    p = ((rng.random(n) * 2.0 - 1.0) * scale).astype(np.float32)
    p[0] = np.float32(1.2345e-6)
    if n > 1:
        p[1] = -0.75
    return p


def _spec_deconstruct(raw):
    """ZIP pre-process from the OpenEXR layout description, one byte at a time:
    t1 takes the bytes at even positions, t2 the odd ones; then each byte
    after the first becomes (t[i] - t[i-1] + 128 + 256) mod 256."""
    raw = list(raw)
    t = [0] * len(raw)
    t1, t2 = 0, (len(raw) + 1) // 2
    s = 0
    while s < len(raw):
        t[t1] = raw[s]
        t1 += 1
        s += 1
        if s < len(raw):
            t[t2] = raw[s]
            t2 += 1
            s += 1
    p = t[0] if t else 0
    for i in range(1, len(t)):
        d = (t[i] - p + 128 + 256) & 0xFF
        p = t[i]
        t[i] = d
    return bytes(t)


@pytest.mark.parametrize("length", [0, 1, 2, 7, 256, 1023])
def test_zip_transform_matches_spec_and_inverts(length):
    raw = np.random.default_rng(length * 31 + 5).integers(0, 256, length, dtype=np.uint8).tobytes()
    enc = F.zip_deconstruct(raw)
    assert enc == _spec_deconstruct(raw)
    assert F.zip_reconstruct(enc) == raw


def test_half_encoding_known_values():
    # round to nearest even at the half ulp, subnormals, the +-65504 clamp
    cases = [
        (1.0, 0x3C00), (-2.0, 0xC000), (65504.0, 0x7BFF), (70000.0, 0x7BFF), (np.inf, 0x7BFF), (-1e6, 0xFBFF),
        (2.0 ** -14, 0x0400), (2.0 ** -24, 0x0001), (2.0 ** -25, 0x0000), (3 * 2.0 ** -25, 0x0002),
        (1.0 + 2.0 ** -11, 0x3C00), (1.0 + 3 * 2.0 ** -11, 0x3C02), (65519.0, 0x7BFF), (0.0, 0x0000),
        (-0.0, 0x8000), (2.0 ** -14 - 2.0 ** -25, 0x0400),
    ]
    v = np.array([c[0] for c in cases], np.float32)
    got = F.half_quantize(v).view(np.uint16)
    assert [int(x) for x in got] == [c[1] for c in cases]
    assert np.isnan(F.half_quantize(np.array([np.nan], np.float32))[0])
    # NaN: sign and top payload bits kept, made quiet
    nans = np.array([0x7F800001, 0xFFC00000, 0x7FA00000, 0x7FC02000, 0xFF800400], np.uint32).view(np.float32)
    got = F.half_quantize(nans).view(np.uint16)
    assert [int(x) for x in got] == [0x7E00, 0xFE00, 0x7F00, 0x7E01, 0xFE00]


def test_header_conforms_to_spec():
    W, H = 3, 37
    plane = np.zeros(W * H, np.float32)
    data = F.exr_bytes(W, H, [("R", F.HALF, plane), ("G", F.HALF, plane), ("B", F.HALF, plane),
                              ("Z", F.FLOAT, plane)])
    version, attrs, offsets, _ = F.exr_parse(data)
    assert struct.unpack_from("<I", data, 0)[0] == 0x01312F76 and version == 2
    assert sorted(attrs) == ["channels", "chromaticities", "compression", "dataWindow", "displayWindow",
                             "lineOrder", "pixelAspectRatio", "screenWindowCenter", "screenWindowWidth"]
    assert attrs["channels"][0] == "chlist" and attrs["chromaticities"][0] == "chromaticities"
    assert attrs["compression"][1] == b"\x03" and attrs["lineOrder"][1] == b"\x00"
    assert attrs["dataWindow"][0] == "box2i" and attrs["displayWindow"][0] == "box2i"
    assert struct.unpack("<f", attrs["pixelAspectRatio"][1])[0] == 1.0
    assert struct.unpack("<f", attrs["screenWindowWidth"][1])[0] == 1.0
    assert len(attrs["screenWindowCenter"][1]) == 8
    assert struct.unpack("<4i", attrs["dataWindow"][1]) == (0, 0, W - 1, H - 1)
    assert attrs["dataWindow"][1] == attrs["displayWindow"][1]
    chl = attrs["channels"][1]
    assert F.exr_chlist(chl) == [("B", 1, 1, 1), ("G", 1, 1, 1), ("R", 1, 1, 1), ("Z", 2, 1, 1)]
    assert chl.endswith(b"\0") and len(chl) == 4 * 2 + 4 * 16 + 1
    assert len(offsets) == 3
    for i, off in enumerate(offsets):
        assert struct.unpack_from("<i", data, off)[0] == 16 * i


def test_half_rgb_round_trip_is_half_quantized():
    W, H = 13, 21
    r, g, b = (_random_plane(W * H, s) for s in (1, 2, 3))
    chans, w, h = F.exr_channels(F.exr_bytes(W, H, [("R", F.HALF, r), ("G", F.HALF, g), ("B", F.HALF, b)]))
    assert (w, h) == (W, H)
    assert [c[0] for c in chans] == ["B", "G", "R"]
    for (name, typ, plane), want in zip(chans, (b, g, r)):
        assert typ == F.HALF
        assert np.array_equal(plane.ravel(), F.half_quantize(want).astype(np.float32))


@pytest.mark.parametrize("w,h", [(9, 33), (1, 1), (1, 5), (5, 37), (33, 16)])
def test_float_channel_round_trip_is_bit_exact(w, h):
    z = _random_plane(w * h, w * 1000 + h, scale=5000.0)
    if z.size > 5:
        z[5] = -1.0
    chans, rw, rh = F.exr_channels(F.exr_bytes(w, h, [("Z", F.FLOAT, z)]))
    assert (rw, rh) == (w, h)
    assert np.array_equal(chans[0][2].ravel().view(np.uint32), z.view(np.uint32))


def test_multilayer_mixed_types():
    W, H = 8, 6
    names = ["R", "G", "B", "albedo.R", "albedo.G", "albedo.B", "normal.X", "normal.Y", "normal.Z"]
    planes = [_random_plane(W * H, 100 + i, 1.0) for i in range(len(names))]
    depth = _random_plane(W * H, 200, 50.0)
    chans, _, _ = F.exr_channels(F.exr_bytes(W, H, [(n, F.HALF, p) for n, p in zip(names, planes)]
                                             + [("Z", F.FLOAT, depth)]))
    got = {c[0]: c for c in chans}
    assert len(got) == 10
    for n, p in zip(names, planes):
        assert got[n][1] == F.HALF and np.array_equal(got[n][2].ravel(), F.half_quantize(p).astype(np.float32))
    assert np.array_equal(got["Z"][2].ravel(), depth)
    assert (got["normal.X"][2] < 0).any() and (got["Z"][2] < 0).any()


def test_incompressible_blocks_fall_back_to_raw():
    W, H = 40, 20
    rng = np.random.default_rng(99)
    bits = rng.integers(0, 1 << 16, W * H * 2, dtype=np.uint32).astype(np.uint16)
    bits = bits[((bits >> 10) & 0x1F) != 0x1F][:W * H]
    plane = bits.view(np.float16).astype(np.float32)
    data = F.exr_bytes(W, H, [("R", F.HALF, plane)])
    _, _, offsets, _ = F.exr_parse(data)
    sizes = [(min(16, H - struct.unpack_from("<i", data, o)[0]), struct.unpack_from("<i", data, o + 4)[0])
             for o in offsets]
    assert any(size == lines * W * 2 for lines, size in sizes)
    chans, _, _ = F.exr_channels(data)
    assert np.array_equal(chans[0][2].ravel(), plane)


def test_half_conversion_clamps_and_keeps_nan():
    plane = np.array([1e6, np.inf, -1e6, np.nan, 70000.0, 65504.0], np.float32)
    chans, _, _ = F.exr_channels(F.exr_bytes(6, 1, [("R", F.HALF, plane)]))
    r = chans[0][2].ravel()
    assert list(r[[0, 1, 2, 4, 5]]) == [65504.0, 65504.0, -65504.0, 65504.0, 65504.0]
    assert np.isnan(r[3])


def test_invalid_channel_sets_rejected():
    p = np.zeros(4, np.float32)
    with pytest.raises(ValueError):
        F.exr_bytes(2, 2, [])
    with pytest.raises(ValueError):
        F.exr_bytes(2, 2, [("R", F.HALF, p), ("R", F.HALF, p)])
    with pytest.raises(ValueError):
        F.exr_bytes(2, 2, [("R", F.HALF, np.zeros(3, np.float32))])
    with pytest.raises(ValueError):
        F.exr_bytes(2, 2, [("", F.HALF, p)])


def test_pfm_header_size_and_row_order():
    data = F.pfm_bytes(np.zeros(3 * 2 * 3, np.float32), 3, 2, 3)
    assert data.startswith(b"PF\n3 2\n-1.0\n") and len(data) == 12 + 3 * 2 * 3 * 4
    one = F.pfm_bytes(np.array([1.0, 2.0], np.float32), 1, 2, 1)   # top row 1, bottom row 2
    assert struct.unpack_from("<2f", one, len(one) - 8) == (2.0, 1.0)


def test_pfm_round_trips_and_rejects_bad_magic():
    rng = np.random.default_rng(42)
    planes = (rng.random(7 * 5 * 3) * 200.0 - 50.0).astype(np.float32)
    planes[::7] = np.float32(1.2345e-6)
    back, w, h, c = F.pfm_read(F.pfm_bytes(planes, 7, 5, 3))
    assert (w, h, c) == (7, 5, 3) and np.array_equal(back, planes)
    gray = (np.arange(24, dtype=np.float32) * 1.5 - 3.0)
    back, w, h, c = F.pfm_read(F.pfm_bytes(gray, 4, 6, 1))
    assert (w, h, c) == (4, 6, 1) and np.array_equal(back, gray)
    with pytest.raises(ValueError):
        F.pfm_read(b"P6\n1 1\n255\nxxx")


def test_pfm_grayscale_matches_pillow():
    from PIL import Image
    gray = np.random.default_rng(3).normal(size=(6, 4)).astype(np.float32)
    im = Image.open(io.BytesIO(F.pfm_bytes(gray.ravel(), 4, 6, 1)))
    assert im.mode == "F" and im.size == (4, 6)
    assert np.array_equal(np.asarray(im, np.float32), gray)


def test_pfm_big_endian_is_read():
    v = np.array([[1.5, -2.0], [3.25, 8.0]], np.float32)
    data = b"Pf\n2 2\n1.0\n" + v[::-1].astype(">f4").tobytes()
    back, w, h, c = F.pfm_read(data)
    assert (w, h, c) == (2, 2, 1) and np.array_equal(back, v.ravel())


def test_png_quantization_known_values():
    vals = np.array([0.0, 1.0, 0.5, 0.999, 1.0 / 256.0, np.nan, np.inf, -1.0, 2.0, 0.00390624], np.float32)
    px = F.png_pixels(np.stack([vals, vals, vals]).reshape(3, 1, -1), vals.size, 1)
    assert list(px[0, :, 0]) == [0, 255, 128, 255, 1, 0, 255, 0, 255, 0]
    assert (px[..., 3] == 255).all()


def test_png_decoder_matches_pillow():
    # a PNG written by Pillow (adaptive filters) decodes to the same pixels
    from PIL import Image
    rgba = np.random.default_rng(5).integers(0, 256, (9, 13, 4), dtype=np.uint8)
    rgba[2:5] = rgba[2]   # repeated rows make the Up/Paeth filters pay
    buf = io.BytesIO()
    Image.fromarray(rgba, "RGBA").save(buf, "PNG", optimize=True)
    assert np.array_equal(F.png_decode(buf.getvalue()), rgba)


def test_library_exports_the_encoders(native):
    L = native.load()
    for name in ("rt_exr_encode", "rt_pfm_encode", "rt_png_encode"):
        assert hasattr(L, name)
    assert zlib.ZLIB_RUNTIME_VERSION   # the encoders link the same system zlib
