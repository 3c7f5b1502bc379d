"""GPU film-file encoders (rt_exr_encode / rt_pfm_encode / rt_png_encode)
against the oracle (oracle/film_oracle.py).

EXR and PFM files must equal the oracle's bytes exactly: the device's
half conversion, interleaving and ZIP pre-transform are integer/byte work,
and both sides deflate with the same system zlib at level 6.  PNG pixels
must equal the oracle's quantization after decoding (the PNG container is
checked by its CRCs).  The half conversion is also checked on every
finite half value and on a million random float bit patterns."""
import numpy as np
import pytest

import film_oracle as F
from raytracer import film_io, native, scenes

pytestmark = pytest.mark.gpu


def _planes(n, k, seed, scale=100.0):
    rng = np.random.default_rng(seed)
    p = ((rng.random((k, n)) * 2.0 - 1.0) * scale).astype(np.float32)
    p[:, 0] = np.float32(1.2345e-6)
    return p


def _same_bytes(got, want):
    """Byte equality with a short message (no pytest diff of large bytes)."""
    if got == want:
        return
    n = min(len(got), len(want))
    a, b = np.frombuffer(got[:n], np.uint8), np.frombuffer(want[:n], np.uint8)
    bad = np.nonzero(a != b)[0]
    first = int(bad[0]) if bad.size else n
    raise AssertionError(f"lengths {len(got)} vs {len(want)}, {bad.size} differing bytes, first at {first}: "
                         f"{got[first:first + 8].hex()} vs {want[first:first + 8].hex()}")


def _exr_case(gpu_ctx, w, h, channels):
    got = film_io.exr_encode(w, h, [film_io.Channel(*c) for c in channels], gpu_ctx)
    want = F.exr_bytes(w, h, channels)
    print(f"exr {w}x{h} x{len(channels)}: {len(got)} bytes", flush=True)
    _same_bytes(got, want)


@pytest.mark.parametrize("w,h", [(13, 21), (1, 1), (1, 5), (5, 37), (33, 16), (9, 33)])
def test_exr_rgb_and_depth_bytes_match_oracle(gpu_ctx, w, h):
    p = _planes(w * h, 4, w * 100 + h)
    _exr_case(gpu_ctx, w, h, [("R", F.HALF, p[0]), ("G", F.HALF, p[1]), ("B", F.HALF, p[2]), ("Z", F.FLOAT, p[3])])


def test_exr_multilayer_bytes_match_oracle(gpu_ctx):
    W, H = 67, 45
    names = ["R", "G", "B", "albedo.R", "albedo.G", "albedo.B", "normal.X", "normal.Y", "normal.Z",
             "variance.R", "variance.G", "variance.B"]
    p = _planes(W * H, len(names) + 1, 7, 3.0)
    chans = [(n, F.HALF, p[i]) for i, n in enumerate(names)] + [("Z", F.FLOAT, p[-1] * 20.0)]
    _exr_case(gpu_ctx, W, H, chans[::-1])   # input order does not matter: stored in ordinal name order


def test_exr_incompressible_and_special_values(gpu_ctx):
    W, H = 40, 20
    rng = np.random.default_rng(99)
    bits = rng.integers(0, 1 << 16, 4 * W * H, dtype=np.uint32).astype(np.uint16)
    bits = bits[((bits >> 10) & 0x1F) != 0x1F][:W * H]
    _exr_case(gpu_ctx, W, H, [("R", F.HALF, bits.view(np.float16).astype(np.float32))])
    special = np.array([1e6, np.inf, -1e6, np.nan, 70000.0, 65504.0, -np.inf, 65520.0, 2.0 ** -25, -2.0 ** -24],
                       np.float32)
    _exr_case(gpu_ctx, special.size, 1, [("R", F.HALF, special), ("Z", F.FLOAT, special)])


def test_exr_half_conversion_exhaustive(gpu_ctx):
    # every finite half value, the midpoints between neighbours (ties go to
    # even), and a million random float32 bit patterns (NaNs compared as NaN)
    h = np.arange(1 << 16, dtype=np.uint32).astype(np.uint16)
    h = h[((h >> 10) & 0x1F) != 0x1F]
    exact = np.sort(h.view(np.float16).astype(np.float32))
    mid = ((exact[:-1].astype(np.float64) + exact[1:].astype(np.float64)) / 2).astype(np.float32)
    rnd = np.random.default_rng(11).integers(0, 1 << 32, 1 << 20, dtype=np.uint64).astype(np.uint32).view(np.float32)
    v = np.concatenate([exact, mid, rnd])
    W = 1024
    H = (v.size + W - 1) // W
    plane = np.zeros(W * H, np.float32)
    plane[:v.size] = v
    data = film_io.exr_encode(W, H, [film_io.Channel("R", film_io.HALF, plane)], gpu_ctx)
    chans, _, _ = F.exr_channels(data)
    raw = F.exr_scanlines(W, H, [("R", F.HALF, plane)])   # the oracle's half bits, NaN payloads included
    want = np.frombuffer(b"".join(raw), "<u2")
    got = np.concatenate([np.asarray(chans[0][2][y], np.float32).astype(np.float16).view(np.uint16) for y in range(H)])
    nan = np.isnan(want.view(np.float16))
    print(f"half values {plane.size}, NaN {int(nan.sum())}", flush=True)
    assert np.array_equal(got[~nan], want[~nan])
    assert np.array_equal(np.isnan(got.view(np.float16)), nan)
    _same_bytes(data, F.exr_bytes(W, H, [("R", F.HALF, plane)]))   # NaN payloads: whole file equal


def test_exr_invalid_channel_sets_raise(gpu_ctx):
    p = np.zeros(4, np.float32)
    for bad in ([], [("R", 1, p), ("R", 1, p)], [("", 1, p)], [("bad name", 1, p)], [("R", 1, np.zeros(3))],
                [("x" * 32, 1, p)], [("R", 3, p)]):
        with pytest.raises(ValueError):
            film_io.exr_encode(2, 2, [film_io.Channel(*c) for c in bad], gpu_ctx)


@pytest.mark.parametrize("w,h,c", [(7, 5, 3), (4, 6, 1), (1, 1, 3), (1, 2, 1), (333, 77, 3)])
def test_pfm_bytes_match_oracle(gpu_ctx, w, h, c, tmp_path):
    p = _planes(w * h * c, 1, w * 7 + h, 50.0)[0]
    got = film_io.pfm_encode(p, w, h, c, gpu_ctx)
    _same_bytes(got, F.pfm_bytes(p, w, h, c))
    path = str(tmp_path / "x.pfm")
    film_io.pfm_write(path, p, w, h, c, gpu_ctx)
    back, rw, rh, rc = film_io.pfm_read(path)
    assert (rw, rh, rc) == (w, h, c) and np.array_equal(back, p)
    with pytest.raises(ValueError):
        film_io.pfm_encode(p, w, h, 2, gpu_ctx)


@pytest.mark.parametrize("w,h", [(1, 1), (13, 7), (640, 361)])
def test_png_pixels_match_oracle(gpu_ctx, w, h):
    rng = np.random.default_rng(w + h)
    d = (rng.random((3, h, w)) * 1.3 - 0.1).astype(np.float32)
    d.ravel()[::97] = np.nan
    d.ravel()[5::101] = np.inf
    d.ravel()[7::103] = -np.inf
    data = film_io.png_encode(d, w, h, gpu_ctx)
    assert np.array_equal(F.png_decode(data), F.png_pixels(d, w, h))


def test_render_to_files_round_trip(gpu_ctx, tmp_path):
    # ExrRenderOutputTests: a rendered beauty written as EXR reads back
    # half-quantized; the AOV files carry the RenderBuffers combinations
    desc = scenes.cornell(32, 32)
    sc = native.Scene(gpu_ctx, desc)
    try:
        res = native.Renderer(sc, 8, 6).render(32, 32, capture_aovs=True)
    finally:
        sc.close()
    buf = res.buffers
    path = str(tmp_path / "out.exr")
    n = film_io.save_exr_beauty(path, buf.beauty, buf, ["albedo", "normal", "depth", "variance", "beauty"], gpu_ctx)
    assert n == 4
    chans, w, h = F.exr_channels(open(path, "rb").read())
    got = {c[0]: c[2] for c in chans}
    assert (w, h) == (32, 32) and len(got) == 3 + 3 + 3 + 1 + 3
    for k, name in enumerate("RGB"):
        assert np.array_equal(got[name], F.half_quantize(buf.beauty[k]).astype(np.float32))
    assert np.array_equal(got["Z"], film_io.combine_depth_halves(buf))
    alb = film_io.combine_halves(buf, buf.albedo_a, buf.albedo_b)
    assert np.array_equal(got["albedo.G"], F.half_quantize(alb[1]).astype(np.float32))
    assert np.array_equal(got["normal.Z"], F.half_quantize(film_io.combine_halves(buf, buf.normal_a, buf.normal_b)[2])
                          .astype(np.float32))
    paths = film_io.save_aovs(buf, ["depth", "variance"], str(tmp_path / "out.png"), "pfm", ctx=gpu_ctx)
    assert paths[0].endswith("out.depth.pfm") and paths[1].endswith("out.variance.pfm")
    d, w, h, c = film_io.pfm_read(paths[0])
    assert c == 1 and np.array_equal(d.reshape(32, 32), film_io.combine_depth_halves(buf))
    v, w, h, c = film_io.pfm_read(paths[1])
    assert c == 3 and np.array_equal(v.reshape(3, 32, 32), film_io.raw_beauty_variance(buf))
    png = str(tmp_path / "out.png")
    film_io.save_image(png, res.pixels, gpu_ctx)
    disp = np.transpose(res.pixels, (2, 0, 1))
    assert np.array_equal(F.png_decode(open(png, "rb").read()), F.png_pixels(disp, 32, 32))
