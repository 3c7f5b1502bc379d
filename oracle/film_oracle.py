"""CPU restatement of the reference's film-file writers, and independent
decoders of the three formats.

TEST INFRASTRUCTURE ONLY — imported by tests/, never by the product.

  exr_bytes       ExrImage.Write            src/RayTracer/Rendering/ExrImage.cs:52-137
                  (header :158-212, half encoding :234-236, ZIP transform
                  :248-266, raw fallback :111-113)
  zip_deconstruct ExrImage.ZipDeconstruct   ExrImage.cs:248-266
  zip_reconstruct ExrImage.ZipReconstruct   ExrImage.cs:272-287
  exr_channels    ExrLoader.LoadChannels    src/RayTracer/Textures/ExrLoader.cs:66 (the
                  decoder is written from the OpenEXR file layout, not from the writer)
  pfm_bytes       PfmImage.Write            src/RayTracer/Rendering/PfmImage.cs:29-65
  pfm_read        PfmImage.Read             PfmImage.cs:71-117
  png_pixels      Program.SaveImage         src/RayTracer/Program.cs:723-737
  png_decode      a PNG decoder (zlib + the five scanline filters), for checking

Pinned by: the reference's ExrImageTests / PfmImageTests expectations
(header attribute set, types and payloads, chlist order and layout, offset
table, ZIP transform against the spec's split-then-predict order, half
clamping and NaN, raw-block fallback on incompressible data, PFM header
text, byte count and bottom-to-top row order), numpy's float16 conversion
(IEEE round to nearest even, subnormals) for the half encoding, and round
trips through the decoders here (tests/test_film_io.py).  zlib level 6 is
.NET's CompressionLevel.Optimal; compressed bytes are a property of the
zlib build, the decoded content is the contract.
"""
import struct
import zlib

import numpy as np

HALF, FLOAT = 1, 2
LINES = 16
HALF_MAX = 65504.0
CHROMATICITIES = (0.6400, 0.3300, 0.3000, 0.6000, 0.1500, 0.0600, 0.3127, 0.3290)


def half_quantize(plane):
    """Math.Clamp(v, -65504, 65504) then the Half cast (ExrImage.cs:234-236):
    numpy's float16 conversion (round to nearest even, subnormals) for every
    non-NaN value; a NaN keeps its sign and the top 10 bits of its payload
    and is made quiet (sign | 0x7e00 | payload >> 13), as System.Half's cast
    does (numpy instead keeps a signalling NaN signalling)."""
    v = np.clip(np.asarray(plane, np.float32), -HALF_MAX, HALF_MAX)
    h = v.astype(np.float16)
    nan = np.isnan(v)
    if nan.any():
        u = v.view(np.uint32)
        bits = h.view(np.uint16).copy()
        bits[nan] = (((u[nan] >> 16) & 0x8000) | 0x7E00 | ((u[nan] >> 13) & 0x3FF)).astype(np.uint16)
        h = bits.view(np.float16)
    return h


def zip_deconstruct(raw):
    """Split (even-index bytes, then odd) THEN delta predictor over the
    reordered buffer, each byte = cur - prev_original + 128."""
    raw = np.frombuffer(bytes(raw), np.uint8)
    t = np.concatenate([raw[0::2], raw[1::2]])
    out = t.copy()
    if len(t) > 1:
        out[1:] = (t[1:].astype(np.int32) - t[:-1].astype(np.int32) + 128) & 0xFF
    return out.tobytes()


def zip_reconstruct(enc):
    """Predictor undo (prefix sum of d - 128) THEN interleave the halves."""
    e = np.frombuffer(bytes(enc), np.uint8).astype(np.int64)
    if len(e) == 0:
        return b""
    d = e.copy()
    d[1:] -= 128
    t = (np.cumsum(d) & 0xFF).astype(np.uint8)
    half = (len(t) + 1) // 2
    raw = np.empty_like(t)
    raw[0::2] = t[:half]
    raw[1::2] = t[half:]
    return raw.tobytes()


def _check_channels(width, height, channels):
    if width <= 0 or height <= 0:
        raise ValueError("invalid EXR dimensions")
    if not channels:
        raise ValueError("EXR requires at least one channel")
    seen = set()
    for name, typ, plane in channels:
        if not name or len(name) > 31 or any(c <= " " or c > "~" for c in name):
            raise ValueError(f"invalid EXR channel name {name!r}")
        if name in seen:
            raise ValueError(f"duplicate EXR channel name {name!r}")
        seen.add(name)
        if np.asarray(plane).size != width * height:
            raise ValueError(f"channel {name!r} plane size")


def _sorted(channels):
    # ordinal order = byte order of the ASCII names
    return sorted(channels, key=lambda c: c[0].encode("ascii"))


def exr_header(width, height, channels):
    chl = b""
    for name, typ, _ in _sorted(channels):
        chl += name.encode("ascii") + b"\0" + struct.pack("<iBBBBii", typ, 0, 0, 0, 0, 1, 1)
    chl += b"\0"
    box = struct.pack("<4i", 0, 0, width - 1, height - 1)

    def attr(name, typ, payload):
        return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(payload)) + payload

    h = b"".join([
        attr("channels", "chlist", chl),
        attr("chromaticities", "chromaticities", struct.pack("<8f", *CHROMATICITIES)),
        attr("compression", "compression", b"\x03"),
        attr("dataWindow", "box2i", box),
        attr("displayWindow", "box2i", box),
        attr("lineOrder", "lineOrder", b"\x00"),
        attr("pixelAspectRatio", "float", struct.pack("<f", 1.0)),
        attr("screenWindowCenter", "v2f", b"\0" * 8),
        attr("screenWindowWidth", "float", struct.pack("<f", 1.0)),
    ])
    return h + b"\0"


def exr_scanlines(width, height, channels):
    """Raw (uncompressed) payload: per scanline, each channel's samples in
    ordinal name order, little-endian half or float."""
    rows = []
    enc = []
    for name, typ, plane in _sorted(channels):
        p = np.asarray(plane, np.float32).reshape(height, width)
        enc.append(half_quantize(p).view(np.uint16).astype("<u2") if typ == HALF else p.astype("<f4"))
    for y in range(height):
        rows.append(b"".join(e[y].tobytes() for e in enc))
    return rows


def exr_bytes(width, height, channels):
    """ExrImage.Write into bytes; channels = [(name, HALF|FLOAT, plane)]."""
    _check_channels(width, height, channels)
    rows = exr_scanlines(width, height, channels)
    header = exr_header(width, height, channels)
    nb = (height + LINES - 1) // LINES
    chunks = []
    for b in range(nb):
        raw = b"".join(rows[b * LINES:min(height, (b + 1) * LINES)])
        z = zlib.compress(zip_deconstruct(raw), 6)
        payload = z if len(z) < len(raw) else raw
        chunks.append(struct.pack("<ii", b * LINES, len(payload)) + payload)
    out = struct.pack("<Ii", 0x01312F76, 2) + header
    off = 8 + len(header) + 8 * nb
    for c in chunks:
        out += struct.pack("<q", off)
        off += len(c)
    return out + b"".join(chunks)


def _cstr(data, pos):
    end = data.index(b"\0", pos)
    return data[pos:end].decode("ascii"), end + 1


def exr_parse(data):
    """Header attributes {name: (type, payload)}, offset table and the byte
    position where the table starts."""
    magic, version = struct.unpack_from("<Ii", data, 0)
    if magic != 0x01312F76:
        raise ValueError("not an OpenEXR file")
    pos = 8
    attrs = {}
    while data[pos] != 0:
        name, pos = _cstr(data, pos)
        typ, pos = _cstr(data, pos)
        (size,) = struct.unpack_from("<i", data, pos)
        pos += 4
        attrs[name] = (typ, data[pos:pos + size])
        pos += size
    pos += 1
    x0, y0, x1, y1 = struct.unpack("<4i", attrs["dataWindow"][1])
    h = y1 - y0 + 1
    nb = (h + LINES - 1) // LINES
    offsets = list(struct.unpack_from(f"<{nb}q", data, pos))
    return version, attrs, offsets, pos


def exr_chlist(payload):
    out = []
    pos = 0
    while payload[pos] != 0:
        name, pos = _cstr(payload, pos)
        typ, plin, _, _, _, xs, ys = struct.unpack_from("<iBBBBii", payload, pos)
        out.append((name, typ, xs, ys))
        pos += 16
    return out


def exr_channels(data):
    """Decodes a scanline ZIP (or uncompressed) EXR: [(name, type, plane
    float32 (H, W))] in file order, plus width and height."""
    version, attrs, offsets, _ = exr_parse(data)
    x0, y0, x1, y1 = struct.unpack("<4i", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    comp = attrs["compression"][1][0]
    chans = exr_chlist(attrs["channels"][1])
    bps = [2 if t == HALF else 4 for _, t, _, _ in chans]
    sb = w * sum(bps)
    planes = [np.empty((h, w), np.float32) for _ in chans]
    lines_per = 1 if comp == 0 else LINES
    for off in offsets:
        y, size = struct.unpack_from("<ii", data, off)
        lines = min(lines_per, h - (y - y0))
        payload = data[off + 8:off + 8 + size]
        expect = lines * sb
        if size < expect:
            if comp != 3:
                raise ValueError("unsupported compression")
            raw = zip_reconstruct(zlib.decompress(payload))
        else:
            raw = payload
        assert len(raw) == expect
        pos = 0
        for ly in range(lines):
            for k, (name, t, _, _) in enumerate(chans):
                n = w * bps[k]
                seg = raw[pos:pos + n]
                pos += n
                v = np.frombuffer(seg, "<f2" if t == HALF else "<f4").astype(np.float32)
                planes[k][y - y0 + ly] = v
    return [(c[0], c[1], planes[k]) for k, c in enumerate(chans)], w, h


def pfm_bytes(planes, width, height, channels):
    if channels not in (1, 3):
        raise ValueError(f"PFM supports 1 or 3 channels, got {channels}")
    p = np.asarray(planes, np.float32).ravel()[:width * height * channels].reshape(channels, height, width)
    body = np.transpose(p, (1, 2, 0))[::-1].astype("<f4").tobytes()
    return f"{'PF' if channels == 3 else 'Pf'}\n{width} {height}\n-1.0\n".encode("ascii") + body


def pfm_read(data):
    """PfmImage.Read on bytes: (planes (C*H*W,) row 0 = top, W, H, C)."""
    toks, pos = [], 0
    ws = b" \n\r\t"
    while len(toks) < 4:
        while data[pos] in ws:
            pos += 1
        start = pos
        while data[pos] not in ws:
            pos += 1
        toks.append(data[start:pos].decode("ascii"))
        pos += 1   # the single terminating whitespace byte
    magic, w, h, scale = toks[0], int(toks[1]), int(toks[2]), float(toks[3])
    if magic not in ("PF", "Pf"):
        raise ValueError(f"not a PFM file (magic {magic!r})")
    c = 3 if magic == "PF" else 1
    v = np.frombuffer(data, "<f4" if scale < 0 else ">f4", w * h * c, pos).astype(np.float32)
    rows = v.reshape(h, w, c)[::-1]
    return np.ascontiguousarray(np.transpose(rows, (2, 0, 1))).ravel(), w, h, c


def png_pixels(display, width, height):
    """RGBA bytes (H, W, 4) of Program.SaveImage: (byte)clamp((int)(clamp(v,0,1)*256), 0, 255)."""
    d = np.asarray(display, np.float32).reshape(3, height, width)
    c = np.clip(d, 0.0, 1.0)
    q = np.where(np.isnan(d), 0, np.clip(np.trunc(c * np.float32(256.0)), 0, 255)).astype(np.uint8)
    out = np.empty((height, width, 4), np.uint8)
    out[..., :3] = np.transpose(q, (1, 2, 0))
    out[..., 3] = 255
    return out


def png_decode(data):
    """8-bit RGBA non-interlaced PNG -> (H, W, 4) uint8; checks every CRC."""
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, ihdr = 8, b"", None
    while True:
        (n,) = struct.unpack_from(">I", data, pos)
        typ = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        (crc,) = struct.unpack_from(">I", data, pos + 8 + n)
        assert zlib.crc32(typ + body) == crc, typ
        pos += 12 + n
        if typ == b"IHDR":
            ihdr = struct.unpack(">IIBBBBB", body)
        elif typ == b"IDAT":
            idat += body
        elif typ == b"IEND":
            break
    w, h, depth, ctype, _, _, interlace = ihdr
    assert (depth, ctype, interlace) == (8, 6, 0)
    raw = zlib.decompress(idat)
    bpp, stride = 4, 4 * w
    out = np.zeros((h, stride), np.uint8)
    prev = np.zeros(stride, np.int32)
    for y in range(h):
        f = raw[y * (stride + 1)]
        line = np.frombuffer(raw, np.uint8, stride, y * (stride + 1) + 1).astype(np.int32)
        if f == 0:
            out[y] = line
            prev = line
            continue
        cur = np.zeros(stride, np.int32)
        for i in range(stride):
            a = cur[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            if f == 0:
                pr = 0
            elif f == 1:
                pr = a
            elif f == 2:
                pr = b
            elif f == 3:
                pr = (a + b) // 2
            else:
                p = a + b - c
                pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
                pr = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
            cur[i] = (line[i] + pr) & 0xFF
        out[y] = cur
        prev = cur
    return out.reshape(h, w, 4)
