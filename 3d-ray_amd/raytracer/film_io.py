"""Film files over the C ABI, shaped like the reference's output writers:

  reference (C#)                                          here
  ExrImage.Write(path, w, h, channels)                    exr_write(path, w, h, channels)      rt_exr_encode
      src/RayTracer/Rendering/ExrImage.cs:52-137
  ExrImage.WriteRgb(path, rgb)            ExrImage.cs:140-152   exr_write_rgb(path, rgb)      rt_exr_encode
  PfmImage.Write(path, planes, w, h, c)   PfmImage.cs:29-65     pfm_write(path, planes, w, h, c)  rt_pfm_encode
  PfmImage.Read(path)                     PfmImage.cs:71-117    pfm_read(path)   (header parse + one read)
  Program.SaveImage(pixels, w, h, path)   Program.cs:723-756    save_image(path, display)     rt_png_encode
  Program.SaveAovs / SaveExrBeauty / AovExrChannels             save_aovs / save_exr_beauty /
      Program.cs:633-721                                        aov_exr_channels
  RenderBuffers.CombineHalves / CombineDepthHalves /            combine_halves / combine_depth_halves /
      RawBeautyVariance   RenderBuffers.cs:108-168              raw_beauty_variance

Planes are plane-major float32 arrays, row 0 = top (FrameBuffer.cs:20-41):
shape (C, H, W) or flat.  The device converts and lays out the pixel
payload, host threads compress it; invalid arguments raise ValueError as
the reference throws ArgumentException.  There is no CPU encoder: a
missing library or device raises.
"""
import ctypes as C
import os
from collections import namedtuple

import numpy as np

from . import abi, native

HALF, FLOAT = abi.RT_EXR_HALF, abi.RT_EXR_FLOAT

# ExrImage.Channel (ExrImage.cs:45-46): name, ExrPixelType, plane of W*H floats
Channel = namedtuple("Channel", "name type plane")

_default_ctx = None


def _ctx(ctx):
    global _default_ctx
    if ctx is not None:
        return ctx
    if _default_ctx is None:
        _default_ctx = native.Context(0)
    return _default_ctx


def _call(ctx, fn, *args):
    """Two-call size protocol of the encoders: query the bound, then encode."""
    L = native.load()
    size = C.c_int64(0)
    rc = getattr(L, fn)(ctx.h, *args, None, 0, C.byref(size))
    if rc == abi.RT_ERR_INVALID_ARGUMENT:
        raise ValueError(L.rt_last_error(ctx.h).decode())
    ctx.check(rc)
    buf = np.empty(max(1, size.value), np.uint8)
    rc = getattr(L, fn)(ctx.h, *args, buf.ctypes.data_as(C.POINTER(C.c_uint8)), buf.size, C.byref(size))
    if rc == abi.RT_ERR_INVALID_ARGUMENT:
        raise ValueError(L.rt_last_error(ctx.h).decode())
    ctx.check(rc)
    return buf[:size.value].tobytes()


def _plane(a, n, what):
    p = np.ascontiguousarray(np.asarray(a, np.float32).ravel())
    if p.size != n:
        raise ValueError(f"{what} has {p.size} floats, expected {n}")
    return p


def exr_encode(width, height, channels, ctx=None):
    """The bytes ExrImage.Write would put in a file."""
    ctx = _ctx(ctx)
    if width <= 0 or height <= 0:
        raise ValueError(f"invalid EXR dimensions {width}x{height}")
    if len(channels) == 0:
        raise ValueError("EXR requires at least one channel")
    keep, arr = [], (abi.rt_exr_channel * len(channels))()
    for i, ch in enumerate(channels):
        name, typ, plane = ch
        p = _plane(plane, width * height, f"channel '{name}' plane")
        nm = str(name).encode("ascii", "replace")
        keep += [p, nm]
        arr[i].name = nm
        arr[i].type = int(typ)
        arr[i].plane = p.ctypes.data_as(C.POINTER(C.c_float))
    return _call(ctx, "rt_exr_encode", width, height, arr, len(channels))


def exr_write(path, width, height, channels, ctx=None):
    """ExrImage.Write: single-part scanline OpenEXR v2, ZIP."""
    data = exr_encode(width, height, channels, ctx)
    with open(path, "wb") as f:
        f.write(data)


def exr_write_rgb(path, rgb, ctx=None):
    """ExrImage.WriteRgb: a 3-channel (3, H, W) buffer as half R, G, B."""
    rgb = np.asarray(rgb, np.float32)
    if rgb.ndim != 3 or rgb.shape[0] != 3:
        raise ValueError(f"exr_write_rgb expects a (3, H, W) buffer, got {rgb.shape}")
    _, h, w = rgb.shape
    exr_write(path, w, h, [Channel("R", HALF, rgb[0]), Channel("G", HALF, rgb[1]), Channel("B", HALF, rgb[2])], ctx)


def pfm_encode(planes, width, height, channels, ctx=None):
    if channels not in (1, 3):
        raise ValueError(f"PFM supports 1 or 3 channels, got {channels}.")
    flat = np.asarray(planes, np.float32).ravel()
    if flat.size < width * height * channels:
        raise ValueError("Plane buffer smaller than width*height*channels.")
    p = np.ascontiguousarray(flat[:width * height * channels])
    return _call(_ctx(ctx), "rt_pfm_encode", p.ctypes.data_as(C.POINTER(C.c_float)), width, height, channels)


def pfm_write(path, planes, width, height, channels, ctx=None):
    """PfmImage.Write: little-endian PFM, rows bottom-to-top."""
    data = pfm_encode(planes, width, height, channels, ctx)
    with open(path, "wb") as f:
        f.write(data)


def pfm_read(path):
    """PfmImage.Read: (planes (C*H*W,) row 0 = top, width, height, channels);
    both byte orders; ValueError for a bad header (InvalidDataException)."""
    with open(path, "rb") as f:
        data = f.read()
    toks, pos = [], 0
    ws = b" \n\r\t"
    for _ in range(4):
        while pos < len(data) and data[pos] in ws:
            pos += 1
        if pos >= len(data):
            raise ValueError("Truncated PFM header.")
        start = pos
        while pos < len(data) and data[pos] not in ws:
            pos += 1
        toks.append(data[start:pos].decode("ascii", "replace"))
        pos += 1
    if toks[0] not in ("PF", "Pf"):
        raise ValueError(f"Not a PFM file (magic '{toks[0]}').")
    c = 3 if toks[0] == "PF" else 1
    try:
        w, h = int(toks[1]), int(toks[2])
    except ValueError:
        raise ValueError("Invalid PFM dimensions.") from None
    if w <= 0 or h <= 0:
        raise ValueError("Invalid PFM dimensions.")
    try:
        scale = float(toks[3])
    except ValueError:
        raise ValueError("Invalid PFM scale.") from None
    if scale == 0.0:
        raise ValueError("Invalid PFM scale.")
    n = w * h * c
    if len(data) - pos < 4 * n:
        raise ValueError("Truncated PFM data.")
    v = np.frombuffer(data, "<f4" if scale < 0 else ">f4", n, pos).astype(np.float32)
    planes = np.ascontiguousarray(np.transpose(v.reshape(h, w, c)[::-1], (2, 0, 1))).ravel()
    return planes, w, h, c


def png_encode(display, width, height, ctx=None):
    p = _plane(display, 3 * width * height, "display buffer")
    return _call(_ctx(ctx), "rt_png_encode", p.ctypes.data_as(C.POINTER(C.c_float)), width, height)


def save_image(path, display, ctx=None):
    """Program.SaveImage for .png (the default): display = tone-mapped
    (3, H, W) planes (rt_film.display / RenderResult.pixels transposed).
    JPEG and BMP go through the reference's image library and are not
    provided."""
    d = np.asarray(display, np.float32)
    if d.ndim == 3 and d.shape[-1] == 3 and d.shape[0] != 3:
        d = np.transpose(d, (2, 0, 1))   # (H, W, 3) pixels
    _, h, w = d.shape
    ext = os.path.splitext(path)[1].lower()
    if ext in (".jpg", ".jpeg", ".bmp"):
        raise NotImplementedError(f"{ext} output is not provided (PNG, EXR and PFM are)")
    data = png_encode(d, w, h, ctx)
    with open(path, "wb") as f:
        f.write(data)


# ── AOV planes and the CLI's output layout ─────────────────────────────────

def combine_halves(buffers, a, b):
    """RenderBuffers.CombineHalves (RenderBuffers.cs:108-123): (nA*A + nB*B)/n."""
    na, nb = buffers.samples_a, buffers.samples_b
    if nb == 0:
        return np.array(a, np.float32, copy=True)
    n = na + nb
    wa, wb = np.float32(na) / np.float32(n), np.float32(nb) / np.float32(n)   # (float)SamplesA / n
    return (wa * np.asarray(a, np.float32) + wb * np.asarray(b, np.float32)).astype(np.float32)


def combine_depth_halves(buffers):
    """RenderBuffers.CombineDepthHalves (RenderBuffers.cs:130-151): a half
    that never hit (-1) defers to the other; both missed keeps -1."""
    if getattr(buffers, "depth_a", None) is None:
        raise ValueError("Depth AOV was not captured.")
    da, db = np.asarray(buffers.depth_a, np.float32), np.asarray(buffers.depth_b, np.float32)
    na, nb = buffers.samples_a, buffers.samples_b
    n = na + nb
    wa = np.float32(na) / np.float32(n) if n > 0 else np.float32(1.0)
    wb = np.float32(1.0) - wa
    both = da if nb == 0 else (wa * da + wb * db).astype(np.float32)
    return np.where((da < 0) & (db < 0), np.float32(-1.0),
                    np.where(da < 0, db, np.where(db < 0, da, both))).astype(np.float32)


def raw_beauty_variance(buffers):
    """RenderBuffers.RawBeautyVariance (RenderBuffers.cs:158-168): ((A-B)/2)^2."""
    d = np.float32(0.5) * (np.asarray(buffers.beauty_a, np.float32) - np.asarray(buffers.beauty_b, np.float32))
    return (d * d).astype(np.float32)


def resolve_aov(buffers, aov, denoised_beauty=None):
    """Program.ResolveAov (Program.cs:652-660)."""
    if aov == "beauty":
        return np.asarray(denoised_beauty if denoised_beauty is not None else buffers.beauty, np.float32)
    if aov == "albedo":
        return combine_halves(buffers, buffers.albedo_a, buffers.albedo_b)
    if aov == "normal":
        return combine_halves(buffers, buffers.normal_a, buffers.normal_b)
    if aov == "depth":
        return combine_depth_halves(buffers)
    if aov == "variance":
        return raw_beauty_variance(buffers)
    raise ValueError(f"unknown AOV '{aov}'")


def aov_exr_channels(aov, planes, embedded):
    """Program.AovExrChannels (Program.cs:701-721): depth = one float Z;
    colour layers half, named R/G/B (standalone or beauty), layer.R/G/B
    embedded, normal.X/Y/Z embedded."""
    p = np.asarray(planes, np.float32)
    if aov == "depth":
        return [Channel("Z", FLOAT, p.reshape(-1))]
    if aov == "beauty" or not embedded:
        names = ("R", "G", "B")
    elif aov == "normal":
        names = ("normal.X", "normal.Y", "normal.Z")
    else:
        names = (f"{aov}.R", f"{aov}.G", f"{aov}.B")
    p = p.reshape(3, -1)
    return [Channel(names[k], HALF, p[k]) for k in range(3)]


def save_aovs(buffers, aovs, output_path, fmt, denoised_beauty=None, ctx=None):
    """Program.SaveAovs (Program.cs:633-650): one file per AOV next to the
    main output, <stem>.<aov>.<pfm|exr>; returns the paths written."""
    stem = os.path.splitext(output_path)[0]
    out = []
    for aov in aovs:
        path = f"{stem}.{aov}.{fmt}"
        fb = resolve_aov(buffers, aov, denoised_beauty)
        h, w = fb.shape[-2], fb.shape[-1]
        if fmt == "exr":
            exr_write(path, w, h, aov_exr_channels(aov, fb, embedded=False), ctx)
        else:
            pfm_write(path, fb, w, h, 1 if aov == "depth" else 3, ctx)
        out.append(path)
    return out


def save_exr_beauty(path, beauty, buffers, embed_aovs=None, ctx=None):
    """Program.SaveExrBeauty (Program.cs:670-691): half R, G, B beauty plus
    the requested AOVs as layers of the same file; returns the number of
    embedded layers (an AOV named beauty is the main layer, skipped)."""
    beauty = np.asarray(beauty, np.float32)
    h, w = beauty.shape[-2], beauty.shape[-1]
    channels = list(aov_exr_channels("beauty", beauty, embedded=True))
    embedded = 0
    for aov in embed_aovs or ():
        if aov == "beauty":
            continue
        channels += aov_exr_channels(aov, resolve_aov(buffers, aov), embedded=True)
        embedded += 1
    exr_write(path, w, h, channels, ctx)
    return embedded
