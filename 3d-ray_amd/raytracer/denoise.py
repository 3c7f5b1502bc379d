"""The film denoiser over the C ABI, shaped like the reference's:

  reference (C#)                                        here
  NforDenoiser.Denoise(buffers, options)                denoise(buffers, DenoiserOptions(...))   rt_denoise
      src/RayTracer/Denoising/NforDenoiser.cs:29-57
  DenoiserOptions { Kind, Quality }                     DenoiserOptions(kind="nfor"|"nlm",
      DenoiserOptions.cs:1-51                                           quality="high"|"fast")

`buffers` is a native.RenderBuffers from a render with capture_aovs=True
(planes (C, H, W), samples_a / samples_b).  The whole pipeline (depth
normalisation, feature prefilter, variance, main filter) runs on the
device; there is no CPU path.
"""
import ctypes as C

import numpy as np

from . import abi, native

_KIND = {"nlm": abi.RT_DENOISE_NLM, "nfor": abi.RT_DENOISE_NFOR}
_QUALITY = {"fast": abi.RT_DENOISE_FAST, "high": abi.RT_DENOISE_HIGH}


class DenoiserOptions:
    """DenoiserOptions: kind "nfor" (default) or "nlm", quality "high"
    (default) or "fast"; selection_margin < 0 = the Sobol sampler's 0.5
    (NforRegression.SelectionMargin)."""

    def __init__(self, kind="nfor", quality="high", selection_margin=-1.0):
        if kind not in _KIND:
            raise ValueError(f"Unexpected denoiser kind {kind!r}.")
        if quality not in _QUALITY:
            raise ValueError(f"Unexpected denoise quality {quality!r}.")
        self.kind, self.quality, self.selection_margin = kind, quality, selection_margin


def _plane(a):
    return None if a is None else np.ascontiguousarray(np.asarray(a, np.float32))


def denoise(buffers, options=None, ctx=None):
    """NforDenoiser.Denoise: the denoised linear beauty, (3, H, W) float32."""
    options = options or DenoiserOptions()
    if getattr(buffers, "albedo_a", None) is None:
        raise RuntimeError("Denoising requires AOV capture (RenderCaptureOptions.Full).")
    if ctx is None:
        from . import film_io
        ctx = film_io._ctx(None)
    h, w = np.asarray(buffers.beauty_a).shape[-2:]
    keep = {k: _plane(getattr(buffers, k)) for k in ("beauty_a", "beauty_b", "albedo_a", "albedo_b", "normal_a",
                                                     "normal_b", "depth_a", "depth_b")}
    film = abi.rt_film()
    for k, a in keep.items():
        setattr(film, k, a.ctypes.data_as(C.POINTER(C.c_float)))
    opt = abi.rt_denoise_options()
    opt.kind = _KIND[options.kind]
    opt.quality = _QUALITY[options.quality]
    opt.selection_margin = float(options.selection_margin)
    opt.samples_a, opt.samples_b = int(buffers.samples_a), int(buffers.samples_b)
    out = np.zeros((3, h, w), np.float32)
    ctx.check(native.load().rt_denoise(ctx.h, w, h, C.byref(film), C.byref(opt),
                                       out.ctypes.data_as(C.POINTER(C.c_float))))
    return out
