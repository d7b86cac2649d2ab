"""Learned denoiser for 1-spp AO frames (host side over the mmhrt C ABI).

The network (4-level U-Net, 3x3 convs 32/48/64/96, ReLU, skip concatenation) runs and trains
in libmmhrt.so with BF16 MFMA kernels. This module handles parameters, the versioned weights
file, the per-pixel feature records and training-pair generation with the renderer itself.
"""
from __future__ import annotations

import ctypes as C
import struct

import numpy as np

from . import _lib as L

# layer table, mirrors include/mmhrt.h (name, c_in full-res, c_in half-res, c_out, relu)
LAYERS = [("enc0a", 16, 0, 32, 1), ("enc0b", 32, 0, 32, 1), ("enc1a", 32, 0, 48, 1), ("enc1b", 48, 0, 48, 1),
          ("enc2a", 48, 0, 64, 1), ("enc2b", 64, 0, 64, 1), ("bota", 64, 0, 96, 1), ("botb", 96, 0, 96, 1),
          ("dec2a", 64, 96, 64, 1), ("dec2b", 64, 0, 64, 1), ("dec1a", 48, 64, 48, 1), ("dec1b", 48, 0, 48, 1),
          ("dec0a", 32, 48, 32, 1), ("dec0b", 32, 0, 32, 1), ("out", 32, 0, 16, 0)]

MAGIC = b"MMHRTDN1"
FORMAT_VERSION = 2  # 2: output channels 1-3 are a colour residual added to albedo * v


def param_count() -> int:
    return sum(9 * (c1 + c2) * co + co for _, c1, c2, co, _ in LAYERS)


def split_params(p: np.ndarray):
    """Canonical flat vector -> list of (W[9, ci, co], b[co])."""
    out, o = [], 0
    for _, c1, c2, co, _ in LAYERS:
        ci = c1 + c2
        w = p[o:o + 9 * ci * co].reshape(9, ci, co)
        o += 9 * ci * co
        b = p[o:o + co]
        o += co
        out.append((w, b))
    assert o == len(p)
    return out


def save_weights(path: str, params: np.ndarray, meta: dict | None = None) -> None:
    """Versioned weights file: magic, format version, layer table, fp32 parameters, JSON meta."""
    import json
    params = np.ascontiguousarray(params, np.float32)
    assert len(params) == param_count()
    head = MAGIC + struct.pack("<II", FORMAT_VERSION, len(LAYERS))
    for name, c1, c2, co, relu in LAYERS:
        head += struct.pack("<8sIIII", name.encode()[:8].ljust(8, b"\0"), c1, c2, co, relu)
    m = json.dumps(meta or {}).encode()
    with open(path, "wb") as f:
        f.write(head + struct.pack("<QI", len(params), len(m)) + m + params.tobytes())


def load_weights(path: str):
    import json
    buf = open(path, "rb").read()
    if buf[:8] != MAGIC:
        raise ValueError("not a denoiser weights file")
    ver, nl = struct.unpack_from("<II", buf, 8)
    if ver != FORMAT_VERSION:
        raise ValueError(f"unsupported weights format version {ver}")
    o = 16
    table = []
    for _ in range(nl):
        name, c1, c2, co, relu = struct.unpack_from("<8sIIII", buf, o)
        table.append((name.rstrip(b"\0").decode(), c1, c2, co, relu))
        o += 24
    if table != LAYERS:
        raise ValueError("weights file has a different layer table")
    n, lm = struct.unpack_from("<QI", buf, o)
    o += 12
    meta = json.loads(buf[o:o + lm].decode())
    o += lm
    params = np.frombuffer(buf, np.float32, count=n, offset=o).copy()
    return params, meta


def features_from_frame(color: np.ndarray, aux: np.ndarray) -> np.ndarray:
    """Frame colour (H,W,4 f32) + aux (H,W,8 f16) -> feature records (H,W,12 f16)."""
    h, w = color.shape[:2]
    f = np.zeros((h, w, 12), np.float16)
    f[..., 0:3] = color[..., 0:3].astype(np.float16)
    f[..., 3:11] = aux
    return f


def box_filter_baseline(aux: np.ndarray, bg=(0.0, 0.0, 0.0), k: int = 5) -> np.ndarray:
    """The non-learned comparison for the network: the 1-spp visibility averaged over the hit
    pixels of a k x k window, times the pixel's albedo; background on misses. aux is a frame's
    MMHRT_FLAG_AUX buffer (H,W,8: albedo rgb, normal xyz, t, visibility)."""
    from scipy.ndimage import uniform_filter
    a = aux.astype(np.float32)
    hit = (a[..., 6] > 0).astype(np.float32)
    v = a[..., 7] * hit
    num = uniform_filter(v, k, mode="constant")
    den = uniform_filter(hit, k, mode="constant")
    vf = np.where(den > 0, num / np.maximum(den, 1e-6), 0.0)
    out = a[..., 0:3] * vf[..., None]
    out[hit == 0] = np.asarray(bg, np.float32)
    return out


class Denoiser:
    """The network state inside one SphereRenderer context."""

    def __init__(self, renderer, seed: int | None = 1234):
        self.r = renderer
        self.lib = renderer.lib
        if seed is not None:
            self._ck(self.lib.mmhrt_denoiser_init(self.r.ctx, seed))

    def _ck(self, code):
        L.check(self.r.ctx, code)

    @property
    def n(self) -> int:
        return param_count()

    def get_params(self) -> np.ndarray:
        p = np.empty(self.n, np.float32)
        self._ck(self.lib.mmhrt_denoiser_get_params(self.r.ctx, p.ctypes.data, self.n))
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        self._ck(self.lib.mmhrt_denoiser_set_params(self.r.ctx, p.ctypes.data, len(p)))

    def save(self, path: str, meta: dict | None = None) -> None:
        save_weights(path, self.get_params(), meta)

    def load(self, path: str) -> dict:
        p, meta = load_weights(path)
        self.set_params(p)
        return meta

    def denoise_frame(self, depth_scale: float) -> None:
        """Denoise the last rendered (MMHRT_FLAG_AUX) frame in place on the device."""
        self._ck(self.lib.mmhrt_denoise(self.r.ctx, depth_scale))

    def denoise(self, feats: np.ndarray, depth_scale: float, bg=(0, 0, 0, 1)) -> np.ndarray:
        feats = np.ascontiguousarray(feats, np.float16)
        if feats.ndim == 3:
            feats = feats[None]
        b, h, w, _ = feats.shape
        out = np.empty((b, h, w, 3), np.float32)
        bgv = (C.c_float * 4)(*bg)
        self._ck(self.lib.mmhrt_denoise_features(self.r.ctx, feats.ctypes.data, b, h, w, depth_scale, bgv,
                                                 out.ctypes.data))
        return out

    def train_step(self, feats: np.ndarray, target: np.ndarray, depth_scale: float, lr: float = 1e-3,
                   apply: bool = True, want_grads: bool = False, bg=(0, 0, 0, 1)):
        feats = np.ascontiguousarray(feats, np.float16)
        target = np.ascontiguousarray(target, np.float32)
        b, h, w, _ = feats.shape
        loss = C.c_float()
        g = np.empty(self.n, np.float32) if want_grads else None
        bgv = (C.c_float * 4)(*bg)
        self._ck(self.lib.mmhrt_denoiser_train_step(self.r.ctx, feats.ctypes.data, target.ctypes.data, b, h, w,
                                                    depth_scale, bgv, lr, 1 if apply else 0, C.byref(loss),
                                                    None if g is None else g.ctypes.data))
        return (float(loss.value), g) if want_grads else float(loss.value)


def psnr(a: np.ndarray, b: np.ndarray, peak: float = 1.0) -> float:
    mse = float(np.mean((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2))
    return float("inf") if mse == 0 else 10.0 * np.log10(peak * peak / mse)


def ssim(a: np.ndarray, b: np.ndarray) -> float:
    """Mean SSIM (Wang et al. 2004) over colour channels with an 8x8 box window, values in [0,1]."""
    from numpy.lib.stride_tricks import sliding_window_view as sw
    c1, c2 = 0.01 ** 2, 0.03 ** 2
    vals = []
    for ch in range(a.shape[-1]):
        x = np.asarray(a[..., ch], np.float64)
        y = np.asarray(b[..., ch], np.float64)
        wx, wy = sw(x, (8, 8))[::4, ::4], sw(y, (8, 8))[::4, ::4]
        mx, my = wx.mean((-1, -2)), wy.mean((-1, -2))
        vx, vy = wx.var((-1, -2)), wy.var((-1, -2))
        cxy = ((wx - mx[..., None, None]) * (wy - my[..., None, None])).mean((-1, -2))
        s = ((2 * mx * my + c1) * (2 * cxy + c2)) / ((mx ** 2 + my ** 2 + c1) * (vx + vy + c2))
        vals.append(s.mean())
    return float(np.mean(vals))
