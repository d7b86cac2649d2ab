"""Host-side renderer object over the mmhrt C ABI.

`SphereRenderer` holds one mmhrt context and mirrors the pieces of MegaMol's renderer module
that the sphere path uses: particle lists in (geocalls::MultiParticleDataCall /
SimpleSphericalParticles, plugins/geometry_calls/include/geometry_calls/SimpleSphericalParticles.h),
camera + parameters (core::view::Camera, plugins/optix_hpg/src/optix/Renderer.cpp:22-39), a
framebuffer out (core::view::CPUFramebuffer). Arrays are numpy; the work runs in libmmhrt.so.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L


@dataclass
class ParticleList:
    """One particle list as SimpleSphericalParticles exposes it (strided views into a buffer)."""

    vert_type: int
    vert: np.ndarray  # any contiguous buffer holding the vertex records
    col_type: int = L.COLDATA_NONE
    col: np.ndarray | None = None
    vert_stride: int = 0
    col_stride: int = 0
    count: int | None = None
    global_radius: float = 0.5
    global_rgba: tuple = (255, 255, 255, 255)
    col_min: float = 0.0
    col_max: float = 1.0
    vert_offset: int = 0  # byte offset of the first vertex inside `vert`
    col_offset: int = 0   # byte offset of the first colour inside `col`

    def to_c(self) -> L.SphereList:
        s = L.SphereList()
        s.vert_type = self.vert_type
        s.col_type = self.col_type
        # unknown types pass through unchanged: the library rejects them with E_INVALID
        vsz = L.VERT_SIZE[self.vert_type] if 0 <= self.vert_type < len(L.VERT_SIZE) else 0
        vs = self.vert_stride or vsz
        n = self.count if self.count is not None else (self.vert.nbytes - self.vert_offset) // max(vs, 1)
        s.vert = self.vert.ctypes.data + self.vert_offset if self.vert is not None else None
        s.col = (self.col.ctypes.data + self.col_offset) if self.col is not None else None
        s.vert_stride = self.vert_stride
        s.col_stride = self.col_stride
        s.count = int(n)
        s.global_radius = float(self.global_radius)
        for i in range(4):
            s.global_rgba[i] = int(self.global_rgba[i])
        s.col_min = float(self.col_min)
        s.col_max = float(self.col_max)
        return s


def look_at(position, target, up=(0.0, 1.0, 0.0), fovy_deg=40.0, aspect=1.0, near=0.01, far=1000.0) -> L.Camera:
    """Camera pose looking from `position` at `target` (float32 arithmetic, orthonormal up)."""
    p = np.asarray(position, np.float32)
    d = np.asarray(target, np.float32) - p
    d = (d / np.float32(np.sqrt(np.float32(d @ d)))).astype(np.float32)
    u = np.asarray(up, np.float32)
    r = np.cross(d, u).astype(np.float32)
    r = (r / np.float32(np.sqrt(np.float32(r @ r)))).astype(np.float32)
    u = np.cross(r, d).astype(np.float32)
    u = (u / np.float32(np.sqrt(np.float32(u @ u)))).astype(np.float32)
    cam = L.Camera()
    for i in range(3):
        cam.position[i] = float(p[i])
        cam.direction[i] = float(d[i])
        cam.up[i] = float(u[i])
    cam.fovy = float(np.float32(np.deg2rad(fovy_deg)))
    cam.aspect = float(aspect)
    cam.near_plane = float(near)
    cam.far_plane = float(far)
    cam.tile_start[0] = cam.tile_start[1] = 0.0
    cam.tile_end[0] = cam.tile_end[1] = 1.0
    return cam


def params(width, height, spp=1, mode=L.MODE_LAMBERT, ao_distance=1e20, max_bounces=0, frame_idx=0,
           accumulate=0, background=(0.0, 0.0, 0.0, 1.0), flags=0) -> L.RenderParams:
    p = L.RenderParams()
    p.width, p.height, p.spp, p.mode = int(width), int(height), int(spp), int(mode)
    p.ao_distance, p.max_bounces, p.frame_idx, p.accumulate = float(ao_distance), int(max_bounces), int(frame_idx), int(accumulate)
    for i in range(4):
        p.background[i] = float(background[i])
    p.flags = int(flags)
    return p


@dataclass
class Frame:
    width: int
    height: int
    rgba8: np.ndarray | None = None
    depth: np.ndarray | None = None
    color: np.ndarray | None = None
    t: np.ndarray | None = None
    hit_id: np.ndarray | None = None
    aux: np.ndarray | None = None
    extra: dict = field(default_factory=dict)


class SphereRenderer:
    """One mmhrt context on one GPU."""

    def __init__(self, device: int = 0):
        self.lib = L.load()
        h = C.c_void_p()
        code = self.lib.mmhrt_create(device, C.byref(h))
        if code != L.OK:
            raise L.MMHRTError(code, (self.lib.mmhrt_last_error(None) or b"").decode())
        self.ctx = h

    def close(self) -> None:
        if getattr(self, "ctx", None):
            self.lib.mmhrt_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _ck(self, code):
        L.check(self.ctx, code)

    # ---- data in
    def set_lists(self, lists: list[ParticleList], data_hash: int = 0, frame_id: int = 0) -> None:
        arr = (L.SphereList * max(len(lists), 1))(*[pl.to_c() for pl in lists])
        self._ck(self.lib.mmhrt_set_spheres(self.ctx, arr, len(lists), data_hash, frame_id))

    def set_spheres(self, xyzr: np.ndarray, rgba: np.ndarray | None = None, global_rgba=(255, 255, 255, 255)) -> None:
        """Packed float32 spheres (N,4) and optional float32 colours (N,4)."""
        xyzr = np.ascontiguousarray(xyzr, np.float32)
        lists = [ParticleList(L.VERTDATA_FLOAT_XYZR, xyzr, global_rgba=global_rgba, count=len(xyzr))]
        if rgba is not None:
            rgba = np.ascontiguousarray(rgba, np.float32)
            lists[0].col_type = L.COLDATA_FLOAT_RGBA
            lists[0].col = rgba
        self.set_lists(lists)
        self._keep = lists

    def set_transfer_function(self, tf: np.ndarray | None, vmin: float = 1.0, vmax: float = 0.0) -> None:
        if tf is None:
            self._ck(self.lib.mmhrt_set_transfer_function(self.ctx, None, 0, vmin, vmax))
            return
        tf = np.ascontiguousarray(tf, np.float32).reshape(-1, 4)
        self._ck(self.lib.mmhrt_set_transfer_function(self.ctx, tf.ctypes.data, len(tf), vmin, vmax))

    def build(self, policy: int = L.BUILD_REBUILD, bits: int = 30) -> None:
        self._ck(self.lib.mmhrt_build(self.ctx, policy, bits))

    # ---- frame out
    def render(self, cam: L.Camera, p: L.RenderParams, outputs=("rgba8", "depth", "color", "t", "hit_id")) -> Frame:
        # outputs cover the render window when one is set
        W, H = (p.window[2], p.window[3]) if p.window[2] > 0 else (p.width, p.height)
        fr = Frame(W, H)
        fb = L.Framebuffer()
        if p.flags & L.FLAG_NO_DOWNLOAD:
            self._ck(self.lib.mmhrt_render(self.ctx, C.byref(cam), C.byref(p), None))
            return fr
        self._alloc(fr, fb, outputs, p.flags & L.FLAG_AUX)
        self._ck(self.lib.mmhrt_render(self.ctx, C.byref(cam), C.byref(p), C.byref(fb)))
        return fr

    def render_async(self, cam: L.Camera, p: L.RenderParams) -> None:
        q = L.RenderParams.from_buffer_copy(p)
        q.flags |= L.FLAG_NO_DOWNLOAD
        self._ck(self.lib.mmhrt_render(self.ctx, C.byref(cam), C.byref(q), None))

    def download(self, width, height, outputs=("rgba8", "depth", "color", "t", "hit_id"), aux=False) -> Frame:
        fr = Frame(width, height)
        fb = L.Framebuffer()
        self._alloc(fr, fb, outputs, aux)
        self._ck(self.lib.mmhrt_download(self.ctx, C.byref(fb)))
        return fr

    def device_framebuffer(self) -> L.DeviceFramebuffer:
        """Device pointers (and the stream) of the last rendered frame, for zero-copy consumers."""
        d = L.DeviceFramebuffer()
        self._ck(self.lib.mmhrt_get_device_framebuffer(self.ctx, C.byref(d)))
        return d

    @staticmethod
    def _alloc(fr: Frame, fb: L.Framebuffer, outputs, aux) -> None:
        W, H = fr.width, fr.height
        spec = {"rgba8": (np.uint32, (H, W)), "depth": (np.float32, (H, W)), "color": (np.float32, (H, W, 4)),
                "t": (np.float32, (H, W)), "hit_id": (np.int32, (H, W))}
        for name in outputs:
            dt, shp = spec[name]
            a = np.empty(shp, dt)
            setattr(fr, name, a)
            setattr(fb, name, a.ctypes.data)
        if aux:
            fr.aux = np.empty((H, W, 8), np.float16)
            fb.aux = fr.aux.ctypes.data

    def sync(self) -> None:
        self._ck(self.lib.mmhrt_sync(self.ctx))

    # ---- batched time series
    def set_timesteps(self, xyzr_steps: np.ndarray, rgba: np.ndarray | None = None,
                      global_rgba=(255, 255, 255, 255)) -> None:
        """xyzr_steps: (T, N, 4) float32, all resident on the device."""
        xs = np.ascontiguousarray(xyzr_steps, np.float32)
        T, N = xs.shape[0], xs.shape[1]
        g = (C.c_uint8 * 4)(*global_rgba)
        rp = None
        if rgba is not None:
            rgba = np.ascontiguousarray(rgba, np.float32)
            rp = rgba.ctypes.data
        self._ck(self.lib.mmhrt_set_timesteps(self.ctx, xs.ctypes.data, T, N, rp, g))

    def render_batch(self, jobs: list[tuple], refit_period: int = 1, bits: int = 30, use_graph: bool = True):
        """jobs: (timestep, camera, params). Returns (list of (rgba8, depth) per job, device ms)."""
        arr = (L.Job * len(jobs))()
        sizes = []
        off = 0
        for i, (t, cam, p) in enumerate(jobs):
            arr[i].timestep = t
            arr[i].camera = cam
            arr[i].params = p
            arr[i].out_offset = off
            w, h = (p.window[2], p.window[3]) if p.window[2] > 0 else (p.width, p.height)
            sizes.append((w, h, off))
            off += w * h
        rgba = np.empty(off, np.uint32)
        depth = np.empty(off, np.float32)
        ms = C.c_float()
        self._ck(self.lib.mmhrt_render_batch(self.ctx, arr, len(jobs), refit_period, bits, 1 if use_graph else 0,
                                             rgba.ctypes.data, depth.ctypes.data, C.byref(ms)))
        out = [(rgba[o:o + w * h].reshape(h, w), depth[o:o + w * h].reshape(h, w)) for (w, h, o) in sizes]
        return out, float(ms.value)

    # ---- introspection
    def info(self):
        n = C.c_uint64()
        nn = C.c_uint64()
        b = (C.c_float * 6)()
        self._ck(self.lib.mmhrt_get_info(self.ctx, C.byref(n), b, C.byref(nn)))
        return int(n.value), np.array(b[:], np.float32), int(nn.value)

    def bvh(self):
        n, _, nn = self.info()
        nodes = np.empty((nn, 16), np.float32)
        ids = np.empty(n, np.uint32)
        keys = np.empty(n, np.uint64)
        self._ck(self.lib.mmhrt_get_bvh(self.ctx, nodes.ctypes.data, ids.ctypes.data, keys.ctypes.data))
        child = nodes[:, 12:14].copy().view(np.int32)
        pairs = nodes[:, 14].copy().view(np.int32)
        height = nodes[:, 15].copy().view(np.int32)
        return nodes[:, :12].copy(), child, ids, keys, height, pairs

    def get_treelets(self) -> dict:
        """The treelet P-k-d trees of the current build (option accel 1), in the oracle's array
        layout: sph (n,4), id (n,), dim (n,), tl_box (T,6), tl_rmax (T,), top_box (L-1,12),
        top_child (L-1,2), plus M, T, L."""
        sizes = np.zeros(3, np.uint32)
        self._ck(self.lib.mmhrt_get_treelets(self.ctx, sizes.ctypes.data, None, None, None, None, None, None))
        M, T, L = (int(x) for x in sizes)
        n = self.info()[0]
        sph = np.empty((n, 4), np.float32)
        ids = np.empty(n, np.uint32)
        dims = np.empty(n, np.uint8)
        tl_box = np.empty((T, 6), np.float32)
        tl_rmax = np.empty(T, np.float32)
        top = np.empty((L - 1, 16), np.float32)
        self._ck(self.lib.mmhrt_get_treelets(self.ctx, None, sph.ctypes.data, ids.ctypes.data, dims.ctypes.data,
                                             tl_box.ctypes.data, tl_rmax.ctypes.data, top.ctypes.data))
        return dict(M=M, T=T, L=L, sph=sph, id=ids, dim=dims, tl_box=tl_box, tl_rmax=tl_rmax,
                    top_box=top[:, :12].copy(), top_child=top[:, 12:14].copy().view(np.int32))

    def timings(self) -> dict:
        t = (C.c_float * 9)()
        self._ck(self.lib.mmhrt_get_timings(self.ctx, t, 9))
        names = ["morton", "sort", "hierarchy", "boxes", "trace", "resolve", "build", "render", "chunk"]
        return dict(zip(names, [float(x) for x in t]))

    def counters(self):
        v, s, r = C.c_uint64(), C.c_uint64(), C.c_uint64()
        self._ck(self.lib.mmhrt_get_counters(self.ctx, C.byref(v), C.byref(s), C.byref(r)))
        return int(v.value), int(s.value), int(r.value)

    def set_option(self, key: str, value: int) -> None:
        self._ck(self.lib.mmhrt_set_option(self.ctx, key.encode(), int(value)))

    def get_option(self, key: str) -> int:
        v = C.c_int64()
        self._ck(self.lib.mmhrt_get_option(self.ctx, key.encode(), C.byref(v)))
        return int(v.value)
