"""CPU oracle for the particle-sphere path — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may import this package,
and only as the checker (or the timed CPU baseline). The product library never uses it.

Wraps oracle/_build/liboracle.so (plain C, built from oracle/oracle.c by oracle/Makefile).
Each C function cites the MegaMol source it restates; see oracle/oracle.c.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"


class OrcCamera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("direction", C.c_float * 3), ("up", C.c_float * 3),
                ("fovy", C.c_float), ("aspect", C.c_float), ("near_plane", C.c_float), ("far_plane", C.c_float),
                ("tile_start", C.c_float * 2), ("tile_end", C.c_float * 2)]


class OrcFrameState(C.Structure):
    _fields_ = [("up", C.c_float * 3), ("right", C.c_float * 3), ("front", C.c_float * 3), ("center", C.c_float * 3),
                ("th", C.c_float), ("rw", C.c_float), ("nearp", C.c_float),
                ("depth_a", C.c_float), ("depth_b", C.c_float), ("depth_c", C.c_float)]


class OrcParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32), ("mode", C.c_int32),
                ("ao_distance", C.c_float), ("frame_idx", C.c_uint32), ("background", C.c_float * 4),
                ("want_aux", C.c_int32), ("row_begin", C.c_uint32), ("row_end", C.c_uint32),
                ("max_bounces", C.c_int32)]


class OrcList(C.Structure):
    _fields_ = [("vert_type", C.c_int32), ("col_type", C.c_int32), ("vert", C.c_void_p), ("col", C.c_void_p),
                ("vert_stride", C.c_uint64), ("col_stride", C.c_uint64), ("count", C.c_uint64),
                ("global_radius", C.c_float), ("global_rgba", C.c_uint8 * 4), ("col_min", C.c_float),
                ("col_max", C.c_float)]


class OrcBvh(C.Structure):
    _fields_ = [("n", C.c_uint64), ("node_box", C.c_void_p), ("node_child", C.c_void_p), ("leaf_id", C.c_void_p),
                ("keys", C.c_void_p), ("leaf_sph", C.c_void_p), ("xyzr_in", C.c_void_p), ("amax", C.c_float)]


class OrcPkd(C.Structure):
    _fields_ = [("n", C.c_uint64), ("M", C.c_uint32), ("T", C.c_uint32), ("top_leaves", C.c_uint32),
                ("sph", C.c_void_p), ("id", C.c_void_p), ("dim", C.c_void_p), ("tl_box", C.c_void_p),
                ("tl_rmax", C.c_void_p), ("top_box", C.c_void_p), ("top_child", C.c_void_p), ("xyzr_in", C.c_void_p),
                ("amax", C.c_float)]


_lib = None


def default_threads() -> int:
    """Worker threads for the multi-threaded oracle paths: the CPUs this process may run on, at
    most 16 (the host share of one GPU on the GPU hosts)."""
    import os
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        lib = C.CDLL(str(LIB))
        f = lib
        f.orc_tea16.restype = C.c_uint32
        f.orc_tea16.argtypes = [C.c_uint32, C.c_uint32]
        f.orc_tea_encrypt.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        f.orc_ao_direction.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p]
        f.orc_lcg.restype = C.c_uint32
        f.orc_lcg.argtypes = [C.POINTER(C.c_uint32)]
        f.orc_rnd.restype = C.c_float
        f.orc_rnd.argtypes = [C.POINTER(C.c_uint32)]
        f.orc_frame_state_from_camera.argtypes = [C.POINTER(OrcCamera), C.POINTER(OrcFrameState)]
        f.orc_primary_ray.argtypes = [C.POINTER(OrcFrameState), C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                      C.c_float, C.c_float, C.c_void_p, C.c_void_p]
        f.orc_intersect_f32.restype = C.c_int
        f.orc_intersect_f32.argtypes = [C.c_void_p, C.c_void_p, C.c_float, C.c_float, C.c_void_p,
                                        C.POINTER(C.c_float)]
        f.orc_intersect_f64.restype = C.c_int
        f.orc_intersect_f64.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.POINTER(C.c_double),
                                        C.POINTER(C.c_double)]
        f.orc_convert_lists.argtypes = [C.POINTER(OrcList), C.c_uint32, C.c_void_p, C.c_uint32, C.c_float, C.c_float,
                                        C.c_int, C.c_void_p, C.c_void_p, C.POINTER(C.c_uint64)]
        f.orc_tf_sample.argtypes = [C.c_void_p, C.c_uint32, C.c_float, C.c_float, C.c_float, C.c_void_p]
        f.orc_render_brute.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(OrcCamera), C.POINTER(OrcParams),
                                       C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        f.orc_primary_f64.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(OrcCamera), C.POINTER(OrcParams), C.c_uint32,
                                      C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p]
        f.orc_bvh_build.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.POINTER(OrcBvh)]
        f.orc_bvh_build_mt.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.c_int, C.POINTER(OrcBvh)]
        f.orc_primary_f64_rects.restype = C.c_int
        f.orc_primary_f64_rects.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(OrcCamera), C.POINTER(OrcParams),
                                            C.c_void_p, C.c_uint32, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_void_p]
        f.orc_bvh_free.argtypes = [C.POINTER(OrcBvh)]
        f.orc_render_bvh.argtypes = [C.POINTER(OrcBvh), C.c_void_p, C.POINTER(OrcCamera), C.POINTER(OrcParams), C.c_int,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p]
        f.orc_bvh_copy_out.argtypes = [C.POINTER(OrcBvh), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        f.orc_blend_frames.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, C.c_void_p,
                                       C.c_void_p, C.c_void_p]
        f.orc_pack_rgba8.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p]
        f.orc_now.restype = C.c_double
        f.orc_pkd_build.restype = C.c_int
        f.orc_pkd_build.argtypes = [C.POINTER(OrcBvh), C.c_uint32, C.c_int, C.POINTER(OrcPkd)]
        f.orc_pkd_free.argtypes = [C.POINTER(OrcPkd)]
        f.orc_render_pkd.argtypes = [C.POINTER(OrcPkd), C.c_void_p, C.POINTER(OrcCamera), C.POINTER(OrcParams), C.c_int,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p]
        f.orc_pkd_subtree_size.restype = C.c_uint32
        f.orc_pkd_subtree_size.argtypes = [C.c_uint32, C.c_uint32]
        _lib = lib
    return _lib


def camera_from(cam) -> OrcCamera:
    """Copy a megamol_amd camera struct (same field layout) into the oracle's struct."""
    o = OrcCamera()
    C.memmove(C.byref(o), C.byref(cam), C.sizeof(OrcCamera))
    return o


def make_params(width, height, spp=1, mode=0, ao_distance=1e20, frame_idx=0, background=(0, 0, 0, 1),
                want_aux=0, max_bounces=0) -> OrcParams:
    p = OrcParams()
    p.width, p.height, p.spp, p.mode = width, height, spp, mode
    p.ao_distance, p.frame_idx, p.want_aux = ao_distance, frame_idx, want_aux
    p.max_bounces = max_bounces
    for i in range(4):
        p.background[i] = background[i]
    return p


def _outs(h, w, aux=False):
    d = dict(rgba8=np.zeros((h, w), np.uint32), depth=np.zeros((h, w), np.float32),
             color=np.zeros((h, w, 4), np.float32), t=np.zeros((h, w), np.float32),
             hit_id=np.zeros((h, w), np.int32))
    d["aux"] = np.zeros((h, w, 8), np.float16) if aux else None
    return d


def _ptr(a):
    return None if a is None else a.ctypes.data


def render_brute(xyzr, rgba, cam, p: OrcParams, aux=False) -> dict:
    lib = load()
    xyzr = np.ascontiguousarray(xyzr, np.float32)
    rgba = np.ascontiguousarray(rgba, np.float32)
    o = _outs(p.height, p.width, aux)
    lib.orc_render_brute(xyzr.ctypes.data, rgba.ctypes.data, len(xyzr), C.byref(camera_from(cam)), C.byref(p),
                         _ptr(o["rgba8"]), _ptr(o["depth"]), _ptr(o["color"]), _ptr(o["t"]), _ptr(o["hit_id"]),
                         _ptr(o["aux"]))
    return o


class Bvh:
    """The oracle's LBVH. threads: build threads (0 = default_threads(), 1 = the serial build;
    both give the same tree)."""

    def __init__(self, xyzr, bits=30, threads=0):
        self.lib = load()
        self.xyzr = np.ascontiguousarray(xyzr, np.float32)
        self.b = OrcBvh()
        t0 = self.lib.orc_now()
        rc = self.lib.orc_bvh_build_mt(self.xyzr.ctypes.data, len(self.xyzr), bits, threads or default_threads(),
                                       C.byref(self.b))
        self.build_s = self.lib.orc_now() - t0
        if rc != 0:
            raise RuntimeError("oracle LBVH build failed")

    def arrays(self):
        n = self.b.n
        m = max(n - 1, 0)
        box = np.zeros((m, 12), np.float32)
        child = np.zeros((m, 2), np.int32)
        ids = np.zeros(n, np.uint32)
        keys = np.zeros(n, np.uint64)
        self.lib.orc_bvh_copy_out(C.byref(self.b), box.ctypes.data, child.ctypes.data, ids.ctypes.data,
                                  keys.ctypes.data)
        return box, child, ids, keys

    def render(self, rgba, cam, p: OrcParams, threads=0, aux=False, outputs=None) -> dict:
        rgba = np.ascontiguousarray(rgba, np.float32)
        o = _outs(p.height, p.width, aux)
        if outputs is not None:
            for k in list(o):
                if k not in outputs and k != "aux":
                    o[k] = None
        cnt = np.zeros(2, np.uint64)
        t0 = self.lib.orc_now()
        self.lib.orc_render_bvh(C.byref(self.b), rgba.ctypes.data, C.byref(camera_from(cam)), C.byref(p),
                                threads or default_threads(),
                                _ptr(o["rgba8"]), _ptr(o["depth"]), _ptr(o["color"]), _ptr(o["t"]),
                                _ptr(o["hit_id"]), _ptr(o["aux"]), cnt.ctypes.data)
        o["seconds"] = self.lib.orc_now() - t0
        o["node_visits"], o["sphere_tests"] = int(cnt[0]), int(cnt[1])
        return o

    def __del__(self):
        try:
            self.lib.orc_bvh_free(C.byref(self.b))
        except Exception:
            pass


class Pkd:
    """The oracle's treelet P-k-d trees over the Morton order of an LBVH build (treelets of M
    consecutive spheres, each a left-balanced k-d tree in heap order; a complete tree of boxes over
    the treelets)."""

    def __init__(self, bvh: "Bvh", M=1024, threads=0):
        self.lib = load()
        self.bvh = bvh  # keeps the LBVH's arrays alive for the lifetime of this object
        self.k = OrcPkd()
        t0 = self.lib.orc_now()
        rc = self.lib.orc_pkd_build(C.byref(bvh.b), M, threads or default_threads(), C.byref(self.k))
        self.build_s = self.lib.orc_now() - t0
        if rc != 0:
            raise ValueError(f"oracle P-k-d build failed (M {M})")

    def arrays(self):
        """sph (n,4), id (n,), dim (n,), tl_box (T,6), tl_rmax (T,), top_box (L-1,12), top_child (L-1,2)."""
        k = self.k
        n, T, L = k.n, k.T, k.top_leaves

        def arr(ptr, shape, dt):
            m = int(np.prod(shape))
            a = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(np.ctypeslib.as_ctypes_type(dt))), (m,))
            return a.reshape(shape).copy()
        return dict(sph=arr(k.sph, (n, 4), np.float32), id=arr(k.id, (n,), np.uint32),
                    dim=arr(k.dim, (n,), np.uint8), tl_box=arr(k.tl_box, (T, 6), np.float32),
                    tl_rmax=arr(k.tl_rmax, (T,), np.float32), top_box=arr(k.top_box, (L - 1, 12), np.float32),
                    top_child=arr(k.top_child, (L - 1, 2), np.int32))

    def render(self, rgba, cam, p: OrcParams, threads=0, aux=False, outputs=None) -> dict:
        rgba = np.ascontiguousarray(rgba, np.float32)
        o = _outs(p.height, p.width, aux)
        if outputs is not None:
            for k in list(o):
                if k not in outputs and k != "aux":
                    o[k] = None
        cnt = np.zeros(2, np.uint64)
        t0 = self.lib.orc_now()
        self.lib.orc_render_pkd(C.byref(self.k), rgba.ctypes.data, C.byref(camera_from(cam)), C.byref(p),
                                threads or default_threads(), _ptr(o["rgba8"]), _ptr(o["depth"]), _ptr(o["color"]),
                                _ptr(o["t"]), _ptr(o["hit_id"]), _ptr(o["aux"]), cnt.ctypes.data)
        o["seconds"] = self.lib.orc_now() - t0
        o["node_visits"], o["sphere_tests"] = int(cnt[0]), int(cnt[1])
        return o

    def __del__(self):
        try:
            self.lib.orc_pkd_free(C.byref(self.k))
        except Exception:
            pass


def pkd_subtree_size(i, m):
    return int(load().orc_pkd_subtree_size(i, m))


def primary_f64(xyzr, cam, p: OrcParams, x0=0, y0=0, x1=None, y1=None):
    lib = load()
    x1 = p.width if x1 is None else x1
    y1 = p.height if y1 is None else y1
    xyzr = np.ascontiguousarray(xyzr, np.float32)
    m = (x1 - x0) * (y1 - y0)
    ids = np.zeros(m, np.int32)
    t = np.zeros(m, np.float64)
    disc = np.zeros(m, np.float64)
    t2 = np.zeros(m, np.float64)
    lib.orc_primary_f64(xyzr.ctypes.data, len(xyzr), C.byref(camera_from(cam)), C.byref(p), x0, y0, x1, y1,
                        ids.ctypes.data, t.ctypes.data, disc.ctypes.data, t2.ctypes.data)
    shp = (y1 - y0, x1 - x0)
    return ids.reshape(shp), t.reshape(shp), disc.reshape(shp), t2.reshape(shp)


def primary_f64_rects(xyzr, cam, p: OrcParams, rects, threads=0):
    """FP64 nearest hit of the first sample's ray for every pixel of the rectangles
    [(x0, y0, x1, y1), ...] (multi-threaded, cone-culled: for scenes of any size). Returns one
    (ids, t, disc, t2) tuple of 2-D arrays per rectangle."""
    lib = load()
    xyzr = np.ascontiguousarray(xyzr, np.float32)
    r = np.ascontiguousarray(np.asarray(rects, np.uint32).reshape(-1, 4))
    npx = int(sum((x1 - x0) * (y1 - y0) for x0, y0, x1, y1 in r.tolist()))
    ids = np.zeros(max(npx, 1), np.int32)
    t = np.zeros(max(npx, 1), np.float64)
    disc = np.zeros(max(npx, 1), np.float64)
    t2 = np.zeros(max(npx, 1), np.float64)
    rc = lib.orc_primary_f64_rects(xyzr.ctypes.data, len(xyzr), C.byref(camera_from(cam)), C.byref(p),
                                   r.ctypes.data, len(r), threads or default_threads(), ids.ctypes.data, t.ctypes.data,
                                   disc.ctypes.data, t2.ctypes.data)
    if rc != 0:
        raise ValueError("rectangle outside the image")
    out, o = [], 0
    for x0, y0, x1, y1 in r.tolist():
        m = (x1 - x0) * (y1 - y0)
        shp = (y1 - y0, x1 - x0)
        out.append((ids[o:o + m].reshape(shp), t[o:o + m].reshape(shp), disc[o:o + m].reshape(shp),
                    t2[o:o + m].reshape(shp)))
        o += m
    return out


def convert_lists(lists, tf=None, tf_min=1.0, tf_max=0.0):
    """Oracle ingest: lists are megamol_amd ParticleList objects."""
    lib = load()
    cl = (OrcList * max(len(lists), 1))()
    total = 0
    for i, pl in enumerate(lists):
        s = pl.to_c()
        C.memmove(C.byref(cl[i]), C.byref(s), C.sizeof(OrcList))
        if s.count and s.vert_type != 0:
            total += s.count
    xyzr = np.zeros((max(total, 1), 4), np.float32)
    rgba = np.zeros((max(total, 1), 4), np.float32)
    n = C.c_uint64()
    tfp, tfn = None, 0
    if tf is not None:
        tf = np.ascontiguousarray(tf, np.float32).reshape(-1, 4)
        tfp, tfn = tf.ctypes.data, len(tf)
    lib.orc_convert_lists(cl, len(lists), tfp, tfn, tf_min, tf_max, 1 if tf_min > tf_max else 0, xyzr.ctypes.data,
                          rgba.ctypes.data, C.byref(n))
    return xyzr[:n.value], rgba[:n.value]


def blend_frames(old_col, new_frame, frame_idx, background):
    """Progressive blend of a new frame (an oracle render dict) into the previous colour buffer
    (raygen.cu:254-268). Returns (colour, rgba8)."""
    lib = load()
    new = np.ascontiguousarray(new_frame["color"], np.float32)
    old = np.ascontiguousarray(old_col if old_col is not None else new, np.float32)
    ids = np.ascontiguousarray(new_frame["hit_id"], np.int32)
    bg = np.asarray(background, np.float32)
    out = np.empty_like(new)
    rgba8 = np.empty(ids.shape, np.uint32)
    lib.orc_blend_frames(old.ctypes.data, new.ctypes.data, ids.ctypes.data, ids.size, frame_idx, bg.ctypes.data,
                         out.ctypes.data, rgba8.ctypes.data)
    return out, rgba8


def pack_rgba8(col):
    lib = load()
    col = np.ascontiguousarray(col, np.float32)
    out = np.empty(col.shape[:-1], np.uint32)
    lib.orc_pack_rgba8(col.ctypes.data, out.size, out.ctypes.data)
    return out


def tea16(a, b):
    return int(load().orc_tea16(a, b))


def tea_encrypt(v0, v1, key, cycles):
    v = np.array([v0, v1], np.uint32)
    k = np.array(key, np.uint32)
    load().orc_tea_encrypt(v.ctypes.data, k.ctypes.data, cycles)
    return int(v[0]), int(v[1])


def ao_direction(normal, seed, smp=0, spp=1):
    """AO direction of sample smp of spp from the sample's stream seed."""
    n = np.asarray(normal, np.float32)
    d = np.zeros(3, np.float32)
    load().orc_ao_direction(n.ctypes.data, seed, smp, spp, d.ctypes.data)
    return d


def lcg_raw(seed, k):
    lib = load()
    s = C.c_uint32(seed)
    out = []
    for _ in range(k):
        lib.orc_lcg(C.byref(s))
        out.append(int(s.value))
    return out


def lcg_stream(seed, k):
    lib = load()
    s = C.c_uint32(seed)
    return [float(lib.orc_rnd(C.byref(s))) for _ in range(k)]


def frame_state(cam) -> OrcFrameState:
    fs = OrcFrameState()
    load().orc_frame_state_from_camera(C.byref(camera_from(cam)), C.byref(fs))
    return fs


def intersect_f32(o, d, sph, tmin=1e-6, tmax=1e20):
    o = np.asarray(o, np.float32)
    d = np.asarray(d, np.float32)
    s = np.asarray(sph, np.float32)
    t = C.c_float()
    hit = load().orc_intersect_f32(o.ctypes.data, d.ctypes.data, tmin, tmax, s.ctypes.data, C.byref(t))
    return (float(t.value) if hit else None)


def intersect_f64(o, d, sph, tmin=1e-6):
    o = np.asarray(o, np.float64)
    d = np.asarray(d, np.float64)
    s = np.asarray(sph, np.float64)
    t = C.c_double()
    disc = C.c_double()
    hit = load().orc_intersect_f64(o.ctypes.data, d.ctypes.data, tmin, s.ctypes.data, C.byref(t), C.byref(disc))
    return (float(t.value) if hit else None), float(disc.value)
