"""GPU: mmhrt_render_noalloc (the render entry point that allocates nothing) through the C ABI.

Two calls in a row, the second on a smaller scene and a smaller frame, each against mmhrt_render
of the same inputs, bit for bit: colour, depth, t, hit ids, RGBA8 and the fp16 features. The
output buffers belong to the caller (allocated here with hipMalloc of the HIP runtime the
library uses, outside the entry point). Also: the entry point rejects an output size that is not
the render window and a missing buffer.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import case, small_camera
from megamol_amd import _lib as L, params, scenes

pytestmark = pytest.mark.gpu


class DeviceFrame:
    """Caller-owned device buffers for one w x h window."""

    def __init__(self, hip, w, h, aux):
        self.hip, self.w, self.h = hip, w, h
        sizes = dict(color=16, depth=4, t=4, hit_id=4, rgba8=4, aux=16 if aux else 0)
        self.ptr = {}
        for k, b in sizes.items():
            p = C.c_void_p()
            if b:
                assert hip.hipMalloc(C.byref(p), C.c_size_t(b * w * h)) == 0
            self.ptr[k] = p
        self.fb = L.DeviceFramebuffer()
        for k, p in self.ptr.items():
            setattr(self.fb, k, p.value)
        self.fb.width, self.fb.height = w, h

    def read(self, key, dtype, shape):
        a = np.empty(shape, dtype)
        assert self.hip.hipMemcpy(a.ctypes.data_as(C.c_void_p), self.ptr[key], C.c_size_t(a.nbytes), 2) == 0
        return a

    def free(self):
        for p in self.ptr.values():
            if p.value:
                self.hip.hipFree(p)


def test_render_noalloc_equals_render(gpu_renderer):
    r = gpu_renderer
    lib = L.load()
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    cases = [(30000, 96, 72, 4, L.MODE_AO, 3), (2000, 40, 30, 1, L.MODE_LAMBERT, 1)]
    for n, w, h, spp, mode, seed in cases:
        case(f"render_noalloc: {n} spheres, {w}x{h}, spp {spp}, mode {mode}")
        xyzr, rgba = scenes.random_spheres(n, seed=seed)
        r.set_spheres(xyzr, rgba)
        r.build(L.BUILD_REBUILD, 30)
        cam = small_camera(w, h)
        p = params(w, h, spp=spp, mode=mode, ao_distance=0.2, background=(0.1, 0.2, 0.3, 1.0), flags=L.FLAG_AUX)
        ref = r.render(cam, p)
        ref_aux = r.download(w, h, outputs=(), aux=True).aux
        d = DeviceFrame(hip, w, h, aux=True)
        try:
            rc = lib.mmhrt_render_noalloc(r.ctx, C.byref(cam), C.byref(p), C.byref(d.fb))
            assert rc == L.OK, lib.mmhrt_last_error(r.ctx)
            assert lib.mmhrt_sync(r.ctx) == L.OK
            np.testing.assert_array_equal(d.read("hit_id", np.int32, (h, w)), ref.hit_id)
            np.testing.assert_array_equal(d.read("t", np.uint32, (h, w)), ref.t.view(np.uint32))
            np.testing.assert_array_equal(d.read("depth", np.uint32, (h, w)), ref.depth.view(np.uint32))
            np.testing.assert_array_equal(d.read("color", np.uint32, (h, w, 4)), ref.color.view(np.uint32))
            np.testing.assert_array_equal(d.read("rgba8", np.uint32, (h, w)), ref.rgba8)
            np.testing.assert_array_equal(d.read("aux", np.uint16, (h, w, 8)), ref_aux.view(np.uint16))
            # wrong output size and a missing buffer are rejected before anything runs
            d.fb.width = w + 1
            assert lib.mmhrt_render_noalloc(r.ctx, C.byref(cam), C.byref(p), C.byref(d.fb)) == L.E_INVALID
            d.fb.width = w
            t_ptr = d.fb.t
            d.fb.t = None
            assert lib.mmhrt_render_noalloc(r.ctx, C.byref(cam), C.byref(p), C.byref(d.fb)) == L.E_INVALID
            d.fb.t = t_ptr
        finally:
            lib.mmhrt_sync(r.ctx)
            d.free()
