"""ctypes binding of libmmhrt.so (the C ABI declared in include/mmhrt.h).

The library is loaded from megamol_amd/lib/. There is no fallback: if the HIP library is
missing or cannot be loaded, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libmmhrt.so"

# status codes
OK, E_INVALID, E_HIP, E_OOM, E_UNSUPPORTED, E_STATE = range(6)
# vertex / colour types (SimpleSphericalParticles.h:27-50)
VERTDATA_NONE, VERTDATA_FLOAT_XYZ, VERTDATA_FLOAT_XYZR, VERTDATA_SHORT_XYZ, VERTDATA_DOUBLE_XYZ = range(5)
(COLDATA_NONE, COLDATA_UINT8_RGB, COLDATA_UINT8_RGBA, COLDATA_FLOAT_RGB, COLDATA_FLOAT_RGBA, COLDATA_FLOAT_I,
 COLDATA_USHORT_RGBA, COLDATA_DOUBLE_I) = range(8)
MODE_LAMBERT, MODE_AO, MODE_BOUNCE = range(3)
FLAG_AUX, FLAG_NO_DOWNLOAD = 1, 2
BUILD_REBUILD, BUILD_REFIT, BUILD_AUTO = range(3)
VERT_SIZE = (0, 12, 16, 6, 24)
COL_SIZE = (0, 3, 4, 12, 16, 4, 8, 8)


class SphereList(C.Structure):
    _fields_ = [
        ("vert_type", C.c_int32), ("col_type", C.c_int32),
        ("vert", C.c_void_p), ("col", C.c_void_p),
        ("vert_stride", C.c_uint64), ("col_stride", C.c_uint64), ("count", C.c_uint64),
        ("global_radius", C.c_float), ("global_rgba", C.c_uint8 * 4),
        ("col_min", C.c_float), ("col_max", C.c_float),
    ]


class Camera(C.Structure):
    _fields_ = [
        ("position", C.c_float * 3), ("direction", C.c_float * 3), ("up", C.c_float * 3),
        ("fovy", C.c_float), ("aspect", C.c_float), ("near_plane", C.c_float), ("far_plane", C.c_float),
        ("tile_start", C.c_float * 2), ("tile_end", C.c_float * 2),
    ]


class RenderParams(C.Structure):
    _fields_ = [
        ("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32), ("mode", C.c_int32),
        ("ao_distance", C.c_float), ("max_bounces", C.c_int32), ("frame_idx", C.c_uint32),
        ("accumulate", C.c_int32), ("background", C.c_float * 4), ("flags", C.c_uint32),
        ("window", C.c_uint32 * 4),
    ]


class Framebuffer(C.Structure):
    _fields_ = [
        ("rgba8", C.c_void_p), ("depth", C.c_void_p), ("color", C.c_void_p), ("t", C.c_void_p),
        ("hit_id", C.c_void_p), ("aux", C.c_void_p),
    ]


class DeviceFramebuffer(C.Structure):
    _fields_ = [
        ("color", C.c_void_p), ("depth", C.c_void_p), ("t", C.c_void_p), ("hit_id", C.c_void_p),
        ("rgba8", C.c_void_p), ("aux", C.c_void_p), ("stream", C.c_void_p), ("width", C.c_uint32),
        ("height", C.c_uint32),
    ]


class Job(C.Structure):
    _fields_ = [("timestep", C.c_uint32), ("reserved", C.c_uint32), ("camera", Camera), ("params", RenderParams),
                ("out_offset", C.c_uint64)]


# name -> (restype, argtypes); must match include/mmhrt.h exactly
SIGNATURES = {
    "mmhrt_create": (C.c_int32, [C.c_int32, C.POINTER(C.c_void_p)]),
    "mmhrt_destroy": (C.c_int32, [C.c_void_p]),
    "mmhrt_last_error": (C.c_char_p, [C.c_void_p]),
    "mmhrt_set_spheres": (C.c_int32, [C.c_void_p, C.POINTER(SphereList), C.c_uint32, C.c_uint64, C.c_uint32]),
    "mmhrt_set_spheres_device": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "mmhrt_set_transfer_function": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_float, C.c_float]),
    "mmhrt_build": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32]),
    "mmhrt_render": (C.c_int32, [C.c_void_p, C.POINTER(Camera), C.POINTER(RenderParams), C.POINTER(Framebuffer)]),
    "mmhrt_download": (C.c_int32, [C.c_void_p, C.POINTER(Framebuffer)]),
    "mmhrt_get_device_framebuffer": (C.c_int32, [C.c_void_p, C.POINTER(DeviceFramebuffer)]),
    "mmhrt_sync": (C.c_int32, [C.c_void_p]),
    "mmhrt_get_info": (C.c_int32, [C.c_void_p, C.POINTER(C.c_uint64), C.c_void_p, C.POINTER(C.c_uint64)]),
    "mmhrt_get_bvh": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "mmhrt_render_noalloc": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "mmhrt_get_treelets": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_void_p]),
    "mmhrt_get_timings": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32]),
    "mmhrt_get_counters": (C.c_int32, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                       C.POINTER(C.c_uint64)]),
    "mmhrt_set_timesteps": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint64, C.c_void_p, C.c_void_p]),
    "mmhrt_render_batch": (C.c_int32, [C.c_void_p, C.POINTER(Job), C.c_uint32, C.c_int32, C.c_int32, C.c_int32,
                                       C.c_void_p, C.c_void_p, C.POINTER(C.c_float)]),
    "mmhrt_denoiser_init": (C.c_int32, [C.c_void_p, C.c_uint64]),
    "mmhrt_denoiser_param_count": (C.c_int32, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "mmhrt_denoiser_get_params": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "mmhrt_denoiser_set_params": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "mmhrt_denoise": (C.c_int32, [C.c_void_p, C.c_float]),
    "mmhrt_denoise_features": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float,
                                           C.c_void_p, C.c_void_p]),
    "mmhrt_denoiser_train_step": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                              C.c_float, C.c_void_p, C.c_float, C.c_int32, C.c_void_p, C.c_void_p]),
    "mmhrt_set_option": (C.c_int32, [C.c_void_p, C.c_char_p, C.c_int64]),
    "mmhrt_get_option": (C.c_int32, [C.c_void_p, C.c_char_p, C.POINTER(C.c_int64)]),
}

_lib = None


class MMHRTError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"mmhrt error {code}: {msg}")
        self.code = code


def load() -> C.CDLL:
    """Load libmmhrt.so (once). Raises if it is missing: there is no CPU fallback."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with megamol_amd.build.build()")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(ctx, code: int) -> None:
    if code != OK:
        lib = load()
        msg = lib.mmhrt_last_error(ctx)
        raise MMHRTError(code, msg.decode() if msg else "")
