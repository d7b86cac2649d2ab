"""CPU tests of the boundary: the C-ABI library loads and exports what include/mmhrt.h declares.

No compute call is made here (there is no GPU in the build container).
"""
import ctypes as C
import re
from pathlib import Path

import pytest

from megamol_amd import _lib as L

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    names = []
    for h in sorted((ROOT / "include").glob("*.h")):
        names += re.findall(r"MMHRT_API\s+[\w\s\*]+?\b(mmhrt_\w+)\s*\(", h.read_text())
    return names


def test_header_declares_the_boundary():
    names = declared_symbols()
    assert len(names) >= 16
    assert {"mmhrt_create", "mmhrt_set_spheres", "mmhrt_build", "mmhrt_render"} <= set(names)


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(L.LIB_PATH))
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_signatures_cover_the_header():
    assert set(declared_symbols()) == set(L.SIGNATURES)


def test_struct_layouts_match_the_header():
    # sizes of the POD structs as laid out by the C compiler for include/mmhrt.h
    import subprocess, tempfile, os
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "mmhrt.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(mmhrt_sphere_list), sizeof(mmhrt_camera), sizeof(mmhrt_render_params),
         sizeof(mmhrt_framebuffer), offsetof(mmhrt_sphere_list, global_rgba), offsetof(mmhrt_render_params, background));
  return 0;
}'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "s.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "s")
        subprocess.run(["gcc", "-I", str(ROOT / "include"), c, "-o", exe], check=True)
        got = list(map(int, subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()))
    want = [C.sizeof(L.SphereList), C.sizeof(L.Camera), C.sizeof(L.RenderParams), C.sizeof(L.Framebuffer),
            L.SphereList.global_rgba.offset, L.RenderParams.background.offset]
    assert got == want


def test_create_without_gpu_fails_cleanly():
    """In a container without a GPU, create must fail with a status code, not crash."""
    import torch  # noqa: F401  (device count query only)

    lib = L.load()
    h = C.c_void_p()
    rc = lib.mmhrt_create(0, C.byref(h))
    if rc == L.OK:  # a GPU is present (running on the GPU host): release it
        lib.mmhrt_destroy(h)
        pytest.skip("a GPU is present")
    assert rc in (L.E_HIP, L.E_INVALID)
    assert b"device" in lib.mmhrt_last_error(None)


def test_null_context_is_rejected():
    lib = L.load()
    assert lib.mmhrt_build(None, 0, 30) == L.E_INVALID
    assert lib.mmhrt_render(None, None, None, None) == L.E_INVALID
    assert lib.mmhrt_destroy(None) == L.E_INVALID
