"""GPU: the MegaMol renderer module (HIPSphereRenderer, compiled against stand-ins of the MegaMol
types) renders an interleaved particle list into a CPU framebuffer identical to the C ABI; the
headless project runner drives it over the GenTestFiles data sets with their projects' camera, and
over a scripted session that exercises the progressive index, against the oracle."""
import subprocess
from pathlib import Path

import pytest

from helpers import case

pytestmark = pytest.mark.gpu
EXE = Path(__file__).resolve().parent.parent / "megamol_amd" / "module" / "_build" / "module_test"


def test_module_renders_like_the_c_abi():
    case("HIPSphereRenderer module driver")
    assert EXE.exists(), "build the module driver with megamol_amd.build.build()"
    res = subprocess.run([str(EXE)], capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "0 pixels differ" in res.stdout


# The view camera of the screenshot projects utils/MMPLD/GenTestFiles.pl writes (:199): the
# camstore settings string, 1280x720, half aperture 0.2618 rad.
GENTEST_CAMSTORE = (
    '{"centre_offset":[0.0,0.0],"convergence_plane":0.0,"eye":0,"far_clipping_plane":12.97671890258789,'
    '"film_gate":[0.0,0.0],"gate_scaling":0,"half_aperture_angle":0.2617993950843811,'
    '"half_disparity":0.02500000037252903,"image_tile":[0,720,1280,0],'
    '"near_clipping_plane":0.012976719066500664,"orientation":[0.17020022869110107,-0.24738462269306183,'
    '-0.0711577907204628,-0.9511932134628296],"position":[4.224766731262207,3.3975491523742676,'
    '7.757389545440674],"projection_type":0,"resolution_gate":[1280,720]}')
HEADLESS = EXE.parent / "headless"
FIXTURES = Path(__file__).resolve().parent / "golden" / "mmpld"


def run_headless(tmp_path, files, script, params=(), bg=(0.0, 0.0, 0.0, 1.0)):
    """Run the headless project runner; returns the frames (per file, per render: index, camera
    struct, background, rgba8, depth)."""
    import numpy as np
    from megamol_amd import _lib as L

    args = [str(HEADLESS), "--camstore", GENTEST_CAMSTORE, "--out", str(tmp_path / "shot"), "--script", script,
            "--bg", ",".join(str(v) for v in bg)]
    for f in files:
        args += ["--mmpld", str(f)]
    for p in params:
        args += ["--param", p]
    res = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    frames = {}
    for line in res.stdout.splitlines():
        w = line.split()
        if not w or w[0] != "frame":
            continue
        k, f, idx, W, H = int(w[1]), int(w[2]), int(w[4]), int(w[6]), int(w[8])
        v = [float.fromhex(x) for x in w[10:23]]
        bgv = tuple(float.fromhex(x) for x in w[24:28])
        cam = L.Camera()
        cam.position[:] = v[0:3]
        cam.direction[:] = v[3:6]
        cam.up[:] = v[6:9]
        cam.fovy, cam.aspect, cam.near_plane, cam.far_plane = v[9:13]
        cam.tile_end[:] = [1.0, 1.0]
        base = tmp_path / f"shot_{k}_{f}"
        rgba = np.fromfile(str(base) + ".rgba", np.uint32).reshape(H, W)
        depth = np.fromfile(str(base) + ".depth", np.float32).reshape(H, W)
        assert (base.parent / (base.name + ".ppm")).stat().st_size == len(f"P6\n{W} {H}\n255\n") + W * H * 3
        frames.setdefault(k, []).append(dict(index=idx, cam=cam, bg=bgv, rgba8=rgba, depth=depth, W=W, H=H))
    return frames


def oracle_chain(xyzr, rgba, frames, settings):
    """The oracle's frames for a sequence of module renders: index 0 renders afresh, index i > 0
    blends the frame of index i into the previous colour (raygen.cu:254-268)."""
    import oracle as O

    col = None
    out = []
    for fr, (spp, mode) in zip(frames, settings):
        p = O.make_params(fr["W"], fr["H"], spp, mode, 4.0, fr["index"], fr["bg"])
        o = O.render_brute(xyzr, rgba, fr["cam"], p)
        if fr["index"] == 0:
            col = o["color"]
            rgba8 = o["rgba8"]
        else:
            col, rgba8 = O.blend_frames(col, o, fr["index"], fr["bg"])
        out.append((rgba8, o["depth"]))
    return out


def test_headless_gentestfiles_projects(tmp_path):
    """The 17 GenTestFiles data sets rendered through MMPLDDataSource -> HIPSphereRenderer with the
    projects' camera at 1280x720, two frames each (the second accumulated at index 1): every
    screenshot equals the oracle's bit for bit."""
    import numpy as np
    import oracle as O
    from megamol_amd import mmpld

    files = sorted(FIXTURES.glob("test_*.mmpld"))
    assert len(files) == 17
    case("headless: 17 GenTestFiles projects, 1280x720, 2 frames each")
    frames = run_headless(tmp_path, files, "render;render")
    for k, f in enumerate(files):
        _, fr = mmpld.read(str(f))
        xyzr, rgba = O.convert_lists(fr[0].lists)
        got = frames[k]
        assert [g["index"] for g in got] == [0, 1], f.name
        d = got[0]["cam"].direction
        pos = np.array(got[0]["cam"].position[:])
        assert np.dot([d[0], d[1], d[2]], -pos / np.linalg.norm(pos)) > 0.9999  # the camera looks at the origin
        want = oracle_chain(xyzr, rgba, got, [(1, 0), (1, 0)])
        for g, (w8, wd) in zip(got, want):
            assert (g["depth"] < 1.0).sum() > 1000, f.name
            np.testing.assert_array_equal(g["rgba8"], w8, err_msg=f.name)
            np.testing.assert_array_equal(g["depth"].view(np.uint32), wd.view(np.uint32), err_msg=f.name)


def test_headless_progressive_index(tmp_path):
    """AbstractRenderer::Render's change detection (AbstractRenderer.cpp:45-108, Renderer.cpp:70-176):
    an unchanged Render advances the progressive index and blends; a parameter, camera,
    background or data change restarts at index 0. Every frame equals the oracle's chain."""
    import numpy as np
    import oracle as O
    from megamol_amd import mmpld

    f = FIXTURES / "test_xyzr_float_rgba_float.mmpld"
    script = ("render;render;render;param spp=2;render;render;move 0.05,0,0;render;bg 0.2,0.3,0.4,1;render;"
              "param accumulate=false;render;rehash;render;param mode=AO;render;render")
    case("headless: progressive index over a scripted session")
    frames = run_headless(tmp_path, [f], script)[0]
    assert [g["index"] for g in frames] == [0, 1, 2, 0, 1, 0, 0, 0, 0, 0, 1]
    _, fr = mmpld.read(str(f))
    xyzr, rgba = O.convert_lists(fr[0].lists)
    # the scripted settings follow the renders one for one
    settings = [(1, 0)] * 3 + [(2, 0)] * 6 + [(2, 1)] * 2
    want = oracle_chain(xyzr, rgba, frames, settings)
    for i, (g, (w8, wd)) in enumerate(zip(frames, want)):
        np.testing.assert_array_equal(g["rgba8"], w8, err_msg=f"render {i}")
        np.testing.assert_array_equal(g["depth"].view(np.uint32), wd.view(np.uint32), err_msg=f"render {i}")
