"""GPU: every traversal variant (LDS stack depth, work stealing, ballot refill) and the render
window give bit-identical frames, equal to the oracle."""
import numpy as np
import pytest

import oracle as O
from helpers import assert_frames_equal, case, oracle_params, small_camera
from megamol_amd import _lib as L, params, scenes

pytestmark = pytest.mark.gpu

VARIANTS = [dict(stack_lds=16, persistent=1, refill=0), dict(stack_lds=8, persistent=0, refill=0),
            dict(stack_lds=32, persistent=1, refill=0), dict(stack_lds=16, persistent=1, refill=1),
            dict(stack_lds=16, persistent=1, refill=32), dict(stack_lds=8, persistent=1, refill=48),
            dict(stack_lds=24, persistent=1, refill=64)]


DEFAULTS = dict(stack_lds=16, persistent=1, refill=-1)  # the library's (refill -1: 16 at 1 spp, else 0)


def set_opts(r, opts):
    for k, v in opts.items():
        r.set_option(k, v)


@pytest.mark.parametrize("mode", [L.MODE_LAMBERT, L.MODE_AO])
def test_variants_bit_identical(gpu_renderer, mode):
    xyzr, rgba = scenes.random_spheres(6000, seed=12, rmin=0.005, rmax=0.03)
    cam = small_camera(101, 67)  # not a multiple of the 8x8 tile
    p = params(101, 67, spp=3, mode=mode, ao_distance=0.2, flags=L.FLAG_AUX)
    gpu_renderer.set_spheres(xyzr, rgba)
    gpu_renderer.build()
    o = O.render_brute(xyzr, rgba, cam, oracle_params(p), aux=True)
    try:
        for v in VARIANTS:
            case(f"variant {v} mode {mode}")
            set_opts(gpu_renderer, v)
            g = gpu_renderer.render(cam, p)
            g.aux = gpu_renderer.download(101, 67, outputs=(), aux=True).aux
            assert_frames_equal(g, o, aux=True, what=str(v))
    finally:
        set_opts(gpu_renderer, DEFAULTS)


@pytest.mark.parametrize("spp", [2, 4, 8, 16, 32, 64, 128])
@pytest.mark.parametrize("mode", [L.MODE_LAMBERT, L.MODE_AO])
def test_sample_lanes_bit_identical(gpu_renderer, mode, spp):
    """The sample-parallel kernel (lanes = samples of a pixel, jitter stream advanced by LCG
    skip-ahead, ordered fold) against the lane-per-pixel kernel and the oracle, for every lane
    group size it serves (L = 2 ... 64 lanes per pixel; 128 spp is two passes of 64)."""
    xyzr, rgba = scenes.random_spheres(5000, seed=14, rmin=0.005, rmax=0.03)
    w, h = (37, 29) if spp >= 64 else (101, 67)
    cam = small_camera(w, h)
    p = params(w, h, spp=spp, mode=mode, ao_distance=0.2, flags=L.FLAG_AUX)
    gpu_renderer.set_spheres(xyzr, rgba)
    gpu_renderer.build()
    frames = {}
    try:
        for sl in (1, 0):
            case(f"sample lanes {sl} mode {mode} spp {spp}")
            gpu_renderer.set_option("sample_lanes", sl)
            g = gpu_renderer.render(cam, p)
            g.aux = gpu_renderer.download(w, h, outputs=(), aux=True).aux
            frames[sl] = g
    finally:
        gpu_renderer.set_option("sample_lanes", 1)
    a, b = frames[1], frames[0]
    for k in ("hit_id", "t", "depth", "color", "rgba8", "aux"):
        np.testing.assert_array_equal(getattr(a, k).view(np.uint8), getattr(b, k).view(np.uint8), err_msg=k)
    o = O.Bvh(xyzr, 30).render(rgba, cam, oracle_params(p), aux=True)
    assert_frames_equal(a, o, aux=True, what=f"sample lanes spp {spp}")


@pytest.mark.parametrize("spp", [1, 3, 64, 128])
@pytest.mark.parametrize("mode", [L.MODE_LAMBERT, L.MODE_AO])
def test_trees_with_and_without_leaf_pairs_render_the_same(gpu_renderer, mode, spp):
    """A tree built with leaf pairs (nodes over two leaves not stored, their visits read the two
    spheres) and one built without: the same frames, bit for bit, through every kernel that the
    frame takes (the sample-lanes kernel has a variant for each kind of tree), equal to the
    oracle."""
    xyzr, rgba = scenes.random_spheres(5000, seed=15, rmin=0.005, rmax=0.03)
    w, h = (37, 29) if spp >= 64 else (101, 67)
    cam = small_camera(w, h)
    p = params(w, h, spp=spp, mode=mode, ao_distance=0.2, flags=L.FLAG_AUX)
    gpu_renderer.set_spheres(xyzr, rgba)
    frames = {}
    try:
        for pairs in (1, 0):
            case(f"leaf pairs {pairs} mode {mode} spp {spp}")
            gpu_renderer.set_option("leaf_pairs", pairs)
            gpu_renderer.build()
            g = gpu_renderer.render(cam, p)
            g.aux = gpu_renderer.download(w, h, outputs=(), aux=True).aux
            frames[pairs] = g
    finally:
        gpu_renderer.set_option("leaf_pairs", 1)
    a, b = frames[1], frames[0]
    for k in ("hit_id", "t", "depth", "color", "rgba8", "aux"):
        np.testing.assert_array_equal(getattr(a, k).view(np.uint8), getattr(b, k).view(np.uint8), err_msg=k)
    o = O.Bvh(xyzr, 30).render(rgba, cam, oracle_params(p), aux=True)
    assert_frames_equal(a, o, aux=True, what=f"leaf pairs spp {spp}")


def test_window_is_a_crop_of_the_full_frame(gpu_renderer):
    xyzr, rgba = scenes.random_spheres(4000, seed=13)
    cam = small_camera(120, 90)
    gpu_renderer.set_spheres(xyzr, rgba)
    gpu_renderer.build()
    full = gpu_renderer.render(cam, params(120, 90, spp=2, mode=L.MODE_AO, ao_distance=0.3))
    for refill in (0, 32):
        gpu_renderer.set_option("refill", refill)
        for (x0, y0, w, h) in [(0, 0, 120, 90), (8, 16, 40, 24), (37, 5, 19, 61), (119, 89, 1, 1)]:
            case(f"window {(x0, y0, w, h)} refill {refill}")
            p = params(120, 90, spp=2, mode=L.MODE_AO, ao_distance=0.3)
            p.window[:] = [x0, y0, w, h]
            win = gpu_renderer.render(cam, p)
            np.testing.assert_array_equal(win.hit_id, full.hit_id[y0:y0 + h, x0:x0 + w])
            np.testing.assert_array_equal(win.color, full.color[y0:y0 + h, x0:x0 + w])
            np.testing.assert_array_equal(win.depth, full.depth[y0:y0 + h, x0:x0 + w])
    gpu_renderer.set_option("refill", -1)
    p = params(120, 90)
    p.window[:] = [100, 0, 40, 10]
    with pytest.raises(L.MMHRTError) as e:
        gpu_renderer.render(cam, p)
    assert e.value.code == L.E_INVALID


def test_fcc_sample_against_fp64(gpu_renderer):
    """Full-size property at scale: on an FCC crystal of 1.4M spheres, sampled pixels' primary hit
    ids equal the FP64 brute-force oracle except at grazing/tie pixels."""
    case("fcc 1.4M spheres vs fp64 on sampled pixels")
    cells = 70
    xyzr, inten = scenes.fcc_crystal(cells, seed=2)
    rgba = np.ones((len(xyzr), 4), np.float32)
    cam = scenes.fcc_camera(cells, width=320, height=180)
    gpu_renderer.set_spheres(xyzr, rgba)
    gpu_renderer.build(L.BUILD_REBUILD, 30)
    g = gpu_renderer.render(cam, params(320, 180, spp=1, mode=L.MODE_LAMBERT))
    # FP64 over a 16x16 block in the middle of the crystal's image
    p = O.make_params(320, 180)
    ids, t64, disc, t2 = O.primary_f64(xyzr, cam, p, 152, 82, 168, 98)
    sub = g.hit_id[82:98, 152:168]
    mism = ids != sub
    for y, x in zip(*np.nonzero(mism)):
        tie = np.isfinite(t2[y, x]) and abs(t2[y, x] - t64[y, x]) <= 1e-5 * t64[y, x]
        assert tie or abs(disc[y, x]) < 1e-4
    assert (sub >= 0).all()
    rel = np.abs(g.t[82:98, 152:168] - t64) / t64
    assert rel[~mism].max() <= 1e-5
