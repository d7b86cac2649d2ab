"""GPU parity: the HIP path through the C ABI against the CPU oracle, bit for bit.

Hit IDs, t, pseudo depth, float colour, RGBA8 and the fp16 feature buffer must be identical to
the oracle's FP32 restatement on the same inputs (integer and FP32 sequences are the same, so
there is no tolerance). The FP64 oracle bounds the FP32 answer itself (see test_oracle_cpu).
"""
import numpy as np
import pytest

import oracle as O
from helpers import assert_frames_equal, case, oracle_params, small_camera
from megamol_amd import ParticleList, _lib as L, look_at, mmpld, params, scenes

pytestmark = pytest.mark.gpu


def render_both(r, xyzr, rgba, cam, p, bits=30, aux=False):
    r.set_spheres(xyzr, rgba)
    r.build(L.BUILD_REBUILD, bits)
    g = r.render(cam, p)
    if aux:
        fr = r.download(p.width, p.height, outputs=(), aux=True)
        g.aux = fr.aux
    o = O.render_brute(xyzr, rgba, cam, oracle_params(p), aux=aux)
    return g, o


@pytest.mark.parametrize("seed", [0, 1])
@pytest.mark.parametrize("mode", [L.MODE_LAMBERT, L.MODE_AO])
@pytest.mark.parametrize("bits", [30, 63])
def test_small_scene_bit_exact(gpu_renderer, seed, mode, bits):
    case(f"small scene seed={seed} mode={mode} bits={bits}")
    xyzr, rgba = scenes.random_spheres(3000, seed=seed)
    xyzr[200:205] = xyzr[7]  # exact duplicates -> ties
    cam = small_camera(96, 72)
    p = params(96, 72, spp=3, mode=mode, ao_distance=0.3, background=(0.1, 0.2, 0.3, 1.0), flags=L.FLAG_AUX)
    g, o = render_both(gpu_renderer, xyzr, rgba, cam, p, bits, aux=True)
    assert (o["hit_id"] >= 0).mean() > 0.5
    assert_frames_equal(g, o, aux=True, what=f"seed={seed} mode={mode} bits={bits}")


def leaf_pair_form(ochild):
    """The oracle's tree in the stored form: a non-root node whose two children are leaves (a leaf
    pair) is not stored; its parent references it as its first leaf and sets the pair flag of
    that side (bit 0 / bit 1). Returns the expected child references, the flags and the mask of
    stored nodes."""
    m = len(ochild)
    is_pair = (ochild[:, 0] < 0) & (ochild[:, 1] < 0)
    is_pair[0] = False
    child = ochild.copy()
    pairs = np.zeros(m, np.int32)
    for side in (0, 1):
        c = ochild[:, side]
        hit = c >= 0
        hit[hit] = is_pair[c[hit]]
        child[hit, side] = ochild[c[hit], 0]
        pairs[hit] |= 1 << side
    # a pair's first child is its first leaf: ~l with l the lower end of its two-leaf range
    return child, pairs, ~is_pair


@pytest.mark.parametrize("chunk,ls,ri,grid,os1,wave,tail,span,dig", [
    (1024, 1, 16, 0, 0, 64, 512, 256, 8), (512, 1, 8, 0, 0, 64, 512, 256, 8), (256, 1, 32, 0, 0, 64, 512, 256, 8),
    (1024, 0, 16, 0, 0, 64, 512, 256, 8), (256, 0, 8, 0, 0, 64, 512, 256, 8), (1024, 1, 32, 3, 0, 64, 512, 256, 8),
    (256, 1, 16, 1, 0, 64, 512, 256, 8), (512, 1, 32, 7, 0, 64, 512, 256, 8), (1024, 1, 32, 0, 1, 64, 512, 256, 8),
    (1024, 1, 16, 0, 1, 64, 512, 256, 8), (512, 1, 8, 5, 1, 64, 512, 256, 8), (1024, 1, 32, 0, 0, 0, 512, 256, 8),
    (256, 1, 16, 0, 0, 32, 1 << 30, 0, 8), (1024, 1, 32, 0, 0, 64, 4096, 0, 8), (1024, 1, 32, 0, 0, 64, 512, 0, 8),
    (256, 1, 16, 0, 0, 64, 512, 1, 8), (1024, 0, 8, 2, 0, 64, 512, 16, 8), (1024, 1, 16, 0, 0, 64, 512, 256, 0),
    (512, 1, 32, 0, 0, 64, 512, 256, 0), (256, 1, 8, 0, 0, 64, 512, 256, 0),
    (1024, 1, 16, 0, 0, 1024, 512, 256, 8), (512, 1, 16, 3, 0, 256, 512, 256, 8),
    (1024, 1, 32, 0, 0, 128, 512, 0, 8)])
def test_lbvh_equals_oracle_lbvh(gpu_renderer, chunk, ls, ri, grid, os1, wave, tail, span, dig):
    """Both chunk passes (level-synchronous rounds, or one climbing thread per leaf) at every
    chunk size, over radix tiles of 8, 16 or 32 keys per thread, with either radix pass
    (histogram + scans + scatter, or single-pass with look-back: os1), give the oracle's tree
    node for node (sorted keys and leaf order included; leaf pairs in their stored form). grid > 0
    caps the persistent chunk workgroups, so that each one runs many chunks through its load
    pipeline; wave is the list size from which the remaining merges climb without rounds, one lane
    per split (one wavefront up to 64, several wavefronts with acquire-release counts above; 0:
    rounds to the top);
    tail is the straddler list size from which one workgroup finishes the tree (1 << 30: it takes
    every straddler from the chunk pass on) when span is 0 (kernel rounds); span > 0 merges the
    straddlers with one workgroup per span of that many chunks before the tail (after the
    level-synchronous chunk pass, which then stores its boundary arrivals per side instead of
    exchanging them, span 1 makes every chunk boundary a span boundary). dig is the radix
    digit width: 8 the 8-bit passes, 0 the wide digits (3 passes of 10 bits for 30-bit keys, 6 of
    11 bits for 63-bit keys; single-pass radix passes always take 8-bit digits)."""
    gpu_renderer.set_option("sort_digit", dig)
    gpu_renderer.set_option("build_chunk", chunk)
    gpu_renderer.set_option("build_ls", ls)
    gpu_renderer.set_option("sort_ri", ri)
    gpu_renderer.set_option("build_grid", grid)
    gpu_renderer.set_option("sort_onesweep", os1)
    gpu_renderer.set_option("build_wave", wave)
    gpu_renderer.set_option("build_tail", tail)
    gpu_renderer.set_option("build_span", span)
    for bits in (30, 63):
        for n, dup in [(1000, False), (50000, True), (2, False), (3, True), (4097, False), (1_200_000, False)]:
            case(f"lbvh n={n} dup={dup} bits={bits} chunk={chunk} ls={ls} ri={ri} grid={grid} onesweep={os1} "
                 f"wave={wave} tail={tail} span={span} digit={dig}")
            g = np.random.default_rng(n + bits)
            xyzr = np.zeros((n, 4), np.float32)
            xyzr[:, :3] = g.random((n, 3)).astype(np.float32)
            xyzr[:, 3] = g.uniform(0.001, 0.01, n).astype(np.float32)
            if dup:
                xyzr[n // 2:, :3] = xyzr[: n - n // 2, :3]
            gpu_renderer.set_spheres(xyzr)
            gpu_renderer.build(L.BUILD_REBUILD, bits)
            box, child, ids, keys, height, pairs = gpu_renderer.bvh()
            ob = O.Bvh(xyzr, bits)
            obox, ochild, oids, okeys = ob.arrays()
            np.testing.assert_array_equal(keys, okeys)
            np.testing.assert_array_equal(ids, oids)
            echild, epairs, stored = leaf_pair_form(ochild)
            np.testing.assert_array_equal(child[stored], echild[stored])
            np.testing.assert_array_equal(pairs[stored], epairs[stored])
            np.testing.assert_array_equal(box[stored].view(np.uint32), obox[stored].view(np.uint32))
            assert gpu_renderer.get_option("leaf_pair_count") == int((~stored).sum())
            assert height[0] >= int(np.ceil(np.log2(n)))
            assert gpu_renderer.get_option("sort_fault") == 0
    gpu_renderer.set_option("sort_onesweep", 0)
    gpu_renderer.set_option("sort_digit", 8)
    gpu_renderer.set_option("build_chunk", 1024)
    gpu_renderer.set_option("build_ls", 1)
    gpu_renderer.set_option("sort_ri", 16)
    gpu_renderer.set_option("build_grid", 0)
    gpu_renderer.set_option("build_wave", 64)
    gpu_renderer.set_option("build_tail", 512)
    gpu_renderer.set_option("build_span", 256)


@pytest.mark.parametrize("ls", [1, 0])
def test_lbvh_without_leaf_pairs_is_the_oracle_tree(gpu_renderer, ls):
    """With leaf_pairs off every node is stored: the oracle's arrays exactly, no pair flags."""
    gpu_renderer.set_option("leaf_pairs", 0)
    gpu_renderer.set_option("build_ls", ls)
    try:
        for n in (3, 4097, 300_000):
            case(f"lbvh without leaf pairs n={n} ls={ls}")
            g = np.random.default_rng(n)
            xyzr = np.zeros((n, 4), np.float32)
            xyzr[:, :3] = g.random((n, 3)).astype(np.float32)
            xyzr[:, 3] = g.uniform(0.001, 0.01, n).astype(np.float32)
            gpu_renderer.set_spheres(xyzr)
            gpu_renderer.build(L.BUILD_REBUILD, 30)
            box, child, ids, keys, height, pairs = gpu_renderer.bvh()
            obox, ochild, oids, okeys = O.Bvh(xyzr, 30).arrays()
            np.testing.assert_array_equal(ids, oids)
            np.testing.assert_array_equal(child, ochild)
            np.testing.assert_array_equal(box.view(np.uint32), obox.view(np.uint32))
            assert not pairs.any()
            assert gpu_renderer.get_option("leaf_pair_count") == 0
    finally:
        gpu_renderer.set_option("leaf_pairs", 1)
        gpu_renderer.set_option("build_ls", 1)


def test_edge_cases(gpu_renderer):
    cam = small_camera(40, 30)
    p = params(40, 30, spp=2, mode=L.MODE_AO, ao_distance=1.0, flags=L.FLAG_AUX)
    cases = {
        "one sphere": np.array([[0.5, 0.5, 0.5, 0.2]], np.float32),
        "two spheres": np.array([[0.4, 0.5, 0.5, 0.1], [0.6, 0.5, 0.5, 0.1]], np.float32),
        "all same centre": np.tile(np.array([[0.5, 0.5, 0.5, 0.2]], np.float32), (64, 1)),
        "camera inside a sphere": np.array([[0.5, 0.5, 2.2, 0.5], [0.5, 0.5, 0.5, 0.2]], np.float32),
        "behind the camera": np.array([[0.5, 0.5, 4.0, 0.3]], np.float32),
        "zero radius": np.array([[0.5, 0.5, 0.5, 0.0], [0.3, 0.5, 0.5, 0.1]], np.float32),
    }
    for name, xyzr in cases.items():
        case(f"edge: {name}")
        rgba = np.linspace(0.1, 1.0, len(xyzr) * 4, dtype=np.float32).reshape(-1, 4)
        g, o = render_both(gpu_renderer, xyzr, rgba, cam, p, 30, aux=True)
        assert_frames_equal(g, o, aux=True, what=name)
        if name == "all same centre":
            assert set(np.unique(g.hit_id)) <= {-1, 0}


def test_mmpld_fixture_lists_through_the_abi(gpu_renderer):
    import glob
    from pathlib import Path
    files = sorted(glob.glob(str(Path(__file__).parent / "golden" / "mmpld" / "*.mmpld")))
    cam = look_at((0.5, 0.5, 6.0), (0, 0, 0), (0, 1, 0), 60.0, 1.0, 0.1, 20)
    p = params(48, 48, spp=1, mode=L.MODE_LAMBERT)
    tf = scenes.transfer_function_256()
    for with_tf in (False, True):
        gpu_renderer.set_transfer_function(tf if with_tf else None)
        for f in files:
            case(f"mmpld {Path(f).name} tf={with_tf}")
            hdr, frames = mmpld.read(f)
            gpu_renderer.set_lists(frames[0].lists)
            gpu_renderer.build()
            g = gpu_renderer.render(cam, p)
            xyzr, rgba = O.convert_lists(frames[0].lists, tf=tf if with_tf else None)
            o = O.render_brute(xyzr, rgba, cam, oracle_params(p))
            assert_frames_equal(g, o, what=Path(f).name)
    gpu_renderer.set_transfer_function(None)


def test_transfer_function_intensity_list(gpu_renderer):
    case("TF intensity list, auto range and explicit range")
    g = np.random.default_rng(5)
    n = 5000
    xyzr = np.zeros((n, 4), np.float32)
    xyzr[:, :3] = g.random((n, 3)).astype(np.float32)
    xyzr[:, 3] = 0.03
    inten = g.uniform(-0.5, 1.5, n).astype(np.float32)
    tf = scenes.transfer_function_256()
    cam = small_camera(64, 64)
    p = params(64, 64)
    for vmin, vmax in [(1.0, 0.0), (0.0, 1.0), (0.25, 0.3)]:
        lst = ParticleList(L.VERTDATA_FLOAT_XYZR, xyzr, L.COLDATA_FLOAT_I, inten, col_min=-0.5, col_max=1.5,
                           count=n)
        gpu_renderer.set_transfer_function(tf, vmin, vmax)
        gpu_renderer.set_lists([lst])
        gpu_renderer.build()
        gr = gpu_renderer.render(cam, p)
        oxyzr, orgba = O.convert_lists([lst], tf=tf, tf_min=vmin, tf_max=vmax)
        o = O.render_brute(oxyzr, orgba, cam, oracle_params(p))
        assert_frames_equal(gr, o, what=f"tf range {vmin},{vmax}")
    gpu_renderer.set_transfer_function(None)


@pytest.mark.parametrize("grid", [0, 2])
def test_refit_equals_rebuild(gpu_renderer, grid):
    """A refit (same keys and tree, boxes from the moved spheres) renders what the oracle renders
    from the moved spheres, and what a rebuild renders; grid 2 runs the refit's chunk pass
    through two persistent workgroups."""
    gpu_renderer.set_option("build_grid", grid)
    xyzr, rgba = scenes.random_spheres(20000, seed=4, rmin=0.005, rmax=0.02)
    cam = small_camera(80, 80)
    p = params(80, 80, spp=2, mode=L.MODE_AO, ao_distance=0.2)
    gpu_renderer.set_spheres(xyzr, rgba)
    gpu_renderer.build(L.BUILD_REBUILD, 30)
    g = np.random.default_rng(1)
    moved = xyzr.copy()
    for step in range(3):
        case(f"refit step {step} grid {grid}")
        moved[:, :3] += g.normal(0, 0.01, (len(moved), 3)).astype(np.float32)
        gpu_renderer.set_spheres(moved, rgba)
        gpu_renderer.build(L.BUILD_REFIT, 30)
        a = gpu_renderer.render(cam, p)
        o = O.render_brute(moved, rgba, cam, oracle_params(p))
        assert_frames_equal(a, o, what=f"refit step {step}")
    gpu_renderer.build(L.BUILD_REBUILD, 30)
    b = gpu_renderer.render(cam, p)
    gpu_renderer.set_option("build_grid", 0)
    np.testing.assert_array_equal(a.hit_id, b.hit_id)
    np.testing.assert_array_equal(a.color, b.color)


def test_progressive_accumulation(gpu_renderer):
    """raygen.cu:254-268: frame k blends into the stored frame with weight 1 / (k + 1), against
    the oracle's blend of the oracle frames; a different window, window origin or frame size
    cannot be blended into (E_STATE); accumulate 0 leaves frame_idx a seed only."""
    case("progressive accumulation")
    xyzr, rgba = scenes.random_spheres(2000, seed=6)
    cam = small_camera(48, 48)
    gpu_renderer.set_spheres(xyzr, rgba)
    gpu_renderer.build()
    bg = (0.1, 0.2, 0.3, 1.0)
    acc = None
    for k in range(4):
        p = params(48, 48, spp=1, mode=L.MODE_AO, ao_distance=0.3, frame_idx=k, accumulate=1, background=bg)
        g = gpu_renderer.render(cam, p)
        o = O.render_brute(xyzr, rgba, cam, O.make_params(48, 48, 1, 1, 0.3, k, bg))
        acc, rgba8 = O.blend_frames(acc, o, k, bg)
        np.testing.assert_array_equal(g.color.view(np.uint32), acc.view(np.uint32), err_msg=f"frame {k}")
        np.testing.assert_array_equal(g.rgba8, rgba8, err_msg=f"frame {k} rgba8")
    # accumulate 0: frame_idx seeds the streams, nothing is blended
    p = params(48, 48, spp=1, mode=L.MODE_AO, ao_distance=0.3, frame_idx=5, accumulate=0, background=bg)
    o = O.render_brute(xyzr, rgba, cam, O.make_params(48, 48, 1, 1, 0.3, 5, bg))
    np.testing.assert_array_equal(gpu_renderer.render(cam, p).color.view(np.uint32), o["color"].view(np.uint32))
    # the stored frame must be the same window of a frame of the same size
    base = params(48, 48, spp=1, frame_idx=0, accumulate=1)
    base.window[:] = [8, 8, 16, 16]
    gpu_renderer.render(cam, base)
    same = params(48, 48, spp=1, frame_idx=1, accumulate=1)
    same.window[:] = [8, 8, 16, 16]
    gpu_renderer.render(cam, same)
    for win, size in [([0, 0, 16, 16], (48, 48)), ([8, 8, 16, 16], (64, 48)), ([8, 8, 20, 16], (48, 48))]:
        q = params(size[0], size[1], spp=1, frame_idx=2, accumulate=1)
        q.window[:] = win
        with pytest.raises(L.MMHRTError) as e:
            gpu_renderer.render(cam, q)
        assert e.value.code == L.E_STATE, (win, size)


def test_error_paths(gpu_renderer):
    from megamol_amd import SphereRenderer
    with SphereRenderer(0) as r:
        cam = small_camera(8, 8)
        with pytest.raises(L.MMHRTError) as e:
            r.render(cam, params(8, 8))
        assert e.value.code == L.E_STATE
        with pytest.raises(L.MMHRTError) as e:
            r.build()
        assert e.value.code == L.E_STATE
        r.set_spheres(np.array([[0, 0, 0, 1]], np.float32))
        with pytest.raises(L.MMHRTError) as e:
            r.build(L.BUILD_REBUILD, 31)
        assert e.value.code == L.E_INVALID
        r.build()
        with pytest.raises(L.MMHRTError) as e:
            r.render(cam, params(0, 0))
        assert e.value.code == L.E_INVALID
        with pytest.raises(L.MMHRTError) as e:
            r.render(cam, params(8, 8, spp=0))
        assert e.value.code == L.E_INVALID
        with pytest.raises(L.MMHRTError) as e:
            r.set_lists([ParticleList(9, np.zeros(16, np.uint8), count=1)])
        assert e.value.code == L.E_INVALID
        # empty and ragged lists are skipped (SphereGeometry.cpp:98-100)
        pts = np.array([[0.1, 0.2, 0.3], [0.4, 0.5, 0.6]], np.float32)
        r.set_lists([ParticleList(L.VERTDATA_FLOAT_XYZ, pts, count=0),
                     ParticleList(L.VERTDATA_NONE, pts, count=2),
                     ParticleList(L.VERTDATA_FLOAT_XYZ, pts, count=2, global_radius=0.05)])
        assert r.info()[0] == 2


@pytest.mark.parametrize("bounces", [1, 3])
def test_bounce_mode_bit_exact(gpu_renderer, bounces):
    case(f"diffuse bounces {bounces}")
    xyzr, rgba = scenes.random_spheres(3000, seed=21)
    cam = small_camera(64, 48)
    p = params(64, 48, spp=2, mode=L.MODE_BOUNCE, max_bounces=bounces)
    g, o = render_both(gpu_renderer, xyzr, rgba, cam, p, 30)
    assert_frames_equal(g, o, what=f"bounces={bounces}")


def test_device_framebuffer_is_the_downloaded_frame(gpu_renderer):
    """The zero-copy device views (CallRender3DCUDA analogue) hold exactly what download copies."""
    import ctypes as C
    case("device framebuffer")
    xyzr, rgba = scenes.random_spheres(3000, seed=8)
    cam = small_camera(70, 50)
    p = params(70, 50, spp=2, mode=L.MODE_AO, ao_distance=0.3, flags=L.FLAG_AUX)
    gpu_renderer.set_spheres(xyzr, rgba)
    gpu_renderer.build()
    fr = gpu_renderer.render(cam, p)
    aux = gpu_renderer.download(70, 50, outputs=(), aux=True).aux
    d = gpu_renderer.device_framebuffer()
    assert (d.width, d.height) == (70, 50) and d.stream and d.aux
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    n = 70 * 50

    def get(ptr, dtype, count):
        out = np.empty(count, dtype)
        assert hip.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0  # hipMemcpyDeviceToHost
        return out

    np.testing.assert_array_equal(get(d.hit_id, np.int32, n).reshape(50, 70), fr.hit_id)
    np.testing.assert_array_equal(get(d.rgba8, np.uint32, n).reshape(50, 70), fr.rgba8)
    np.testing.assert_array_equal(get(d.color, np.float32, 4 * n).reshape(50, 70, 4).view(np.uint32),
                                  fr.color.view(np.uint32))
    np.testing.assert_array_equal(get(d.depth, np.float32, n).reshape(50, 70).view(np.uint32),
                                  fr.depth.view(np.uint32))
    np.testing.assert_array_equal(get(d.aux, np.uint16, 8 * n).reshape(50, 70, 8), aux.view(np.uint16))


def test_allocation_failure_is_reported():
    """An allocation failure (injected from the host with the fail_alloc hook) returns E_OOM with
    the buffer named, leaves the context usable, and the next attempt succeeds."""
    from megamol_amd import SphereRenderer
    case("injected allocation failure")
    xyzr, rgba = scenes.random_spheres(5000, seed=3)
    cam = small_camera(32, 24)
    p = params(32, 24, spp=1, mode=L.MODE_AO, ao_distance=0.2)
    with SphereRenderer(0) as r:
        for k in (1, 2, 3):  # the first, second or third allocation of the upload fails
            r.set_option("fail_alloc", k)
            with pytest.raises(L.MMHRTError) as e:
                r.set_spheres(xyzr, rgba)
                r.build(L.BUILD_REBUILD, 30)
                r.render(cam, p)
            assert e.value.code == L.E_OOM and "injected" in str(e.value)
            assert r.get_option("fail_alloc") == 0
        r.set_spheres(xyzr, rgba)
        r.set_option("fail_alloc", 1)  # the build's first allocation
        with pytest.raises(L.MMHRTError) as e:
            r.build(L.BUILD_REBUILD, 30)
        assert e.value.code == L.E_OOM
        r.build(L.BUILD_REBUILD, 30)
        g = r.render(cam, p)
        o = O.render_brute(xyzr, rgba, cam, oracle_params(p))
        assert_frames_equal(g, o, what="after injected failures")
