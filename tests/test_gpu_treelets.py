"""GPU parity of the treelet P-k-d trees (option accel 1), through the C ABI.

  * the device build equals the oracle's structure array for array (spheres, particle ids and
    split axes in treelet heap order, treelet boxes and radii, the top tree of boxes) for treelet
    sizes 256, 512 and 1024, sphere counts from 1 to past a whole number of treelets, 30- and
    63-bit keys and coincident spheres;
  * frames traced through it equal the oracle bit for bit (Lambert, AO with the feature buffer,
    diffuse bounces; the lane-per-pixel kernel at 1 and 3 spp, the sample-lanes kernel at 4 and
    64 spp) and equal the LBVH frames of the same scene;
  * a batched launch over timesteps equals its single renders.
"""
import numpy as np
import pytest

import oracle as O
from helpers import assert_frames_equal, case, oracle_params, small_camera
from megamol_amd import _lib as L, params, scenes

pytestmark = pytest.mark.gpu


@pytest.fixture
def tl_renderer(gpu_renderer):
    gpu_renderer.set_option("accel", 1)
    yield gpu_renderer
    gpu_renderer.set_option("accel", 0)
    gpu_renderer.set_option("treelet", 1024)


def structure_equal(g, k):
    o = k.arrays()
    assert (g["M"], g["T"], g["L"]) == (k.k.M, k.k.T, k.k.top_leaves)
    for key in ("sph", "tl_box", "tl_rmax", "top_box"):
        np.testing.assert_array_equal(g[key].view(np.uint32), o[key].view(np.uint32), err_msg=key)
    for key in ("id", "dim", "top_child"):
        np.testing.assert_array_equal(g[key], o[key], err_msg=key)


@pytest.mark.parametrize("M", [256, 512, 1024])
@pytest.mark.parametrize("n", [1, 300, 1024, 1025, 100_003])
def test_treelet_build_equals_oracle(tl_renderer, M, n):
    r = tl_renderer
    case(f"treelets: build M={M} n={n}")
    xyzr, rgba = scenes.random_spheres(n, seed=n % 97)
    r.set_option("treelet", M)
    r.set_spheres(xyzr, rgba)
    r.build(L.BUILD_REBUILD, 30)
    structure_equal(r.get_treelets(), O.Pkd(O.Bvh(xyzr, 30), M))


def test_treelet_build_coincident_63bit(tl_renderer):
    r = tl_renderer
    case("treelets: coincident spheres, 63-bit keys")
    rng = np.random.default_rng(4)
    xyzr = np.repeat(rng.random((20000, 4)).astype(np.float32), 3, axis=0)
    xyzr[:, 3] = 0.002 + 0.004 * xyzr[:, 3]
    xyzr[::5, 1] = xyzr[3, 1]
    r.set_option("treelet", 512)
    r.set_spheres(xyzr, None)
    r.build(L.BUILD_REBUILD, 63)
    structure_equal(r.get_treelets(), O.Pkd(O.Bvh(xyzr, 63), 512))


@pytest.mark.parametrize("spp", [1, 3, 4, 64])
@pytest.mark.parametrize("mode", [L.MODE_LAMBERT, L.MODE_AO])
def test_treelet_frames_equal_oracle(tl_renderer, spp, mode):
    r = tl_renderer
    case(f"treelets: frame spp={spp} mode={mode}")
    xyzr, rgba = scenes.random_spheres(20000, seed=21)
    xyzr[300:306] = xyzr[11]  # coincident spheres: equal t, lowest id wins
    w, h = (40, 30) if spp == 64 else (96, 72)
    cam = small_camera(w, h)
    p = params(w, h, spp=spp, mode=mode, ao_distance=0.2, background=(0.1, 0.2, 0.3, 1.0), flags=L.FLAG_AUX)
    r.set_option("treelet", 256)
    r.set_spheres(xyzr, rgba)
    r.build(L.BUILD_REBUILD, 30)
    g = r.render(cam, p)
    g.aux = r.download(w, h, outputs=(), aux=True).aux
    k = O.Pkd(O.Bvh(xyzr, 30), 256)
    o = k.render(rgba, cam, oracle_params(p), aux=True)
    assert (o["hit_id"] >= 0).mean() > 0.4
    assert_frames_equal(g, o, aux=True, what=f"treelets spp={spp} mode={mode}")
    # the LBVH frame of the same scene is the same frame
    r.set_option("accel", 0)
    r.build(L.BUILD_REBUILD, 30)
    b = r.render(cam, p)
    np.testing.assert_array_equal(b.hit_id, g.hit_id)
    np.testing.assert_array_equal(b.color.view(np.uint32), g.color.view(np.uint32))


def test_treelet_bounce_mode(tl_renderer):
    r = tl_renderer
    case("treelets: diffuse bounces")
    xyzr, rgba = scenes.random_spheres(8000, seed=5)
    cam = small_camera(64, 48)
    p = params(64, 48, spp=2, mode=L.MODE_BOUNCE)
    p.max_bounces = 3
    r.set_option("treelet", 512)
    r.set_spheres(xyzr, rgba)
    r.build(L.BUILD_REBUILD, 30)
    g = r.render(cam, p)
    o = O.render_brute(xyzr, rgba, cam, oracle_params(p))
    assert_frames_equal(g, o, what="treelets bounce")


def test_treelet_clusters_63bit(tl_renderer):
    r = tl_renderer
    case("treelets: clustered scene, 63-bit keys, 1-spp AO")
    xyzr = scenes.clusters(400_000, n_clusters=50, box=10.0, mu_sigma=-1.0, s_sigma=0.3, mu_r=-4.5, s_r=0.3)
    rgba = np.random.default_rng(2).random((len(xyzr), 4)).astype(np.float32)
    rgba[:, 3] = 1.0
    cam = small_camera(160, 120, pos=(5.0, 5.0, 22.0), target=(5.0, 5.0, 5.0))
    p = params(160, 120, spp=1, mode=L.MODE_AO, ao_distance=1.0)
    r.set_option("treelet", 1024)
    r.set_spheres(xyzr, rgba)
    r.build(L.BUILD_REBUILD, 63)
    g = r.render(cam, p)
    o = O.Bvh(xyzr, 63).render(rgba, cam, oracle_params(p))
    assert (o["hit_id"] >= 0).mean() > 0.05
    assert_frames_equal(g, o, what="treelets clusters")


def test_treelet_batch_equals_single(tl_renderer):
    r = tl_renderer
    case("treelets: batched timesteps")
    x0, rgba = scenes.random_spheres(30000, seed=8)
    xs = np.stack([x0] + [scenes.time_series_step(x0, k, sigma=0.01) for k in range(2)])
    cam = small_camera(64, 48)
    jobs = [(t, cam, params(64, 48, spp=1, mode=L.MODE_AO, ao_distance=0.2)) for t in range(3)]
    r.set_option("treelet", 512)
    r.set_timesteps(xs, rgba)
    out, _ = r.render_batch(jobs, refit_period=2, use_graph=True)
    for t in range(3):
        r.set_spheres(xs[t], rgba)
        r.build(L.BUILD_REBUILD, 30)
        f = r.render(cam, jobs[t][2], outputs=("rgba8", "depth"))
        np.testing.assert_array_equal(out[t][0], f.rgba8, err_msg=f"timestep {t}")
        np.testing.assert_array_equal(out[t][1].view(np.uint32), f.depth.view(np.uint32))
