"""GPU parity at the full sizes of the sphere path's workloads.

  uniform cube   1M spheres, r 0.002, 1024x1024, Lambert 1 spp: the whole frame bit for bit
                 against the oracle's LBVH renderer, every pixel against the FP64 oracle
  FCC crystal    49,948,672 spheres, 1920x1080, AO distance 4.0: sampled rows of the 1-spp
                 feature frame and of the 64-spp frame (trees with and without leaf pairs) bit
                 for bit against the oracle; FP64 tiles
  time series    64 timesteps x 4M spheres, 1-6 cameras of 640x480 ... 1920x1080 per step in
                 one graph launch with refits: jobs of the first, middle and last timestep and
                 one of every resolution equal their single renders; one job equals the oracle
  clusters       500M spheres in 10k Gaussian clusters, 63-bit keys, 3840x2160: a 64x64 window
                 against the FP64 oracle; orbit batch equal to single renders

Each test runs the HIP library through the C ABI; the oracle is the checker only.
"""
import time

import numpy as np
import pytest

import oracle as O
from helpers import assert_frames_equal, case, fp64_check, oracle_params
from megamol_amd import ParticleList, _lib as L, params, scenes

pytestmark = pytest.mark.gpu


def tiles(x0, y0, x1, y1, s=16):
    return [(x, y, min(x + s, x1), min(y + s, y1)) for y in range(y0, y1, s) for x in range(x0, x1, s)]


def test_uniform_cube_1m_full_frame(gpu_renderer):
    r = gpu_renderer
    case("uniform cube 1M, 1024x1024, Lambert: scene")
    xyzr, rgba8 = scenes.uniform_cube(1_000_000)
    lists = [ParticleList(L.VERTDATA_FLOAT_XYZR, xyzr, L.COLDATA_UINT8_RGBA, rgba8, count=len(xyzr))]
    cam = scenes.uniform_cube_camera(1024, 1024)
    p = params(1024, 1024, spp=1, mode=L.MODE_LAMBERT)
    case("uniform cube 1M: GPU render")
    r.set_lists(lists)
    r.build(L.BUILD_REBUILD, 30)
    g = r.render(cam, p)
    case("uniform cube 1M: oracle LBVH render")
    xo, ro = O.convert_lists(lists)
    o = O.Bvh(xo, 30).render(ro, cam, oracle_params(p))
    assert (o["hit_id"] >= 0).mean() > 0.5
    assert_frames_equal(g, o, what="uniform cube 1M")
    case("uniform cube 1M: every pixel against FP64")
    m, n = fp64_check(xo, cam, 1024, 1024, g.hit_id, g.t, tiles(0, 0, 1024, 1024), what="uniform cube")
    assert m <= 1e-5 * n, (m, n)


@pytest.fixture(scope="module")
def fcc50m():
    case("FCC 49.9M: scene")
    xyzr, inten = scenes.fcc_crystal(232)
    tf = scenes.transfer_function_256()
    lists = [ParticleList(L.VERTDATA_FLOAT_XYZR, xyzr, L.COLDATA_FLOAT_I, inten, count=len(xyzr), col_min=0.0,
                          col_max=1.0)]
    cam = scenes.fcc_camera(232, width=1920, height=1080)
    case("FCC 49.9M: oracle colours and LBVH")
    xo, ro = O.convert_lists(lists, tf=tf)
    bvh = O.Bvh(xo, 30)
    return dict(lists=lists, tf=tf, cam=cam, xo=xo, ro=ro, bvh=bvh)


ROWS_1080 = [37, 181, 333, 470, 541, 666, 802, 1049]


def oracle_rows(bvh, rgba, cam, p, rows, aux=False):
    """Oracle frame with only `rows` rendered (the LBVH renderer's row range, one row at a time)."""
    out = None
    for y in rows:
        q = oracle_params(p)
        q.row_begin, q.row_end = y, y + 1
        o = bvh.render(rgba, cam, q, aux=aux)
        if out is None:
            out = o
        else:
            for k in ("rgba8", "depth", "color", "t", "hit_id") + (("aux",) if aux else ()):
                out[k][y] = o[k][y]
    return out


def test_fcc50m_denoiser_input_frame(gpu_renderer, fcc50m):
    """The denoised frame's trace (1-spp AO with the feature buffer, tree with leaf pairs): sampled
    rows bit for bit, FP64 tiles across the image."""
    r, S = gpu_renderer, fcc50m
    case("FCC 49.9M: upload, build with leaf pairs, 1-spp AO + features")
    r.set_transfer_function(S["tf"])
    r.set_lists(S["lists"])
    r.set_option("leaf_pairs", 1)
    r.build(L.BUILD_REBUILD, 30)
    p = params(1920, 1080, spp=1, mode=L.MODE_AO, ao_distance=4.0, flags=L.FLAG_AUX)
    g = r.render(S["cam"], p)
    g.aux = r.download(1920, 1080, outputs=(), aux=True).aux
    r.set_transfer_function(None)
    case("FCC 49.9M: oracle rows (1 spp)")
    o = oracle_rows(S["bvh"], S["ro"], S["cam"], p, ROWS_1080, aux=True)
    assert (g.hit_id >= 0).mean() > 0.3
    assert_frames_equal(g, o, aux=True, what="fcc 1spp rows", rows=ROWS_1080)
    case("FCC 49.9M: FP64 tiles")
    rects = [(x, y, x + 16, y + 16) for x, y in [(952, 532), (200, 100), (1700, 900), (480, 760), (1400, 300),
                                                 (96, 1000), (1880, 40), (1100, 640)]]
    m, n = fp64_check(S["xo"], S["cam"], 1920, 1080, g.hit_id, g.t, rects, what="fcc 1spp")
    case(f"FCC 49.9M: FP64 mismatches {m} of {n} sampled pixels")


@pytest.mark.parametrize("pairs", [0, 1])
def test_fcc50m_plain_64spp_frame(gpu_renderer, fcc50m, pairs):
    """The plain frame (64-spp AO, one AO ray per sample; the sample-lanes kernel with the full
    64-lane fold) on trees with and without leaf pairs: sampled rows bit for bit."""
    r, S = gpu_renderer, fcc50m
    case(f"FCC 49.9M: build (leaf pairs {pairs}), 64-spp AO")
    r.set_transfer_function(S["tf"])
    r.set_lists(S["lists"])
    r.set_option("leaf_pairs", pairs)
    try:
        r.build(L.BUILD_REBUILD, 30)
        p = params(1920, 1080, spp=64, mode=L.MODE_AO, ao_distance=4.0)
        g = r.render(S["cam"], p)
    finally:
        r.set_option("leaf_pairs", 1)
        r.set_transfer_function(None)
    rows = ROWS_1080
    case("FCC 49.9M: oracle rows (64 spp)")
    o = oracle_rows(S["bvh"], S["ro"], S["cam"], p, rows)
    assert_frames_equal(g, o, what=f"fcc 64spp rows pairs={pairs}", rows=rows)


def test_time_series_64x4m_batch(gpu_renderer):
    """64 timesteps x 4M spheres with their 1-6 cameras per step, one graph launch with refits
    (period 4): sampled jobs equal their single renders, one job equals the oracle."""
    r = gpu_renderer
    case("time series 64 x 4M: scene")
    t0 = time.time()
    xs = np.empty((64, 4_000_000, 4), np.float32)
    xs[0] = scenes.time_series_initial()
    for k in range(63):
        xs[k + 1] = scenes.time_series_step(xs[k], k)
    rgba = np.random.default_rng(7).random((4_000_000, 4)).astype(np.float32)
    rgba[:, 3] = 1.0
    jobs = []
    for t in range(64):
        for cam, w, h in scenes.time_series_cameras(t):
            jobs.append((t, cam, params(w, h, spp=1, mode=L.MODE_AO, ao_distance=4.0)))
    case(f"time series: {len(jobs)} jobs, generated in {time.time() - t0:.1f}s; batch launch")
    r.set_timesteps(xs, rgba)
    out, ms = r.render_batch(jobs, refit_period=4, use_graph=True)
    assert ms > 0
    picks = [k for k, (t, _, _) in enumerate(jobs) if t in (0, 32, 63)]
    seen = set()
    for k, (t, _, p) in enumerate(jobs):
        if (p.width, p.height) not in seen:
            seen.add((p.width, p.height))
            picks.append(k)
    assert len(seen) == len(scenes.RESOLUTIONS)
    for k in sorted(set(picks)):
        t, cam, p = jobs[k]
        case(f"time series: job {k} (timestep {t}, {p.width}x{p.height}) single render")
        r.set_spheres(xs[t], rgba)
        r.build(L.BUILD_REBUILD, 30)
        f = r.render(cam, p, outputs=("rgba8", "depth"))
        np.testing.assert_array_equal(out[k][0], f.rgba8, err_msg=f"job {k} rgba8")
        np.testing.assert_array_equal(out[k][1].view(np.uint32), f.depth.view(np.uint32), err_msg=f"job {k} depth")
    k = min(range(len(jobs)), key=lambda j: (jobs[j][2].width * jobs[j][2].height, -jobs[j][0]))
    t, cam, p = jobs[k]
    case(f"time series: job {k} (timestep {t}) against the oracle")
    o = O.Bvh(xs[t], 30).render(rgba, cam, oracle_params(p))
    np.testing.assert_array_equal(out[k][0], o["rgba8"])
    np.testing.assert_array_equal(out[k][1].view(np.uint32), o["depth"].view(np.uint32))


def test_clusters_500m(gpu_renderer):
    """500M clustered spheres with 63-bit keys at 3840x2160: the primary hits of a 64x64 window
    against the FP64 brute-force oracle over all spheres; two orbit cameras in one batched graph
    launch equal their single renders."""
    r = gpu_renderer
    case("clusters 500M: scene")
    t0 = time.time()
    xyzr = scenes.clusters(500_000_000)
    case(f"clusters 500M: generated in {time.time() - t0:.1f}s; upload + 63-bit build")
    W, H = 3840, 2160
    cam = scenes.clusters_camera(width=W, height=H)
    r.set_spheres(xyzr, None, (200, 200, 220, 255))
    r.build(L.BUILD_REBUILD, 63)
    x0, y0 = 1888, 1048
    p = params(W, H, spp=1, mode=L.MODE_LAMBERT)
    p.window[:] = [x0, y0, 64, 64]
    g = r.render(cam, p, outputs=("hit_id", "t"))
    assert (g.hit_id >= 0).mean() > 0.2, (g.hit_id >= 0).mean()
    case("clusters 500M: FP64 over all spheres for the window")
    m, n = fp64_check(xyzr, cam, W, H, g.hit_id, g.t, tiles(x0, y0, x0 + 64, y0 + 64), x_off=x0, y_off=y0,
                      what="clusters window")
    case(f"clusters 500M: FP64 mismatches {m} of {n}")
    case("clusters 500M: orbit batch of 2 cameras")
    cams = scenes.orbit_cameras(120, width=W, height=H)[::60]
    po = params(W, H, spp=1, mode=L.MODE_AO, ao_distance=4.0)
    r.set_timesteps(xyzr[None], None, (200, 200, 220, 255))
    out, _ = r.render_batch([(0, c, po) for c in cams], refit_period=1, bits=63, use_graph=True)
    r.build(L.BUILD_REBUILD, 63)
    for i, c in enumerate(cams):
        case(f"clusters 500M: orbit camera {i} single render")
        f = r.render(c, po, outputs=("rgba8", "depth"))
        np.testing.assert_array_equal(out[i][0], f.rgba8, err_msg=f"orbit {i}")
        np.testing.assert_array_equal(out[i][1].view(np.uint32), f.depth.view(np.uint32), err_msg=f"orbit {i}")
