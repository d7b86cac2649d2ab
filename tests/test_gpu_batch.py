"""GPU: batched time series (timesteps x cameras in one hipGraph) equals individual renders."""
import numpy as np
import pytest

from helpers import case
from megamol_amd import _lib as L, look_at, params, scenes

pytestmark = pytest.mark.gpu


def series(n=20000, steps=6):
    x0 = scenes.time_series_initial(n, box=20.0)
    xs = [x0]
    for k in range(steps - 1):
        xs.append(scenes.time_series_step(xs[-1], k, sigma=0.2))
    return np.stack(xs)


def jobs_for(steps):
    g = np.random.default_rng(0)
    jobs = []
    for t in range(steps):
        for c in range(1 + t % 3):
            w, h = [(64, 48), (80, 60), (40, 72)][c]
            ang = g.uniform(0, 2 * np.pi)
            pos = (10 + 45 * np.cos(ang), 14, 10 + 45 * np.sin(ang))
            cam = look_at(pos, (10, 10, 10), (0, 1, 0), 40.0, w / h, 1.0, 200.0)
            p = params(w, h, spp=1 + c, mode=L.MODE_AO if t % 2 else L.MODE_LAMBERT, ao_distance=1.5,
                       background=(0.2, 0.1, 0.0, 1.0))
            if c == 2:
                p.window[:] = [4, 8, 32, 40]
            jobs.append((t, cam, p))
    return jobs


def test_batch_equals_single_renders(gpu_renderer):
    r = gpu_renderer
    xs = series()
    rgba = np.random.default_rng(1).random((xs.shape[1], 4)).astype(np.float32)
    jobs = jobs_for(len(xs))
    # individual reference renders
    ref = []
    for t, cam, p in jobs:
        r.set_spheres(xs[t], rgba)
        r.build(L.BUILD_REBUILD, 30)
        f = r.render(cam, p, outputs=("rgba8", "depth"))
        ref.append((f.rgba8, f.depth))
    r.set_timesteps(xs, rgba)
    # batch_launch 1: every timestep's structure kept, one launch per mode over the whole series
    # (the jobs alternate Lambert and AO by timestep: two launches); 0: one launch per timestep
    for one in (1, 0):
        r.set_option("batch_launch", one)
        for use_graph in (False, True):
            for period in (1, 3, 100):
                case(f"batch launch={one} graph={use_graph} refit_period={period}")
                for rep in range(2):  # a graph launched again must give the same frames
                    out, ms = r.render_batch(jobs, refit_period=period, use_graph=use_graph)
                    assert ms > 0
                    for k, ((a, d), (ra, rd)) in enumerate(zip(out, ref)):
                        np.testing.assert_array_equal(a, ra, err_msg=f"job {k} rgba launch={one} graph={use_graph} "
                                                                     f"period={period}")
                        np.testing.assert_array_equal(d.view(np.uint32), rd.view(np.uint32), err_msg=f"job {k} depth")
    r.set_option("batch_launch", 1)


def test_batch_errors(gpu_renderer):
    from megamol_amd import SphereRenderer
    with SphereRenderer(0) as r:
        cam = look_at((0, 0, 5), (0, 0, 0))
        with pytest.raises(L.MMHRTError) as e:
            r.render_batch([(0, cam, params(8, 8))])
        assert e.value.code == L.E_STATE
        r.set_timesteps(series(500, 2))
        with pytest.raises(L.MMHRTError) as e:
            r.render_batch([(5, cam, params(8, 8))])
        assert e.value.code == L.E_INVALID
        with pytest.raises(L.MMHRTError) as e:
            r.render_batch([(0, cam, params(8, 8, mode=L.MODE_BOUNCE))])
        assert e.value.code == L.E_INVALID


def test_batch_keeps_its_own_colours_and_the_context_state(gpu_renderer):
    """set_timesteps, then set_spheres with another set, then render_batch: the batch renders the
    timesteps with their own colours; afterwards the context renders its own spheres again (after
    the build that the batch's last-timestep tree requires) and equals the oracle."""
    import oracle as O
    from helpers import assert_frames_equal, oracle_params
    r = gpu_renderer
    xs = series(5000, 3)
    rgba_ts = np.random.default_rng(2).random((xs.shape[1], 4)).astype(np.float32)
    other, rgba_o = scenes.random_spheres(7000, seed=31, lo=5.0, hi=15.0, rmin=0.1, rmax=0.4)
    jobs = jobs_for(3)
    ref = []
    for t, cam, p in jobs:
        r.set_spheres(xs[t], rgba_ts)
        r.build(L.BUILD_REBUILD, 30)
        ref.append(r.render(cam, p, outputs=("rgba8", "depth")))
    case("batch after set_spheres of another set")
    r.set_timesteps(xs, rgba_ts)
    r.set_spheres(other, rgba_o)
    out, _ = r.render_batch(jobs, refit_period=2, use_graph=True)
    for k, ((a, d), f) in enumerate(zip(out, ref)):
        np.testing.assert_array_equal(a, f.rgba8, err_msg=f"job {k}")
        np.testing.assert_array_equal(d.view(np.uint32), f.depth.view(np.uint32), err_msg=f"job {k}")
    assert r.info()[0] == len(other)
    cam = look_at((10, 10, 40), (10, 10, 10), (0, 1, 0), 40.0, 1.0, 1.0, 200.0)
    p = params(40, 40, spp=2, mode=L.MODE_AO, ao_distance=2.0)
    with pytest.raises(L.MMHRTError) as e:
        r.render(cam, p)
    assert e.value.code == L.E_STATE
    r.build(L.BUILD_REBUILD, 30)
    g = r.render(cam, p)
    assert_frames_equal(g, O.render_brute(other, rgba_o, cam, oracle_params(p)), what="context spheres after batch")
