"""CPU tests: pin the oracle against published vectors, closed forms and independent methods.

These tests never touch the GPU. They make the oracle trustworthy before it judges the HIP path.
"""
import math
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from helpers import small_camera
from megamol_amd import look_at, mmpld, scenes

GOLD = Path(__file__).resolve().parent / "golden"


# ------------------------------------------------------------------ RNG
def test_lcg_numerical_recipes_vector():
    # "An even quicker generator", Numerical Recipes in C 2nd ed. sec. 7.1: idum from 0
    want = [0x3C6EF35F, 0x47502932, 0xD1CCF6E9, 0xAAF95334, 0x6252E503, 0x9F2EC686, 0x57FE6C2D, 0xA3D95FA8,
            0x81FDBEE7, 0x94F0AF1A, 0xCBF633B1]
    assert O.lcg_raw(0, len(want)) == want
    # rnd() = low 24 bits / 2^24 (random.h:60-62)
    r = O.lcg_stream(0, 4)
    assert r == [float(np.float32((w & 0xFFFFFF) / 16777216.0)) for w in want[:4]]


def test_tea_published_vector():
    # TEA (Wheeler & Needham 1994), 32 cycles, zero key, zero plaintext -> 41EA3A0A 94BAA940
    assert O.tea_encrypt(0, 0, [0, 0, 0, 0], 32) == (0x41EA3A0A, 0x94BAA940)


def test_tea16_is_tea_with_reference_key():
    key = [0xA341316C, 0xC8013EA4, 0xAD90777D, 0x7E95761E]  # random.h:39-40
    for a, b in [(0, 0), (1, 0), (12345, 7), (2 ** 32 - 1, 99)]:
        assert O.tea16(a, b) == O.tea_encrypt(a, b, key, 16)[0]
    # decorrelation: neighbouring pixel seeds differ in about half their bits
    bits = [bin(O.tea16(i, 0) ^ O.tea16(i + 1, 0)).count("1") for i in range(256)]
    assert 14 < np.mean(bits) < 18


# --------------------------------------------------------------- camera
def test_frame_state_closed_form():
    cam = look_at((0, 0, 5), (0, 0, 0), (0, 1, 0), fovy_deg=50.0, aspect=1.5, near=0.5, far=20.0)
    fs = O.frame_state(cam)
    assert fs.nearp == 100.0
    assert math.isclose(fs.th, math.tan(math.radians(25.0)) * 100.0, rel_tol=1e-6)
    assert math.isclose(fs.rw, fs.th * 1.5, rel_tol=1e-6)
    assert np.allclose(list(fs.right), [1, 0, 0], atol=1e-7)  # cross(dir, up) with dir=-z, up=y
    n, f = 0.5, 20.0
    assert math.isclose(fs.depth_a, -(f + n) / (f - n), rel_tol=1e-6)
    assert math.isclose(fs.depth_c, -2 * f * n / (f - n), rel_tol=1e-6)


def test_primary_ray_geometry():
    W, H = 640, 480
    cam = look_at((1, 2, 3), (1, 2, 0), (0, 1, 0), fovy_deg=60.0, aspect=W / H)
    fs = O.frame_state(cam)
    import ctypes as C
    lib = O.load()
    o = np.zeros(3, np.float32)
    d = np.zeros(3, np.float32)
    # centre of the image plane: u = v = 0 -> the ray is the view direction
    lib.orc_primary_ray(C.byref(fs), W, H, W // 2, H // 2, 0.0, 0.0, o.ctypes.data, d.ctypes.data)
    assert np.allclose(d, [0, 0, -1], atol=1e-6) and np.allclose(o, [1, 2, 3])
    # bottom-left corner (y = 0 is the bottom row): vertical angle = fovy / 2, looking down
    lib.orc_primary_ray(C.byref(fs), W, H, 0, 0, 0.0, 0.0, o.ctypes.data, d.ctypes.data)
    assert d[1] < 0 and d[0] < 0
    vert = math.degrees(math.atan2(-d[1], -d[2]))
    horiz = math.degrees(math.atan2(-d[0], -d[2]))
    assert math.isclose(vert, 30.0, abs_tol=1e-4)
    assert math.isclose(math.tan(math.radians(horiz)), math.tan(math.radians(30.0)) * W / H, rel_tol=1e-5)
    assert math.isclose(float(np.linalg.norm(d)), 1.0, abs_tol=1e-6)


def test_pseudo_depth_maps_near_to_0_and_far_to_1():
    # raygen.cu:262-264: 0.5 * ((C / t) - A) + 0.5 is the GL depth of distance t
    n, f = 2.0, 11.0
    for t_hit, want in [(n, 0.0), (f, 1.0)]:
        # a sphere whose front is at distance t_hit along the optical axis
        xyzr = np.array([[0, 0, -(t_hit + 1.0), 1.0]], np.float32)
        cam = look_at((0, 0, 0), (0, 0, -1), (0, 1, 0), fovy_deg=1e-3, aspect=1.0, near=n, far=f)
        o = O.render_brute(xyzr, np.ones((1, 4), np.float32), cam, O.make_params(1, 1))
        assert o["hit_id"][0, 0] == 0
        assert math.isclose(float(o["t"][0, 0]), t_hit, rel_tol=1e-6)
        assert math.isclose(float(o["depth"][0, 0]), want, abs_tol=1e-5)


# --------------------------------------------------------- intersection
@pytest.mark.parametrize("h", [0.0, 0.3, 0.6, 0.9, 0.999])
def test_intersection_closed_form(h):
    # ray parallel to -z at lateral offset h from a unit sphere at the origin: t = 5 - sqrt(1 - h^2)
    t = O.intersect_f32([h, 0, 5], [0, 0, -1], [0, 0, 0, 1])
    assert t is not None and math.isclose(t, 5 - math.sqrt(1 - h * h), rel_tol=2e-6)
    t64, disc = O.intersect_f64([h, 0, 5], [0, 0, -1], [0, 0, 0, 1])
    assert math.isclose(t64, 5 - math.sqrt(1 - h * h), rel_tol=1e-12)
    assert math.isclose(disc, 1 - h * h, abs_tol=1e-12)


def test_intersection_special_cases():
    assert O.intersect_f32([1, 0, 5], [0, 0, -1], [0, 0, 0, 1]) == 5.0  # tangent: delta == 0 is a hit
    assert O.intersect_f32([1.0001, 0, 5], [0, 0, -1], [0, 0, 0, 1]) is None  # just outside
    assert O.intersect_f32([0, 0, 0.5], [0, 0, -1], [0, 0, 0, 1]) is None  # origin inside (sphere.cuh:34-36)
    assert O.intersect_f32([0, 0, -5], [0, 0, -1], [0, 0, 0, 1]) is None  # sphere behind
    assert O.intersect_f32([0, 0, 5], [0, 0, -1], [0, 0, 0, 1], tmin=4.0) is None  # t == tmin excluded
    assert O.intersect_f32([0, 0, 5], [0, 0, -1], [0, 0, 0, 1], tmax=4.0) is None  # t == tmax excluded
    assert O.intersect_f32([0, 0, 5], [0, 0, -1], [0, 0, 0, 1], tmax=4.0001) == 4.0


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_fp32_matches_fp64_except_grazing_and_ties(seed):
    xyzr, rgba = scenes.random_spheres(1500, seed=seed, rmin=0.01, rmax=0.06)
    cam = small_camera(96, 96)
    p = O.make_params(96, 96)
    f32 = O.render_brute(xyzr, rgba, cam, p)
    ids, t64, disc, t2 = O.primary_f64(xyzr, cam, p)
    mism = ids != f32["hit_id"]
    for y, x in zip(*np.nonzero(mism)):
        grazing = abs(disc[y, x]) < 1e-4 or (f32["hit_id"][y, x] < 0 or ids[y, x] < 0)
        tie = np.isfinite(t2[y, x]) and abs(t2[y, x] - t64[y, x]) <= 1e-5 * t64[y, x]
        assert grazing or tie, (y, x)
    assert mism.mean() <= 1e-3
    hit = (ids >= 0) & ~mism
    rel = np.abs(f32["t"][hit].astype(np.float64) - t64[hit]) / t64[hit]
    assert rel.max() <= 1e-5


# ----------------------------------------------------------------- LBVH
def radix_tree_topdown(keys: np.ndarray):
    """Independent construction: the binary radix tree over sorted (key, index) pairs, built
    top-down by splitting each range at its highest differing bit (Karras 2012 sec. 3 defines the
    LBVH as exactly this tree); internal node for range [i, j] is numbered by its split end
    adjacent to the range boundary, as in the paper (node index = i if it is the left end of its
    parent's split, else j)."""
    n = len(keys)
    ak = [(int(k) << 32) | i for i, k in enumerate(keys)]  # augmented keys, unique
    child = {}

    def split(first, last):
        a, b = ak[first], ak[last]
        common = 96 - (a ^ b).bit_length()
        s = first
        step = last - first
        while True:
            step = (step + 1) >> 1
            new = s + step
            if new < last:
                pre = 96 - (a ^ ak[new]).bit_length()
                if pre > common:
                    s = new
            if step <= 1:
                break
        return s

    def ref(first, last, is_left_of_parent):
        if first == last:
            return ~first
        return first if not is_left_of_parent else last

    stack = [(0, n - 1, 0)]
    while stack:
        first, last, idx = stack.pop()
        g = split(first, last)
        left = ref(first, g, True)
        right = ref(g + 1, last, False)
        child[idx] = (left, right)
        if first != g:
            stack.append((first, g, left))
        if g + 1 != last:
            stack.append((g + 1, last, right))
    return np.array([child[i] for i in range(n - 1)], np.int32)


@pytest.mark.parametrize("n,dups,bits", [(2, False, 30), (7, False, 30), (300, False, 30), (300, True, 30),
                                         (1000, True, 63), (513, False, 63)])
def test_lbvh_topology_matches_radix_tree(n, dups, bits):
    g = np.random.default_rng(n)
    xyzr = np.zeros((n, 4), np.float32)
    xyzr[:, :3] = g.random((n, 3)).astype(np.float32)
    if dups:  # many identical centres -> equal Morton keys -> index tie-break
        xyzr[n // 2:, :3] = xyzr[: n - n // 2, :3]
        xyzr[::7, :3] = xyzr[0, :3]
    xyzr[:, 3] = 0.01
    b = O.Bvh(xyzr, bits)
    box, child, ids, keys = b.arrays()
    assert np.all(np.diff(keys.astype(np.float64)) >= 0)
    same = keys[1:] == keys[:-1]
    assert np.all(ids[1:][same] > ids[:-1][same]), "equal keys keep the original order (stable)"
    assert sorted(ids.tolist()) == list(range(n))
    np.testing.assert_array_equal(child, radix_tree_topdown(keys))


def test_lbvh_boxes_contain_their_subtrees():
    xyzr, _ = scenes.random_spheres(2000, seed=9)
    b = O.Bvh(xyzr, 30)
    box, child, ids, keys = b.arrays()
    leaf = xyzr[ids]
    lo = np.full((len(child), 3), np.inf, np.float32)
    hi = np.full((len(child), 3), -np.inf, np.float32)

    def sub(r):
        if r < 0:
            s = leaf[~r]
            return s[:3] - s[3], s[:3] + s[3]
        return lo[r], hi[r]

    order = np.argsort([-1])  # placeholder to keep numpy import used
    # post-order evaluation
    post, stack = [], [0]
    while stack:
        i = stack.pop()
        post.append(i)
        for r in child[i]:
            if r >= 0:
                stack.append(r)
    for i in reversed(post):
        for c in range(2):
            l, h = sub(child[i][c])
            np.testing.assert_array_equal(box[i, 6 * c:6 * c + 3], l)
            np.testing.assert_array_equal(box[i, 6 * c + 3:6 * c + 6], h)
        lo[i] = np.minimum(box[i, 0:3], box[i, 6:9])
        hi[i] = np.maximum(box[i, 3:6], box[i, 9:12])
    del order


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("bits", [30, 63])
def test_lbvh_traversal_equals_brute_force(mode, bits):
    xyzr, rgba = scenes.random_spheres(2500, seed=11)
    xyzr[100:110] = xyzr[0]  # exact duplicates: ties resolve to the lowest id
    cam = small_camera(80, 60)
    p = O.make_params(80, 60, spp=2, mode=mode, ao_distance=0.25, want_aux=1)
    a = O.render_brute(xyzr, rgba, cam, p, aux=True)
    b = O.Bvh(xyzr, bits).render(rgba, cam, p, threads=4, aux=True)
    for k in ("hit_id", "t", "depth", "color", "rgba8"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    np.testing.assert_array_equal(a["aux"].view(np.uint16), b["aux"].view(np.uint16))
    assert not np.isin(np.arange(101, 110), a["hit_id"]).any(), "duplicates of sphere 0 never win"


# ------------------------------------------------------------------- AO
def test_ao_direction_is_cosine_weighted():
    n = np.array([0.3, -0.5, 0.8], np.float32)
    n /= np.linalg.norm(n)
    cos = np.array([float(O.ao_direction(n, O.tea16(i, 77)) @ n) for i in range(20000)])
    assert cos.min() >= -1e-6
    assert abs(cos.mean() - 2 / 3) < 0.01  # E[cos] of a cosine-weighted hemisphere
    # cos^2 is uniform on [0,1] for a cosine-weighted hemisphere
    hist, _ = np.histogram(cos ** 2, bins=10, range=(0, 1))
    assert np.all(np.abs(hist / len(cos) - 0.1) < 0.01)


def _concentric(u1, u2):
    """Shirley-Chiu concentric map in float64: disk radius and azimuth of a unit-square point."""
    a, b = 2 * u1 - 1, 2 * u2 - 1
    if abs(a) > abs(b):
        return abs(a), (np.pi / 4) * (b / a) + (np.pi if a < 0 else 0.0)
    return abs(b), np.pi / 2 - (np.pi / 4) * (a / b) + (np.pi if b < 0 else 0.0)


@pytest.mark.parametrize("normal", [(0.3, -0.5, 0.8), (0, 0, 1), (0, 0, -1), (1, 0, 0), (-0.6, 0.0, -0.8)])
def test_ao_direction_closed_forms(normal):
    """Per sample, with the jitter taken from the LCG stream: the polar angle satisfies
    cos = sqrt(1 - r^2) (Malley over the concentric map, r = max(|a|, |b|)) and the azimuth
    difference of two samples equals that of their concentric-map angles, whatever the tangent
    basis; stratum s covers [s % g, s % g + 1) x [s / g, s / g + 1) / g."""
    n = np.array(normal, np.float32)
    n /= np.linalg.norm(n)
    spp, g = 18, 4  # 16 stratified samples and two unstratified ones
    # any unit vector in the tangent plane serves as the azimuth origin
    e = np.cross(n.astype(np.float64), [1.0, 0, 0] if abs(n[0]) < 0.9 else [0, 1.0, 0])
    e /= np.linalg.norm(e)
    f = np.cross(n.astype(np.float64), e)
    az, phi = [], []
    for seed in (3, 99):
        for s in range(spp):
            d = O.ao_direction(n, seed + 1000 * s, s, spp).astype(np.float64)
            r1, r2 = O.lcg_stream(seed + 1000 * s, 2)
            u1, u2 = (((s % g) + r1) / g, ((s // g) + r2) / g) if s < g * g else (r1, r2)
            r, ph = _concentric(u1, u2)
            assert abs(np.linalg.norm(d) - 1) < 1e-6
            assert abs(d @ n - np.sqrt(max(0.0, 1 - r * r))) < 2e-6, (s, d @ n, r)
            if r > 1e-2:  # the azimuth of a direction next to the pole is ill-conditioned
                az.append(np.arctan2(d @ f, d @ e))
                phi.append(ph)
    az, phi = np.array(az), np.array(phi)
    assert len(az) > 30
    diff = np.angle(np.exp(1j * ((az - az[0]) - (phi - phi[0]))))
    assert np.abs(diff).max() < 1e-4, np.abs(diff).max()


def test_ao_strata_reduce_variance():
    """Stratified cosine-weighted samples: E[cos] = 2/3 per pixel estimate with far lower spread
    than independent samples (the point of the strata)."""
    n = np.array([0.0, 0.6, 0.8], np.float32)
    spp = 64
    strat, indep = [], []
    for pix in range(100):
        c1 = [O.ao_direction(n, O.tea16(pix, s), s, spp) @ n for s in range(spp)]
        c2 = [O.ao_direction(n, O.tea16(pix, s), 0, 1) @ n for s in range(spp)]
        strat.append(np.mean(c1))
        indep.append(np.mean(c2))
    # per-pixel spread is about 0.013 (stratified) against 0.029 (independent): 4 sigma of the mean
    assert abs(np.mean(strat) - 2 / 3) < 5e-3
    assert np.std(strat) < 0.6 * np.std(indep), (np.std(strat), np.std(indep))


def test_ao_isolated_sphere_is_unoccluded():
    xyzr = np.array([[0.5, 0.5, 0.5, 0.3]], np.float32)
    rgba = np.array([[0.2, 0.4, 0.6, 1.0]], np.float32)
    cam = small_camera(32, 32)
    for frame in range(4):  # 1 spp: a hit pixel's only sample is the one that hit
        o = O.render_brute(xyzr, rgba, cam, O.make_params(32, 32, spp=1, mode=1, ao_distance=10.0, frame_idx=frame,
                                                          want_aux=1), aux=True)
        hit = o["hit_id"] >= 0
        assert hit.sum() > 50
        assert np.all(o["aux"][hit][:, 7] == 1.0)
        np.testing.assert_array_equal(o["color"][hit][:, :3], np.broadcast_to(rgba[0, :3], (hit.sum(), 3)))


def test_ao_enclosed_point_is_occluded():
    # a small sphere in the middle of a closed shell of large spheres; the camera sits inside
    # the shell, so every AO ray from the small sphere hits the shell within ao_distance
    shell = []
    for d in [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]:
        shell.append([3 * d[0], 3 * d[1], 3 * d[2], 2.6])
    xyzr = np.array([[0, 0, 0, 0.1]] + shell, np.float32)
    rgba = np.ones((len(xyzr), 4), np.float32)
    cam = look_at((0, 0, 0.25), (0, 0, 0), (0, 1, 0), 20.0, 1.0, 0.01, 10)
    o = O.render_brute(xyzr, rgba, cam, O.make_params(16, 16, spp=8, mode=1, ao_distance=100.0, want_aux=1),
                       aux=True)
    small = o["hit_id"] == 0
    assert small.sum() > 10
    assert np.all(o["aux"][small][:, 7] == 0.0)


# ------------------------------------------------------------ ingest / TF
def test_tf_closed_form_linear_ramp():
    n = 256
    tf = np.repeat(np.linspace(0, 1, n, dtype=np.float32)[:, None], 4, 1)
    out = np.zeros(4, np.float32)
    import ctypes as C
    lib = O.load()
    for v in [0.0, 0.1, 0.37, 0.5, 0.9]:
        lib.orc_tf_sample(tf.ctypes.data, n, C.c_float(0.0), C.c_float(1.0), C.c_float(v), out.ctypes.data)
        assert math.isclose(float(out[0]), v * n / (n - 1), abs_tol=2e-6)
    lib.orc_tf_sample(tf.ctypes.data, n, C.c_float(0.0), C.c_float(1.0), C.c_float(1.0), out.ctypes.data)
    assert out[0] == 1.0  # at the top of the range: last entry (AddParticleColors.cpp:28-29)
    lib.orc_tf_sample(tf.ctypes.data, n, C.c_float(0.0), C.c_float(1.0), C.c_float(-1.0), out.ctypes.data)
    assert np.all(out == 0.0)  # base < 0 -> zero (AddParticleColors.cpp:25-26)
    lib.orc_tf_sample(tf.ctypes.data, n, C.c_float(0.0), C.c_float(1.0), C.c_float(np.nan), out.ctypes.data)
    assert np.all(out == 0.0)


def test_mmpld_fixtures_ingest():
    e = np.load(GOLD / "mmpld" / "expected.npz")
    files = sorted((GOLD / "mmpld").glob("*.mmpld"))
    assert len(files) == 17
    for f in files:
        hdr, frames = mmpld.read(str(f))
        assert hdr.version == 102 and hdr.frame_count == 1 and hdr.bbox == (-2, -2, -2, 2, 2, 2)
        assert len(frames[0].lists) == 4
        xyzr, rgba = O.convert_lists(frames[0].lists)
        np.testing.assert_array_equal(xyzr, e[f.stem + "_xyzr"], err_msg=f.stem)
        np.testing.assert_array_equal(rgba, e[f.stem + "_rgba"], err_msg=f.stem)


def test_mmpld_roundtrip_versions_and_errors(tmp_path):
    g = np.random.default_rng(0)
    xyz = g.random((50, 3))
    for ver in (100, 102, 103):
        p = tmp_path / f"v{ver}.mmpld"
        ls = mmpld.ListSpec(2, 5, xyz, radius=g.random(50), colors=g.random((50, 4)))
        mmpld.write(str(p), [[ls], [ls]], version=ver)
        hdr, frames = mmpld.read(str(p))
        assert hdr.version == ver and len(frames) == 2
        xyzr, rgba = O.convert_lists(frames[1].lists)
        np.testing.assert_array_equal(xyzr[:, :3], xyz.astype(np.float32))
    bad = tmp_path / "bad.mmpld"
    raw = bytearray(open(tmp_path / "v103.mmpld", "rb").read())
    raw[0:5] = b"XXXXX"
    bad.write_bytes(bytes(raw))
    with pytest.raises(mmpld.MMPLDError, match="id wrong"):
        mmpld.read(str(bad))
    raw = bytearray(open(tmp_path / "v103.mmpld", "rb").read())
    raw[6:8] = (99).to_bytes(2, "little")
    bad.write_bytes(bytes(raw))
    with pytest.raises(mmpld.MMPLDError, match="version wrong"):
        mmpld.read(str(bad))
    raw = bytearray(open(tmp_path / "v103.mmpld", "rb").read())
    raw[8:12] = (0).to_bytes(4, "little")
    bad.write_bytes(bytes(raw))
    with pytest.raises(mmpld.MMPLDError, match="any frame"):
        mmpld.read(str(bad))


def test_bounce_closed_forms():
    # raygen.cu:122-155. An isolated sphere: a bounce direction N + u (|u| < 1) leaves the convex
    # sphere, so one bounce always misses: colour = albedo * 0.8; with a zero-length bounce budget
    # the depth cap gives 0.1 -- here max_bounces = 1 reaches the cap only after a second hit.
    xyzr = np.array([[0.5, 0.5, 0.5, 0.3]], np.float32)
    rgba = np.array([[0.2, 0.4, 0.6, 1.0]], np.float32)
    cam = small_camera(32, 32)
    for mb in (1, 2, 5):
        o = O.render_brute(xyzr, rgba, cam, O.make_params(32, 32, spp=1, mode=2, max_bounces=mb))
        hit = o["hit_id"] >= 0
        assert hit.sum() > 50
        np.testing.assert_allclose(o["color"][hit][:, :3], np.broadcast_to(rgba[0, :3] * np.float32(0.8),
                                                                             (hit.sum(), 3)), rtol=0, atol=0)
    # two touching spheres facing each other: some bounces hit the neighbour; a budget of 1
    # then ends at the 0.1 cap, a budget of 2 continues
    xyzr = np.array([[0.3, 0.5, 0.5, 0.2], [0.7, 0.5, 0.5, 0.2]], np.float32)
    rgba = np.ones((2, 4), np.float32)
    o1 = O.render_brute(xyzr, rgba, cam, O.make_params(32, 32, spp=4, mode=2, max_bounces=1))
    vals = {round(float(v), 5) for v in o1["color"][o1["hit_id"] >= 0][:, 0]}
    # per-sample values are 0.8 (escaped or primary miss) or 0.1 (capped); 4-sample averages lie on that grid
    grid = {round((k * 0.1 + (4 - k) * 0.8) / 4, 5) for k in range(5)}
    assert vals <= grid and len(vals) > 1


def _delta(keys, i, bits):
    """delta(i) of the sorted keys as the LBVH defines it: the common prefix of keys i and i + 1,
    key width + common prefix of the indices when the keys are equal, -1 past either end."""
    n = len(keys)
    if i < 0 or i >= n - 1:
        return -1
    w = 32 if bits == 30 else 64
    a, b = int(keys[i]), int(keys[i + 1])
    if a == b:
        return w + (32 - (i ^ (i + 1)).bit_length())
    return w - (a ^ b).bit_length()


@pytest.mark.parametrize("n,dups,bits", [(3, False, 30), (300, False, 30), (300, True, 30), (2000, True, 30),
                                         (1000, True, 63)])
def test_leaf_pair_count_rule_matches_the_tree(n, dups, bits):
    """The rule the device counts leaf pairs with (leaves l and l + 1 both attach to split l when
    delta(l) exceeds delta(l - 1) and delta(l + 1)) against the oracle tree's non-root nodes
    whose two children are leaves; and every such node covers two adjacent leaves."""
    g = np.random.default_rng(n + 1)
    xyzr = np.zeros((n, 4), np.float32)
    xyzr[:, :3] = g.random((n, 3)).astype(np.float32)
    if dups:
        xyzr[n // 2:, :3] = xyzr[: n - n // 2, :3]
    xyzr[:, 3] = 0.01
    _, child, _, keys = O.Bvh(xyzr, bits).arrays()
    d = [_delta(keys, i, bits) for i in range(-1, n)]  # d[i + 1] = delta(i)
    rule = sum(1 for l in range(n - 1) if d[l + 1] > d[l] and d[l + 1] > d[l + 2])
    both = (child[:, 0] < 0) & (child[:, 1] < 0)
    both[0] = False  # the root is always stored
    assert rule == int(both.sum())
    for i in np.nonzero(both)[0]:
        assert ~child[i, 1] == ~child[i, 0] + 1


# ------------------------------------------------------- Lambert and miss shading
def _sample_rays(cam, W, H, x, y, spp, frame=0):
    """The jittered primary rays of pixel (x, y): sample s takes draws 2s and 2s + 1 of the
    pixel's stream tea16(y * W + x, frame) (raygen.cu:182,241-242)."""
    import ctypes as C
    lib = O.load()
    fs = O.frame_state(cam)
    draws = O.lcg_stream(O.tea16(y * W + x, frame), 2 * spp)
    out = []
    for s in range(spp):
        o = np.zeros(3, np.float32)
        d = np.zeros(3, np.float32)
        lib.orc_primary_ray(C.byref(fs), W, H, x, y, C.c_float(draws[2 * s]), C.c_float(draws[2 * s + 1]),
                            o.ctypes.data, d.ctypes.data)
        out.append((o.astype(np.float64), d.astype(np.float64)))
    return out


def _perp(d):
    e = np.cross(d, [0.0, 1.0, 0.0] if abs(d[1]) < 0.9 else [1.0, 0.0, 0.0])
    return e / np.linalg.norm(e)


@pytest.mark.parametrize("h", [0.0, 0.3, 0.6, 0.9])
def test_lambert_shading_closed_form(h):
    """raygen.cu:135-142 with maxBounces 0: a sample that hits a sphere of albedo A at lateral
    offset h (in radii) from its centre line gets A * (0.2 + 0.6 |N.d|), |N.d| = sqrt(1 - h^2),
    with N the face-forwarded unit normal; h = 0 is the head-on ray, A * 0.8. The normal kept in
    the feature buffer faces the ray."""
    W, H, x, y = 8, 6, 3, 2
    cam = look_at((0.2, -0.1, 4.0), (0.0, 0.0, 0.0), (0, 1, 0), 30.0, W / H, 0.1, 50.0)
    (o, d), = _sample_rays(cam, W, H, x, y, 1)
    r, dist = 0.5, 3.0
    e = _perp(d)
    c = o + dist * d + h * r * e
    xyzr = np.array([[*c, r]], np.float32)
    alb = np.array([[0.25, 0.5, 1.0, 1.0]], np.float32)
    o1 = O.render_brute(xyzr, alb, cam, O.make_params(W, H, want_aux=1), aux=True)
    assert o1["hit_id"][y, x] == 0
    cos = math.sqrt(1.0 - h * h)
    want = alb[0, :3].astype(np.float64) * (0.2 + 0.6 * cos)
    np.testing.assert_allclose(o1["color"][y, x, :3], want, rtol=2e-5, atol=0)
    assert o1["color"][y, x, 3] == 1.0
    # the normal written to the features: (P - c) / r with P on the near side of the sphere
    cf = xyzr[0, :3].astype(np.float64)
    t = (cf - o) @ d - math.sqrt(max(0.0, r * r - float(np.sum(((cf - o) - ((cf - o) @ d) * d) ** 2))))
    n_want = (o + t * d - cf) / r
    n_got = o1["aux"][y, x, 3:6].astype(np.float64)
    np.testing.assert_allclose(n_got, n_want, atol=2e-3)  # fp16 storage
    assert n_got @ d < 0.0
    assert math.isclose(float(o1["t"][y, x]), t, rel_tol=1e-5)


def test_partial_coverage_averages_hit_shading_with_miss_white():
    """raygen.cu:131-133,250,265-268: a sample that misses every sphere contributes
    col * 0.8 = (0.8, 0.8, 0.8, 1) to the pixel average; only a pixel that no sample hit takes
    the background (and pseudo depth 1)."""
    W, H, x, y, spp = 4, 4, 1, 2, 2
    cam = look_at((0.0, 0.0, 5.0), (0.0, 0.0, 0.0), (0, 1, 0), 40.0, 1.0, 0.1, 50.0)
    (o, d0), (_, d1) = _sample_rays(cam, W, H, x, y, spp)
    dist = 4.0
    ang = math.acos(min(1.0, float(d0 @ d1)))
    r = 0.25 * dist * math.sin(ang)  # covers sample 0's ray, far from sample 1's
    assert r > 1e-3
    xyzr = np.array([[*(o + dist * d0), r]], np.float32)
    alb = np.array([[0.9, 0.3, 0.6, 1.0]], np.float32)
    bg = (0.1, 0.2, 0.3, 0.4)
    f = O.render_brute(xyzr, alb, cam, O.make_params(W, H, spp=spp, background=bg))
    assert f["hit_id"][y, x] == 0
    assert O.intersect_f32(o, d1, xyzr[0]) is None
    want = (alb[0, :3].astype(np.float64) * 0.8 + 0.8) / 2.0
    np.testing.assert_allclose(f["color"][y, x, :3], want, rtol=2e-5, atol=0)
    assert f["color"][y, x, 3] == 1.0
    assert f["depth"][y, x] < 1.0
    # every other pixel: no sample hits -> background, depth 1, no id, t = inf
    miss = np.ones((H, W), bool)
    miss[y, x] = False
    assert (f["hit_id"][miss] == -1).all() and (f["depth"][miss] == 1.0).all()
    np.testing.assert_array_equal(f["color"][miss], np.broadcast_to(np.array(bg, np.float32), (miss.sum(), 4)))
    assert np.isinf(f["t"][miss]).all()


def test_progressive_blend_is_the_running_mean():
    """raygen.cu:254-260: blending frame k with weight 1 / (k + 1) into the previous result makes
    the running mean of the frames (to FP32 rounding); a pixel that the newest frame misses takes
    the background whatever was accumulated."""
    g = np.random.default_rng(3)
    P, K = 64, 12
    frames = [g.random((P, 4)).astype(np.float32) for _ in range(K)]
    ids = np.zeros(P, np.int32)
    bg = (0.5, 0.25, 0.125, 1.0)
    acc = None
    for k, fr in enumerate(frames):
        acc, rgba8 = O.blend_frames(acc, {"color": fr, "hit_id": ids}, k, bg)
        np.testing.assert_allclose(acc, np.mean(frames[:k + 1], axis=0), rtol=0, atol=3e-6)
    np.testing.assert_array_equal(rgba8, O.pack_rgba8(acc))
    ids[5] = -1
    acc2, _ = O.blend_frames(acc, {"color": frames[0], "hit_id": ids}, K, bg)
    np.testing.assert_array_equal(acc2[5], np.array(bg, np.float32))
    # RGBA8: clamp to [0, 1], then round half up per channel, R in the low byte
    px = O.pack_rgba8(np.array([[0.0, 1.0, 0.5, 2.0], [-1.0, 0.00196, 0.00197, 1.0]], np.float32))
    assert px[0] == (0 | 255 << 8 | 128 << 16 | 255 << 24)
    assert px[1] == (0 | 0 << 8 | 1 << 16 | 255 << 24)


# ------------------------------------------------------------ scale checks of the oracle itself
@pytest.mark.parametrize("bits", [30, 63])
def test_parallel_lbvh_build_is_the_serial_tree(bits):
    """The multi-threaded oracle build (per-thread radix histograms, offsets in digit-then-thread
    order, parallel Karras nodes, atomic second-arrival boxes) gives the serial build's arrays."""
    g = np.random.default_rng(bits)
    n = 300_000
    xyzr = np.zeros((n, 4), np.float32)
    xyzr[:, :3] = g.random((n, 3)).astype(np.float32)
    xyzr[n // 3:2 * n // 3, :3] = xyzr[:n // 3, :3]  # duplicate keys
    xyzr[:, 3] = g.uniform(0.001, 0.004, n).astype(np.float32)
    a, b = O.Bvh(xyzr, bits, threads=1).arrays(), O.Bvh(xyzr, bits, threads=8).arrays()
    for u, v in zip(a, b):
        np.testing.assert_array_equal(u.view(np.uint8), v.view(np.uint8))


def test_cube_frame_fp32_oracle_against_fp64_everywhere():
    """The uniform-cube workload (1M spheres, r 0.002, 1024x1024, Lambert): the oracle's FP32 LBVH
    render against the FP64 brute-force nearest hit at every pixel. Mismatches are grazing or
    near-tie pixels only and at most 1e-5 of the image; t within 1e-5 relative elsewhere."""
    from helpers import fp64_check, oracle_params
    from megamol_amd import ParticleList, _lib as L, params
    xyzr, rgba8 = scenes.uniform_cube(1_000_000)
    lists = [ParticleList(L.VERTDATA_FLOAT_XYZR, xyzr, L.COLDATA_UINT8_RGBA, rgba8, count=len(xyzr))]
    xo, ro = O.convert_lists(lists)
    cam = scenes.uniform_cube_camera(1024, 1024)
    p = params(1024, 1024)
    o = O.Bvh(xo, 30).render(ro, cam, oracle_params(p))
    rects = [(x, y, x + 16, y + 16) for y in range(0, 1024, 16) for x in range(0, 1024, 16)]
    m, n = fp64_check(xo, cam, 1024, 1024, o["hit_id"], o["t"], rects, what="cube oracle")
    assert n == 1024 * 1024 and m <= 1e-5 * n, m
