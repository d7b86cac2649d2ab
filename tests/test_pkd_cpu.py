"""CPU checks of the oracle's treelet P-k-d trees (the compact structure for very large scenes).

Pinned by what the structure must satisfy, computed here independently of the oracle's code:
  * every treelet holds exactly the spheres of its run of the Morton order (the LBVH's leaf order);
  * the k-d property: along a node's split axis no sphere of its left subtree lies above the node's
    sphere and none of its right subtree below it (subtrees found by walking heap indices);
  * the split axis is the widest extent of the node's cell (the treelet's sphere bounds cut at the
    ancestors' coordinates);
  * the top tree's boxes hold their treelets' boxes, and absent children only pad the last level;
  * renders through the trees equal the brute-force oracle bit for bit (both tie rules: equal t
    goes to the lowest particle id), in Lambert and AO mode, for treelet sizes from 2 to 1024 and
    for scenes with coincident spheres, fewer spheres than one treelet and one sphere past a
    whole number of treelets.
"""
import numpy as np
import pytest

import oracle as O
from helpers import small_camera
from megamol_amd import scenes


def subtree(i, m):
    out, todo = [], [i]
    while todo:
        j = todo.pop()
        if j < m:
            out.append(j)
            todo += [2 * j + 1, 2 * j + 2]
    return out


def widest(box):
    e = [box[3 + a] - box[a] for a in range(3)]
    d = 0
    for a in (1, 2):
        if e[a] > e[d]:
            d = a
    return d


def check_structure(xyzr, bits, M):
    b = O.Bvh(xyzr, bits)
    _, _, leaf_ids, _ = b.arrays()
    k = O.Pkd(b, M)
    a = k.arrays()
    n = len(xyzr)
    T = (n + M - 1) // M
    assert k.k.T == T
    for t in range(T):
        lo, hi = t * M, min(n, (t + 1) * M)
        m = hi - lo
        ids = a["id"][lo:hi]
        assert sorted(ids.tolist()) == sorted(leaf_ids[lo:hi].tolist()), t
        sph = a["sph"][lo:hi]
        np.testing.assert_array_equal(sph, xyzr[ids])
        box = np.concatenate([(sph[:, :3] - sph[:, 3:]).min(0), (sph[:, :3] + sph[:, 3:]).max(0)])
        np.testing.assert_array_equal(a["tl_box"][t], box.astype(np.float32))
        assert a["tl_rmax"][t] == sph[:, 3].max()
        cells = {0: box.copy()}
        for i in range(m):
            d = int(a["dim"][lo + i])
            assert d == widest(cells[i]), (t, i)
            c = sph[i, d]
            left, right = subtree(2 * i + 1, m), subtree(2 * i + 2, m)
            if left:
                assert sph[left, d].max() <= c, (t, i)
            if right:
                assert sph[right, d].min() >= c, (t, i)
            lb, rb = cells[i].copy(), cells[i].copy()
            lb[3 + d] = rb[d] = c
            cells[2 * i + 1], cells[2 * i + 2] = lb, rb
            assert len(subtree(i, m)) == O.pkd_subtree_size(i, m)
    L = k.k.top_leaves
    assert L >= max(T, 2) and L & (L - 1) == 0 and L // 2 < max(T, 2)
    # every treelet appears once, as the child of the top node its heap index names
    seen = []
    for h in range(L - 1):
        for c in range(2):
            ch = 2 * h + 1 + c
            ref = a["top_child"][h, c]
            bx = a["top_box"][h, 6 * c:6 * c + 6]
            if ch >= L - 1:
                tl = ch - (L - 1)
                if tl < T:
                    assert ref == ~tl
                    np.testing.assert_array_equal(bx, a["tl_box"][tl])
                    seen.append(tl)
                else:
                    assert ref == np.iinfo(np.int32).min
            else:
                leaves = [x - (L - 1) for x in subtree(ch, 2 * L - 1) if x >= L - 1 and x - (L - 1) < T]
                if not leaves:
                    assert ref == np.iinfo(np.int32).min
                    continue
                assert ref == ch
                np.testing.assert_array_equal(bx[:3], a["tl_box"][leaves, :3].min(0))
                np.testing.assert_array_equal(bx[3:], a["tl_box"][leaves, 3:].max(0))
    assert sorted(seen) == list(range(T))
    return b, k


@pytest.mark.parametrize("M", [2, 8, 64, 1024])
def test_pkd_structure_random(M):
    xyzr, _ = scenes.random_spheres(3000 + 7, seed=11)
    check_structure(xyzr, 30, M)


def test_pkd_structure_coincident_and_63bit():
    rng = np.random.default_rng(5)
    xyzr = np.repeat(rng.random((300, 4)).astype(np.float32), 3, axis=0)  # triples of equal spheres
    xyzr[:, 3] = 0.01 + 0.02 * xyzr[:, 3]
    xyzr[::7, 0] = xyzr[0, 0]  # many equal x coordinates across the scene
    check_structure(xyzr, 63, 16)


def _render_equal(xyzr, rgba, M, bits=30, spp=2, w=48, h=36):
    cam = small_camera(w, h)
    b = O.Bvh(xyzr, bits)
    k = O.Pkd(b, M)
    for mode in (0, 1):
        p = O.make_params(w, h, spp, mode, 0.25)
        ref = O.render_brute(xyzr, rgba, cam, p, aux=True)
        got = k.render(rgba, cam, p, aux=True)
        for key in ("hit_id", "rgba8"):
            np.testing.assert_array_equal(got[key], ref[key], err_msg=f"{key} M={M} mode={mode}")
        for key in ("t", "depth", "color"):
            np.testing.assert_array_equal(got[key].view(np.uint32), ref[key].view(np.uint32),
                                          err_msg=f"{key} M={M} mode={mode}")
        np.testing.assert_array_equal(got["aux"].view(np.uint16), ref["aux"].view(np.uint16))
    return k


@pytest.mark.parametrize("M,n", [(2, 1500), (16, 1500), (256, 1500), (1024, 1500), (1024, 1), (64, 65),
                                 (64, 128), (1024, 700)])
def test_pkd_render_equals_brute_force(M, n):
    xyzr, rgba = scenes.random_spheres(n, seed=M + n)
    _render_equal(xyzr, rgba, M)


def test_pkd_render_ties():
    """Coincident spheres (equal t everywhere they are hit): the lowest particle id wins."""
    rng = np.random.default_rng(9)
    base = np.zeros((400, 4), np.float32)
    base[:, :3] = rng.random((400, 3)).astype(np.float32)
    base[:, 3] = 0.04
    xyzr = np.concatenate([base, base[::-1], base])
    rgba = rng.random((len(xyzr), 4)).astype(np.float32)
    _render_equal(xyzr, rgba, 32)


def test_pkd_render_clusters_63bit():
    xyzr = scenes.clusters(20000, n_clusters=20, box=10.0, mu_sigma=-1.0, s_sigma=0.3, mu_r=-4.0, s_r=0.3)
    rgba = np.random.default_rng(2).random((len(xyzr), 4)).astype(np.float32)
    cam = small_camera(40, 30, pos=(5.0, 5.0, 22.0), target=(5.0, 5.0, 5.0))
    b = O.Bvh(xyzr, 63)
    k = O.Pkd(b, 256)
    p = O.make_params(40, 30, 1, 1, 2.0)
    ref = b.render(rgba, cam, p)
    got = k.render(rgba, cam, p)
    assert (ref["hit_id"] >= 0).mean() > 0.05
    np.testing.assert_array_equal(got["hit_id"], ref["hit_id"])
    np.testing.assert_array_equal(got["color"].view(np.uint32), ref["color"].view(np.uint32))


def test_pkd_rejects_bad_treelet_size():
    xyzr, _ = scenes.random_spheres(100, seed=1)
    b = O.Bvh(xyzr, 30)
    for M in (0, 1, 3, 100):
        with pytest.raises(ValueError):
            O.Pkd(b, M)
