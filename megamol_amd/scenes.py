"""Seeded synthetic sphere scenes (host side, numpy).

Workloads of the sphere path:
  uniform cube   1M spheres, xyz ~ U[0,1)^3, radius 0.002, random RGBA8 colours, 1024x1024,
                 camera at (0.5,0.5,2.5) looking at the cube centre
  FCC crystal    4*n^3 spheres on a face-centred cubic lattice (cell edge 1.0, n = 232 gives
                 49,948,672), jitter N(0,0.05) per axis, radius U[0.35,0.5], intensity U[0,1]
                 coloured through a 256-entry RGBA transfer function, 1920x1080, AO tmax 4.0
  time series    4M spheres uniform in a 200^3 box, radius 0.5, Gaussian steps sigma 0.05
  clusters       spheres from Gaussian clusters with lognormal sigma and radius

Random numbers come from numpy's PCG64 with a fixed seed per field, so a scene is the same on
every machine. Positions are computed in float64 and stored as float32.
"""
from __future__ import annotations

import numpy as np

from .renderer import look_at


def _rngs(seed: int, k: int):
    ss = np.random.SeedSequence(seed)
    return [np.random.Generator(np.random.PCG64(s)) for s in ss.spawn(k)]


def uniform_cube(n: int = 1_000_000, radius: float = 0.002, seed: int = 1):
    g_pos, g_col = _rngs(seed, 2)
    xyzr = np.empty((n, 4), np.float32)
    xyzr[:, :3] = g_pos.random((n, 3), dtype=np.float64).astype(np.float32)
    xyzr[:, 3] = np.float32(radius)
    rgba8 = g_col.integers(0, 256, size=(n, 4), dtype=np.uint8)
    return xyzr, rgba8


def uniform_cube_camera(width=1024, height=1024):
    return look_at((0.5, 0.5, 2.5), (0.5, 0.5, 0.5), (0, 1, 0), fovy_deg=40.0, aspect=width / height, near=0.1,
                   far=10.0)


def fcc_count(cells: int) -> int:
    return 4 * cells ** 3


def fcc_crystal(cells: int = 232, spacing: float = 1.0, jitter: float = 0.05, rmin: float = 0.35, rmax: float = 0.5,
                seed: int = 2, chunk_cells: int = 8):
    """FCC lattice, returns xyzr float32 (N,4) and intensity float32 (N,)."""
    n = fcc_count(cells)
    g_j, g_r, g_i = _rngs(seed, 3)
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]], np.float64) * spacing
    xyzr = np.empty((n, 4), np.float32)
    inten = np.empty(n, np.float32)
    per_slab = 4 * cells * cells
    j = np.arange(cells, dtype=np.float64)
    yy, zz = np.meshgrid(j, j, indexing="ij")
    yz = np.stack([yy.ravel(), zz.ravel()], 1) * spacing  # (cells^2, 2)
    for x0 in range(0, cells, chunk_cells):
        xs = np.arange(x0, min(cells, x0 + chunk_cells), dtype=np.float64) * spacing
        m = len(xs) * per_slab
        base = np.empty((len(xs), cells * cells, 4, 3), np.float64)
        base[..., 0] = xs[:, None, None]
        base[..., 1] = yz[None, :, None, 0]
        base[..., 2] = yz[None, :, None, 1]
        base += basis[None, None, :, :]
        base = base.reshape(-1, 3)
        base += g_j.normal(0.0, jitter, size=(m, 3))
        o = x0 * per_slab
        xyzr[o:o + m, :3] = base.astype(np.float32)
        xyzr[o:o + m, 3] = g_r.uniform(rmin, rmax, size=m).astype(np.float32)
        inten[o:o + m] = g_i.random(m, dtype=np.float64).astype(np.float32)
    return xyzr, inten


def transfer_function_256() -> np.ndarray:
    """A fixed 256-entry RGBA ramp (dark blue -> teal -> yellow), float32 (256,4)."""
    t = np.linspace(0.0, 1.0, 256, dtype=np.float64)
    r = 0.15 + 0.85 * t ** 1.5
    g = 0.20 + 0.70 * np.sin(np.pi * 0.5 * t)
    b = 0.55 + 0.35 * np.cos(np.pi * t) * (1 - t)
    a = np.ones_like(t)
    return np.stack([r, g, b, a], 1).astype(np.float32)


def fcc_camera(cells: int = 232, spacing: float = 1.0, width=1920, height=1080, fovy_deg=40.0):
    c = 0.5 * cells * spacing
    d = np.array([0.55, 0.40, 1.0])
    d /= np.linalg.norm(d)
    dist = 1.3 * c / np.sin(np.deg2rad(fovy_deg) / 2)
    pos = c + d * dist
    return look_at(tuple(pos), (c, c, c), (0, 1, 0), fovy_deg=fovy_deg, aspect=width / height, near=1.0,
                   far=4.0 * dist)


def time_series_initial(n: int = 4_000_000, box: float = 200.0, seed: int = 3):
    (g,) = _rngs(seed, 1)
    xyzr = np.empty((n, 4), np.float32)
    xyzr[:, :3] = (g.random((n, 3)) * box).astype(np.float32)
    xyzr[:, 3] = np.float32(0.5)
    return xyzr


def time_series_step(xyzr: np.ndarray, step: int, sigma: float = 0.05, seed: int = 3) -> np.ndarray:
    """Step k -> k+1: Gaussian displacement N(0, sigma) per axis (float32 result)."""
    (g,) = _rngs(seed * 1000003 + step, 1)
    out = xyzr.copy()
    out[:, :3] = (xyzr[:, :3].astype(np.float64) + g.normal(0.0, sigma, size=(len(xyzr), 3))).astype(np.float32)
    return out


RESOLUTIONS = [(640, 480), (800, 600), (1024, 768), (1280, 720), (1600, 900), (1920, 1080)]


def time_series_cameras(step: int, box: float = 200.0, seed: int = 3):
    """1-6 cameras per timestep at mixed resolutions, orbiting the box."""
    (g,) = _rngs(seed * 7919 + step, 1)
    k = int(g.integers(1, 7))
    out = []
    c = box / 2
    for i in range(k):
        w, h = RESOLUTIONS[int(g.integers(0, len(RESOLUTIONS)))]
        ang = g.uniform(0, 2 * np.pi)
        el = g.uniform(-0.4, 0.6)
        dist = g.uniform(1.6, 2.4) * box
        pos = (c + dist * np.cos(ang) * np.cos(el), c + dist * np.sin(el), c + dist * np.sin(ang) * np.cos(el))
        out.append((look_at(pos, (c, c, c), (0, 1, 0), fovy_deg=40.0, aspect=w / h, near=1.0, far=4 * dist), w, h))
    return out


def clusters(n: int, n_clusters: int = 10_000, box: float = 1e4, mu_sigma: float = 3.0, s_sigma: float = 0.5,
             mu_r: float = -0.5, s_r: float = 0.3, seed: int = 5, with_cluster=False):
    """n spheres drawn from n_clusters Gaussian clusters: centres uniform in a box^3 cube,
    sigma lognormal(mu_sigma, s_sigma), radius lognormal(mu_r, s_r); xyzr float32 (N,4), sorted by
    cluster. with_cluster also returns each sphere's cluster index (int32)."""
    g_c, g_s, g_a, g_p, g_r = _rngs(seed, 5)
    centres = g_c.random((n_clusters, 3)) * box
    sig = g_s.lognormal(mu_sigma, s_sigma, n_clusters)
    counts = g_a.multinomial(n, np.full(n_clusters, 1.0 / n_clusters))
    xyzr = np.empty((n, 4), np.float32)
    cl = np.empty(n, np.int32) if with_cluster else None
    o = 0
    step = 1 << 24
    a_all = np.repeat(np.arange(n_clusters, dtype=np.int32), counts)
    for o in range(0, n, step):
        a = a_all[o:o + step]
        m = len(a)
        xyzr[o:o + m, :3] = (centres[a] + g_p.normal(size=(m, 3)) * sig[a, None]).astype(np.float32)
        xyzr[o:o + m, 3] = g_r.lognormal(mu_r, s_r, m).astype(np.float32)
        if with_cluster:
            cl[o:o + m] = a
    return (xyzr, cl) if with_cluster else xyzr


def clusters_camera(box: float = 1e4, width=3840, height=2160, fovy_deg=40.0):
    """Looking at the centre of the cluster box from outside one corner region."""
    c = np.full(3, 0.5 * box)
    d = np.array([0.35, 0.25, 1.0])
    d /= np.linalg.norm(d)
    dist = 0.9 * box
    pos = c + d * dist
    return look_at(tuple(pos), tuple(c), (0, 1, 0), fovy_deg=fovy_deg, aspect=width / height, near=1.0,
                   far=4.0 * box)


def orbit_cameras(k: int, box: float = 1e4, width=3840, height=2160, fovy_deg=40.0):
    """k cameras on a circle around the box centre (the orbit path of the clustered scene)."""
    c = np.full(3, 0.5 * box)
    out = []
    for i in range(k):
        a = 2.0 * np.pi * i / k
        pos = c + 0.9 * box * np.array([np.sin(a) * 0.95, 0.25, np.cos(a) * 0.95])
        out.append(look_at(tuple(pos), tuple(c), (0, 1, 0), fovy_deg=fovy_deg, aspect=width / height, near=1.0,
                           far=4.0 * box))
    return out


def random_spheres(n: int, seed: int = 0, lo=0.0, hi=1.0, rmin=0.01, rmax=0.05):
    g = np.random.Generator(np.random.PCG64(seed))
    xyzr = np.empty((n, 4), np.float32)
    xyzr[:, :3] = g.uniform(lo, hi, (n, 3)).astype(np.float32)
    xyzr[:, 3] = g.uniform(rmin, rmax, n).astype(np.float32)
    rgba = g.random((n, 4)).astype(np.float32)
    rgba[:, 3] = 1.0
    return xyzr, rgba
