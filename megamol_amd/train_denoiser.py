"""Train the learned denoiser on pairs that the renderer itself generates.

A pair is a crop of a 1920x1080 frame of a seeded FCC crystal scene seen from a random camera:
the 1-spp AO colour with its feature buffer (albedo, normal, linear depth, visibility) as
input, and the same crop rendered with many samples (a different random stream) as target.
Crops are rendered through the render window, so each one is bit-identical to that region of
the full frame. Linear depth is divided by the camera's distance to the crystal centre.

Usage: python -m megamol_amd.train_denoiser --pairs 2000 --steps 3000 --out weights.bin
"""
from __future__ import annotations

import argparse
import sys
import time

import numpy as np

from . import _lib as L
from . import scenes
from .denoiser import Denoiser, psnr, save_weights
from .renderer import ParticleList, SphereRenderer, look_at, params


def log(msg):
    print(f"[train] {msg}", file=sys.stderr, flush=True)


def load_fcc(r: SphereRenderer, cells: int, seed: int):
    xyzr, inten = scenes.fcc_crystal(cells, seed=seed)
    r.set_transfer_function(scenes.transfer_function_256())
    r.set_lists([ParticleList(L.VERTDATA_FLOAT_XYZR, xyzr, L.COLDATA_FLOAT_I, inten, count=len(xyzr),
                              col_min=0.0, col_max=1.0)])
    r.build(L.BUILD_REBUILD, 30)
    return len(xyzr)


def random_camera(g: np.random.Generator, cells: int, width=1920, height=1080, fovy=40.0):
    """A camera outside the crystal looking at a point near its centre, from far (whole crystal
    in view) to close (a few hundred spheres across)."""
    c = 0.5 * cells
    full = 1.3 * c / np.sin(np.deg2rad(fovy) / 2)
    d = g.normal(size=3)
    d[1] = abs(d[1]) * 0.8
    d /= np.linalg.norm(d)
    dist = full * g.uniform(0.45, 1.1)
    target = np.full(3, c) + g.uniform(-0.25, 0.25, 3) * c
    pos = target + d * dist
    cam = look_at(tuple(pos), tuple(target), (0, 1, 0), fovy_deg=fovy, aspect=width / height, near=1.0,
                  far=4 * full)
    return cam, float(np.linalg.norm(pos - np.full(3, c)))


def render_pair(r, cam, depth_scale, frame, x0, y0, crop, spp_target, ao_distance=4.0, width=1920, height=1080):
    p1 = params(width, height, spp=1, mode=L.MODE_AO, ao_distance=ao_distance, frame_idx=frame, flags=L.FLAG_AUX)
    p1.window[:] = [x0, y0, crop, crop]
    fr = r.render(cam, p1, outputs=("color",))
    aux = r.download(crop, crop, outputs=(), aux=True).aux
    feats = np.zeros((crop, crop, 12), np.float16)
    feats[..., 0:3] = fr.color[..., 0:3]
    feats[..., 3:11] = aux
    feats[..., 9] = (aux[..., 6].astype(np.float32) / depth_scale).astype(np.float16)
    pt = params(width, height, spp=spp_target, mode=L.MODE_AO, ao_distance=ao_distance,
                frame_idx=1_000_000 + frame)
    pt.window[:] = [x0, y0, crop, crop]
    tgt = r.render(cam, pt, outputs=("color",)).color[..., 0:3].copy()
    return feats, tgt


def make_pairs(r, n, cells, seed, crop=128, spp_target=256, min_hit=0.25, width=1920, height=1080):
    g = np.random.default_rng(seed)
    feats = np.zeros((n, crop, crop, 12), np.float16)
    tg = np.zeros((n, crop, crop, 3), np.float32)
    k = 0
    tries = 0
    while k < n:
        cam, dscale = random_camera(g, cells, width, height)
        for _ in range(8):  # several crops per camera
            if k >= n:
                break
            x0 = int(g.integers(0, width - crop + 1)) // 8 * 8
            y0 = int(g.integers(0, height - crop + 1)) // 8 * 8
            f, t = render_pair(r, cam, dscale, seed * 100000 + tries, x0, y0, crop, spp_target, width=width,
                               height=height)
            tries += 1
            if (f[..., 9] > 0).mean() < min_hit:
                continue
            feats[k], tg[k] = f, t
            k += 1
            if k % 250 == 0:
                log(f"{k}/{n} pairs")
    return feats, tg


def train(dn: Denoiser, feats, targets, steps, batch=8, lr=2e-3, lr_min=1e-4, seed=0, log_every=250):
    g = np.random.default_rng(seed)
    n = len(feats)
    losses = []
    t0 = time.time()
    for s in range(steps):
        idx = g.choice(n, batch, replace=False)
        f = feats[idx]
        t = targets[idx]
        if g.random() < 0.5:  # flips keep the statistics and double the data
            f, t = f[:, :, ::-1], t[:, :, ::-1]
        cur = lr_min + 0.5 * (lr - lr_min) * (1 + np.cos(np.pi * s / max(steps - 1, 1)))
        losses.append(dn.train_step(np.ascontiguousarray(f), np.ascontiguousarray(t), 1.0, lr=float(cur)))
        if log_every and (s % log_every == 0 or s == steps - 1):
            log(f"step {s} loss {np.mean(losses[-log_every:]):.6f} lr {cur:.2e} ({time.time() - t0:.1f}s)")
    return losses


def evaluate(dn: Denoiser, feats, targets):
    """PSNR of raw 1-spp colour and of the denoised colour against the targets."""
    out = dn.denoise(feats, 1.0)
    raw = feats[..., 0:3].astype(np.float32)
    return psnr(raw, targets), psnr(out, targets)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--cells", type=int, default=232)
    ap.add_argument("--pairs", type=int, default=2000)
    ap.add_argument("--eval-pairs", type=int, default=64)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--spp-target", type=int, default=256)
    ap.add_argument("--crop", type=int, default=128)
    ap.add_argument("--seeds", type=str, default="21,22")
    ap.add_argument("--eval-seed", type=int, default=29)
    ap.add_argument("--out", type=str, default="")
    args = ap.parse_args()
    r = SphereRenderer(0)
    dn = Denoiser(r, seed=1234)
    seeds = [int(s) for s in args.seeds.split(",")]
    parts_f, parts_t = [], []
    t0 = time.time()
    for i, s in enumerate(seeds):
        load_fcc(r, args.cells, s)
        k = args.pairs // len(seeds) + (1 if i < args.pairs % len(seeds) else 0)
        f, t = make_pairs(r, k, args.cells, s, args.crop, args.spp_target)
        parts_f.append(f)
        parts_t.append(t)
        log(f"scene seed {s}: {k} pairs ({time.time() - t0:.1f}s)")
    feats, targets = np.concatenate(parts_f), np.concatenate(parts_t)
    load_fcc(r, args.cells, args.eval_seed)
    ef, et = make_pairs(r, args.eval_pairs, args.cells, args.eval_seed, args.crop, args.spp_target)
    before = evaluate(dn, ef, et)
    losses = train(dn, feats, targets, args.steps, batch=args.batch, lr=args.lr)
    after = evaluate(dn, ef, et)
    log(f"held-out PSNR: raw 1spp {after[0]:.2f} dB, untrained net {before[1]:.2f} dB, trained {after[1]:.2f} dB")
    if args.out:
        save_weights(args.out, dn.get_params(), {"pairs": len(feats), "steps": args.steps, "batch": args.batch,
                                                 "spp_target": args.spp_target,
                                                 "crop": args.crop, "seeds": seeds, "final_loss": losses[-1],
                                                 "heldout_psnr_raw": after[0], "heldout_psnr_denoised": after[1]})
        log(f"saved {args.out}")
    r.close()


if __name__ == "__main__":
    main()
