#!/usr/bin/env python3
"""Benchmark of the sphere ray-tracing path on one MI355X.

Workload (default, "fcc"): 49,948,672 spheres on a jittered FCC lattice (232^3 cells, radius
U[0.35,0.5], intensity through a 256-entry transfer function), 1920x1080, ambient occlusion with
64 samples per pixel (1 AO ray per sample, AO distance 4.0), LBVH rebuilt every frame. A "step"
is one frame: rebuild + render, with the spheres already resident in HBM.

Prints one JSON line. `value` is the denoised 1-spp frame (LBVH rebuild + 1-spp AO trace with
features + U-Net denoise). `roofline` is measured live for that frame's dominant kernel, the
hierarchy chunk pass of the build: algorithmic bytes per launch (per leaf a 4-byte particle id,
a 4-byte key and the 16-byte gathered sphere in, the 16-byte leaf sphere out; 64 bytes per
stored node) divided by its average duration from HIP events on the library's stream.
`roofline_plain_trace` does the same for the plain 64-spp frame's trace kernel (64 B per node
visit, 32 B per leaf-pair visit, 16 B per leaf sphere fetched or hit sphere shaded, 32 B per
pixel written). `cpu_baseline` times the CPU oracle (multi-threaded LBVH renderer with the same
rules) on a bounded sample of the same frame.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE summaries (committed with the commands that made them):
# the DRAM bytes per launch of the chunk pass in the denoised frame and of the trace kernel of the
# plain 64-spp frame, for the 50M FCC workload
DRAM_PROFILE_CHUNK = os.path.join(ROOT, "profiles", "k_bu_chunk_ls_fcc50m_dram.md")
DRAM_PROFILE = os.path.join(ROOT, "profiles", "k_trace_sl_fcc50m_ao64_dram.md")
# rocprofv3 --pmc SQ_INSTS_VALU summary of the plain frame's trace kernel (the same workload)
VALU_PROFILE = os.path.join(ROOT, "profiles", "k_trace_sl_fcc50m_ao64_valu.md")
# VALU issue roof: 256 CUs x 4 SIMDs, one wave64 instruction per SIMD per 4 cycles, 2.4 GHz peak clock
VALU_ISSUE_PEAK = 256 * 4 * 2.4e9 / 4


def dram_traffic(workload_key, path=DRAM_PROFILE, field="dram_bytes_per_launch"):
    """A per-launch figure (DRAM bytes by default) recorded in a committed counter summary for this
    workload, or None."""
    try:
        txt = open(path).read()
    except OSError:
        return None
    if workload_key not in txt:
        return None
    for line in txt.splitlines():
        if line.startswith(field + ":"):
            return int(line.split(":")[1].strip())
    return None


def trace_kernel_label(r, spp, mode):
    """Name of the traversal kernel the library dispatches for this frame (mirrors render_frame)."""
    sl = (r.get_option("sample_lanes") and r.get_option("refill") == 0 and spp > 1
          and ((spp <= 64 and spp & (spp - 1) == 0) or spp % 64 == 0))
    name = "k_trace_sl" if sl else "k_trace"
    return f"{name} ({mode}, {spp} spp)"


# Traversal ablation: one choice changed at a time from the first row of each frame (every
# variant renders the same image bit for bit). Frames: the 64-spp AO frame and the denoised frame's 1-spp
# AO frame with the feature buffer.
ABLATION = {
    "64spp": [
        ("sample lanes, work stealing, LDS stack 16 (library default), 8x8 tiles", {"stack_lds": 16, "refill": 0}),
        ("LDS stack 8 entries", {"stack_lds": 8, "refill": 0}),
        ("LDS stack 24 entries", {"stack_lds": 24, "refill": 0}),
        ("LDS stack 32 entries", {"stack_lds": 32, "refill": 0}),
        ("4x16 tiles", {"stack_lds": 16, "refill": 0, "tile_w": 4}),
        ("lane per pixel (samples in a loop)", {"stack_lds": 16, "refill": 0, "sample_lanes": 0}),
        ("lane per pixel, static tile order (no work stealing)",
         {"stack_lds": 16, "refill": 0, "sample_lanes": 0, "persistent": 0}),
        ("lane per pixel + ballot compaction (refill at 32 idle lanes)",
         {"stack_lds": 16, "sample_lanes": 0, "refill": 32}),
        ("sample lanes, 2 workgroups per CU", {"stack_lds": 16, "refill": 0, "blocks_per_cu": 2}),
        ("tree with leaf pairs (sample-lanes kernel with the pair branch)",
         {"stack_lds": 16, "refill": 0, "leaf_pairs": 1}),
    ],
    "1spp": [
        ("lane per pixel, work stealing, LDS stack 16, 8x8 tiles, no compaction", {"stack_lds": 16, "refill": 0}),
        ("static tile order (no work stealing)", {"stack_lds": 16, "refill": 0, "persistent": 0}),
        ("ballot compaction (refill at 16 idle lanes; library default)", {"stack_lds": 16, "refill": 16}),
        ("ballot compaction (refill at 32 idle lanes)", {"stack_lds": 16, "refill": 32}),
        ("LDS stack 8 entries", {"stack_lds": 8, "refill": 0}),
        ("LDS stack 32 entries", {"stack_lds": 32, "refill": 0}),
        ("4 workgroups per CU", {"stack_lds": 16, "refill": 0, "blocks_per_cu": 4}),
        ("tree without leaf pairs", {"stack_lds": 16, "refill": 0, "leaf_pairs": 0}),
    ],
}
# the tree each frame's rows trace unless a row says otherwise (the autotuner's choice for it)
ABLATION_PAIRS = {"64spp": 0, "1spp": 1}


def run_ablation(r, cam, args):
    """Trace time (HIP events) and Mrays/s of each ABLATION variant on one LBVH of the scene (rebuilt
    when a row asks for the other tree: with or without leaf pairs)."""
    from megamol_amd import _lib as L, params
    keys = sorted({k for fr in ABLATION.values() for _, o in fr for k in o} - {"leaf_pairs"})
    defaults = {k: r.get_option(k) for k in keys}
    pairs_default = r.get_option("leaf_pairs")
    built = {"pairs": None}

    def tree(pairs):
        if built["pairs"] != pairs:
            r.set_option("leaf_pairs", pairs)
            r.build(L.BUILD_REBUILD, args.bits)
            built["pairs"] = pairs
    frames = {
        "64spp": params(args.width, args.height, spp=args.spp, mode=L.MODE_AO, ao_distance=args.ao_distance,
                        background=(0.0, 0.0, 0.0, 1.0), flags=L.FLAG_NO_DOWNLOAD),
        "1spp": params(args.width, args.height, spp=1, mode=L.MODE_AO, ao_distance=args.ao_distance,
                       background=(0.0, 0.0, 0.0, 1.0), flags=L.FLAG_AUX | L.FLAG_NO_DOWNLOAD),
    }
    rows = []
    for fname, variants in ABLATION.items():
        p = frames[fname]
        tree(ABLATION_PAIRS[fname])
        r.set_option("count_visits", 1)
        r.render_async(cam, p)
        _, _, rays = r.counters()
        r.set_option("count_visits", 0)
        ref_id = r.download(args.width, args.height, outputs=("hit_id",)).hit_id.copy()
        for name, opts in variants:
            for k in keys:
                r.set_option(k, defaults[k])
            for k, v in opts.items():
                if k != "leaf_pairs":
                    r.set_option(k, v)
            tree(opts.get("leaf_pairs", ABLATION_PAIRS[fname]))
            r.render_async(cam, p)
            r.sync()
            same = bool(np.array_equal(r.download(args.width, args.height, outputs=("hit_id",)).hit_id, ref_id))
            tr = []
            for _ in range(args.steps):
                r.render_async(cam, p)
                tr.append(r.timings()["trace"])
            ms = float(np.median(tr))
            rows.append({"frame": fname, "variant": name, "options": opts, "trace_ms": round(ms, 3),
                         "mrays_per_s": round(rays / (ms * 1e-3) / 1e6, 1), "same_hit_ids": same})
            log(f"{fname} | {name}: {ms:.3f} ms, {rays / (ms * 1e-3) / 1e6:.0f} Mrays/s")
    for k in keys:
        r.set_option(k, defaults[k])
    r.set_option("leaf_pairs", pairs_default)
    print(json.dumps({"ablation": rows, "workload": f"{args.config}, {args.width}x{args.height}, AO distance "
                                                 f"{args.ao_distance}"}), flush=True)


BUILD_SWEEP = {"build_chunk": (1024, 512, 256), "sort_ri": (8, 16, 32), "sort_digit": (0, 8),
               "build_wave": (64, 0, 128, 256, 1024), "build_span": (256, 64, 128, 512, 1024),
               "leaf_pairs": (1, 0)}


def run_build_sweep(r, args):
    """LBVH build time (HIP events, median of `steps` rebuilds) for each build option changed one at a time from
    the library defaults, on the loaded scene."""
    from megamol_amd import _lib as L
    defaults = {k: r.get_option(k) for k in BUILD_SWEEP}
    rows = []

    def measure():
        r.build(L.BUILD_REBUILD, args.bits)
        r.sync()
        t = []
        for _ in range(max(3, args.steps)):
            r.build(L.BUILD_REBUILD, args.bits)
            t.append(r.timings())
        return {k: round(float(np.median([x[k] for x in t])), 3) for k in ("build", "morton", "sort", "boxes", "chunk")}
    rows.append({"options": dict(defaults), **measure()})
    log(f"build defaults {defaults}: {rows[-1]}")
    for k, vals in BUILD_SWEEP.items():
        for v in vals:
            if v == defaults[k]:
                continue
            r.set_option(k, v)
            rows.append({"options": {k: v}, **measure()})
            log(f"build {k}={v}: {rows[-1]}")
            r.set_option(k, defaults[k])
    print(json.dumps({"build_sweep": rows, "spheres": r.info()[0], "bits": args.bits}), flush=True)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def scenes_mod():
    from megamol_amd import scenes
    return scenes


def make_scene(args):
    from megamol_amd import ParticleList, _lib as L, scenes

    if args.config == "fcc":
        t0 = time.time()
        xyzr, inten = scenes.fcc_crystal(args.cells)
        log(f"generated {len(xyzr)} FCC spheres in {time.time() - t0:.1f}s")
        tf = scenes.transfer_function_256()
        lists = [ParticleList(L.VERTDATA_FLOAT_XYZR, xyzr, L.COLDATA_FLOAT_I, inten, count=len(xyzr),
                              col_min=0.0, col_max=1.0)]
        cam = scenes.fcc_camera(args.cells, width=args.width, height=args.height)
        return dict(lists=lists, tf=tf, xyzr=xyzr, inten=inten, cam=cam, workload=(
            f"FCC crystal {len(xyzr)} spheres (cells {args.cells}), {args.width}x{args.height}, AO {args.spp} spp, "
            f"AO distance {args.ao_distance}, LBVH rebuild per frame"))
    if args.config == "clusters":
        t0 = time.time()
        n = args.n if args.n != 1_000_000 else 500_000_000
        xyzr, cl = scenes.clusters(n, with_cluster=True)
        inten = (cl.astype(np.float32) * np.float32(1.0 / 10_000))
        del cl
        log(f"generated {len(xyzr)} clustered spheres in {time.time() - t0:.1f}s")
        tf = scenes.transfer_function_256()
        lists = [ParticleList(L.VERTDATA_FLOAT_XYZR, xyzr, L.COLDATA_FLOAT_I, inten, count=len(xyzr),
                              col_min=0.0, col_max=1.0)]
        cam = scenes.clusters_camera(width=args.width, height=args.height)
        return dict(lists=lists, tf=tf, xyzr=xyzr, inten=inten, cam=cam, workload=(
            f"clustered scene {len(xyzr)} spheres (10k Gaussian clusters in a 1e4 box), {args.width}x{args.height}, "
            f"AO {args.spp} spp, AO distance {args.ao_distance}, LBVH rebuild per frame"))
    xyzr, rgba8 = scenes.uniform_cube(args.n)
    lists = [ParticleList(L.VERTDATA_FLOAT_XYZR, xyzr, L.COLDATA_UINT8_RGBA, rgba8, count=len(xyzr))]
    cam = scenes.uniform_cube_camera(args.width, args.height)
    return dict(lists=lists, tf=None, xyzr=xyzr, rgba8=rgba8, cam=cam, workload=(
        f"uniform cube {len(xyzr)} spheres, {args.width}x{args.height}, mode {args.mode}, {args.spp} spp"))


def cpu_baseline(scene, args, r):
    """CPU oracle on a bounded sample of the same frame: the full LBVH build (multi-threaded) and
    `cpu_rows` rows spread evenly over the image, each timed `cpu_repeats` times (median).

    Threads: every CPU this process may run on (its affinity mask), which on the GPU hosts is the
    16-CPU share of one GPU; the machine's total count is reported beside it."""
    import oracle as O

    threads = O.default_threads()
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = os.cpu_count() or 1
    xyzr, rgba = O.convert_lists(scene["lists"], tf=scene["tf"])
    builds = []
    b = None
    for _ in range(2):
        b = None
        b = O.Bvh(xyzr, args.bits, threads=threads)
        builds.append(b.build_s)
    build_s = float(np.median(builds))
    nrows = max(1, min(args.cpu_rows, args.height))
    rows = [int((k + 0.5) * args.height / nrows) for k in range(nrows)]
    p = O.make_params(args.width, args.height, args.spp, 1 if args.mode == "ao" else 0, args.ao_distance)
    runs = []
    for _ in range(max(1, args.cpu_repeats)):
        tot = 0.0
        for y in rows:
            p.row_begin, p.row_end = y, y + 1
            tot += b.render(rgba, scene["cam"], p, threads=threads, outputs=("hit_id",))["seconds"]
        runs.append(tot)
    trace_s = float(np.median(runs))
    sample_px = nrows * args.width
    frac = sample_px / (args.width * args.height)
    frame_ms = 1e3 * (build_s + trace_s / frac)
    return dict(value=round(frame_ms, 1), unit="ms/frame (projected from sample)", cores=threads,
                kind="port",
                sample=(f"oracle LBVH build over all {len(xyzr)} spheres (median of {len(builds)}: {build_s:.2f}s, "
                        f"{threads} threads) + {nrows} rows spread over the frame x {args.width} px x {args.spp} spp "
                        f"(median of {len(runs)} runs: {trace_s:.2f}s, {threads} threads); full frame projected as "
                        f"build + trace/{frac:.4f}; CPUs: {affinity} in this process's affinity mask, "
                        f"{os.cpu_count()} on the machine"),
                build_ms=round(1e3 * build_s, 1), trace_ms_projected=round(1e3 * trace_s / frac, 1),
                trace_seconds=trace_s, build_seconds=build_s, sample_pixels=sample_px)


def denoised_leg(r, scene, args, cam, p_plain, plain_cfg=None):
    """1-spp AO + learned denoiser, timed like the plain frame, and the image quality of both
    against a 256-spp frame of the same view (independent random stream)."""
    from megamol_amd import _lib as L, params
    from megamol_amd.denoiser import Denoiser, psnr, ssim

    dnz = Denoiser(r, seed=1234)
    wpath = args.weights
    trained = "loaded"
    meta = None
    if wpath and os.path.exists(wpath):
        try:
            meta = dnz.load(wpath)
        except ValueError as e:
            log(f"weights file {wpath} unusable ({e})")
    if meta is None:
        from megamol_amd import train_denoiser as T
        log("no denoiser weights: training a small model on this scene's pairs (other cameras)")
        f, t = T.make_pairs(r, args.train_pairs, args.cells, 99, crop=128, spp_target=256)
        r.build(L.BUILD_REBUILD, args.bits)
        T.train(dnz, f, t, args.train_steps, log_every=0)
        meta, trained = {"pairs": args.train_pairs, "steps": args.train_steps}, "trained in bench"
    lo, hi = scene["xyzr"][:, :3].min(0), scene["xyzr"][:, :3].max(0)
    centre = 0.5 * (lo + hi)
    depth_scale = float(np.linalg.norm(np.array(cam.position[:]) - centre))
    p1 = params(args.width, args.height, spp=1, mode=L.MODE_AO, ao_distance=args.ao_distance,
                flags=L.FLAG_AUX | L.FLAG_NO_DOWNLOAD)

    def frame():
        r.build(L.BUILD_REBUILD, args.bits)
        r.render_async(cam, p1)
        dnz.denoise_frame(depth_scale)

    tuned1 = None
    if not args.no_autotune and not args.option:
        # the 1-spp frame has its own best tree (leaf pairs) and launch (lane-per-pixel kernel,
        # refill threshold)
        from megamol_amd import autotune
        pairs1, pairs1_ms = autotune.tune_tree(r, cam, p1, args.bits)
        r.build(L.BUILD_REBUILD, args.bits)
        cfg1, best1, _ = autotune.tune_frame(r, cam, p1)
        cfg1["leaf_pairs"] = pairs1
        tuned1 = {"config": cfg1, "trace_ms": round(best1, 3),
                  "leaf_pairs_build_plus_trace_ms": {str(k): round(v, 3) for k, v in pairs1_ms.items()}}
        log(f"autotuned 1-spp launch: {cfg1} ({best1:.3f} ms trace)")

    for _ in range(args.warmup):
        frame()
    r.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        frame()
    r.sync()
    ms = 1e3 * (time.perf_counter() - t0) / args.steps
    # the pieces: build + 1-spp trace without the denoiser
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r.build(L.BUILD_REBUILD, args.bits)
        r.render_async(cam, p1)
    r.sync()
    ms_nodn = 1e3 * (time.perf_counter() - t0) / args.steps
    # per-kernel durations of the frame's build and trace (HIP events), one frame at a time
    kt = {"build": [], "chunk": [], "trace": []}
    for _ in range(args.steps):
        r.build(L.BUILD_REBUILD, args.bits)
        r.render_async(cam, p1)
        t = r.timings()
        for k in kt:
            kt[k].append(t[k])
    kms = {k: float(np.mean(v)) for k, v in kt.items()}
    n = len(scene["xyzr"])
    pairs = r.get_option("leaf_pair_count")
    stored = n - 1 - pairs
    # the chunk pass reads per leaf its particle id, its key and its sphere (gathered), and writes
    # the leaf sphere and every stored node (64 bytes; leaf pairs are not stored)
    chunk_bytes = (4 + 4 + 16) * n + 16 * n + 64 * stored
    chunk = {"kernel": "k_bu_chunk_ls (hierarchy chunk pass)", "ms": round(kms["chunk"], 3),
             "algorithmic_bytes": int(chunk_bytes), "leaf_pairs": int(pairs), "stored_nodes": int(stored),
             "build_ms": round(kms["build"], 3), "trace_ms": round(kms["trace"], 3)}
    log(f"denoised frame kernels: {chunk}")
    # image quality against 256 spp (frame index 7777: independent of both frames' streams)
    W, H = args.width, args.height

    def colour(p):
        q = L.RenderParams.from_buffer_copy(p)
        q.flags &= ~L.FLAG_NO_DOWNLOAD
        return r.render(cam, q, outputs=("color",)).color[..., :3].copy()

    ref = colour(params(W, H, spp=256, mode=L.MODE_AO, ao_distance=args.ao_distance, frame_idx=7777))
    plain = colour(params(W, H, spp=args.spp, mode=L.MODE_AO, ao_distance=args.ao_distance))
    # quality against cost of the plain path: PSNR of 1 ... 128 spp against the 256-spp frame, and
    # each frame's time (LBVH rebuild + trace, the plain frame's tree and launch)
    curve = []
    if not args.no_curve:
        # the plain frame's launch and tree (its autotuned configuration; refill back to the
        # library default: compaction for 1-spp frames only)
        saved = {k: r.get_option(k) for k in (plain_cfg or {})}
        saved["refill"] = r.get_option("refill")
        for k, v in (plain_cfg or {}).items():
            r.set_option(k, v)
        r.set_option("refill", -1)
        for spp_k in (1, 2, 4, 8, 16, 32, 64, 128):
            pk = params(W, H, spp=spp_k, mode=L.MODE_AO, ao_distance=args.ao_distance, flags=L.FLAG_NO_DOWNLOAD)
            for _ in range(2):
                r.build(L.BUILD_REBUILD, args.bits)
                r.render_async(cam, pk)
            r.sync()
            t0 = time.perf_counter()
            reps = max(2, min(args.steps, 6))
            for _ in range(reps):
                r.build(L.BUILD_REBUILD, args.bits)
                r.render_async(cam, pk)
            r.sync()
            ms_k = 1e3 * (time.perf_counter() - t0) / reps
            img = colour(params(W, H, spp=spp_k, mode=L.MODE_AO, ao_distance=args.ao_distance))
            curve.append({"spp": spp_k, "ms": round(ms_k, 3), "psnr_db": round(psnr(img, ref), 2)})
            log(f"plain {spp_k} spp: {ms_k:.2f} ms/frame, {curve[-1]['psnr_db']} dB")
        for k, v in saved.items():
            r.set_option(k, v)
        r.build(L.BUILD_REBUILD, args.bits)
    raw = colour(params(W, H, spp=1, mode=L.MODE_AO, ao_distance=args.ao_distance, flags=L.FLAG_AUX))
    raw_aux = r.download(W, H, outputs=(), aux=True).aux
    from megamol_amd.denoiser import box_filter_baseline
    box = box_filter_baseline(raw_aux, (0.0, 0.0, 0.0))
    dnz.denoise_frame(depth_scale)
    den = r.download(W, H, outputs=("color",)).color[..., :3].copy()
    res = {"ms": round(ms, 3), "ms_without_denoiser": round(ms_nodn, 3),
           "denoiser_ms": round(ms - ms_nodn, 3),
           "plain_over_denoised": None,
           "psnr_db_vs_256spp": {"denoised_1spp": round(psnr(den, ref), 2), "plain": round(psnr(plain, ref), 2),
                                 "raw_1spp": round(psnr(raw, ref), 2), "box5_filter_1spp": round(psnr(box, ref), 2)},
           "ssim_vs_256spp": {"denoised_1spp": round(ssim(den, ref), 4), "plain": round(ssim(plain, ref), 4),
                              "raw_1spp": round(ssim(raw, ref), 4), "box5_filter_1spp": round(ssim(box, ref), 4)},
           "weights": trained, "weights_meta": meta, "autotune_1spp": tuned1, "chunk_pass": chunk}
    if args.footprint > 0:
        res["footprint"] = footprint_variant(r, cam, args, ref, raw_aux, den, ms, colour)
    if curve:
        # the plain frame of equal quality: PSNR interpolated linearly in log2(spp) between the two
        # measured counts around the denoised frame's PSNR, and the time the same way
        dpsnr = res["psnr_db_vs_256spp"]["denoised_1spp"]
        eq = None
        for a, b in zip(curve, curve[1:]):
            if a["psnr_db"] <= dpsnr <= b["psnr_db"]:
                w = (dpsnr - a["psnr_db"]) / max(b["psnr_db"] - a["psnr_db"], 1e-9)
                spp_eq = 2.0 ** (np.log2(a["spp"]) + w)
                ms_eq = a["ms"] + w * (b["ms"] - a["ms"])
                eq = {"spp": round(float(spp_eq), 2), "plain_ms": round(float(ms_eq), 3),
                      "plain_over_denoised": round(float(ms_eq) / ms, 2), "between_spp": [a["spp"], b["spp"]]}
                break
        if eq is None and curve[0]["psnr_db"] >= dpsnr:
            eq = {"spp": 1, "plain_ms": curve[0]["ms"], "plain_over_denoised": round(curve[0]["ms"] / ms, 2),
                  "note": "the 1-spp plain frame already reaches the denoised PSNR"}
        res["plain_psnr_curve"] = curve
        res["equal_psnr"] = eq
    log(f"denoised frame {ms:.2f} ms (no denoiser {ms_nodn:.2f}); PSNR {res['psnr_db_vs_256spp']}")
    if args.save_images:
        # centre crops for offline error analysis: the 256-spp mean albedo (AO distance ~0: every
        # sample unoccluded) and the 1-spp feature records next to the colour images
        alb = colour(params(W, H, spp=256, mode=L.MODE_AO, ao_distance=1e-7, frame_idx=7777))
        from megamol_amd.denoiser import features_from_frame
        feats = features_from_frame(np.concatenate([raw, np.ones_like(raw[..., :1])], -1), raw_aux)
        c = args.crop
        y0, x0 = (H - c) // 2, (W - c) // 2
        sl = (slice(y0, y0 + c), slice(x0, x0 + c))
        np.savez_compressed(args.save_images, ref=ref[sl].astype(np.float16), plain=plain[sl].astype(np.float16),
                            raw=raw[sl].astype(np.float16), den=den[sl].astype(np.float16),
                            alb256=alb[sl].astype(np.float16), feats=feats[sl])
    return res


def footprint_variant(r, cam, args, ref, raw_aux, den, ms, colour):
    """The albedo-aliasing ceiling of a one-sample frame and a variant past it: a primary-only
    pass of K jittered samples per pixel (AO distance ~0, so every sample on a sphere is
    unoccluded and the pixel colour is the mean over the samples of the albedo, or the black
    background on a miss: the pixel's footprint albedo with its coverage) replaces the one-sample
    albedo the denoised frame is modulated with. The network is not retrained: its output is
    divided by the sample-0 albedo (the visibility it estimates) and multiplied by the footprint
    albedo; sample-0 misses take the mean estimated visibility of the hit pixels in their 5x5
    window. Reported beside the strict 1-spp frame with the footprint pass's cost added."""
    from scipy.ndimage import uniform_filter
    from megamol_amd import _lib as L, params
    from megamol_amd.denoiser import psnr
    W, H, K = args.width, args.height, args.footprint
    pf = params(W, H, spp=K, mode=L.MODE_AO, ao_distance=1e-7, frame_idx=4242, flags=L.FLAG_NO_DOWNLOAD)
    for _ in range(2):
        r.render_async(cam, pf)
    r.sync()
    t0 = time.perf_counter()
    reps = max(2, min(args.steps, 6))
    for _ in range(reps):
        r.render_async(cam, pf)
    r.sync()
    fp_ms = 1e3 * (time.perf_counter() - t0) / reps
    fp = colour(pf)
    # the reference's footprint albedo (256 samples) and, where it is nonzero, the reference's
    # visibility (the 256-spp AO colour over it)
    alb256 = colour(params(W, H, spp=256, mode=L.MODE_AO, ao_distance=1e-7, frame_idx=7777))
    a = raw_aux.astype(np.float32)
    hit = a[..., 6] > 0
    alb1 = a[..., 0:3] * hit[..., None]
    eps = 1e-4
    vref = np.where(alb256 > eps, ref / np.maximum(alb256, eps), 0.0)
    vnet = np.where(alb1 > eps, den / np.maximum(alb1, eps), 0.0).mean(-1)
    vnet = np.clip(vnet, 0.0, 1.0) * hit
    num = uniform_filter(vnet, 5, mode="constant")
    cnt = uniform_filter(hit.astype(np.float32), 5, mode="constant")
    vfill = np.where(cnt > 0, num / np.maximum(cnt, 1e-6), 0.0)
    v = np.where(hit, vnet, vfill)
    variant = fp * v[..., None]
    out = {"samples": K, "footprint_pass_ms": round(fp_ms, 3),
           "frame_ms": round(ms + fp_ms, 3),
           "psnr_db_vs_256spp": {
               "one_sample_albedo_x_reference_visibility": round(psnr(alb1 * vref, ref), 2),
               "footprint_albedo_x_reference_visibility": round(psnr(fp * vref, ref), 2),
               "variant_footprint_albedo_x_network_visibility": round(psnr(variant, ref), 2),
               "strict_denoised_1spp": round(psnr(den, ref), 2)}}
    log(f"footprint variant: {out}")
    return out


def pipelined_frames(scene, args, cam, opts_plain, opts_dn, weights, depth_scale):
    """Frames in flight: K contexts on the one GPU, each with the scene resident and its own stream,
    take the frames of a sequence in turn, so that one frame's LBVH build (latency-bound rounds)
    runs beside another's trace and denoiser. Returns ms per frame for the plain 64-spp frame and
    the denoised 1-spp frame, each frame still a full rebuild + trace (+ denoise)."""
    from megamol_amd import SphereRenderer, _lib as L, params
    from megamol_amd.denoiser import Denoiser

    K = args.pipelined
    ctxs = []
    for _ in range(K):
        r = SphereRenderer(0)
        if scene["tf"] is not None:
            r.set_transfer_function(scene["tf"])
        r.set_lists(scene["lists"])
        d = Denoiser(r, seed=1234)
        if weights:
            d.load(weights)
        ctxs.append((r, d))
    p64 = params(args.width, args.height, spp=args.spp, mode=L.MODE_AO, ao_distance=args.ao_distance,
                 background=(0.0, 0.0, 0.0, 1.0), flags=L.FLAG_NO_DOWNLOAD)
    p1 = params(args.width, args.height, spp=1, mode=L.MODE_AO, ao_distance=args.ao_distance,
                flags=L.FLAG_AUX | L.FLAG_NO_DOWNLOAD)
    out = {"frames_in_flight": K}
    for name, p, opts, dn in (("plain", p64, opts_plain, False), ("denoised", p1, opts_dn, True)):
        for r, _ in ctxs:
            for k, v in (opts or {}).items():
                r.set_option(k, v)

        def frame(i):
            r, d = ctxs[i % K]
            r.build(L.BUILD_REBUILD, args.bits)
            r.render_async(cam, p)
            if dn:
                d.denoise_frame(depth_scale)
        for i in range(2 * K):
            frame(i)
        for r, _ in ctxs:
            r.sync()
        n = max(args.steps, 4) * K
        t0 = time.perf_counter()
        for i in range(n):
            frame(i)
        for r, _ in ctxs:
            r.sync()
        out[f"{name}_ms_per_frame"] = round(1e3 * (time.perf_counter() - t0) / n, 3)
        log(f"pipelined {name}: {out[f'{name}_ms_per_frame']} ms/frame with {K} frames in flight")
    out["plain_over_denoised"] = round(out["plain_ms_per_frame"] / out["denoised_ms_per_frame"], 2)
    for r, _ in ctxs:
        r.close()
    return out


def structure_compare(r, args, cam):
    """The LBVH against the treelet P-k-d trees on the same scene: build time and 1-spp AO trace
    time (HIP events), node visits per ray, and the device memory each structure keeps beyond
    the spheres (both keep the spheres in their own order, 16 B, and a particle id, 4 B, per
    sphere)."""
    from megamol_amd import _lib as L, params

    W, H = args.width, args.height
    p1 = params(W, H, spp=1, mode=L.MODE_AO, ao_distance=args.ao_distance, flags=L.FLAG_NO_DOWNLOAD)
    n = r.info()[0]
    out, ids = {}, {}
    for name, accel in (("lbvh", 0), ("treelets", 1)):
        r.set_option("accel", accel)
        bt, tt = [], []
        for i in range(max(2, args.steps) + 1):
            r.build(L.BUILD_REBUILD, args.bits)
            r.render_async(cam, p1)
            t = r.timings()
            if i:
                bt.append(t["build"])
                tt.append(t["trace"])
        r.set_option("count_visits", 1)
        r.render_async(cam, p1)
        visits, tests, rays = r.counters()
        r.set_option("count_visits", 0)
        ids[name] = r.download(W, H, outputs=("hit_id",)).hit_id.copy()
        trace_ms = float(np.median(tt))
        if accel == 0:
            pairs = r.get_option("leaf_pair_count")
            mem = {"node_slots_bytes": 64 * (n - 1), "stored_node_bytes": 64 * (n - 1 - pairs)}
        else:
            M = r.get_option("treelet")
            T = (n + M - 1) // M
            Lt = 2
            while Lt < T:
                Lt *= 2
            mem = {"top_node_bytes": 64 * (Lt - 1), "split_axis_bytes": n, "treelet_record_bytes": 28 * T}
        mem["total_beyond_spheres_and_ids"] = (mem.get("node_slots_bytes", 0) + mem.get("top_node_bytes", 0)
                                              + mem.get("split_axis_bytes", 0) + mem.get("treelet_record_bytes", 0))
        out[name] = {"build_ms": round(float(np.median(bt)), 3), "trace_1spp_ms": round(trace_ms, 3),
                     "mrays_per_s": round(rays / (trace_ms * 1e-3) / 1e6, 1), "visits_per_ray": round(visits / rays, 2),
                     "tests_per_ray": round(tests / rays, 2), "memory": mem}
        log(f"structure {name}: {out[name]}")
    r.set_option("accel", 0)
    out["same_hit_ids"] = bool(np.array_equal(ids["lbvh"], ids["treelets"]))
    return out


def run_series(args):
    """Batched time series: timesteps x 1-6 cameras (640x480 .. 1920x1080) in one hipGraph launch
    (every timestep's structure, then one trace launch over all jobs), refit/rebuild policy and
    launch knobs picked by the autotuner; the same series with one trace launch per timestep is
    timed beside it."""
    from megamol_amd import SphereRenderer, _lib as L, autotune, params, scenes

    t0 = time.time()
    x = scenes.time_series_initial(args.series_n)
    steps = [x]
    for k in range(args.series_t - 1):
        steps.append(scenes.time_series_step(steps[-1], k))
    xs = np.stack(steps)
    del steps
    log(f"generated {args.series_t} x {args.series_n} spheres in {time.time() - t0:.1f}s")
    jobs = []
    mode = L.MODE_AO if args.mode == "ao" else L.MODE_LAMBERT
    for t in range(args.series_t):
        for cam, w, h in scenes.time_series_cameras(t):
            jobs.append((t, cam, params(w, h, spp=args.spp, mode=mode, ao_distance=args.ao_distance)))
    px = sum(p.width * p.height for _, _, p in jobs)
    r = SphereRenderer(0)
    r.set_timesteps(xs, None, (200, 200, 220, 255))
    cfg = {"refit_period": 1}
    if args.autotune:
        cfg, _, meta = autotune.tune_batch(r, jobs)
        log(f"autotuned: {cfg}")
    for _ in range(args.warmup):
        r.render_batch(jobs, refit_period=cfg["refit_period"], use_graph=True)
    ms_all = []
    for _ in range(args.steps):
        _, ms = r.render_batch(jobs, refit_period=cfg["refit_period"], use_graph=True)
        ms_all.append(ms)
    ms = float(np.median(ms_all))
    # the same series with one trace launch per timestep (right after its build) instead of one
    # launch over all timesteps
    r.set_option("batch_launch", 0)
    ms_per_ts = []
    for _ in range(max(1, args.steps)):
        _, m2 = r.render_batch(jobs, refit_period=cfg["refit_period"], use_graph=True)
        ms_per_ts.append(m2)
    r.set_option("batch_launch", 1)
    ms_ts = float(np.median(ms_per_ts))
    log(f"one launch over all timesteps {ms:.2f} ms, one launch per timestep {ms_ts:.2f} ms")
    rays = px * args.spp * (2 if mode == L.MODE_AO else 1)
    out = {"metric": "batched series ms per launch (timesteps x cameras in one hipGraph)", "value": round(ms, 3),
           "unit": "ms/launch", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded Gaussian-step time series)",
           "config": {"workload": f"time series {args.series_t} x {args.series_n} spheres, {len(jobs)} jobs, "
                                  f"{px} pixels, mode {args.mode}, {args.spp} spp",
                      "autotuned": cfg},
           "frames": len(jobs), "ms_per_frame": round(ms / len(jobs), 3),
           "launch_per_timestep_ms": round(ms_ts, 3),
           "mrays_per_s_upper": round(rays / (ms * 1e-3) / 1e6, 1)}
    print(json.dumps(out), flush=True)
    r.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", choices=["fcc", "cube", "series", "clusters"], default="fcc")
    ap.add_argument("--orbit", type=int, default=0,
                    help="clusters config: also render an orbit of this many cameras at 1 spp in one batched launch")
    ap.add_argument("--series-n", type=int, default=4_000_000)
    ap.add_argument("--series-t", type=int, default=64)
    ap.add_argument("--autotune", action="store_true", help="series config: tune refit period and launch knobs")
    ap.add_argument("--footprint", type=int, default=8,
                    help="samples of the primary-only footprint pass of the labelled variant (0: skip it)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="fcc/cube configs: keep the library's default launch knobs instead of tuning them")
    ap.add_argument("--no-denoise", action="store_true")
    ap.add_argument("--no-curve", action="store_true", help="skip the plain path's PSNR-vs-spp curve")
    ap.add_argument("--structures", action="store_true",
                    help="also compare the LBVH with the treelet P-k-d trees (build, 1-spp AO trace, memory)")
    ap.add_argument("--weights", type=str,
                    default=os.path.join(ROOT, "megamol_amd", "weights", "denoiser_fcc.bin"))
    ap.add_argument("--train-pairs", type=int, default=256)
    ap.add_argument("--train-steps", type=int, default=600)
    ap.add_argument("--save-images", type=str, default="", help="save centre crops of the denoised leg's images")
    ap.add_argument("--crop", type=int, default=512)
    ap.add_argument("--cells", type=int, default=232)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--mode", choices=["ao", "lambert"], default="ao")
    ap.add_argument("--ao-distance", type=float, default=4.0)
    ap.add_argument("--bits", type=int, default=30)
    ap.add_argument("--cpu-rows", type=int, default=12, help="CPU baseline: rows spread evenly over the frame")
    ap.add_argument("--cpu-repeats", type=int, default=5, help="CPU baseline: timed runs of the rows (median)")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--option", action="append", default=[], help="key=value library option")
    ap.add_argument("--ablation", action="store_true",
                    help="only trace the 64-spp and 1-spp AO frames once per traversal variant (one choice "
                         "changed at a time) and print their Mrays/s")
    ap.add_argument("--pipelined", type=int, default=0, metavar="K",
                    help="also time both frames with K frames in flight (K contexts on the GPU, one stream each)")
    ap.add_argument("--build-sweep", action="store_true",
                    help="only time the LBVH build with each build option changed one at a time")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB five times (calibrates the HBM "
                         "byte counters of a profiler against a kernel of known traffic)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        n = args.copy_probe * (1 << 20) // 4
        src = torch.arange(n, dtype=torch.int32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": 4 * n, "copy_probe_bytes_written": 4 * n, "launches": 5}))
        return
    if args.config == "cube" and args.width == 1920:
        args.width = args.height = 1024
    if args.config == "clusters" and args.width == 1920:
        args.width, args.height = 3840, 2160
    if args.config == "clusters" and args.bits == 30:
        args.bits = 63  # 30-bit codes put thousands of clustered spheres on one key (10-unit cells)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the sphere path runs on a single MI355X in this build "
              "(multi-GPU image-tile sharding is designed, not built). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from megamol_amd import SphereRenderer, _lib as L, params

    if args.config == "series":
        if args.spp == 64:
            args.spp = 1
        run_series(args)
        return
    scene = make_scene(args)
    r = SphereRenderer(0)
    for kv in args.option:
        k, v = kv.split("=")
        r.set_option(k, int(v))
    if scene["tf"] is not None:
        r.set_transfer_function(scene["tf"])
    t0 = time.time()
    r.set_lists(scene["lists"])
    log(f"uploaded + converted in {time.time() - t0:.2f}s")
    mode = L.MODE_AO if args.mode == "ao" else L.MODE_LAMBERT
    p = params(args.width, args.height, spp=args.spp, mode=mode, ao_distance=args.ao_distance,
               background=(0.0, 0.0, 0.0, 1.0), flags=L.FLAG_NO_DOWNLOAD)
    cam = scene["cam"]
    if args.ablation:
        run_ablation(r, cam, args)
        r.close()
        return
    if args.build_sweep:
        run_build_sweep(r, args)
        r.close()
        return

    def frame():
        r.build(L.BUILD_REBUILD, args.bits)
        r.render_async(cam, p)

    tuned = None
    if not args.no_autotune and not args.option:
        from megamol_amd import autotune
        pairs, pairs_ms = autotune.tune_tree(r, cam, p, args.bits)
        r.build(L.BUILD_REBUILD, args.bits)
        cfg, best, cache = autotune.tune_frame(r, cam, p)
        cfg["leaf_pairs"] = pairs
        tuned = {"config": cfg, "trace_ms": round(best, 3),
                 "leaf_pairs_build_plus_trace_ms": {str(k): round(v, 3) for k, v in pairs_ms.items()},
                 "candidates_trace_ms": {k: round(v, 3) for k, v in cache.items()}}
        log(f"autotuned launch: {cfg} ({best:.2f} ms trace)")

    for i in range(args.warmup):
        frame()
        r.sync()
        log(f"warmup frame {i} done: {r.timings()}")
    r.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        frame()
    r.sync()
    t1 = time.perf_counter()
    ms = 1e3 * (t1 - t0) / args.steps
    log(f"timed {args.steps} frames: {ms:.2f} ms/frame")

    # per-kernel durations (HIP events on the library stream), one frame at a time
    tr, bu, so = [], [], []
    stages = {k: [] for k in ("morton", "sort", "hierarchy", "boxes")}
    for _ in range(args.steps):
        frame()
        t = r.timings()
        tr.append(t["trace"])
        bu.append(t["build"])
        so.append(t["sort"])
        for k in stages:
            stages[k].append(t[k])
    trace_ms = float(np.mean(tr))
    build_ms = float(np.mean(bu))
    build_stages = {k: round(float(np.mean(v)), 3) for k, v in stages.items()}
    log(f"kernel times: trace {trace_ms:.3f} ms, build {build_ms:.3f} ms {build_stages}")

    # instrumented frame: node visits, sphere tests, rays (algorithmic bytes of the trace kernel)
    r.set_option("count_visits", 1)
    frame()
    visits, tests, rays = r.counters()
    overflow = r.get_option("stack_overflow")
    pair_visits, leaf_fetches = r.get_option("pair_visits"), r.get_option("leaf_fetches")
    r.set_option("count_visits", 0)
    fr = r.download(args.width, args.height, outputs=("hit_id",))
    hit_px = int((fr.hit_id >= 0).sum())
    P = args.width * args.height
    hit_samples = rays - P * args.spp if mode == L.MODE_AO else hit_px * args.spp
    # a stored node is 64 bytes, a leaf pair's two spheres 32, a leaf child's sphere 16; shading
    # reads each hit sphere (16) and the frame writes 32 bytes per pixel
    algo_bytes = (64 * (visits - pair_visits) + 32 * pair_visits + 16 * leaf_fetches
                  + 16 * max(hit_samples, 0) + 32 * P)
    achieved = algo_bytes / (trace_ms * 1e-3) / 1e9
    kernel_name = trace_kernel_label(r, args.spp, args.mode)
    traffic = (dram_traffic(f"cells {args.cells}, {args.width}x{args.height}, spp {args.spp}, mode {args.mode}")
               if args.config == "fcc" else None)
    mrays = rays / (ms * 1e-3) / 1e6
    log(f"visits {visits} (leaf pairs {pair_visits}) tests {tests} leaf fetches {leaf_fetches} rays {rays} "
        f"overflow {overflow} hit px {hit_px}")
    split = None
    if mode == L.MODE_AO:
        # the same primary rays alone (Lambert shading): the AO rays' share by difference
        pl = L.RenderParams.from_buffer_copy(p)
        pl.mode = L.MODE_LAMBERT
        r.set_option("count_visits", 1)
        r.render_async(cam, pl)
        pv, pt, pr = r.counters()
        r.set_option("count_visits", 0)
        ao_rays = max(rays - pr, 1)
        split = {"primary": {"rays": int(pr), "visits_per_ray": round(pv / max(pr, 1), 2),
                             "tests_per_ray": round(pt / max(pr, 1), 2)},
                 "ao": {"rays": int(ao_rays), "visits_per_ray": round((visits - pv) / ao_rays, 2),
                        "tests_per_ray": round((tests - pt) / ao_rays, 2)}}
        log(f"per-ray work: {split}")

    dn = None
    if mode == L.MODE_AO and args.config in ("fcc", "clusters") and not args.no_denoise:
        dn = denoised_leg(r, scene, args, cam, p, tuned["config"] if tuned else None)
        dn["plain_over_denoised"] = round(ms / dn["ms"], 2)
        dn["target_8x_at_38db"] = bool(dn["plain_over_denoised"] >= 8.0 and
                                       dn["psnr_db_vs_256spp"]["denoised_1spp"] >= 38.0)

    pipelined = None
    if args.pipelined > 1 and dn is not None:
        lo, hi = scene["xyzr"][:, :3].min(0), scene["xyzr"][:, :3].max(0)
        ds = float(np.linalg.norm(np.array(cam.position[:]) - 0.5 * (lo + hi)))
        opts_dn = (dn.get("autotune_1spp") or {}).get("config")
        pipelined = pipelined_frames(scene, args, cam, tuned["config"] if tuned else None, opts_dn,
                                     args.weights if args.weights and os.path.exists(args.weights) else None, ds)

    structures = None
    if args.structures:
        structures = structure_compare(r, args, cam)

    orbit = None
    if args.config == "clusters" and args.orbit > 0:
        # the orbit path: every camera at 1 spp in one batched persistent launch captured in a
        # hipGraph (one LBVH build for the static scene, then all traversal jobs)
        cams = scenes_mod().orbit_cameras(args.orbit, width=args.width, height=args.height)
        r.set_timesteps(scene["xyzr"][None], None, (200, 200, 220, 255))
        po = params(args.width, args.height, spp=1, mode=L.MODE_AO, ao_distance=args.ao_distance)
        jobs = [(0, c, po) for c in cams]
        r.render_batch(jobs, refit_period=1, use_graph=True)  # warm-up: allocations, graph
        _, oms = r.render_batch(jobs, refit_period=1, use_graph=True)
        orbit = {"frames": args.orbit, "spp": 1, "ms_per_launch": round(oms, 3),
                 "ms_per_frame": round(oms / args.orbit, 3),
                 "note": "one launch: LBVH build of the timestep + all orbit frames, hipGraph"}
        log(f"orbit: {orbit}")

    cpu = None
    if not args.skip_cpu:
        log("CPU baseline ...")
        cpu = cpu_baseline(scene, args, r)
        cpu_mrays = (cpu["sample_pixels"] * (rays / P)) / cpu["trace_seconds"] / 1e6
        cpu["sample"] += f"; {cpu_mrays:.2f} Mrays/s"
        cpu["mrays_per_s"] = round(cpu_mrays, 2)
        for k in ("trace_seconds", "build_seconds", "sample_pixels"):
            cpu.pop(k)
        log(f"CPU baseline: {cpu}")

    # the plain frame's trace kernel against the roof it runs into: VALU instruction issue (its node
    # fetches are mostly L2 / MALL hits, so algorithmic bytes over HBM would exceed 1)
    valu = (dram_traffic(f"cells {args.cells}, 1920x1080, spp {args.spp}, mode {args.mode}", VALU_PROFILE,
                         "valu_instructions_per_launch") if args.config == "fcc" else None)
    plain_trace_roof = {"kernel": kernel_name, "algorithmic_bytes": int(algo_bytes),
                        "node_and_sphere_gbs": round(achieved, 1), "dram_traffic": traffic,
                        "dram_frac_of_hbm": (round(traffic / (trace_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                                             if traffic else None)}
    if valu:
        ach = valu / (trace_ms * 1e-3)
        plain_trace_roof.update({"bound": "valu_issue", "achieved": round(ach / 1e9, 1),
                                 "peak": round(VALU_ISSUE_PEAK / 1e9, 1), "unit": "G wave-instructions/s",
                                 "frac": round(ach / VALU_ISSUE_PEAK, 4), "valu_instructions": int(valu)})
    value = dn["ms"] if dn else ms
    # the host boundary, outside `value` (whose inputs are resident in HBM): the MultiParticleDataCall
    # lists uploaded again (pageable host memory -> device, decoded there; what a data change costs)
    # and the CPU framebuffer (RGBA8 + depth) downloaded, each timed on the wall clock around a
    # synchronised call
    host_io = None
    if args.config == "fcc":
        ups, downs = [], []
        for _ in range(3):
            t0 = time.perf_counter()
            r.set_lists(scene["lists"], data_hash=1)
            ups.append((time.perf_counter() - t0) * 1e3)
        r.build(L.BUILD_REBUILD, args.bits)
        r.render(cam, p, outputs=())
        for _ in range(5):
            t0 = time.perf_counter()
            r.download(args.width, args.height, outputs=("rgba8", "depth"))
            downs.append((time.perf_counter() - t0) * 1e3)
        up, down = float(np.median(ups)), float(np.median(downs))
        nbytes = sum(int(l.vert.nbytes) + (int(l.col.nbytes) if l.col is not None else 0) for l in scene["lists"])
        host_io = {"upload_ms": round(up, 2), "upload_bytes": nbytes, "upload_gb_per_s": round(nbytes / up / 1e6, 1),
                   "download_ms": round(down, 3), "download_bytes": args.width * args.height * 8,
                   "frame_with_upload_and_download_ms": round(value + up + down, 2),
                   "frame_with_download_ms": round(value + down, 3)}
        log(f"host boundary: {host_io}")
    # roofline of the dominant kernel of the frame that `value` times: the chunk pass of the
    # denoised frame's build (the plain frame's trace kernel when there is no denoised frame)
    if dn:
        cp = dn["chunk_pass"]
        ach = cp["algorithmic_bytes"] / (cp["ms"] * 1e-3) / 1e9
        ctraffic = (dram_traffic(f"cells {args.cells}, leaf pairs on", DRAM_PROFILE_CHUNK)
                    if args.config == "fcc" else None)
        roofline = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": ctraffic,
                    "dram_frac": (round(ctraffic / (cp["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if ctraffic else None),
                    "kernel": cp["kernel"], "algorithmic_bytes": cp["algorithmic_bytes"]}
    else:
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": kernel_name,
                    "algorithmic_bytes": int(algo_bytes)}
    out = {
        "metric": "ms/frame 50M spheres 1080p AO: denoised 1spp vs plain 64spp @>=38dB; Mrays/s",
        "value": round(value, 3),
        "unit": "ms/frame",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(value, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded FCC lattice generated on the host, resident in HBM before timing)",
        "config": {"workload": scene["workload"],
                   "frame": ("denoised 1-spp AO (LBVH rebuild + 1-spp trace with features + U-Net denoise)"
                             if dn else f"plain {args.spp}-spp (LBVH rebuild + trace)"),
                   "spheres": int(len(scene["xyzr"])), "width": args.width, "height": args.height,
                   "spp": args.spp, "morton_bits": args.bits},
        "plain_ms_per_frame": round(ms, 3),
        "denoised": dn,
        "mrays_per_s": round(mrays, 1),
        "rays_per_frame": int(rays),
        "build_ms": round(build_ms, 3),
        "build_stages_ms": build_stages,
        "trace_ms": round(trace_ms, 3),
        "trace_work": split,
        "autotune": tuned,
        "roofline": roofline,
        "roofline_plain_trace": plain_trace_roof,
        "cpu_baseline": cpu,
        "orbit": orbit,
        "structures": structures,
        "pipelined": pipelined,
        "host_io": host_io,
    }
    print(json.dumps(out), flush=True)
    r.close()


if __name__ == "__main__":
    main()
