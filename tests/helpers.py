"""Shared helpers for the tests (scene set-up, case banners, FP64 checks)."""
import ctypes as C
import sys

import numpy as np

import oracle as O
from megamol_amd import look_at


def case(msg: str) -> None:
    """Name the case before a GPU step, flushed, so that a crash still leaves a trace. Also goes
    to the process's real stderr, past pytest's capture, as a progress line of long tests."""
    sys.stdout.write(f"[case] {msg}\n")
    sys.stdout.flush()
    try:
        sys.__stderr__.write(f"[case] {msg}\n")
        sys.__stderr__.flush()
    except (AttributeError, ValueError):
        pass


def small_camera(w=64, h=48, pos=(0.5, 0.5, 2.2), target=(0.5, 0.5, 0.5), fovy=40.0):
    return look_at(pos, target, (0, 1, 0), fovy, w / h, 0.1, 10.0)


def oracle_params(p):
    """Oracle params from a megamol_amd RenderParams."""
    return O.make_params(p.width, p.height, p.spp, p.mode, p.ao_distance, p.frame_idx, tuple(p.background),
                         1 if (p.flags & 1) else 0, p.max_bounces)


def assert_frames_equal(g, o, aux=False, what="", rows=None):
    """Bit-exact comparison of a GPU frame with an oracle frame (rows: only these row indices)."""
    sel = slice(None) if rows is None else rows
    np.testing.assert_array_equal(g.hit_id[sel], o["hit_id"][sel], err_msg=f"hit_id {what}")
    np.testing.assert_array_equal(g.t[sel].view(np.uint32), o["t"][sel].view(np.uint32), err_msg=f"t {what}")
    np.testing.assert_array_equal(g.depth[sel].view(np.uint32), o["depth"][sel].view(np.uint32),
                                  err_msg=f"depth {what}")
    np.testing.assert_array_equal(g.color[sel].view(np.uint32), o["color"][sel].view(np.uint32),
                                  err_msg=f"color {what}")
    np.testing.assert_array_equal(g.rgba8[sel], o["rgba8"][sel], err_msg=f"rgba8 {what}")
    if aux:
        np.testing.assert_array_equal(g.aux[sel].view(np.uint16), o["aux"][sel].view(np.uint16),
                                      err_msg=f"aux {what}")


def sample_ray(cam, W, H, x, y, frame=0, sample=0):
    """The jittered primary ray of a sample: draws 2s and 2s + 1 of the pixel's stream
    tea16(y * W + x, frame) (raygen.cu:182,241-242), as float64 arrays."""
    lib = O.load()
    fs = O.frame_state(cam)
    draws = O.lcg_stream(O.tea16(y * W + x, frame), 2 * sample + 2)
    o = np.zeros(3, np.float32)
    d = np.zeros(3, np.float32)
    lib.orc_primary_ray(C.byref(fs), W, H, x, y, C.c_float(draws[2 * sample]), C.c_float(draws[2 * sample + 1]),
                        o.ctypes.data, d.ctypes.data)
    return o.astype(np.float64), d.astype(np.float64)


def fp64_check(xyzr, cam, W, H, gpu_ids, gpu_t, rects, x_off=0, y_off=0, frame=0, what=""):
    """Primary hit ids of a 1-spp frame against the FP64 brute-force oracle over the pixels of
    `rects` ((x0, y0, x1, y1) in frame pixels; gpu_ids / gpu_t hold the frame from pixel
    (x_off, y_off) on). A mismatch is allowed only where FP64 itself cannot separate the
    answers: the FP32 winner's sphere is a near tie of the FP64 winner (FP64 distances within
    1e-5 relative) or either sphere is grazed (|normalised discriminant| < 1e-4). Depth agrees
    within 1e-5 relative wherever the ids agree. Returns (mismatches, pixels)."""
    p = O.make_params(W, H, frame_idx=frame)
    res = O.primary_f64_rects(xyzr, cam, p, rects)
    mism = total = 0
    for (x0, y0, x1, y1), (ids, t64, disc, t2) in zip(rects, res):
        g = gpu_ids[y0 - y_off:y1 - y_off, x0 - x_off:x1 - x_off]
        gt = gpu_t[y0 - y_off:y1 - y_off, x0 - x_off:x1 - x_off].astype(np.float64)
        bad = ids != g
        for yy, xx in zip(*np.nonzero(bad)):
            o, d = sample_ray(cam, W, H, x0 + xx, y0 + yy, frame)
            reasons = []
            if ids[yy, xx] >= 0 and abs(disc[yy, xx]) < 1e-4:
                reasons.append("fp64 winner grazed")
            if g[yy, xx] >= 0:
                tg, dg = O.intersect_f64(o, d, xyzr[g[yy, xx]].astype(np.float64))
                if abs(dg) < 1e-4:
                    reasons.append("fp32 winner grazed")
                if tg is not None and ids[yy, xx] >= 0 and abs(tg - t64[yy, xx]) <= 1e-5 * t64[yy, xx]:
                    reasons.append("near tie")
            assert reasons, (f"{what}: pixel ({x0 + xx}, {y0 + yy}) gpu {g[yy, xx]} fp64 {ids[yy, xx]} "
                             f"t64 {t64[yy, xx]} disc {disc[yy, xx]}")
        ok = ~bad & (ids >= 0)
        if ok.any():
            rel = np.abs(gt[ok] - t64[ok]) / t64[ok]
            assert rel.max() <= 1e-5, f"{what}: depth rel err {rel.max()}"
        mism += int(bad.sum())
        total += bad.size
    return mism, total
