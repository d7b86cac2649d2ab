"""Shared helpers for the tests (scene set-up, case banners)."""
import sys

import numpy as np

import oracle as O
from megamol_amd import look_at


def case(msg: str) -> None:
    """Name the case before a GPU step, flushed, so that a crash still leaves a trace."""
    sys.stdout.write(f"[case] {msg}\n")
    sys.stdout.flush()


def small_camera(w=64, h=48, pos=(0.5, 0.5, 2.2), target=(0.5, 0.5, 0.5), fovy=40.0):
    return look_at(pos, target, (0, 1, 0), fovy, w / h, 0.1, 10.0)


def oracle_params(p):
    """Oracle params from a megamol_amd RenderParams."""
    return O.make_params(p.width, p.height, p.spp, p.mode, p.ao_distance, p.frame_idx, tuple(p.background),
                         1 if (p.flags & 1) else 0, p.max_bounces)


def assert_frames_equal(g, o, aux=False, what=""):
    """Bit-exact comparison of a GPU frame with an oracle frame."""
    np.testing.assert_array_equal(g.hit_id, o["hit_id"], err_msg=f"hit_id {what}")
    np.testing.assert_array_equal(g.t.view(np.uint32), o["t"].view(np.uint32), err_msg=f"t {what}")
    np.testing.assert_array_equal(g.depth.view(np.uint32), o["depth"].view(np.uint32), err_msg=f"depth {what}")
    np.testing.assert_array_equal(g.color.view(np.uint32), o["color"].view(np.uint32), err_msg=f"color {what}")
    np.testing.assert_array_equal(g.rgba8, o["rgba8"], err_msg=f"rgba8 {what}")
    if aux:
        np.testing.assert_array_equal(g.aux.view(np.uint16), o["aux"].view(np.uint16), err_msg=f"aux {what}")
