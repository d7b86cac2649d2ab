"""CPU tests of the autotuner's search (coordinate descent over the library knobs)."""
import ctypes as C

from megamol_amd import _lib as L, autotune


class FakeRenderer:
    def __init__(self):
        self.opts = {}

    def set_option(self, k, v):
        self.opts[k] = v


def test_coordinate_descent_finds_separable_optimum():
    best = {"tile_w": 16, "stack_lds": 24, "blocks_per_cu": 4, "refill": 32}

    def measure(r):  # separable convex cost with its minimum at `best`
        return 10.0 + sum(abs(r.opts[k] - v) / (1 + v) for k, v in best.items())

    r = FakeRenderer()
    cfg, ms, cache = autotune.tune(measure, r, rounds=3, reps=1)
    assert cfg == best
    assert abs(ms - 10.0) < 1e-9
    assert r.opts == best  # the winning configuration is left applied
    assert len(cache) < 4 * 4 * 3 * 4  # fewer evaluations than the full grid


def test_job_struct_matches_header_layout():
    import os
    import subprocess
    import tempfile
    from pathlib import Path
    root = Path(__file__).resolve().parent.parent
    src = ('#include <stdio.h>\n#include <stddef.h>\n#include "mmhrt.h"\n'
           'int main(void){printf("%zu %zu %zu\\n", sizeof(mmhrt_job), offsetof(mmhrt_job, params),'
           ' offsetof(mmhrt_job, out_offset));return 0;}\n')
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "j.c"), "w").write(src)
        subprocess.run(["gcc", "-I", str(root / "include"), os.path.join(d, "j.c"), "-o", os.path.join(d, "j")],
                       check=True)
        got = list(map(int, subprocess.run([os.path.join(d, "j")], capture_output=True, text=True).stdout.split()))
    assert got == [C.sizeof(L.Job), L.Job.params.offset, L.Job.out_offset.offset]
