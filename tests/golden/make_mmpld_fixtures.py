"""Write the 17 MMPLD format fixtures under tests/golden/mmpld/ and their expected values.

The matrix and particle values restate utils/MMPLD/GenTestFiles.pl:17-117 of the MegaMol tree
(4 particles with radii 0.5/0.2/0.3/0.4 and colours white/red/green/blue scaled by the format's
range, intensities 255/64/128/192 in [0, 255]; then 3 edge lists of 5 global-colour particles
with radius 0.1; bounding box overridden to +-2; frame time 1.23). The Perl generator itself
needs the CPAN module Readonly, which this image lacks, so the files are written with
megamol_amd.mmpld at version 102, the version MMPLD.pm writes (utils/MMPLD/MMPLD.pm:73).

versions/ holds the first fixture's lists written at versions 100 and 103 (no timestamp; per-list
boxes) and as a two-frame version 103 file, so that the version-dependent fields and the seek table
are read back by an independent parser too (tests/golden/make_mmpldinfo_vectors.py).

expected.npz holds, per fixture, the float32 xyzr and rgba that SphereGeometry::assertData
(plugins/optix_hpg/src/optix/SphereGeometry.cpp:94-146) derives, written from the closed form
above (no renderer code involved).

Run: python tests/golden/make_mmpld_fixtures.py
"""
from pathlib import Path
import sys

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parents[2]))
from megamol_amd import mmpld  # noqa: E402

OUT = Path(__file__).resolve().parent / "mmpld"

# GenTestFiles vertex/colour names -> MMPLD file codes (MMPLD.pm:173-199)
VCODE = {"xyz_float": 1, "xyzr_float": 2, "xyz_double": 4}
CCODE = {"none": 0, "int_float": 3, "rgba_byte": 2, "rgb_float": 4, "rgba_float": 5, "rgba_ushort": 6,
         "int_double": 7}
TESTS = [  # GenTestFiles.pl:95-113 (vertex, colour, colscale)
    ("xyz_float", "rgb_float", 1.0), ("xyz_double", "rgb_float", 1.0), ("xyzr_float", "rgb_float", 1.0),
    ("xyz_float", "rgba_float", 1.0), ("xyz_double", "rgba_float", 1.0), ("xyzr_float", "rgba_float", 1.0),
    ("xyz_float", "rgba_byte", 255.0), ("xyz_double", "rgba_byte", 255.0), ("xyzr_float", "rgba_byte", 255.0),
    ("xyz_float", "int_float", 1.0), ("xyz_double", "int_float", 1.0), ("xyz_double", "int_double", 1.0),
    ("xyzr_float", "int_float", 1.0), ("xyz_float", "none", 1.0), ("xyz_double", "rgba_ushort", 65535.0),
    ("xyz_double", "none", 1.0), ("xyzr_float", "none", 1.0),
]
POS = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float64)
RAD = np.array([0.5, 0.2, 0.3, 0.4])
COL = np.array([[1, 1, 1, 1], [1, 0, 0, 1], [0, 1, 0, 1], [0, 0, 1, 1]], np.float64)
INT = np.array([255.0, 64.0, 128.0, 192.0])
EDGE_X = np.array([2.0, 1.8, 1.6, 1.4, 1.2])
EDGE_COLS = [(255, 0, 0, 255), (0, 255, 0, 255), (0, 0, 255, 255)]


def edge_lists():
    xs = EDGE_X
    m = len(xs)
    pts = [np.stack([xs, np.full(m, -2.0), np.full(m, -2.0)], 1),
           np.stack([np.full(m, -2.0), xs, np.full(m, -2.0)], 1),
           np.stack([np.full(m, -2.0), np.full(m, -2.0), xs], 1)]
    return [mmpld.ListSpec(1, 0, p, global_radius=0.1, global_rgba=c) for p, c in zip(pts, EDGE_COLS)]


def main():
    OUT.mkdir(parents=True, exist_ok=True)
    expected = {}
    for vn, cn, scale in TESTS:
        vc, cc = VCODE[vn], CCODE[cn]
        cols = COL * scale
        if cn == "rgb_float":
            cols = cols[:, :3]
        first = mmpld.ListSpec(vc, cc, POS, radius=RAD if vc == 2 else None, colors=cols, intensity=INT,
                               global_radius=0.1, global_rgba=(255, 255, 0, 255), imin=0.0, imax=255.0)
        name = f"test_{vn}_{cn}"
        mmpld.write(str(OUT / f"{name}.mmpld"), [[first] + edge_lists()], version=102, timestamps=[1.23],
                    bbox=(-2, -2, -2, 2, 2, 2), clipbox=(-2.5, -2.5, -2.5, 2.5, 2.5, 2.5))
        # expected sphere data per SphereGeometry::assertData
        xyzr = np.zeros((19, 4), np.float32)
        rgba = np.zeros((19, 4), np.float32)
        xyzr[:4, :3] = POS
        xyzr[:4, 3] = RAD if vc == 2 else 0.1
        if cn in ("none",):
            rgba[:4] = np.array([255, 255, 0, 255]) / np.float32(255.0)
        elif cn in ("int_float", "int_double"):
            rgba[:4] = np.float32(192.0) / np.float32(255.0)  # reader's default global colour, alpha 255
            rgba[:4, 3] = 1.0
        elif cn == "rgb_float":
            rgba[:4, :3] = COL[:, :3]
            rgba[:4, 3] = 1.0
        else:  # rgba_float / rgba_byte / rgba_ushort: Get_f is a plain cast, no normalisation
            rgba[:4] = COL * scale
        k = 4
        for li, c in enumerate(EDGE_COLS):
            for x in EDGE_X:
                p = [-2.0, -2.0, -2.0]
                p[li] = x
                xyzr[k, :3] = p
                xyzr[k, 3] = 0.1
                rgba[k] = np.array(c, np.float32) / np.float32(255.0)
                k += 1
        expected[name + "_xyzr"] = xyzr
        expected[name + "_rgba"] = rgba
    np.savez(OUT / "expected.npz", **expected)
    print(f"wrote {len(TESTS)} fixtures to {OUT}")
    write_versions()


def write_versions():
    """The first GenTestFiles list set at versions 100 and 103, and as two frames at 103."""
    vd = OUT / "versions"
    vd.mkdir(exist_ok=True)
    first = mmpld.ListSpec(1, 4, POS, colors=COL[:, :3], global_radius=0.1, imin=0.0, imax=255.0)
    second = mmpld.ListSpec(2, 3, POS[::-1].copy(), radius=RAD[::-1].copy(), intensity=INT, imin=10.0, imax=200.0,
                            box=(-0.5, -0.25, 0.0, 1.5, 1.25, 1.0))
    for ver in (100, 103):
        mmpld.write(str(vd / f"v{ver}.mmpld"), [[first] + edge_lists()], version=ver, timestamps=[1.23],
                    bbox=(-2, -2, -2, 2, 2, 2), clipbox=(-2.5, -2.5, -2.5, 2.5, 2.5, 2.5))
    mmpld.write(str(vd / "v103_two_frames.mmpld"), [[first] + edge_lists(), [second, first]], version=103,
                timestamps=[0.5, 2.75], bbox=(-2, -2, -2, 2, 2, 2), clipbox=(-2.5, -2.5, -2.5, 2.5, 2.5, 2.5))
    print(f"wrote 3 version fixtures to {vd}")


if __name__ == "__main__":
    main()
