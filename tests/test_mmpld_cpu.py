"""CPU tests of the MMPLD reader (MMPLDDataSource::Frame::SetData,
plugins/moldyn/src/io/MMPLDDataSource.cpp:61-216, header :373-399).

The fixtures are written by this repository's writer; what pins the byte layout independently
is MegaMol's own MMPLD dumper, utils/MMPLD/mmpldinfo.py, whose reading of the same files is
kept in tests/golden/mmpld/mmpldinfo.json (made by tests/golden/make_mmpldinfo_vectors.py).
"""
import json
import struct
from pathlib import Path

import numpy as np
import pytest

from megamol_amd import _lib as L, mmpld

FIX = Path(__file__).resolve().parent / "golden" / "mmpld"

VERT_NAME = {L.VERTDATA_NONE: "VERTDATA_NONE", L.VERTDATA_FLOAT_XYZ: "VERTDATA_FLOAT_XYZ",
             L.VERTDATA_FLOAT_XYZR: "VERTDATA_FLOAT_XYZR", L.VERTDATA_SHORT_XYZ: "VERTDATA_SHORT_XYZ",
             L.VERTDATA_DOUBLE_XYZ: "VERTDATA_DOUBLE_XYZ"}
COL_NAME = {L.COLDATA_NONE: "COLDATA_NONE", L.COLDATA_UINT8_RGB: "COLDATA_UINT8_RGB",
            L.COLDATA_UINT8_RGBA: "COLDATA_UINT8_RGBA", L.COLDATA_FLOAT_RGB: "COLDATA_FLOAT_RGB",
            L.COLDATA_FLOAT_RGBA: "COLDATA_FLOAT_RGBA", L.COLDATA_FLOAT_I: "COLDATA_FLOAT_I",
            L.COLDATA_USHORT_RGBA: "COLDATA_USHORT_RGBA", L.COLDATA_DOUBLE_I: "COLDATA_DOUBLE_I"}
VERT_FMT = {L.VERTDATA_FLOAT_XYZ: ("<f4", 3), L.VERTDATA_FLOAT_XYZR: ("<f4", 4), L.VERTDATA_SHORT_XYZ: ("<u2", 3),
            L.VERTDATA_DOUBLE_XYZ: ("<f8", 3)}
COL_FMT = {L.COLDATA_UINT8_RGB: ("u1", 3), L.COLDATA_UINT8_RGBA: ("u1", 4), L.COLDATA_FLOAT_RGB: ("<f4", 3),
           L.COLDATA_FLOAT_RGBA: ("<f4", 4), L.COLDATA_FLOAT_I: ("<f4", 1), L.COLDATA_USHORT_RGBA: ("<u2", 4),
           L.COLDATA_DOUBLE_I: ("<f8", 1)}


def view(buf, offset, stride, count, fmt):
    """count records of `fmt` at `offset` every `stride` bytes (the strided view a list exposes)."""
    dt = np.dtype({"names": ["v"], "formats": [(fmt[0], (fmt[1],))], "offsets": [0], "itemsize": stride})
    # the last record's view ends with its own field, not a full stride: pad the copy
    return np.frombuffer(buf + bytes(stride), dt, count=count, offset=offset)["v"].astype(np.float64)


def r6(a):
    """The 6-decimal rendering mmpldinfo prints."""
    return [float(f"{x:f}") for x in np.ravel(a)]


REF = json.loads((FIX / "mmpldinfo.json").read_text())


@pytest.mark.parametrize("name", sorted(REF))
def test_reader_agrees_with_the_reference_parser(name):
    want = REF[name]
    assert want["problems"] == [], want["problems"]  # seek table and frame sizes consistent
    hdr, frames = mmpld.read(str(FIX / name))
    assert hdr.version == want["version"] and hdr.frame_count == want["frame_count"]
    assert r6(hdr.bbox) == want["bbox"] and r6(hdr.clipbox) == want["clipbox"]
    assert len(frames) == len(want["frames"])
    for fr, wf in zip(frames, want["frames"]):
        assert hdr.seek[wf["index"]] == wf["offset"]
        if wf["timestamp"] is not None:
            assert r6([fr.timestamp]) == [wf["timestamp"]]
        else:
            assert hdr.version < 102
        assert len(fr.lists) == wf["list_count"] == len(wf["lists"])
        buf = fr.data.tobytes()
        for pl, wl in zip(fr.lists, wf["lists"]):
            assert VERT_NAME[pl.vert_type] == wl["vert"]
            assert COL_NAME[pl.col_type] == wl["col"]
            assert pl.vert_stride == wl["stride"] and pl.count == wl["count"]
            if "global_radius" in wl:
                assert r6([pl.global_radius]) == [wl["global_radius"]]
            if "global_color" in wl:
                # mmpldinfo prints the alpha byte of the file; the reader keeps alpha 255 as
                # SetGlobalColour(r, g, b) does (MMPLDDataSource.cpp:167-168)
                assert list(pl.global_rgba[:3]) == wl["global_color"][:3]
            if "intensity_range" in wl:
                assert r6([pl.col_min, pl.col_max]) == wl["intensity_range"]
            pos = view(buf, pl.vert_offset, pl.vert_stride, pl.count, VERT_FMT[pl.vert_type])
            cols = (view(buf, pl.col_offset, pl.col_stride, pl.count, COL_FMT[pl.col_type])
                    if pl.col_type != L.COLDATA_NONE else None)
            assert len(wl["particles"]) == pl.count
            for k, wp in enumerate(wl["particles"]):
                assert r6(pos[k, :3]) == wp["pos"]
                if "rad" in wp:
                    assert r6([pos[k, 3]]) == [wp["rad"]]
                if wp["col"] is None:
                    assert cols is None
                else:
                    assert r6(cols[k]) == wp["col"]


def _frame_file(lists_bytes, n_lists, version=102):
    frame = (struct.pack("<f", 0.0) if version >= 102 else b"") + struct.pack("<I", n_lists) + lists_bytes
    head = b"MMPLD\x00" + struct.pack("<HI", version, 1) + struct.pack("<12f", *([-1] * 3 + [1] * 3) * 2)
    start = len(head) + 16
    return head + struct.pack("<2Q", start, start + len(frame)) + frame


def test_raw_codes_decide_which_header_fields_follow(tmp_path):
    """MMPLDDataSource.cpp:160-178 reads the global radius, the global colour and the intensity
    range by the raw codes of the file, also for a list whose vertex code is 0 (its colour type
    then reads as NONE, :153-156) or whose colour code is unknown (:149-152). The list after
    such a list must be read from the right offset."""
    pts = np.array([[0.5, 1.5, 2.5], [3.0, 4.0, 5.0]], "<f4")
    lists = b""
    # vertex code 0, colour code 3: no radius, the 8-byte intensity range, no records
    lists += struct.pack("<BB", 0, 3) + struct.pack("<2f", -1.0, 7.0) + struct.pack("<Q", 2)
    # vertex code 1, unknown colour code 9: global radius, neither global colour nor range
    lists += struct.pack("<BB", 1, 9) + struct.pack("<f", 0.25) + struct.pack("<Q", 2) + pts.tobytes()
    # vertex code 0, colour code 0: the 4-byte global colour
    lists += struct.pack("<BB", 0, 0) + bytes([10, 20, 30, 40]) + struct.pack("<Q", 0)
    # unknown vertex code 7 with colour code 7: no radius, the intensity range, NONE types
    lists += struct.pack("<BB", 7, 7) + struct.pack("<2f", 2.0, 3.0) + struct.pack("<Q", 0)
    # a normal list last: FLOAT_XYZ + UINT8_RGBA
    lists += struct.pack("<BB", 1, 2) + struct.pack("<f", 0.5) + struct.pack("<Q", 2)
    lists += b"".join(p.tobytes() + bytes([1, 2, 3, 4]) for p in pts)
    f = tmp_path / "odd.mmpld"
    f.write_bytes(_frame_file(lists, 5))
    hdr, frames = mmpld.read(str(f))
    a, b, c, d, e = frames[0].lists
    assert (a.vert_type, a.col_type, a.count, a.vert_stride) == (L.VERTDATA_NONE, L.COLDATA_NONE, 2, 0)
    assert (a.col_min, a.col_max) == (-1.0, 7.0) and a.global_rgba == (192, 192, 192, 255)
    assert (b.vert_type, b.col_type, b.vert_stride, b.global_radius) == (L.VERTDATA_FLOAT_XYZ, L.COLDATA_NONE, 12,
                                                                         0.25)
    assert b.global_rgba == (192, 192, 192, 255) and (b.col_min, b.col_max) == (0.0, 1.0)
    buf = frames[0].data.tobytes()
    np.testing.assert_array_equal(view(buf, b.vert_offset, 12, 2, ("<f4", 3)), pts)
    assert (c.vert_type, c.col_type, c.global_rgba) == (L.VERTDATA_NONE, L.COLDATA_NONE, (10, 20, 30, 255))
    assert (d.vert_type, d.col_type, d.global_radius, d.col_min, d.col_max) == (L.VERTDATA_NONE, L.COLDATA_NONE,
                                                                                0.05, 2.0, 3.0)
    assert (e.vert_type, e.col_type, e.vert_stride, e.count) == (L.VERTDATA_FLOAT_XYZ, L.COLDATA_UINT8_RGBA, 16, 2)
    np.testing.assert_array_equal(view(buf, e.vert_offset, 16, 2, ("<f4", 3)), pts)
    np.testing.assert_array_equal(view(buf, e.col_offset, 16, 2, ("u1", 4)), [[1, 2, 3, 4]] * 2)
