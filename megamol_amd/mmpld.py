"""MMPLD particle files (versions 100-103): reader and writer.

Format as parsed by MMPLDDataSource (plugins/moldyn/src/io/MMPLDDataSource.cpp):
  header (:373-399): "MMPLD\\0", u16 version (100..103), u32 frame count, 6 f32 bounding box,
                     6 f32 clip box, u64 seek table with frame count + 1 entries
  frame (:61-216):   [f32 timestamp if version >= 102], u32 list count, then per list
                     u8 vertex code, u8 colour code,
                     [f32 global radius if vertex code in {1,3,4}],
                     [4 u8 global colour if colour code == 0 (alpha byte skipped, alpha = 255)],
                     [2 f32 intensity range if colour code in {3,7}],
                     (these three are decided by the raw codes, even for a list whose vertex
                     code is 0 or whose colour code is unknown; only the record sizes and the
                     in-memory types fall back to NONE)
                     u64 particle count, [6 f32 list box if version >= 103],
                     count interleaved records of vertex bytes then colour bytes,
                     [version 101: u32 cluster count, u64 byte size, cluster bytes]
File colour codes differ from the in-memory enum (SimpleSphericalParticles.h:41-50): file 3 is
FLOAT_I, 4 FLOAT_RGB, 5 FLOAT_RGBA. Lists come back as zero-copy strided views into the frame
bytes, like the reference's SetVertexData/SetColourData on the frame buffer (:206-207).
"""
from __future__ import annotations

import struct
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .renderer import ParticleList

VERT_SIZE_FILE = {0: 0, 1: 12, 2: 16, 3: 6, 4: 24}
COL_SIZE_FILE = {0: 0, 1: 3, 2: 4, 3: 4, 4: 12, 5: 16, 6: 8, 7: 8}
# file colour code -> in-memory ColourDataType (MMPLDDataSource.cpp:117-156)
COL_FILE_TO_MEM = {0: L.COLDATA_NONE, 1: L.COLDATA_UINT8_RGB, 2: L.COLDATA_UINT8_RGBA, 3: L.COLDATA_FLOAT_I,
                   4: L.COLDATA_FLOAT_RGB, 5: L.COLDATA_FLOAT_RGBA, 6: L.COLDATA_USHORT_RGBA, 7: L.COLDATA_DOUBLE_I}
COL_MEM_TO_FILE = {v: k for k, v in COL_FILE_TO_MEM.items()}


class MMPLDError(ValueError):
    pass


@dataclass
class Header:
    version: int
    frame_count: int
    bbox: tuple
    clipbox: tuple
    seek: list


def read_header(buf: bytes) -> Header:
    if len(buf) < 60:
        raise MMPLDError("file too short for an MMPLD header")
    if buf[:6] != b"MMPLD\x00":
        raise MMPLDError("MMPLD file header id wrong")
    (ver,) = struct.unpack_from("<H", buf, 6)
    if ver < 100 or ver > 103:
        raise MMPLDError("MMPLD file header version wrong")
    (frames,) = struct.unpack_from("<I", buf, 8)
    if frames == 0:
        raise MMPLDError("MMPLD file does not contain any frame information")
    bbox = struct.unpack_from("<6f", buf, 12)
    clip = struct.unpack_from("<6f", buf, 36)
    need = 60 + 8 * (frames + 1)
    if len(buf) < need:
        raise MMPLDError("Unable to read MMPLD file header")
    seek = list(struct.unpack_from(f"<{frames + 1}Q", buf, 60))
    return Header(ver, frames, bbox, clip, seek)


@dataclass
class Frame:
    timestamp: float
    lists: list  # ParticleList, views into `data`
    data: np.ndarray  # frame bytes (kept alive)
    list_boxes: list


def read_frame(buf: bytes, hdr: Header, index: int) -> Frame:
    if not 0 <= index < hdr.frame_count:
        raise MMPLDError("frame index out of range")
    a, b = hdr.seek[index], hdr.seek[index + 1]
    if not (0 <= a <= b <= len(buf)):
        raise MMPLDError("corrupt seek table")
    data = np.frombuffer(buf, np.uint8, count=b - a, offset=a).copy()
    raw = data.tobytes()
    p = 0
    ts = float(index)
    if hdr.version >= 102:
        (ts,) = struct.unpack_from("<f", raw, p)
        p += 4
    (plc,) = struct.unpack_from("<I", raw, p)
    p += 4
    lists, boxes = [], []
    for _ in range(plc):
        raw_vt, raw_ct = raw[p], raw[p + 1]
        p += 2
        # types and record sizes: unknown codes read as NONE (size 0), and a list without
        # vertices has no colours either (:89-157)
        vt = raw_vt if raw_vt in VERT_SIZE_FILE else 0
        ct = raw_ct if (vt != 0 and raw_ct in COL_SIZE_FILE) else 0
        vs, cs = VERT_SIZE_FILE[vt], COL_SIZE_FILE[ct]
        # which header fields follow is decided by the raw codes (:160-178)
        grad = 0.05
        if raw_vt in (1, 3, 4):
            (grad,) = struct.unpack_from("<f", raw, p)
            p += 4
        gcol = (192, 192, 192, 255)
        cmin, cmax = 0.0, 1.0
        if raw_ct == 0:
            gcol = (raw[p], raw[p + 1], raw[p + 2], 255)
            p += 4
        elif raw_ct in (3, 7):
            cmin, cmax = struct.unpack_from("<2f", raw, p)
            p += 8
        (cnt,) = struct.unpack_from("<Q", raw, p)
        p += 8
        box = None
        if hdr.version >= 103:
            box = struct.unpack_from("<6f", raw, p)
            p += 24
        stride = vs + cs
        if p + stride * cnt > len(raw):
            raise MMPLDError("list data runs past the frame end")
        pl = ParticleList(vert_type=vt, vert=data, col_type=COL_FILE_TO_MEM[ct], col=data if ct else None,
                          vert_stride=stride, col_stride=stride, count=int(cnt), global_radius=grad,
                          global_rgba=gcol, col_min=cmin, col_max=cmax, vert_offset=p, col_offset=p + vs)
        lists.append(pl)
        boxes.append(box)
        p += stride * cnt
        if hdr.version == 101:
            (ncl,) = struct.unpack_from("<I", raw, p)
            p += 4
            (sz,) = struct.unpack_from("<Q", raw, p)
            p += 8 + sz
    return Frame(ts, lists, data, boxes)


def read(path: str):
    buf = open(path, "rb").read()
    hdr = read_header(buf)
    return hdr, [read_frame(buf, hdr, i) for i in range(hdr.frame_count)]


@dataclass
class ListSpec:
    """One list to write: vertex code and colour code are FILE codes."""
    vert_code: int
    col_code: int
    xyz: np.ndarray  # (n,3) float64 values; stored as float/short/double per vert_code
    radius: np.ndarray | None = None
    colors: np.ndarray | None = None  # (n,k) values in the storage type's range
    intensity: np.ndarray | None = None
    global_radius: float = 0.1
    global_rgba: tuple = (255, 255, 255, 255)
    imin: float = 0.0
    imax: float = 1.0
    box: tuple | None = None


def _records(ls: ListSpec) -> bytes:
    n = len(ls.xyz)
    vdt = {1: [("p", "<f4", 3)], 2: [("p", "<f4", 3), ("r", "<f4")], 3: [("p", "<u2", 3)],
           4: [("p", "<f8", 3)]}[ls.vert_code]
    cdt = {0: [], 1: [("c", "u1", 3)], 2: [("c", "u1", 4)], 3: [("c", "<f4")], 4: [("c", "<f4", 3)],
           5: [("c", "<f4", 4)], 6: [("c", "<u2", 4)], 7: [("c", "<f8")]}[ls.col_code]
    rec = np.zeros(n, dtype=np.dtype(vdt + cdt))
    rec["p"] = ls.xyz
    if ls.vert_code == 2:
        rec["r"] = ls.radius
    if ls.col_code in (3, 7):
        rec["c"] = ls.intensity
    elif ls.col_code:
        rec["c"] = ls.colors
    return rec.tobytes()


def write(path: str, frames: list[list[ListSpec]], version: int = 103, timestamps=None, bbox=None, clipbox=None):
    """Write frames (each a list of ListSpec). bbox/clipbox default to the data bounds."""
    if not 100 <= version <= 103 or version == 101:
        raise MMPLDError("writer supports versions 100, 102, 103")
    out = bytearray()
    frame_bytes = []
    lo = np.full(3, np.inf)
    hi = np.full(3, -np.inf)
    rmax = 0.0
    for fi, lists in enumerate(frames):
        fb = bytearray()
        if version >= 102:
            fb += struct.pack("<f", float(timestamps[fi]) if timestamps else float(fi))
        fb += struct.pack("<I", len(lists))
        for ls in lists:
            fb += struct.pack("<BB", ls.vert_code, ls.col_code)
            if ls.vert_code in (1, 3, 4):
                fb += struct.pack("<f", ls.global_radius)
                rmax = max(rmax, ls.global_radius)
            if ls.col_code == 0:
                fb += bytes(ls.global_rgba[:4])
            elif ls.col_code in (3, 7):
                fb += struct.pack("<2f", ls.imin, ls.imax)
            fb += struct.pack("<Q", len(ls.xyz))
            if len(ls.xyz):
                lo = np.minimum(lo, ls.xyz.min(0))
                hi = np.maximum(hi, ls.xyz.max(0))
                if ls.radius is not None:
                    rmax = max(rmax, float(np.max(ls.radius)))
            if version >= 103:
                box = ls.box
                if box is None:
                    box = (tuple(ls.xyz.min(0)) + tuple(ls.xyz.max(0))) if len(ls.xyz) else (0.0,) * 6
                fb += struct.pack("<6f", *box)
            fb += _records(ls)
        frame_bytes.append(bytes(fb))
    if bbox is None:
        bbox = tuple(lo) + tuple(hi)
    if clipbox is None:
        clipbox = tuple(np.array(bbox[:3]) - rmax) + tuple(np.array(bbox[3:]) + rmax)
    out += b"MMPLD\x00" + struct.pack("<HI", version, len(frames))
    out += struct.pack("<6f", *bbox) + struct.pack("<6f", *clipbox)
    pos = len(out) + 8 * (len(frames) + 1)
    seek = []
    for fb in frame_bytes:
        seek.append(pos)
        pos += len(fb)
    seek.append(pos)
    out += struct.pack(f"<{len(seek)}Q", *seek)
    for fb in frame_bytes:
        out += fb
    with open(path, "wb") as f:
        f.write(out)
