"""Parse the MMPLD fixtures with the reference's own reader and keep what it reads as vectors.

The fixtures under tests/golden/mmpld/ are written by megamol_amd.mmpld, so a reader test on
them alone would only show that the writer and the reader agree. This script runs MegaMol's
stdlib-only MMPLD dumper, utils/MMPLD/mmpldinfo.py (read from the reference tree in this build
container only; it never travels with the repository), over every fixture with `-vv --head all`,
parses what it prints (header, every list header, every particle at 6 decimals) and writes
tests/golden/mmpld/mmpldinfo.json. tests/test_mmpld_cpu.py compares megamol_amd.mmpld.read with
those values.

mmpldinfo.py prints the file's clipping box where it means the list box of a version-103 list
(its `box` variable), so list boxes are not taken from it; they are pinned by the particle
records that follow them (a box of the wrong length shifts every particle).

Run: python tests/golden/make_mmpldinfo_vectors.py [path/to/mmpldinfo.py]
"""
import json
import re
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
FIX = HERE / "mmpld"
OUT = FIX / "mmpldinfo.json"
TOOL = Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference/utils/MMPLD/mmpldinfo.py")

NUM = r"[-+]?(?:\d+\.?\d*(?:[eE][-+]?\d+)?|inf|nan)"


def nums(s):
    return [float(x) for x in re.findall(NUM, s)]


def parse(text):
    d = {"frames": [], "problems": []}
    cur_frame = cur_list = None
    for line in text.splitlines():
        t = line.strip()
        if t.startswith("mmpld version"):
            a, b = t.split()[-1].split(".")
            d["version"] = 100 + int(b) if a == "1" else None
        elif t.startswith("Number of frames:"):
            d["frame_count"] = int(t.split(":")[1])
        elif t.startswith("Bounding box:"):
            d["bbox"] = nums(t.split(":", 1)[1])
        elif t.startswith("Clipping box:"):
            d["clipbox"] = nums(t.split(":", 1)[1])
        elif t.startswith("Frame #"):
            m = re.match(r"Frame # (\d+) \(offset (\d+) = 0x[0-9a-f]+\) (?:\((" + NUM + r")\))?- (\d+) list", t)
            cur_frame = {"index": int(m.group(1)), "offset": int(m.group(2)),
                         "timestamp": float(m.group(3)) if m.group(3) is not None else None,
                         "list_count": int(m.group(4)), "lists": []}
            d["frames"].append(cur_frame)
        elif re.match(r"#\d+: VERTDATA_", t):
            v, c = t.split(":", 1)[1].split(",")
            cur_list = {"vert": v.strip(), "col": c.strip(), "particles": []}
            cur_frame["lists"].append(cur_list)
        elif t.endswith("per particle"):
            cur_list["stride"] = int(t.split()[0])
        elif t.startswith("global radius:"):
            cur_list["global_radius"] = nums(t)[0]
        elif t.startswith("global color:"):
            cur_list["global_color"] = [int(x) for x in nums(t)]
        elif t.startswith("intensity color range:"):
            cur_list["intensity_range"] = nums(t)
        elif re.match(r"\d+ particles?$", t):
            cur_list["count"] = int(t.split()[0])
        elif t.startswith("pos ="):
            m = re.match(r"pos = \(([^)]*)\)(?:, rad = (" + NUM + r"))?(?:, col = \(([^)]*)\)|, no color)$", t)
            p = {"pos": nums(m.group(1)), "col": nums(m.group(3)) if m.group(3) is not None else None}
            if m.group(2) is not None:
                p["rad"] = float(m.group(2))
            cur_list["particles"].append(p)
        elif ("error" in t or "warning" in t or "differs" in t) and not t.startswith("File:"):
            d["problems"].append(t)
    return d


def main():
    files = sorted(FIX.glob("*.mmpld")) + sorted((FIX / "versions").glob("*.mmpld"))
    out = {}
    for f in files:
        res = subprocess.run([sys.executable, str(TOOL), "-vv", "--head", "all", str(f)], capture_output=True,
                             text=True, timeout=60)
        if res.returncode != 0:
            raise SystemExit(f"{TOOL} failed on {f}: {res.stderr}")
        out[str(f.relative_to(FIX))] = parse(res.stdout)
    OUT.write_text(json.dumps(out, indent=1, sort_keys=True) + "\n")
    print(f"wrote {OUT} ({len(out)} files)")


if __name__ == "__main__":
    main()
