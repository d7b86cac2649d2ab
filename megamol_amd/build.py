"""Build the HIP library libmmhrt.so in-tree for gfx950.

Every HIP source under megamol_amd/csrc is compiled with hipcc to an object file (in parallel),
then linked into megamol_amd/lib/libmmhrt.so. The library is what travels to the GPU host; the
Python package only loads it.

Flags that matter for results:
  -ffp-contract=off                       no implicit multiply-add fusion, so the FP32 ray and
                                          intersection sequences round exactly as written
  -fhip-fp32-correctly-rounded-divide-sqrt  IEEE division and square root
  -Xclang -target-feature -Xclang -packed-fp32-ops   disable packed FP32 instructions
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "lib"
OBJ = PKG / "_build"
ARCH = os.environ.get("MMHRT_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fvisibility=hidden",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-value", "-Wno-unused-result",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or Path(cand).exists()):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass warns that it does not know the packed-fp32 feature: harmless
    err = "\n".join(l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}\n{err}")
    if err.strip():
        sys.stderr.write(err + "\n")


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile all HIP sources and link libmmhrt.so. Always recompiles."""
    LIB.mkdir(exist_ok=True)
    OBJ.mkdir(exist_ok=True)
    hipcc = _hipcc()
    srcs = sorted(CSRC.glob("*.hip"))
    inc = ["-I", str(PKG.parent / "include"), "-I", str(CSRC)]

    def compile_one(src: Path) -> Path:
        obj = OBJ / (src.stem + ".o")
        cmd = [hipcc, *HIP_FLAGS, *inc, "-c", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)
        return obj

    n = jobs or min(len(srcs), os.cpu_count() or 4, 8)
    with ThreadPoolExecutor(max_workers=n) as ex:
        objs = list(ex.map(compile_one, srcs))
    out = LIB / "libmmhrt.so"
    tmp = LIB / "libmmhrt.so.tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)])
    os.replace(tmp, out)
    build_module_test()
    return out


def build_module_test() -> Path:
    """The MegaMol renderer module compiled against the in-repo stand-ins of the MegaMol types,
    with a driver that renders through it and the headless project runner (host C++ only, links
    libmmhrt.so)."""
    mod = PKG / "module"
    outdir = mod / "_build"
    outdir.mkdir(exist_ok=True)
    exe = outdir / "module_test"
    inc = ["-I", str(mod / "mmstub"), "-I", str(mod), "-I", str(PKG.parent / "include")]
    link = ["-L", str(LIB), "-lmmhrt", "-Wl,-rpath,$ORIGIN/../../lib"]
    _run(["g++", "-std=c++17", "-O2", "-Wall", *inc, str(mod / "module_test.cpp"), str(mod / "HIPSphereRenderer.cpp"),
          *link, "-o", str(exe)])
    # the headless project runner: MMPLD source -> renderer module -> screenshots
    _run(["g++", "-std=c++17", "-O2", "-Wall", *inc, str(mod / "headless.cpp"), str(mod / "MMPLDSource.cpp"),
          str(mod / "HIPSphereRenderer.cpp"), *link, "-o", str(outdir / "headless")])
    return exe


if __name__ == "__main__":
    p = build(verbose="-v" in sys.argv)
    print(p)
