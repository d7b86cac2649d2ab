"""CPU: AddressSanitizer and UndefinedBehaviorSanitizer builds of the host code that runs without a
GPU: the CPU oracle (all its entry points on small scenes, tests/native/oracle_sanitize.c) and the
renderer module's MMPLD data source (tests/native/mmpld_sanitize.cpp: the 17 format fixtures, the
version fixtures, every truncation and corrupted headers). A sanitizer report fails the run."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]


def _build(tmp_path, cmd):
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not found")
def test_oracle_under_asan_ubsan(tmp_path):
    exe = tmp_path / "oracle_san"
    _build(tmp_path, ["gcc", "-std=c11", *SAN, "-ffp-contract=off", "-I", str(ROOT / "oracle"),
                      str(ROOT / "tests" / "native" / "oracle_sanitize.c"), str(ROOT / "oracle" / "oracle.c"),
                      "-pthread", "-lm", "-o", str(exe)])
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600,
                         env={"ASAN_OPTIONS": "detect_leaks=1", "UBSAN_OPTIONS": "print_stacktrace=1"})
    assert res.returncode == 0, res.stdout + res.stderr
    assert "ok" in res.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not found")
def test_mmpld_source_under_asan_ubsan(tmp_path):
    exe = tmp_path / "mmpld_san"
    mod = ROOT / "megamol_amd" / "module"
    _build(tmp_path, ["g++", "-std=c++17", *SAN, "-I", str(mod / "mmstub"), "-I", str(mod),
                      str(ROOT / "tests" / "native" / "mmpld_sanitize.cpp"), str(mod / "MMPLDSource.cpp"),
                      "-o", str(exe)])
    gold = ROOT / "tests" / "golden" / "mmpld"
    files = []
    for f in sorted(gold.glob("*.mmpld")) + sorted((gold / "versions").glob("*.mmpld")):
        c = tmp_path / f.name
        shutil.copy(f, c)
        files.append(str(c))
    assert len(files) >= 20
    res = subprocess.run([str(exe), *files], capture_output=True, text=True, timeout=600,
                         env={"ASAN_OPTIONS": "detect_leaks=1", "UBSAN_OPTIONS": "print_stacktrace=1"})
    assert res.returncode == 0, res.stdout + res.stderr
    assert res.stdout.count("frame 0:") == len(files)
