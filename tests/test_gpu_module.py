"""GPU: the MegaMol renderer module (HIPSphereRenderer, compiled against stand-ins of the MegaMol
types) renders an interleaved particle list into a CPU framebuffer identical to the C ABI."""
import subprocess
from pathlib import Path

import pytest

from helpers import case

pytestmark = pytest.mark.gpu
EXE = Path(__file__).resolve().parent.parent / "megamol_amd" / "module" / "_build" / "module_test"


def test_module_renders_like_the_c_abi():
    case("HIPSphereRenderer module driver")
    assert EXE.exists(), "build the module driver with megamol_amd.build.build()"
    res = subprocess.run([str(EXE)], capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "0 pixels differ" in res.stdout
