import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; runs the HIP library")
    config.addinivalue_line("markers", "slow: larger CPU workload")


def _ensure_built():
    """Tests load what is built; build only when nothing is there."""
    lib = ROOT / "megamol_amd" / "lib" / "libmmhrt.so"
    if not lib.exists():
        from megamol_amd import build as B

        B.build()
    orc = ROOT / "oracle" / "_build" / "liboracle.so"
    if not orc.exists():
        import oracle

        oracle.build()


_ensure_built()


@pytest.fixture(scope="session")
def gpu_renderer():
    from megamol_amd import SphereRenderer

    r = SphereRenderer(0)
    yield r
    r.close()
