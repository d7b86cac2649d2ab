"""megamol_amd — MI355X-native particle-sphere ray tracing behind MegaMol's sphere-renderer API.

The work runs in libmmhrt.so (hand-written HIP for gfx950, C ABI in include/mmhrt.h); this
package only loads and drives it.
"""
from . import _lib
from .renderer import Frame, ParticleList, SphereRenderer, look_at, params

__all__ = ["SphereRenderer", "ParticleList", "Frame", "look_at", "params", "_lib"]
