"""vare_amd -- MI355X (gfx950) FLIP step on the two-level sparse grid behind a C ABI.

``libvareflip.so`` (built from ``vare_amd/csrc``) holds all compute; this package only provides
ctypes bindings (``_lib``) and a mirror of the reference ``FLIPMethod`` API (``flip``).
"""
from . import _lib
from .flip import FLIPMethod, Grid, Params, Particles

__all__ = ["FLIPMethod", "Grid", "Params", "Particles", "_lib"]
