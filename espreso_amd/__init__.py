"""espreso_amd: MI355X-native Total FETI dual-loop core.

The compute path is libespreso_amd.so (HIP kernels for gfx950 behind the C
ABI in include/espreso_amd.h).  This package only loads it; there is no CPU
fallback: if the library is missing or no GPU is present, calls raise.
"""
from .problems import ALL as PROBLEMS  # noqa: F401

__all__ = ["FETISolver", "GridFETI", "lib", "LibraryNotBuilt", "PROBLEMS"]


def __getattr__(name):
    if name in ("FETISolver", "GridFETI", "lib", "LibraryNotBuilt", "Options", "Stats"):
        from . import core
        return getattr(core, name)
    raise AttributeError(name)
