"""MI355X-native implementation of gds2Para's simulate-mode hot path.

The compute lives in libgds2para.so (hand-written HIP for gfx950 behind the C
ABI in include/gds2para.h); this package is its Python face: ctypes bindings,
synthetic layouts, the solve driver and the request-serving loop.
"""
from . import _capi  # noqa: F401
