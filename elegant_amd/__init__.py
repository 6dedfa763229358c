"""MI355X-native FP64 6D tracking for elegant's per-particle hot path.

The GPU library (libelegant_hip.so: HIP kernels for gfx950 + a C host layer)
is driven through its C ABI (include/elegant_hip.h); this package holds the
ctypes bindings, the element constructors with elegant's defaults, and the
benchmark lattices."""
from . import _abi
from . import elements
from . import lattices

__all__ = ["_abi", "elements", "lattices", "native"]
