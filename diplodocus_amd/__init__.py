"""diplodocus_amd — MI355X-native BVH build-and-trace engine.

The product is libdpl.so (hand-written HIP kernels for gfx950 behind the C ABI in
include/dpl/dpl.h). This package holds its sources (csrc/), a ctypes binding (_lib),
the host-side mirror of the reference's acceleration interface (accel) and the seeded
workload generators (scenes).
"""
from . import scenes  # noqa: F401  (numpy only)

__all__ = ["scenes", "load"]


def load():
    """Imports the binding; raises ImportError when libdpl.so has not been built."""
    from . import accel, _lib  # noqa: F401

    return accel
