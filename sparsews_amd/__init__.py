"""sparsews_amd -- MI355X (gfx950) sparse-workspace kernels.

The compute path is libspws.so (hand-written HIP, built by
``python -m sparsews_amd.build``) behind the C ABI in include/spws/spws.h.
This package holds the host-side mirror of the reference interface
(``taco``), the device-resident API (``device``), the input generators
(``gen``) and the level-format builders (``formats``).
"""
from ._lib import SpwsError, lib  # noqa: F401

__all__ = ["SpwsError", "lib"]
