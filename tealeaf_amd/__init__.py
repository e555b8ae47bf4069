"""tealeaf_amd: an MI355X (gfx950) kernel suite for TeaLeaf's implicit solve.

The product is the HIP shared library ``tealeaf_amd/_lib/libtealeaf_hip.so``
(C ABI in ``include/tealeaf_hip.h``) plus the ``hip_cdna4`` glue that plugs it
into the reference's own C host.  This Python package is the host-side
mirror used by tests and the benchmark: ctypes bindings (``abi``), device
buffers (``device``), the problem deck (``deck``), the seeded block-density
input generator (``blocks``), the build recipe (``build``) and the native
timestep driver (``problem``).
"""
from .abi import TealeafError, lib  # noqa: F401

__all__ = ["TealeafError", "lib"]
