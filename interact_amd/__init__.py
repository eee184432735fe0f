"""interact_amd: MI355X-native dense rsf_solve hot path (librsfg).

The product is the C-ABI library ``librsfg.so`` (``include/rsfg.h``) built
from ``csrc/`` for gfx950.  ``lib`` holds its ctypes bindings, ``model``
the host-side mirror of the rsf_solve driver, ``faults`` the seeded
fault/friction generators of the benchmark configurations.
"""
from . import lib  # noqa: F401

__all__ = ["lib", "model", "faults"]
