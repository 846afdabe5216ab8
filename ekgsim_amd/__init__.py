"""MI355X-native ekgsim simulation core.

The compute path is the gfx950 library lib/libekgsim.so (C ABI in
include/ekg.h); this package binds it with ctypes and carries the seeded
synthetic workload generator.  Importing the package does not load the
library; the first use does, and fails loudly if it is not built.
"""
from . import synth  # noqa: F401
from ._capi import (  # noqa: F401
    Context, EkgError, Heart, NHOOD, Server, lead_basis, lib, move_leads, num_steps,
)

__all__ = ["Context", "Heart", "Server", "EkgError", "NHOOD", "lead_basis", "move_leads",
           "num_steps", "lib", "synth"]
