"""MI355X-native rigid-body stepper: the reference engine's per-substep hot path
(narrow phase, graph-coloured soft-step contact solver, integration, finalize) as
gfx950 kernels behind the C ABI declared in include/natsel_gpu.h.

`natsel_amd.world.World` mirrors the reference's World API over that ABI.
"""
from .capi import load_library, library_path  # noqa: F401
