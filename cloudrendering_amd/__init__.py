"""cloudrendering_amd: MI355X (gfx950) volumetric path tracer behind the reference's `vpt` module surface.

`cloudrendering_amd.vpt` is the drop-in for the reference's PyTorch module; the compute path is libcrvpt.so
(hand-written HIP kernels for gfx950 behind the C ABI in include/crvpt.h).
"""
from . import vpt  # noqa: F401

__all__ = ["vpt"]
