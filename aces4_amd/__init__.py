"""aces4_amd -- the SIP block-contraction hot path (contract and permute
super-instructions) on MI355X / gfx950.

The product is the C-ABI library libsipgpu.so (include/sipgpu.h); this
package holds its HIP sources (csrc/), the build script, ctypes bindings
(sipgpu) and the synthetic coupled-cluster workloads (workloads).
"""
from . import sipgpu  # noqa: F401

__all__ = ["sipgpu"]
