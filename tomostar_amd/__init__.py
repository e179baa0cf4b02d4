"""tomostar_amd: TomoSTAR's GPU hot path (batched eikonal, ray back-trace, LSQR) on MI355X.

The compute is hand-written HIP for gfx950 in ``csrc/``, built into ``libtomostar_hip.so`` and exposed
through the C ABI of ``include/tomostar_hip.h``; the Python modules are thin host mirrors of the
reference's solver interfaces over that ABI.
"""
from .grid import GridDefinition  # noqa: F401

__all__ = ["GridDefinition"]
