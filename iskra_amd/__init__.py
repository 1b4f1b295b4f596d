"""iskra_amd -- MI355X (gfx950) electrostatic particle-in-cell step.

The compute path is libiskra_hip.so (hand-written HIP kernels behind the C ABI
of include/iskra.h).  The Python modules mirror the reference's Julia call
surface (RegularGrids, FiniteDifferenceMethod, ParticleInCell, Chemistry) on
top of that ABI.
"""
from ._lib import Context, IskraError, lib  # noqa: F401

__all__ = ["Context", "IskraError", "lib"]
