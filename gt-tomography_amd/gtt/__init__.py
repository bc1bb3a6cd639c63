"""gt-tomography on MI355X: Python view of the gfx950 cone-beam projector.

``capi`` binds the C ABI (include/cbct.h, libcbct.so); ``operators`` binds the
C++ operator/solver layer (libgtt.so) that mirrors the reference's
cuConebeamProjectionOperator / CBSubsetOperator / osSARTSolver /
osMOMSolverD / FDK surface; ``phantom`` holds the seeded test phantoms and
scan geometries.
"""
from . import capi, phantom  # noqa: F401
