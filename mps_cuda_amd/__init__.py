"""MI355X (gfx950) implementation of the explicit MPS time step.

The compute path is the HIP library ``libmps_hip.so`` (C ABI in
include/mps.h); ``ffi`` binds it, ``configs`` defines the workloads and
``mps`` mirrors the reference's ``class MPS`` on top of it.
"""
from . import ffi  # noqa: F401
