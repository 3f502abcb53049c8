"""Host-side Python harness for the MI355X explicit elastic-wave step library.

`setup`  -- driver-side problem setup (problem.in, meshes, boundary arrays).
`cases`  -- the benchmark / test workloads.
`lib`    -- ctypes binding of libdamian_hip.so (imported lazily: it fails
            loudly when the HIP library has not been built).
"""
from . import setup  # noqa: F401

__all__ = ["setup"]
