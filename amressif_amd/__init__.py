"""MI355X (gfx950) implementation of AMRESSIF's fractional-step time step.

Submodules:
  api        ctypes bindings of include/amressif.h (loads libamressif_hip.so; raises if missing)
  workloads  synthetic inputs of the cavity workloads (host numpy)
  build      in-tree hipcc build of libamressif_hip.so
"""
__all__ = ["api", "workloads", "build"]
