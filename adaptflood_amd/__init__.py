"""MI355X-native local-inertial flood step (the hot path of AdaptFlood).

Layers:
  _lib        ctypes binding of libadaptflood_amd.so (include/adaptflood.h)
  kernels     the reference's per-function call surface (FP64, reference layout)
  domain      fused single-kernel step for one scenario or a ragged batch
  serve       continuous-batching request loop over a device-resident pool
  tune        autotuner for the tile shape, kernel variant and tile order
  formats     .par / .dem.ascii / .bci / .bdy / rain / stage readers, the two
              ESRI-ASCII writers and the NetCDF writer
  scenario    host-side scenario description
  generators  seeded synthetic workloads
"""
from .scenario import PointBC, Scenario  # noqa: F401

__all__ = ["PointBC", "Scenario"]
