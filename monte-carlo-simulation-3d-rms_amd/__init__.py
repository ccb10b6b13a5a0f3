"""rms_amd -- MI355X-native random-walk engine with the file contract of the
Realistic Microstructure Simulator (RMS) CUDA engine.

The package directory name is not a Python identifier; load it with
`load_package()` from __graft_entry__ (or importlib) as ``rms_amd``.

Modules:
  engine    Engine / Plan / Geometry over librms_amd.so (the HIP engine)
  fileio    setupRMS-style input writer, readRMS / readRMS_pgse readers
  analysis  akc_mom cumulants, batch-means standard errors
  geometry  synthetic planes, cylinders, beaded/undulating axons, bundles
  autograd  torch.autograd.Function over the adjoint pass
  autotune  launch-shape and packing search
  workloads the five benchmark workloads as job lists
  build     hipcc recipe
"""
from . import _capi  # noqa: F401
from . import workloads  # noqa: F401
from .engine import Engine, Geometry, Plan, nval, split_records  # noqa: F401

__all__ = ["Engine", "Geometry", "Plan", "nval", "split_records"]
