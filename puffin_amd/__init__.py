"""MI355X (gfx950) implementation of Puffin's per-step integrator.

Modules:
  params     module constants (puff_params) and scaling relations
  gpu        ctypes front end of libpuffin_gpu.so (the C ABI)
  undulator  the split-step schedule of one undulator module
  configs    seeded synthetic workloads
"""
from .params import FP64, MIXED, ModuleParams  # noqa: F401
