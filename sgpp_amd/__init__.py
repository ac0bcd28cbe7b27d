"""MI355X-native multi-evaluation (mult / multTranspose) of SG++'s datadriven
path: gfx950 HIP kernels behind SG++'s OperationMultipleEval interface.

Layout:
  csrc/            HIP kernels, device runtime, C++ mirror of the SG++
                   interface (grid engine, operator, factory) and the C-ABI
  capi.py          ctypes binding of libsgpp_hip.so (include/sgpp_hip.h)
  workloads.py     seeded synthetic workloads
  build.py         in-tree build of libsgpp_hip.so
"""
__version__ = "1.0.0"
