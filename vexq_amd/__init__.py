"""vexq_amd — MI355X (gfx950) execution of vexq's data-parallel hot path.

Layout:
  include/vexq_gpu.h      the C ABI (the drop-in boundary)
  vexq_amd/csrc/          HIP kernels + the C ABI implementation
  vexq_amd/gpu.py         ctypes host binding of libvexqgpu.so
  vexq_amd/host/          C++ mirror of vexq's exec.Operator interface
  vexq_amd/datagen.py     seeded synthetic tables and the workload queries
"""
from . import abi  # noqa: F401

__all__ = ["abi"]
