"""parabiofvm_amd — paraBioFVM's per-timestep solver loop (secretion, LOD Thomas sweeps, Dirichlet,
gradient) as hand-written HIP kernels for MI355X (gfx950), behind a C ABI (include/parabiofvm_hip.h).

The compute library is libparabiofvm_hip.so, built in-tree from parabiofvm_amd/csrc; the C++ drop-in
facade of the reference's solver classes lives in parabiofvm_amd/cpp.  This Python package is the
test and benchmark harness over the same C ABI.
"""
from ._native import BFVM_SWEEP_EXACT, BFVM_SWEEP_PARTITIONED, BFVM_SWEEP_STAGED, BfvmError, load  # noqa: F401
from .solver import Batch, DeviceSolver, grid_shape  # noqa: F401
