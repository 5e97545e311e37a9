"""dpgsolver_amd -- MI355X-native DG residual evaluation behind DPGSolver's entry points.

The compute path is the HIP library ``libdgx.so`` (hand-written gfx950 kernels behind a
C ABI, declared in ``include/dgx.h``).  This package holds the mesh and state generators
(``mesh.py``: structured QUAD/HEX, variable-order and mixed TRI/QUAD meshes, the benchmark
workloads) and the ctypes binding of the C ABI (``_lib.py``: ``Context`` for one
discretisation, graphs, the implicit solve, the learned preconditioner).
"""
from . import mesh  # noqa: F401

__all__ = ["mesh"]
