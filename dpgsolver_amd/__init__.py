"""dpgsolver_amd -- MI355X-native DG residual evaluation behind DPGSolver's entry points.

The compute path is the HIP library ``libdgx.so`` (hand-written gfx950 kernels behind a
C ABI, declared in ``include/dgx.h``).  This package holds the host-side mirror of the
reference's solver interface (``solver.py``), the mesh/state generators (``mesh.py``)
and the ctypes binding (``_lib.py``).
"""
from . import mesh  # noqa: F401

__all__ = ["mesh"]
