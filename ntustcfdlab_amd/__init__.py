"""MI355X (gfx950) implementation of the per-time-step hot path of the
HSUAN221/NTUSTCFDLAB projection-method Navier-Stokes solver.

    csrc/        HIP kernels and the C ABI (libntust_hip.so, include/ntust.h)
    fortran/     iso_c_binding interface, drop-in shims, Fortran driver
    hip.py       Python host mirror of the reference interface (ctypes)
    grid.py      host-side set-up: metrics, BC codes, initial fields
    build.py     build recipes
"""
from .grid import Grid, bc_codes  # noqa: F401

__all__ = ["Grid", "bc_codes"]
