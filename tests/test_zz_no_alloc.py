"""The allocation-free force integral (ntust_virtual_force_integrator_noalloc)
against the allocating entry point (ntust_virtual_force_integrator), bit for
bit, on two calls in a row, the second on a smaller grid: the static scratch
then still holds the first call's line sums where the second grid's layout
needs zeros (rows 0 and ny+1 of the z-line sums, entries 0 and nx+1 of the
y-line sums), so a missing clear shows as a different coefficient."""
import ctypes as ct

import numpy as np
import pytest

from conftest import stretched_grid
from ntustcfdlab_amd import hiptime
from ntustcfdlab_amd.hip import ProjectionSolver

pytestmark = pytest.mark.gpu


def _rt():
    rt = hiptime._rt()
    rt.hipMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t]
    rt.hipFree.argtypes = [ct.c_void_p]
    rt.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
    return rt


def to_device(rt, a):
    p = ct.c_void_p()
    assert rt.hipMalloc(ct.byref(p), a.nbytes) == 0
    assert rt.hipMemcpy(p, a.ctypes.data_as(ct.c_void_p), a.nbytes, 1) == 0   # H2D
    return p


def force_fields(g, rng, j0, j1):
    """FY, FZ of a body spanning all of x and rows j0..j1 (halo planes 0, as
    the projection leaves them), with exact zeros inside it so the skip tests
    of the line sums (virtualForceIntegrator.f90:41-75) take both branches."""
    FY, FZ = g.halo(), g.halo()
    k0, k1 = 2, g.nz - 1
    for a in (FY, FZ):
        blk = rng.uniform(-1.0, 1.0, (g.nx, j1 - j0 + 1, k1 - k0 + 1))
        blk[rng.random(blk.shape) < 0.15] = 0.0
        a[2:g.nx + 2, j0 + 1:j1 + 2, k0 + 1:k1 + 2] = blk
    return FY, FZ


@pytest.mark.parametrize("repeat", [0])
def test_noalloc_force_integral_matches_entry_point(repeat, hip_lib):
    rt = _rt()
    rng = np.random.default_rng(19)
    # the body touches j = 1 on the first grid and j = ny on the second one,
    # where the line sums read the rows outside the array
    for shape, (j0, j1) in (((20, 14, 9), (1, 11)), ((12, 9, 7), (3, 9))):
        g = stretched_grid(*shape)
        s = ProjectionSolver(g)
        ptrs = []
        try:
            FY, FZ = force_fields(g, rng, j0, j1)
            cd, cl = s.virtualForceIntegrator(FY, FZ, 1.3)
            dfy, dfz = to_device(rt, FY), to_device(rt, FZ)
            dc = to_device(rt, np.full(2, np.nan))
            ptrs += [dfy, dfz, dc]
            s.virtualForceIntegrator_noalloc(dfy.value, dfz.value, 1.3, dc.value)
            got = np.empty(2)
            assert rt.hipMemcpy(got.ctypes.data_as(ct.c_void_p), dc, 16, 2) == 0   # D2H
            print(f"{shape}: cDrag {cd!r} / {got[0]!r}, cLift {cl!r} / {got[1]!r}", flush=True)
            assert cd != 0.0 and cl != 0.0
            assert np.array_equal(got.view(np.uint64), np.array([cd, cl]).view(np.uint64))
        finally:
            for p in ptrs:
                rt.hipFree(p)
            s.close()


def test_noalloc_force_integral_refuses_oversized_grid(hip_lib):
    from ntustcfdlab_amd.grid import Grid
    from ntustcfdlab_amd.hip import NtustError
    g = Grid(1500, 400, 2)   # 2 (nx+2)(ny+3) > 2^20
    s = ProjectionSolver(g)
    try:
        with pytest.raises(NtustError, match="allocation-free"):
            s.virtualForceIntegrator_noalloc(1, 1, 1.0, 1)
    finally:
        s.close()
