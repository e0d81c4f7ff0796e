"""The fused red-black sweep (one launch per sweep on LDS tiles, two copies
of the colour planes) against the two-pass sweep (one launch per colour, in
place): same sweep count, same last change, same pressure, bit for bit, in
FP64 and in the FP32-sweep solver, on grids whose extents are not
multiples of the tile."""
import numpy as np
import pytest

from conftest import rand_halo, stretched_grid
from ntustcfdlab_amd.grid import Grid, bc_codes, noise_ic
from ntustcfdlab_amd.hip import (QUICK_CENTRE, SOR_MIXED, SOR_REDBLACK, SWEEP_FUSED,
                                 SWEEP_TWO_PASS, ProjectionSolver)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(40, 36, 44), (33, 9, 17), (64, 16, 8), (7, 5, 3), (70, 18, 90)])
@pytest.mark.parametrize("mode", [SOR_REDBLACK, SOR_MIXED])
@pytest.mark.parametrize("itmax,zeta", [(1.0, 0.0), (2.0, 0.0), (65.0, 0.0), (400.0, 1e-7)])
def test_fused_equals_two_pass(shape, mode, itmax, zeta, hip_lib):
    g = stretched_grid(*shape)
    s = ProjectionSolver(g)
    rng = np.random.default_rng(sum(shape) + int(itmax))
    us, vs, ws = (rand_halo(g, rng, 1e-2) for _ in range(3))
    p0 = rand_halo(g, rng, 1e-3)
    out = {}
    try:
        for variant in (SWEEP_TWO_PASS, SWEEP_FUSED):
            s.set_sweep_variant(variant)
            p = p0.copy(order="F")
            pre = g.interior()
            ik, pm, _ = s.gauss_seidel(us, vs, ws, p, pre, 1e-3, 1.5, zeta, itmax, mode)
            out[variant] = (ik, pm, p, pre)
    finally:
        s.close()
    a, b = out[SWEEP_TWO_PASS], out[SWEEP_FUSED]
    assert a[0] == b[0]
    assert a[1] == b[1]
    assert np.array_equal(a[2].view(np.uint64), b[2].view(np.uint64))
    assert np.array_equal(a[3].view(np.uint64), b[3].view(np.uint64))


@pytest.mark.parametrize("mode", [SOR_REDBLACK, SOR_MIXED])
def test_fused_resident_steps_equal_two_pass(mode, hip_lib):
    g = Grid(48, 40, 36)
    u0, v0, w0 = noise_ic(g, 1e-3, seed=12)
    res = {}
    for variant in (SWEEP_TWO_PASS, SWEEP_FUSED):
        s = ProjectionSolver(g, bc_codes())
        try:
            s.set_sweep_variant(variant)
            u, v, w, p = u0.copy(order="F"), v0.copy(order="F"), w0.copy(order="F"), g.halo()
            s.boundary_conditions(u, v, w, p)
            s.configure(dt=5e-4, nu=1e-3, Cs=0.1, LES=1, scheme=QUICK_CENTRE, omega=1.5,
                        zeta=1e-10, itmax=300.0, sor=mode)
            s.upload(u, v, w, p, g.halo())
            ik, pm, cd, cl, _ = s.step(4)
            out = [g.halo() for _ in range(4)]
            s.download(*out)
            res[variant] = (ik, pm, out)
        finally:
            s.close()
    a, b = res[SWEEP_TWO_PASS], res[SWEEP_FUSED]
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    for x, y in zip(a[2], b[2]):
        assert np.array_equal(x.view(np.uint64), y.view(np.uint64))
