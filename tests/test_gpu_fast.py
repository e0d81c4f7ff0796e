"""The fast arithmetic modes of the resident step's predictor and eddy
viscosity (ntust_set_arithmetic) against the exact mode, which the other
GPU tests pin bit for bit to the oracle:
  * ARITH_FAST (FP64, reciprocal tables, fused multiply-adds): one step with
    a fixed sweep count (zeta = 0) agrees with the exact step to 1e-12 of
    the state's magnitude, on stretched grids, with and without LES;
  * ARITH_FP32 (the increment u* - u and nu_t in FP32): the same step within
    1e-7 of the state's magnitude;
  * the whole FP32 pattern -- FP32 predictor, eddy viscosity and
    SOR sweeps inside FP64 defect correction -- on the 128^3 LES cavity for
    three steps against GS.f90's own order in FP64 (the lexicographic
    solver, the oracle's bits): velocity within 1e-6."""
import numpy as np
import pytest

from conftest import stretched_grid
from ntustcfdlab_amd import configs
from ntustcfdlab_amd.grid import Grid, bc_codes, noise_ic
from ntustcfdlab_amd.hip import ARITH_EXACT, ARITH_FAST, ARITH_FP32, QUICK_CENTRE, \
    SOR_LEXICOGRAPHIC, SOR_MIXED, SOR_REDBLACK, NtustError, ProjectionSolver

pytestmark = pytest.mark.gpu


def run(g, arith, les, state, steps=1, zeta=0.0, itmax=40.0, sor=SOR_REDBLACK, dt=1e-3,
        nu=1e-2):
    s = ProjectionSolver(g, bc_codes())
    try:
        s.set_arithmetic(arith)
        s.configure(dt=dt, nu=nu, Cs=0.17, LES=les, scheme=QUICK_CENTRE, omega=1.5,
                    zeta=zeta, itmax=itmax, sor=sor, lx=1.0, force_every=1)
        s.upload(*[a.copy(order="F") for a in state], g.halo())
        ik, *_ = s.step(steps)
        out = [g.halo() for _ in range(4)]
        s.download(*out)
    finally:
        s.close()
    return out, list(ik)


@pytest.mark.parametrize("les", [0, 1])
@pytest.mark.parametrize("shape", [(14, 12, 10), (33, 9, 17), (5, 2, 3), (7, 3, 2), (70, 6, 37)])
def test_fast_modes_against_exact_step(shape, les, hip_lib):
    """(5, 2, 3), (7, 3, 2), (70, 6, 37): rows and planes that leave the
    eddy viscosity's 4-row, 32-plane thread segments and the predictor's
    64-wide rows partly empty."""
    g = stretched_grid(*shape)
    u, v, w = noise_ic(g, 0.3, seed=sum(shape) + les)
    p = g.halo()
    s = ProjectionSolver(g, bc_codes())
    s.boundary_conditions(u, v, w, p)
    s.close()
    state = [u, v, w, p]
    ref, ik0 = run(g, ARITH_EXACT, les, state)
    for arith, tol in ((ARITH_FAST, 1e-12), (ARITH_FP32, 1e-7)):
        got, ik = run(g, arith, les, state)
        assert ik == ik0
        for a, b, name in zip(got, ref, "uvwp"):
            e = np.abs(a - b).max() / np.abs(b).max()
            print(f"{shape} LES {les} mode {arith}: {name} {e:.2e}", flush=True)
            assert e <= tol, (name, e)


def test_set_arithmetic_refuses_unknown_mode(hip_lib):
    g = Grid(4, 4, 4)
    s = ProjectionSolver(g)
    try:
        with pytest.raises(NtustError, match="arithmetic"):
            s.set_arithmetic(7)
    finally:
        s.close()


def test_fp32_pattern_les_cavity_128_vs_reference_order(hip_lib):
    w = configs.les_cavity(128)
    g = Grid(128, 128, 128)
    u, v, ww = noise_ic(g, w["noise"], w["seed"])
    p = g.halo()
    s = ProjectionSolver(g, bc_codes())
    s.boundary_conditions(u, v, ww, p)
    s.close()
    state = [u, v, ww, p]
    kw = dict(steps=3, zeta=w["zeta"], itmax=w["itmax"], dt=w["dt"], nu=1.0 / w["Re"])
    ref, ik_ref = run(g, ARITH_EXACT, 1, state, sor=SOR_LEXICOGRAPHIC, **kw)
    got, ik = run(g, ARITH_FP32, 1, state, sor=SOR_MIXED, **kw)
    vel = max(np.abs(x).max() for x in ref[:3])
    ev = max(np.abs(a - b).max() for a, b in zip(got[:3], ref[:3])) / vel
    ep = np.abs(got[3] - ref[3]).max() / np.abs(ref[3]).max()
    print(f"FP32 predictor + LES + sweeps vs lexicographic FP64, 3 steps at 128^3: velocity "
          f"{ev:.2e}, pressure {ep:.2e}, sweeps {ik} / {ik_ref}", flush=True)
    assert ev <= 1e-6
