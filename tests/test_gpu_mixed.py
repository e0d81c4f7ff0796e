"""FP32 sweeps inside FP64 defect correction (NTUST_SOR_MIXED) against the
FP64 red-black solver it accelerates, and against the reference ordering.

The mixed solver performs the FP64 red-black iteration on the correction
e = p - p0 in FP32, so after the same number of sweeps its pressure must
agree with the FP64 solver to FP32 rounding of the correction, and the
velocity after full time steps to the 1e-6 relative bar of the precision
pattern."""
import numpy as np
import pytest

from conftest import rand_halo, stretched_grid
from ntustcfdlab_amd.grid import Grid, bc_codes, noise_ic
from ntustcfdlab_amd.hip import QUICK_CENTRE, SOR_LEXICOGRAPHIC, SOR_MIXED, SOR_REDBLACK, \
    ProjectionSolver

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


# FP32 sweeps between the FP64 corrections of the mixed solve: the first
# segment of a solve is 64 sweeps, the later ones SEGMENT
FIRST_SEGMENT = 64
SEGMENT = 256


@pytest.mark.parametrize("itmax", [1.0, 63.0, 255.0, 256.0, 257.0, 600.0])
def test_mixed_solve_tracks_fp64_iterates(itmax, hip_lib):
    """From p = 0 with a smooth large-scale source the solve is still far
    from converged at itmax, so the last sweep's max change is a meaningful
    value (not round-off) and must agree too, to FP32 resolution.  From
    p = 0 the FP32 correction of the first segment holds the whole
    solution, and its rounding stays in p once folded: the iterates agree
    to m * 2^-24 relative with m the sweeps of that segment (at most
    FIRST_SEGMENT; 27 * 2^-24 measured at itmax 64).  The later segments,
    up to SEGMENT sweeps long, carry only the change since the last fold
    and add little (4.4 * 2^-24 at itmax 257).  Time steps, which start
    from the previous pressure, meet 1e-6 on the velocity
    (test_mixed_steps_velocity_error)."""
    g = stretched_grid(40, 36, 44)
    s = ProjectionSolver(g)
    rng = np.random.default_rng(int(itmax))
    us, vs, ws = (rand_halo(g, rng, 1e-2) for _ in range(3))
    x = np.linspace(0, np.pi, g.halo_shape[0])[:, None, None]
    us += np.asfortranarray(np.broadcast_to(np.sin(x), g.halo_shape))
    p = g.halo()
    out = {}
    try:
        for mode in (SOR_REDBLACK, SOR_MIXED):
            q = p.copy(order="F")
            pre = g.interior()
            ik, pm, _ = s.gauss_seidel(us, vs, ws, q, pre, 1e-3, 1.5, 1e-30, itmax, mode)
            out[mode] = (ik, pm, q)
    finally:
        s.close()
    assert out[SOR_REDBLACK][0] == out[SOR_MIXED][0] == int(itmax)
    m = min(int(itmax), FIRST_SEGMENT)
    assert rel(out[SOR_MIXED][2], out[SOR_REDBLACK][2]) < max(m, 16) * 2.0 ** -24
    # still far above FP32 resolution of the correction
    assert out[SOR_REDBLACK][1] > 1e-7 * np.abs(out[SOR_REDBLACK][2]).max()
    # the change is a difference of FP32 iterates of the correction, whose
    # size in a cold start approaches p's: it agrees to 0.1 %, or to the
    # FP32 resolution of p where the change is that small (0.37 * 2^-24
    # max|p| measured at itmax 257, inside the second segment)
    pmax = np.abs(out[SOR_REDBLACK][2]).max()
    assert abs(out[SOR_MIXED][1] - out[SOR_REDBLACK][1]) <= max(
        1e-3 * out[SOR_REDBLACK][1], 2.0 ** -24 * pmax)


@pytest.mark.parametrize("shape", [(33, 9, 17), (7, 5, 3), (20, 31, 13)])
@pytest.mark.parametrize("itmax", [65.0, 257.0, 530.0])
def test_mixed_tracks_fp64_on_odd_shapes(shape, itmax, hip_lib):
    """Odd extents exercise the colour-split layout's row ends and the
    interior tests of the correction kernel (absorb + residual); the FP32
    solve must still track the FP64 one to one segment's FP32 rounding."""
    g = stretched_grid(*shape)
    s = ProjectionSolver(g)
    rng = np.random.default_rng(sum(shape) + int(itmax))
    us, vs, ws = (rand_halo(g, rng, 1e-2) for _ in range(3))
    p0 = rand_halo(g, rng, 1e-3)
    out = {}
    try:
        for mode in (SOR_REDBLACK, SOR_MIXED):
            q = p0.copy(order="F")
            pre = g.interior()
            ik, pm, _ = s.gauss_seidel(us, vs, ws, q, pre, 1e-3, 1.5, 1e-30, itmax, mode)
            out[mode] = (ik, pm, q)
    finally:
        s.close()
    assert out[SOR_REDBLACK][0] == out[SOR_MIXED][0] == int(itmax)
    assert rel(out[SOR_MIXED][2], out[SOR_REDBLACK][2]) < SEGMENT * 2.0 ** -24


@pytest.mark.parametrize("stop_at", [100, 200, 300, 600])
def test_mixed_stop_inside_a_segment(stop_at, hip_lib):
    """The gated loop of a first solve checks the stop on the host after
    sweeps 128, 256, 384, ..., while the segments end at 64, 320, 576, ...,
    so a segment is queued in parts, and a stop inside a segment must still
    fold that segment's correction into p.  zeta is the FP64 solver's
    change at sweep stop_at, so both solvers stop there (an earlier sweep
    whose change is below zeta may end the solve a little sooner): at 100
    and 200 inside the first and second parts of the segment 65..320, where
    the host closes the segment after the stop; at 300 in a part that runs
    on to the segment's end (closed on the device side, the next segment's
    correction skipped); at 600 in the first part of the segment
    577..832.  The mixed pressure must track the FP64 pressure after the
    same number of sweeps."""
    g = stretched_grid(30, 26, 22)
    rng = np.random.default_rng(stop_at)
    us, vs, ws = (rand_halo(g, rng, 1e-2) for _ in range(3))
    p0 = rand_halo(g, rng, 1e-3)
    s = ProjectionSolver(g)
    try:
        q = p0.copy(order="F")
        _, probe, _ = s.gauss_seidel(us, vs, ws, q, g.interior(), 1e-3, 1.5, 0.0,
                                     float(stop_at), SOR_REDBLACK)
    finally:
        s.close()
    zeta = probe * (1.0 + 1e-6)

    def solve(mode, zeta, itmax):
        s = ProjectionSolver(g)
        try:
            q = p0.copy(order="F")
            ik, _, _ = s.gauss_seidel(us, vs, ws, q, g.interior(), 1e-3, 1.5, zeta, itmax, mode)
        finally:
            s.close()
        return ik, q

    ik64, _ = solve(SOR_REDBLACK, zeta, 5000.0)
    ik32, p32 = solve(SOR_MIXED, zeta, 5000.0)
    print(f"stop_at {stop_at}: FP64 {ik64}, mixed {ik32}", flush=True)
    assert stop_at - 40 <= ik64 <= stop_at and ik64 % 128 != 0
    assert abs(ik32 - ik64) <= 1
    _, ref = solve(SOR_REDBLACK, 0.0, float(ik32))
    assert rel(p32, ref) < SEGMENT * 2.0 ** -24


def test_mixed_stops_where_fp64_stops(hip_lib):
    """zeta reached well before itmax: same sweep count (the change per sweep
    is far from zeta's resolution limit in FP32 here)."""
    g = stretched_grid(16, 14, 12)
    s = ProjectionSolver(g)
    rng = np.random.default_rng(5)
    us, vs, ws = (rand_halo(g, rng, 1e-2) for _ in range(3))
    p = g.halo()
    res = []
    try:
        for mode in (SOR_REDBLACK, SOR_MIXED):
            q = p.copy(order="F")
            ik, pm, _ = s.gauss_seidel(us, vs, ws, q, g.interior(), 1e-3, 1.5, 1e-6, 5000.0, mode)
            res.append((ik, q))
    finally:
        s.close()
    assert abs(res[0][0] - res[1][0]) <= 1
    assert rel(res[1][1], res[0][1]) < 1e-5


def test_mixed_steps_velocity_error(hip_lib):
    """Resident LES cavity steps: mixed vs FP64 red-black velocity within 1e-6."""
    g = Grid(32, 32, 32)
    u0, v0, w0 = noise_ic(g, 1e-3, seed=4)
    finals = {}
    for mode in (SOR_REDBLACK, SOR_MIXED):
        s = ProjectionSolver(g, bc_codes())
        try:
            u, v, w, p = u0.copy(order="F"), v0.copy(order="F"), w0.copy(order="F"), g.halo()
            s.boundary_conditions(u, v, w, p)
            s.configure(dt=5e-4, nu=1e-3, Cs=0.1, LES=1, scheme=QUICK_CENTRE, omega=1.5,
                        zeta=1e-10, itmax=400.0, sor=mode)
            s.upload(u, v, w, p, g.halo())
            ik, *_ = s.step(5)
            out = [g.halo() for _ in range(4)]
            s.download(*out)
            finals[mode] = (ik, out)
        finally:
            s.close()
    a, b = finals[SOR_MIXED][1], finals[SOR_REDBLACK][1]
    vel = max(np.abs(x).max() for x in b[:3])
    err = max(np.abs(x - y).max() for x, y in zip(a[:3], b[:3])) / vel
    assert err < 1e-6, err
