"""The speculative sweep batches of the red-black solvers against the gated
sweeps.  A solver's first solve gates every sweep on the device (the sweep
reads the previous sweep's maximum change and returns at once when GS.f90's
loop would have stopped); later solves run a batch of 7/8 of the previous
solve's sweep count, and after a solve that ran to itmax the tail up to
itmax, without the gate, and replay from a copy when the loop would have
stopped inside.  Same operations, same count, same bits: each solve of a
sequence on one solver must equal, bit for bit, the same solve done first
on a fresh solver (gated), in FP64 and in the FP32-sweep solver, on grids
whose extents are not multiples of the thread blocks."""
import numpy as np
import pytest

from conftest import rand_halo, stretched_grid
from ntustcfdlab_amd.hip import SOR_MIXED, SOR_REDBLACK, ProjectionSolver

pytestmark = pytest.mark.gpu


def gated(g, us, vs, ws, p0, zeta, itmax, mode, omega=1.5):
    """The solve as the first one of a fresh solver: every sweep gated."""
    s = ProjectionSolver(g)
    try:
        p = p0.copy(order="F")
        pre = g.interior()
        ik, pm, _ = s.gauss_seidel(us, vs, ws, p, pre, 1e-3, omega, zeta, itmax, mode)
    finally:
        s.close()
    return ik, pm, p, pre


def sequence(g, us, vs, ws, p0, calls, mode):
    """The solves one after the other on one solver (the second and later
    ones speculative)."""
    s = ProjectionSolver(g)
    out = []
    try:
        for zeta, itmax in calls:
            p = p0.copy(order="F")
            pre = g.interior()
            ik, pm, _ = s.gauss_seidel(us, vs, ws, p, pre, 1e-3, 1.5, zeta, itmax, mode)
            out.append((ik, pm, p, pre))
    finally:
        s.close()
    return out


def assert_same(a, b):
    assert a[0] == b[0] and a[1] == b[1]
    assert np.array_equal(a[2].view(np.uint64), b[2].view(np.uint64))
    assert np.array_equal(a[3].view(np.uint64), b[3].view(np.uint64))


@pytest.mark.parametrize("shape", [(40, 36, 44), (33, 9, 17), (64, 16, 8), (7, 5, 3), (70, 18, 90)])
@pytest.mark.parametrize("mode", [SOR_REDBLACK, SOR_MIXED])
def test_speculative_equals_gated(shape, mode, hip_lib):
    """Two itmax-capped solves (the second runs its first batch and tail
    speculatively: 7/8 of 800 sweeps, for the FP32-sweep solver rounded
    down to its last segment end, 576, then the tail), then a solve that
    the stopping rule ends."""
    g = stretched_grid(*shape)
    rng = np.random.default_rng(sum(shape))
    us, vs, ws = (rand_halo(g, rng, 1e-2) for _ in range(3))
    p0 = rand_halo(g, rng, 1e-3)
    calls = [(0.0, 800.0), (0.0, 800.0), (1e-7, 800.0)]
    seq = sequence(g, us, vs, ws, p0, calls, mode)
    for (zeta, itmax), a in zip(calls, seq):
        assert_same(a, gated(g, us, vs, ws, p0, zeta, itmax, mode))


@pytest.mark.parametrize("mode", [SOR_REDBLACK, SOR_MIXED])
def test_speculative_first_batch_and_replay(mode, hip_lib):
    """A sequence of solves whose counts drop (the first batch runs past
    the stop and is replayed), hold and rise (the gated tail after the
    batch) must give the gated solves' bits, solve by solve."""
    g = stretched_grid(26, 22, 20)
    rng = np.random.default_rng(31)
    us, vs, ws = (rand_halo(g, rng, 1e-2) for _ in range(3))
    p0 = rand_halo(g, rng, 1e-3)
    _, pm300, _, _ = gated(g, us, vs, ws, p0, 0.0, 300.0, mode)
    z = pm300 * (1.0 + 1e-9)
    calls = [(0.0, 700.0), (z, 700.0), (z, 700.0), (0.0, 700.0), (z, 700.0), (0.0, 90.0)]
    seq = sequence(g, us, vs, ws, p0, calls, mode)
    iks = [r[0] for r in seq]
    assert iks[0] == 700 and iks[3] == 700 and iks[5] == 90
    assert 250 <= iks[1] <= 310 and iks[1] < 0.875 * 700     # replayed
    for (zeta, itmax), a in zip(calls, seq):
        assert_same(a, gated(g, us, vs, ws, p0, zeta, itmax, mode))


@pytest.mark.parametrize("mode", [SOR_REDBLACK, SOR_MIXED])
def test_speculative_tail_replay(mode, hip_lib):
    """After a solve that ran to itmax, the next solve runs its first batch
    (7/8 of itmax) and then the tail up to itmax without the device gate,
    from a copy of the colour planes.  A stop inside the tail restores the
    planes from that copy and replays the tail's sweeps up to the stop.
    zeta is the change of a gated probe's sweep 650, so the stop lands
    between 7/8 * 700 and 700."""
    g = stretched_grid(26, 22, 20)
    rng = np.random.default_rng(37)
    us, vs, ws = (rand_halo(g, rng, 1e-2) for _ in range(3))
    p0 = rand_halo(g, rng, 1e-3)
    _, pm650, _, _ = gated(g, us, vs, ws, p0, 0.0, 650.0, mode)
    z = pm650 * (1.0 + 1e-9)
    calls = [(0.0, 700.0), (z, 700.0), (0.0, 700.0), (z, 700.0)]
    seq = sequence(g, us, vs, ws, p0, calls, mode)
    iks = [r[0] for r in seq]
    print(f"mode {mode}: sweeps {iks}", flush=True)
    assert iks[0] == 700 and iks[2] == 700
    for q in (iks[1], iks[3]):
        assert 0.875 * 700 < q < 700                          # stopped inside the tail
    for (zeta, itmax), a in zip(calls, seq):
        assert_same(a, gated(g, us, vs, ws, p0, zeta, itmax, mode))


@pytest.mark.parametrize("mode", [SOR_REDBLACK, SOR_MIXED])
def test_speculative_tail_beyond_the_batch_limit(mode, hip_lib):
    """itmax above the 65536 sweeps one speculative batch or tail may hold:
    the second solve runs its batch, a tail cut at the limit (for the
    FP32-sweep solver at the last segment end below it, 65344) and the
    gated sweeps from there to itmax; both solves must equal a fresh
    solver's gated solve bit for bit.  zeta < 0: no sweep's change is at or
    below it, so every solve runs to itmax even where the tiny grid's
    iteration reaches a fixed point."""
    g = stretched_grid(7, 5, 3)
    rng = np.random.default_rng(41)
    us, vs, ws = (rand_halo(g, rng, 1e-2) for _ in range(3))
    p0 = rand_halo(g, rng, 1e-3)
    calls = [(-1.0, 70000.0), (-1.0, 70000.0)]
    seq = sequence(g, us, vs, ws, p0, calls, mode)
    assert [r[0] for r in seq] == [70000, 70000]
    ref = gated(g, us, vs, ws, p0, -1.0, 70000.0, mode)
    for a in seq:
        assert_same(a, ref)
