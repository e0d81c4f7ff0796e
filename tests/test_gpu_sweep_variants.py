"""The fused red-black sweep (one launch per sweep on LDS tiles, two copies
of the colour planes) and the graph-replayed two-pass sweep (the sweep
index kept on the device, whole replays past the stopping sweep) against
the two-pass sweep launched from the stream (one launch per colour, in
place): same sweep count, same last change, same pressure, bit for bit, in
FP64 and in the FP32-sweep solver, on grids whose extents are not
multiples of the tile."""
import numpy as np
import pytest

from conftest import rand_halo, stretched_grid
from ntustcfdlab_amd.grid import Grid, bc_codes, noise_ic
from ntustcfdlab_amd.hip import (QUICK_CENTRE, SOR_MIXED, SOR_REDBLACK, SWEEP_FUSED,
                                 SWEEP_GRAPH, SWEEP_TILE, SWEEP_TWO_PASS, ProjectionSolver)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(40, 36, 44), (33, 9, 17), (64, 16, 8), (7, 5, 3), (70, 18, 90)])
@pytest.mark.parametrize("mode", [SOR_REDBLACK, SOR_MIXED])
@pytest.mark.parametrize("itmax,zeta", [(1.0, 0.0), (2.0, 0.0), (65.0, 0.0), (400.0, 1e-7)])
@pytest.mark.parametrize("variant", [SWEEP_FUSED, SWEEP_GRAPH, SWEEP_TILE])
def test_variant_equals_two_pass(shape, mode, itmax, zeta, variant, hip_lib):
    g = stretched_grid(*shape)
    s = ProjectionSolver(g)
    rng = np.random.default_rng(sum(shape) + int(itmax))
    us, vs, ws = (rand_halo(g, rng, 1e-2) for _ in range(3))
    p0 = rand_halo(g, rng, 1e-3)
    out = {}
    try:
        for v in (SWEEP_TWO_PASS, variant):
            s.set_sweep_variant(v)
            p = p0.copy(order="F")
            pre = g.interior()
            ik, pm, _ = s.gauss_seidel(us, vs, ws, p, pre, 1e-3, 1.5, zeta, itmax, mode)
            out[v] = (ik, pm, p, pre)
    finally:
        s.close()
    a, b = out[SWEEP_TWO_PASS], out[variant]
    assert a[0] == b[0]
    assert a[1] == b[1]
    assert np.array_equal(a[2].view(np.uint64), b[2].view(np.uint64))
    assert np.array_equal(a[3].view(np.uint64), b[3].view(np.uint64))


@pytest.mark.parametrize("mode", [SOR_REDBLACK, SOR_MIXED])
@pytest.mark.parametrize("variant", [SWEEP_FUSED, SWEEP_GRAPH, SWEEP_TILE])
def test_resident_steps_equal_two_pass(mode, variant, hip_lib):
    g = Grid(48, 40, 36)
    u0, v0, w0 = noise_ic(g, 1e-3, seed=12)
    res = {}
    for var in (SWEEP_TWO_PASS, variant):
        s = ProjectionSolver(g, bc_codes())
        try:
            s.set_sweep_variant(var)
            u, v, w, p = u0.copy(order="F"), v0.copy(order="F"), w0.copy(order="F"), g.halo()
            s.boundary_conditions(u, v, w, p)
            s.configure(dt=5e-4, nu=1e-3, Cs=0.1, LES=1, scheme=QUICK_CENTRE, omega=1.5,
                        zeta=1e-10, itmax=300.0, sor=mode)
            s.upload(u, v, w, p, g.halo())
            ik, pm, cd, cl, _ = s.step(4)
            out = [g.halo() for _ in range(4)]
            s.download(*out)
            res[var] = (ik, pm, out)
        finally:
            s.close()
    a, b = res[SWEEP_TWO_PASS], res[variant]
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    for x, y in zip(a[2], b[2]):
        assert np.array_equal(x.view(np.uint64), y.view(np.uint64))


@pytest.mark.parametrize("mode", [SOR_REDBLACK, SOR_MIXED])
def test_graph_replays_follow_changed_arguments(mode, hip_lib):
    """One solver context, solves whose captured arguments change from call
    to call (omega, itmax, zeta) and repeat: each must
    match the stream-launched sweeps of a separate context."""
    g = stretched_grid(30, 22, 26)
    rng = np.random.default_rng(77)
    us, vs, ws = (rand_halo(g, rng, 1e-2) for _ in range(3))
    p0 = rand_halo(g, rng, 1e-3)
    calls = [(1.5, 0.0, 40.0), (1.3, 0.0, 40.0), (1.5, 1e-8, 500.0), (1.5, 0.0, 40.0),
             (1.5, 0.0, 130.0), (1.2, 0.0, 7.0), (1.2, 0.0, 7.0)]
    res = {}
    for v in (SWEEP_TWO_PASS, SWEEP_GRAPH):
        s = ProjectionSolver(g)
        try:
            s.set_sweep_variant(v)
            out = []
            for omega, zeta, itmax in calls:
                p = p0.copy(order="F")
                pre = g.interior()
                ik, pm, _ = s.gauss_seidel(us, vs, ws, p, pre, 1e-3, omega, zeta, itmax, mode)
                out.append((ik, pm, p, pre))
            res[v] = out
        finally:
            s.close()
    for a, b in zip(res[SWEEP_TWO_PASS], res[SWEEP_GRAPH]):
        assert a[0] == b[0] and a[1] == b[1]
        assert np.array_equal(a[2].view(np.uint64), b[2].view(np.uint64))
        assert np.array_equal(a[3].view(np.uint64), b[3].view(np.uint64))


@pytest.mark.parametrize("mode", [SOR_REDBLACK, SOR_MIXED])
def test_speculative_first_batch_and_replay(mode, hip_lib):
    """The two-pass solve runs its first batch (7/8 of the previous solve's
    sweep count) without the device gate and replays from the initial state
    when the loop would have stopped inside it.  A sequence of solves whose
    counts drop (a replay), hold (none) and rise (the gated tail) must give
    the same bits as the gated graph replay, solve by solve."""
    g = stretched_grid(26, 22, 20)
    rng = np.random.default_rng(31)
    us, vs, ws = (rand_halo(g, rng, 1e-2) for _ in range(3))
    p0 = rand_halo(g, rng, 1e-3)
    probe = ProjectionSolver(g)
    try:
        probe.set_sweep_variant(SWEEP_GRAPH)
        p = p0.copy(order="F")
        _, pm300, _ = probe.gauss_seidel(us, vs, ws, p, g.interior(), 1e-3, 1.5, 0.0, 300.0,
                                         mode)
    finally:
        probe.close()
    z = pm300 * (1.0 + 1e-9)
    calls = [(0.0, 700.0), (z, 700.0), (z, 700.0), (0.0, 700.0), (z, 700.0), (0.0, 90.0)]
    res = {}
    for v in (SWEEP_TWO_PASS, SWEEP_GRAPH):
        s = ProjectionSolver(g)
        try:
            s.set_sweep_variant(v)
            out = []
            for zeta, itmax in calls:
                p = p0.copy(order="F")
                pre = g.interior()
                ik, pm, _ = s.gauss_seidel(us, vs, ws, p, pre, 1e-3, 1.5, zeta, itmax, mode)
                out.append((ik, pm, p, pre))
            res[v] = out
        finally:
            s.close()
    iks = [r[0] for r in res[SWEEP_TWO_PASS]]
    assert iks[0] == 700 and iks[3] == 700 and iks[5] == 90
    assert 250 <= iks[1] <= 310 and iks[1] < 0.875 * 700     # replayed
    for a, b in zip(res[SWEEP_TWO_PASS], res[SWEEP_GRAPH]):
        assert a[0] == b[0] and a[1] == b[1]
        assert np.array_equal(a[2].view(np.uint64), b[2].view(np.uint64))
        assert np.array_equal(a[3].view(np.uint64), b[3].view(np.uint64))


@pytest.mark.parametrize("mode", [SOR_REDBLACK, SOR_MIXED])
def test_speculative_tail_replay(mode, hip_lib):
    """After a solve that ran to itmax, the two-pass solve runs its first
    batch (7/8 of itmax) and then the tail up to itmax without the device
    gate, from a copy of the colour planes.  A stop inside the tail restores
    the planes from that copy and replays the tail's sweeps up to the stop.
    zeta is the change of a probe's sweep 650, so the stop lands between
    7/8 * 700 and 700: every solve must give the gated graph replay's bits."""
    g = stretched_grid(26, 22, 20)
    rng = np.random.default_rng(37)
    us, vs, ws = (rand_halo(g, rng, 1e-2) for _ in range(3))
    p0 = rand_halo(g, rng, 1e-3)
    probe = ProjectionSolver(g)
    try:
        probe.set_sweep_variant(SWEEP_GRAPH)
        p = p0.copy(order="F")
        _, pm650, _ = probe.gauss_seidel(us, vs, ws, p, g.interior(), 1e-3, 1.5, 0.0, 650.0,
                                         mode)
    finally:
        probe.close()
    z = pm650 * (1.0 + 1e-9)
    calls = [(0.0, 700.0), (z, 700.0), (0.0, 700.0), (z, 700.0)]
    res = {}
    for v in (SWEEP_TWO_PASS, SWEEP_GRAPH):
        s = ProjectionSolver(g)
        try:
            s.set_sweep_variant(v)
            out = []
            for zeta, itmax in calls:
                p = p0.copy(order="F")
                pre = g.interior()
                ik, pm, _ = s.gauss_seidel(us, vs, ws, p, pre, 1e-3, 1.5, zeta, itmax, mode)
                out.append((ik, pm, p, pre))
            res[v] = out
        finally:
            s.close()
    iks = [r[0] for r in res[SWEEP_TWO_PASS]]
    print(f"mode {mode}: sweeps {iks}", flush=True)
    assert iks[0] == 700 and iks[2] == 700
    for q in (iks[1], iks[3]):
        assert 0.875 * 700 < q < 700                          # stopped inside the tail
    for a, b in zip(res[SWEEP_TWO_PASS], res[SWEEP_GRAPH]):
        assert a[0] == b[0] and a[1] == b[1]
        assert np.array_equal(a[2].view(np.uint64), b[2].view(np.uint64))
        assert np.array_equal(a[3].view(np.uint64), b[3].view(np.uint64))
