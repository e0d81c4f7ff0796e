"""Discrete adjoint of the resident step (ntust_adjoint_*).

The adjoint has no counterpart in the reference; it is checked against the
forward step it differentiates, which the other GPU tests pin bit for bit
to the CPU restatement of the reference:
  * dot-product test of one step: <lambda, J dx> = <J^T lambda, dx>, with
    J dx the central difference of the FP64 forward step (exact for the
    bilinear QUICK / upwind predictor up to rounding; O(h^2) with LES);
  * the gradient of a time-averaged kinetic energy with respect to the body
    force, from the checkpointed reverse pass, against central finite
    differences of the forward along seeded random directions;
  * checkpoint spacing and the per-step entry point give the same result.
The pressure solve runs a fixed sweep count where the test needs the
forward map to be smooth (zeta = 0, itmax = k), else the adjoint is still
exact but a perturbation could change the sweep count.
"""
import zlib

import numpy as np
import pytest

from conftest import rand_halo, stretched_grid
from ntustcfdlab_amd.grid import Grid, bc_codes, gaussian_blob, noise_ic
from ntustcfdlab_amd.hip import QUICK_CENTRE, SOR_MIXED, SOR_REDBLACK, UPWIND, ProjectionSolver

pytestmark = pytest.mark.gpu


def dot(a, b):
    return float(sum(np.vdot(x, y) for x, y in zip(a, b)))


def solver_for(g, bc, *, scheme, les, sor=SOR_REDBLACK, zeta=0.0, itmax=30.0, dt=1e-3, nu=1e-2):
    s = ProjectionSolver(g, bc)
    s.configure(dt=dt, nu=nu, Cs=0.17, LES=les, scheme=scheme, omega=1.4, zeta=zeta,
                itmax=itmax, sor=sor, force_every=0)
    return s


def one_step(s, state, eta):
    u, v, w, p = (a.copy(order="F") for a in state)
    s.upload(u, v, w, p, eta)
    s.step(1)
    out = [s.grid.halo() for _ in range(4)]
    s.download(*out)
    return out


CASES = [
    # scheme, LES, BC codes (W E S N B F x u v w), ETA
    ("quick-les-cavity", QUICK_CENTRE, 1, None, False),
    ("quick-mixed-bc-eta", QUICK_CENTRE, 0,
     [0, 1, 2, 1, 0, 1, 2, 2, 0, 0, 1, 2, 1, 1, 0, 2, 0, 1], True),
    ("upwind-cavity", UPWIND, 0, None, False),
    ("quick-les-neumann-eta", QUICK_CENTRE, 1,
     [1, 1, 1, 0, 0, 0, 1, 0, 1, 2, 2, 2, 0, 1, 1, 1, 2, 0], True),
]


@pytest.mark.parametrize("name,scheme,les,bc,with_eta", CASES, ids=[c[0] for c in CASES])
def test_adjoint_step_dot_product(name, scheme, les, bc, with_eta, hip_lib):
    g = stretched_grid(14, 12, 10)
    bc = bc_codes() if bc is None else bc
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    state = [rand_halo(g, rng, 0.5) for _ in range(3)] + [rand_halo(g, rng, 1e-3)]
    eta = g.halo()
    if with_eta:
        eta[...] = np.where(rng.uniform(size=eta.shape) < 0.2, rng.uniform(size=eta.shape), 0.0)
    dx = [rand_halo(g, rng, 1.0) for _ in range(3)] + [rand_halo(g, rng, 1e-3)]
    lam = [rand_halo(g, rng, 1.0) for _ in range(4)]
    s = solver_for(g, bc, scheme=scheme, les=les)
    try:
        h = 1e-7
        plus = one_step(s, [x + h * d for x, d in zip(state, dx)], eta)
        minus = one_step(s, [x - h * d for x, d in zip(state, dx)], eta)
        jdx = [(a - b) / (2 * h) for a, b in zip(plus, minus)]
        u, v, w, p = (a.copy(order="F") for a in state)
        s.upload(u, v, w, p, eta)
        s.adjoint_set(*lam)
        ik = s.adjoint_step()
        jtl = s.adjoint_get()
        # the state is left as it was
        after = [g.halo() for _ in range(4)]
        s.download(*after)
    finally:
        s.close()
    assert ik == 30
    for a, b in zip(after, state):
        assert np.array_equal(a, b)
    lhs, rhs = dot(lam, jdx), dot(jtl, dx)
    assert abs(lhs - rhs) <= 1e-6 * max(abs(lhs), abs(rhs)), (lhs, rhs)


def test_adjoint_pressure_transpose_with_converging_sweeps(hip_lib):
    """zeta > 0: the sweep count comes from the convergence test; a small
    perturbation keeps it, and the transposed sweeps match it."""
    g = stretched_grid(12, 10, 14)
    rng = np.random.default_rng(11)
    state = [rand_halo(g, rng, 0.3) for _ in range(3)] + [g.halo()]
    eta = g.halo()
    dx = [rand_halo(g, rng, 1.0) for _ in range(3)] + [rand_halo(g, rng, 1e-4)]
    lam = [rand_halo(g, rng, 1.0) for _ in range(4)]
    s = solver_for(g, bc_codes(), scheme=QUICK_CENTRE, les=1, zeta=1e-9, itmax=500.0)
    try:
        h = 1e-7
        plus = one_step(s, [x + h * d for x, d in zip(state, dx)], eta)
        minus = one_step(s, [x - h * d for x, d in zip(state, dx)], eta)
        jdx = [(a - b) / (2 * h) for a, b in zip(plus, minus)]
        s.upload(*[a.copy(order="F") for a in state], eta)
        s.adjoint_set(*lam)
        ik = s.adjoint_step()
        jtl = s.adjoint_get()
    finally:
        s.close()
    assert 1 < ik < 500
    lhs, rhs = dot(lam, jdx), dot(jtl, dx)
    assert abs(lhs - rhs) <= 1e-6 * max(abs(lhs), abs(rhs)), (lhs, rhs)


def cavity_setup(n=20, seed=3):
    g = Grid(n, n - 2, n - 4)
    u, v, w = noise_ic(g, 1e-2, seed=seed)
    p = g.halo()
    return g, [u, v, w, p]


def forward_J(s, g, state, eta, f, nsteps, navg):
    """J = mean KE after each of the last navg steps, by plain stepping."""
    s.set_body_force(*f)
    u, v, w, p = (a.copy(order="F") for a in state)
    s.upload(u, v, w, p, eta)
    tot = 0.0
    for n in range(nsteps):
        s.step(1)
        if n >= nsteps - navg:
            tot += s.kinetic_energy()
    return tot / navg


@pytest.mark.parametrize("les,zeta", [(1, 0.0), (1, 1e-9), (0, 0.0)])
def test_adjoint_run_gradient_matches_central_fd(les, zeta, hip_lib):
    g, state = cavity_setup()
    eta = g.halo()
    N, K, M = 12, 4, 5
    s = ProjectionSolver(g, bc_codes())
    try:
        s.configure(dt=2e-3, nu=1.0 / 400.0, Cs=0.1, LES=les, scheme=QUICK_CENTRE, omega=1.5,
                    zeta=zeta, itmax=60.0, sor=SOR_REDBLACK, force_every=0)
        s.boundary_conditions(*state)
        f = gaussian_blob(g, 0.1, 0.1, (0.4, 0.55, 0.45))
        s.set_body_force(*f)
        s.upload(*[a.copy(order="F") for a in state], eta)
        J, ik, ms = s.adjoint_run(N, K, M)
        gf = s.adjoint_gradient()
        # J equals the plain forward's
        J0 = forward_J(s, g, state, eta, f, N, M)
        assert J == J0
        rng = np.random.default_rng(7)
        for _ in range(3):
            d = [g.halo() for _ in range(3)]
            for a in d:
                a[2:-2, 2:-2, 2:-2] = rng.standard_normal(a[2:-2, 2:-2, 2:-2].shape)
            h = 1e-6
            jp = forward_J(s, g, state, eta, [x + h * y for x, y in zip(f, d)], N, M)
            jm = forward_J(s, g, state, eta, [x - h * y for x, y in zip(f, d)], N, M)
            fd = (jp - jm) / (2 * h)
            ad = dot(gf, d)
            assert abs(fd - ad) <= 1e-5 * abs(ad), (fd, ad)
    finally:
        s.close()
    assert ik.min() >= 1 and ik.max() <= 60


def test_checkpoint_spacing_and_step_entry_agree(hip_lib):
    g, state = cavity_setup(16, seed=9)
    eta = g.halo()
    N, M = 6, 3
    s = ProjectionSolver(g, bc_codes())
    try:
        s.configure(dt=2e-3, nu=1.0 / 400.0, Cs=0.1, LES=1, scheme=QUICK_CENTRE, omega=1.5,
                    zeta=1e-10, itmax=80.0, sor=SOR_REDBLACK, force_every=0)
        s.boundary_conditions(*state)
        f = gaussian_blob(g, 0.1, 0.15, (0.5, 0.5, 0.5))
        s.set_body_force(*f)
        res = {}
        for K in (1, 2, 4, 6):
            s.upload(*[a.copy(order="F") for a in state], eta)
            J, ik, _ = s.adjoint_run(N, K, M)
            res[K] = (J, ik, s.adjoint_gradient(), s.adjoint_get())
        # the same reverse pass by hand: checkpoints of every step, then
        # ntust_adjoint_step from each
        s.upload(*[a.copy(order="F") for a in state], eta)
        s.checkpoint_reserve(N + 1)
        for n in range(N):
            s.checkpoint_save(n)
            s.step(1)
        s.checkpoint_save(N)
        s.adjoint_set()
        for n in range(N - 1, -1, -1):
            s.checkpoint_load(n + 1)
            if n >= N - M:
                s.adjoint_add_ke(1.0 / M)
            s.checkpoint_load(n)
            s.adjoint_step()
        manual = (s.adjoint_gradient(), s.adjoint_get())
    finally:
        s.close()
    ref = res[6]
    scale = max(np.abs(x).max() for x in ref[2])
    for K, (J, ik, gr, lam) in res.items():
        assert J == ref[0] and np.array_equal(ik, ref[1])
        assert max(np.abs(a - b).max() for a, b in zip(gr, ref[2])) <= 1e-10 * scale
    assert max(np.abs(a - b).max() for a, b in zip(manual[0], ref[2])) <= 1e-10 * scale
    lsc = max(np.abs(x).max() for x in ref[3])
    assert max(np.abs(a - b).max() for a, b in zip(manual[1], ref[3])) <= 1e-10 * lsc


def test_body_force_step_matches_oracle(hip_lib):
    """The forward with a body force, bit for bit against the CPU restatement."""
    from oracle import OracleRun
    g = stretched_grid(16, 12, 10)
    rng = np.random.default_rng(21)
    u, v, w = noise_ic(g, 1e-2, seed=21)
    f = [rand_halo(g, rng, 0.5) for _ in range(3)]
    r = OracleRun(g, dt=1e-3, nu=1e-2, Cs=0.1, omega=1.5, zeta=1e-8, itmax=200.0, LES=1,
                  scheme=0, sor=1, bc=bc_codes())
    r.f["u"][...], r.f["v"][...], r.f["w"][...] = u, v, w
    r.f["fu"], r.f["fv"], r.f["fw"] = f
    r.apply_bc()
    state = [r.f[n].copy(order="F") for n in ("u", "v", "w", "p")]
    iks = [r.step()[0] for _ in range(3)]
    s = ProjectionSolver(g, bc_codes())
    try:
        s.configure(dt=1e-3, nu=1e-2, Cs=0.1, LES=1, scheme=QUICK_CENTRE, omega=1.5, zeta=1e-8,
                    itmax=200.0, sor=SOR_REDBLACK, force_every=1)
        s.set_body_force(*f)
        s.upload(*state, g.halo())
        ik, *_ = s.step(3)
        out = [g.halo() for _ in range(4)]
        s.download(*out)
    finally:
        s.close()
    assert list(ik) == iks
    for a, n in zip(out, ("u", "v", "w", "p")):
        assert np.array_equal(a, r.f[n]), n


@pytest.mark.parametrize("n,itmax", [(16, 70.0), (13, 70.0), (16, 65.0), (12, 1.0),
                                     (16, 400.0)])
def test_mixed_forward_adjoint_matches_fp64(n, itmax, hip_lib):
    """With a fixed sweep count the FP32-sweep forward runs the same linear
    iteration, so its adjoint gradient (FP32 transposed sweeps with FP64
    folds) agrees with the FP64 one; n = 13 gives odd extents (13, 11, 9).
    The transposed half-sweeps after the first run in segments of 128, then
    of 512: 70 sweeps give a segment ending with a red update and a last one
    of 11 ending with the black update that does not count, 65 a last
    segment of that black update alone, 1 sweep nothing but it, 400 the
    segments 128, 512 and a last one of 159."""
    g, state = cavity_setup(n, seed=5)
    eta = g.halo()
    out = {}
    for sor in (SOR_REDBLACK, SOR_MIXED):
        s = ProjectionSolver(g, bc_codes())
        try:
            s.configure(dt=2e-3, nu=1.0 / 400.0, Cs=0.1, LES=1, scheme=QUICK_CENTRE, omega=1.5,
                        zeta=0.0, itmax=itmax, sor=sor, force_every=0)
            st = [a.copy(order="F") for a in state]
            s.boundary_conditions(*st)
            s.set_body_force(*gaussian_blob(g, 0.1, 0.1, (0.5, 0.5, 0.5)))
            s.upload(*st, eta)
            J, ik, _ = s.adjoint_run(6, 3, 3)
            out[sor] = (J, ik, s.adjoint_gradient())
        finally:
            s.close()
    a, b = out[SOR_MIXED], out[SOR_REDBLACK]
    assert np.array_equal(a[1], b[1])
    assert abs(a[0] - b[0]) <= 1e-6 * abs(b[0])
    scale = max(np.abs(x).max() for x in b[2])
    assert max(np.abs(x - y).max() for x, y in zip(a[2], b[2])) <= 1e-5 * scale


def test_checkpoint_plan(hip_lib):
    """The spacing plan: every step kept when it fits (one segment, no
    recomputation), ceil(sqrt(n)) when the reserve leaves no room; the
    store sizes it reports; argument errors; and the planned spacing giving
    the gradient of the square-root one."""
    from ntustcfdlab_amd.hip import NtustError
    g, state = cavity_setup(12, seed=9)
    nh = (g.nx + 4) * (g.ny + 4) * (g.nz + 4)
    fsz = 2 * (g.nx + 2) * (g.ny + 2) + 2 * (g.nx + 2)
    f = gaussian_blob(g, 0.1, 0.1, (0.5, 0.5, 0.5))
    s = ProjectionSolver(g, bc_codes())
    try:
        s.configure(dt=2e-3, nu=1.0 / 400.0, Cs=0.1, LES=1, scheme=QUICK_CENTRE, omega=1.5,
                    zeta=0.0, itmax=20.0, sor=SOR_REDBLACK, force_every=0)
        assert s.adjoint_checkpoint_plan(9, 1e9) == (9, 8.0 * (4 * nh + 10 * 3 * nh + 4 * nh))
        assert s.adjoint_checkpoint_plan(9, 1e18) == (3, 8.0 * (3 * 4 * nh + 4 * 3 * nh + 4 * nh))
        assert s.adjoint_checkpoint_plan(1, 1e18) == (1, 8.0 * (4 * nh + 2 * 3 * nh + 4 * nh))
        s.adjoint_objective(0.0, 1.0, 0.0, eta_gradient=True)
        k, nb = s.adjoint_checkpoint_plan(9, 1e9)
        assert k == 9 and nb == 8.0 * (4 * nh + 10 * 3 * nh + 4 * nh + 9 * fsz + 9 * 3 * nh + nh)
        s.adjoint_objective(1.0, 0.0, 0.0)
        for bad in ((0, 0.0), (4, -1.0), (4, float("nan")), (4, float("inf"))):
            with pytest.raises(NtustError):
                s.adjoint_checkpoint_plan(*bad)
        out = []
        for every in (s.adjoint_checkpoint_plan(9, 1e9)[0], s.adjoint_checkpoint_plan(9, 1e18)[0]):
            st = [a.copy(order="F") for a in state]
            s.boundary_conditions(*st)
            s.set_body_force(*f)
            s.upload(*st, g.halo())
            J, ik, _ = s.adjoint_run(9, every, 3)
            out.append((J, ik, s.adjoint_gradient()))
    finally:
        s.close()
    (J1, ik1, g1), (J2, ik2, g2) = out
    assert J1 == J2 and np.array_equal(ik1, ik2)
    scale = max(np.abs(x).max() for x in g2)
    assert max(np.abs(x - y).max() for x, y in zip(g1, g2)) <= 1e-10 * scale


def test_adjoint_rejects_lexicographic(hip_lib):
    from ntustcfdlab_amd.hip import SOR_LEXICOGRAPHIC, NtustError
    g = Grid(8, 8, 8)
    s = ProjectionSolver(g, bc_codes())
    try:
        s.configure(dt=1e-3, nu=0.1, sor=SOR_LEXICOGRAPHIC)
        s.upload(*[g.halo() for _ in range(5)])
        with pytest.raises(NtustError, match="red-black"):
            s.adjoint_run(2, 1, 1)
    finally:
        s.close()
