"""Drag / lift objectives and the ETA gradient of the discrete adjoint
(ntust_adjoint_objective, ntust_adjoint_eta_gradient).

No counterpart in the reference; checked against the forward it
differentiates (which the other GPU tests pin to the CPU restatement):
  * J of ntust_adjoint_run equals the mean of the drag / lift / kinetic
    energy that plain stepping reports;
  * dJ/df (body force) against central differences of the forward along
    seeded random directions, for C_D, C_L and a mixed objective;
  * dJ/dETA against central differences along random directions supported
    on partial cells (0 < ETA < 1, so no line sum of the force integral
    changes between zero and non-zero: its skip test makes C_D and C_L jump
    there, virtualForceIntegrator.f90:41-75);
  * dJ/dAOA by the chain rule through the cast ETA against central
    differences of the forward at two angles (kinetic-energy objective,
    which has no such jumps).
The pressure solve runs a fixed sweep count (zeta = 0), so the forward map
is smooth in the perturbations.
"""
import numpy as np
import pytest

from conftest import stretched_grid
from ntustcfdlab_amd.grid import Grid, bc_codes, gaussian_blob, noise_ic, tanh_nodes
from ntustcfdlab_amd.hip import QUICK_CENTRE, SOR_REDBLACK, ProjectionSolver

pytestmark = pytest.mark.gpu

N, K, M = 8, 3, 4


def dot(a, b):
    return float(sum(np.vdot(x, y) for x, y in zip(a, b)))


def box_eta(g):
    """A solid block with fractional faces (0 < ETA < 1 on its skin)."""
    eta = g.halo()
    core = (slice(7, 12), slice(6, 11), slice(8, 14))   # halo indices
    eta[core] = 1.0
    eta[6, 6:11, 8:14] = 0.35
    eta[12, 6:11, 8:14] = 0.6
    eta[7:12, 5, 8:14] = 0.45
    eta[7:12, 11, 8:14] = 0.7
    eta[7:12, 6:11, 7] = 0.25
    eta[7:12, 6:11, 14] = 0.8
    return eta


def make_solver(g):
    s = ProjectionSolver(g, bc_codes())
    s.configure(dt=2e-3, nu=1.0 / 400.0, Cs=0.1, LES=1, scheme=QUICK_CENTRE, omega=1.5,
                zeta=0.0, itmax=40.0, sor=SOR_REDBLACK, lx=1.0, force_every=1)
    return s


def setup():
    g = stretched_grid(18, 16, 20)
    u, v, w = noise_ic(g, 5e-2, seed=5)
    w += 0.3          # a mean stream through the block, so drag and lift are O(1)
    return g, [u, v, w, g.halo()]


def forward_J(s, state, eta, f, wts):
    """mean over the last M of N steps of w_ke*KE + w_cd*C_D + w_cl*C_L"""
    s.set_body_force(*f)
    s.upload(*[a.copy(order="F") for a in state], eta)
    tot = 0.0
    for n in range(N):
        ik, pm, cd, cl, rc = s.step(1)
        if n >= N - M:
            if wts[0] != 0.0:
                tot += wts[0] * s.kinetic_energy()
            tot += wts[1] * float(cd[0]) + wts[2] * float(cl[0])
    return tot * (1.0 / M)


@pytest.mark.parametrize("wts", [(0.0, 1.0, 0.0), (0.0, 0.0, 1.0), (0.5, 2.0, -1.5)],
                         ids=["drag", "lift", "mixed"])
def test_force_objective_gradient_matches_central_fd(wts, hip_lib):
    g, state = setup()
    eta = box_eta(g)
    s = make_solver(g)
    try:
        s.boundary_conditions(*state)
        f = gaussian_blob(g, 0.1, 0.15, (0.5, 0.5, 0.4))
        s.set_body_force(*f)
        s.adjoint_objective(*wts)
        s.upload(*[a.copy(order="F") for a in state], eta)
        J, ik, ms = s.adjoint_run(N, K, M)
        gf = s.adjoint_gradient()
        J0 = forward_J(s, state, eta, f, wts)
        assert abs(J - J0) <= 1e-13 * abs(J0), (J, J0)
        assert abs(J0) > 1e-6
        rng = np.random.default_rng(17)
        for _ in range(3):
            d = [g.halo() for _ in range(3)]
            for a in d:
                a[2:-2, 2:-2, 2:-2] = rng.standard_normal(a[2:-2, 2:-2, 2:-2].shape)
            h = 1e-6
            jp = forward_J(s, state, eta, [x + h * y for x, y in zip(f, d)], wts)
            jm = forward_J(s, state, eta, [x - h * y for x, y in zip(f, d)], wts)
            fd = (jp - jm) / (2 * h)
            ad = dot(gf, d)
            assert abs(fd - ad) <= 1e-5 * abs(ad), (fd, ad)
    finally:
        s.close()
    assert (ik == 40).all()


def test_eta_gradient_matches_central_fd(hip_lib):
    g, state = setup()
    eta = box_eta(g)
    wts = (1.0, 1.0, 0.5)
    s = make_solver(g)
    try:
        s.boundary_conditions(*state)
        f = [g.halo() for _ in range(3)]
        s.set_body_force(*f)
        s.adjoint_objective(*wts, eta_gradient=True)
        s.upload(*[a.copy(order="F") for a in state], eta)
        J, ik, ms = s.adjoint_run(N, K, M)
        ge = s.adjoint_eta_gradient()
        rng = np.random.default_rng(23)
        partial = (eta > 0.0) & (eta < 1.0)
        assert partial.sum() > 100
        for _ in range(3):
            d = np.asfortranarray(np.where(partial, rng.standard_normal(eta.shape), 0.0))
            h = 1e-6
            jp = forward_J(s, state, np.asfortranarray(eta + h * d), f, wts)
            jm = forward_J(s, state, np.asfortranarray(eta - h * d), f, wts)
            fd = (jp - jm) / (2 * h)
            ad = float(np.vdot(ge, d))
            assert abs(fd - ad) <= 1e-5 * abs(ad), (fd, ad)
    finally:
        s.close()


def test_eta_gradient_same_from_step_entry(hip_lib):
    """ntust_adjoint_step accumulates the same dJ/dETA as the run (KE
    objective after the last step only)."""
    g, state = setup()
    eta = box_eta(g)
    s = make_solver(g)
    try:
        s.configure(dt=2e-3, nu=1.0 / 400.0, Cs=0.1, LES=1, scheme=QUICK_CENTRE, omega=1.5,
                    zeta=0.0, itmax=40.0, sor=SOR_REDBLACK, lx=1.0, force_every=0)
        s.boundary_conditions(*state)
        s.set_body_force(None, None, None)
        s.adjoint_objective(1.0, 0.0, 0.0, eta_gradient=True)
        n = 3
        s.upload(*[a.copy(order="F") for a in state], eta)
        J, _, _ = s.adjoint_run(n, 1, 1)
        ge_run = s.adjoint_eta_gradient()
        # by hand: states at the start of each step, then reverse steps
        s.upload(*[a.copy(order="F") for a in state], eta)
        starts = []
        for _ in range(n):
            x = [g.halo() for _ in range(4)]
            s.download(*x)
            starts.append(x)
            s.step(1)
        s.adjoint_set()
        s.adjoint_add_ke(1.0)
        for x in reversed(starts):
            s.upload(*x, eta)
            s.adjoint_step()
        ge_step = s.adjoint_eta_gradient()
    finally:
        s.close()
    scale = np.abs(ge_run).max()
    assert scale > 0.0
    assert np.abs(ge_run - ge_step).max() <= 1e-12 * scale


def small_naca_case():
    from ntustcfdlab_amd.airfoil import naca4_polygon
    g = Grid(4, 48, 48, lx=0.25, ly=8.0, lz=8.0, Y=tanh_nodes(48, 8.0, 4.0, 5.0),
             Z=tanh_nodes(48, 8.0, 4.0, 5.0))
    bc = bc_codes(BackWall_w="Dirichlet", FrontWall_u="Neumann", FrontWall_v="Neumann",
                  FrontWall_w="Neumann", NorthWall_u="Neumann", NorthWall_w="Neumann",
                  SouthWall_u="Neumann", SouthWall_w="Neumann", WestWall_v="Neumann",
                  WestWall_w="Neumann", EastWall_v="Neumann", EastWall_w="Neumann")
    return g, bc, naca4_polygon(poly=400)


@pytest.mark.parametrize("objective", ["ke", "drag"])
def test_aoa_gradient_by_chain_rule(objective, hip_lib):
    """dJ/dAOA = sum_m dJ/dX_m . dX_m/dAOA through the airfoil's markers (the
    marker direct forcing) against the central differences of J itself at
    +-0.01 and +-0.005 degrees, on a small NACA 0012 case: J the kinetic
    energy or C_D averaged over the last 3 of 6 steps.  The difference at
    0.01 degrees has a truncation error of about 1e-4 of the derivative, so
    the 1e-4 bound is taken against the extrapolation (4 D(h/2) - D(h)) / 3
    (error O(h^4)); D(0.01) itself is held to 5e-4."""
    from ntustcfdlab_amd.markers import airfoil_marker_set
    g, bc, (az, ay) = small_naca_case()
    s = ProjectionSolver(g, bc)
    aoa, h = 8.0, 0.01
    wke, wcd = (1.0, 0.0) if objective == "ke" else (0.0, 1.0)
    mk = airfoil_marker_set(g, az, ay, aoa, 4.0, 4.0, 0.05)
    try:
        s.configure(dt=2e-3, nu=1.0 / 500.0, Cs=0.1, LES=1, scheme=QUICK_CENTRE, omega=1.5,
                    zeta=0.0, itmax=60.0, sor=SOR_REDBLACK, lx=0.25, force_every=1)
        u, v, w, p = g.halo(), g.halo(), g.halo(), g.halo()
        w[...] = 1.0
        s.boundary_conditions(u, v, w, p)
        state = [u, v, w, p]
        eta = g.halo()

        def J_at(a):
            m = airfoil_marker_set(g, az, ay, a, 4.0, 4.0, 0.05, mk["h"], mk["hn"])
            s.set_markers(m["zm"], m["ym"], m["wm"], m["h"])
            s.upload(*[x.copy(order="F") for x in state], eta)
            tot = 0.0
            for n in range(6):
                ik, pm, cd, cl, _ = s.step(1)
                if n >= 3:
                    tot += wke * s.kinetic_energy() + wcd * float(cd[0])
            return tot / 3
        s.set_markers(mk["zm"], mk["ym"], mk["wm"], mk["h"])
        s.adjoint_objective(wke, wcd, 0.0, eta_gradient=True)
        s.upload(*[x.copy(order="F") for x in state], eta)
        J, _, _ = s.adjoint_run(6, 2, 3)
        gz, gy = s.adjoint_marker_gradient()
        chain = float(np.dot(gz, mk["dzm"]) + np.dot(gy, mk["dym"]))
        fd = (J_at(aoa + h) - J_at(aoa - h)) / (2 * h)
        fd2 = (J_at(aoa + h / 2) - J_at(aoa - h / 2)) / h
    finally:
        s.close()
    rich = (4.0 * fd2 - fd) / 3.0
    print(f"{objective}: dJ/dAOA chain {chain:.10e}, central differences {fd:.10e} (rel "
          f"{abs(chain - fd) / abs(fd):.2e}), {fd2:.10e}, extrapolated {rich:.10e} (rel "
          f"{abs(chain - rich) / abs(rich):.2e})", flush=True)
    assert abs(chain - rich) <= 1e-4 * abs(rich), (chain, rich)
    assert abs(chain - fd) <= 5e-4 * abs(fd), (chain, fd)


def test_aoa_gradient_through_sampled_eta_is_first_order(hip_lib):
    """The volume-fraction route: dJ/dAOA ~ sum dJ/dETA (ETA(a+h) -
    ETA(a-h)) / 2h with two casts.  RayCasting samples each cell on 50 x 50
    sub-cells (airfoil.f90:860-876), so ETA is piecewise constant in AOA and
    this chain rule is first order in the change of ETA per cell: checked to
    5e-2 on the same case (kinetic energy)."""
    g, bc, (az, ay) = small_naca_case()
    s = ProjectionSolver(g, bc)
    # h small enough that the surface moves a small fraction of a cell
    aoa, h = 8.0, 0.1
    try:
        s.configure(dt=2e-3, nu=1.0 / 500.0, Cs=0.1, LES=1, scheme=QUICK_CENTRE, omega=1.5,
                    zeta=0.0, itmax=60.0, sor=SOR_REDBLACK, lx=0.25, force_every=0)
        u, v, w, p = g.halo(), g.halo(), g.halo(), g.halo()
        w[...] = 1.0
        s.boundary_conditions(u, v, w, p)
        state = [u, v, w, p]

        def cast(a):
            e = g.halo()
            s.ray_casting(az, ay, a, 4.0, 4.0, ETA=e)
            return e

        def J_at(e):
            s.upload(*[x.copy(order="F") for x in state], e)
            tot = 0.0
            for n in range(6):
                s.step(1)
                if n >= 3:
                    tot += s.kinetic_energy()
            return tot / 3

        e0, ep, em = cast(aoa), cast(aoa + h), cast(aoa - h)
        assert np.abs(ep - em).max() < 0.1 and (ep != em).sum() > 50
        s.adjoint_objective(1.0, 0.0, 0.0, eta_gradient=True)
        s.upload(*[x.copy(order="F") for x in state], e0)
        J, _, _ = s.adjoint_run(6, 2, 3)
        ge = s.adjoint_eta_gradient()
        chain = float(np.vdot(ge, (ep - em) / (2 * h)))
        fd = (J_at(ep) - J_at(em)) / (2 * h)
    finally:
        s.close()
    assert abs(chain - fd) <= 5e-2 * abs(fd), (chain, fd)


def test_objective_and_eta_gradient_errors(hip_lib):
    from ntustcfdlab_amd.hip import NtustError
    g = stretched_grid(8, 8, 8)
    s = make_solver(g)
    try:
        with pytest.raises(NtustError):
            s.adjoint_objective(float("nan"), 1.0, 0.0)
        s.adjoint_objective(1.0, 0.0, 0.0, eta_gradient=False)
        with pytest.raises(NtustError):
            s.adjoint_eta_gradient()
    finally:
        s.close()
