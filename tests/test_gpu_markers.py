"""Lagrangian-marker direct forcing (ntust_set_markers; no counterpart in
the reference) on the GPU:
  * resident steps of an airfoil in uniform inflow forced by its markers,
    bit for bit against the oracle's restatement (oracle/ntust_markers.c):
    sweep counts, force coefficients, state, centred velocity;
  * the adjoint of a step with markers: dot-product test against the
    central difference of the forward step;
  * dC_D/dAOA of the large airfoil configuration (512 x 384 x 128, the
    FP32-sweep forward) by the chain rule through the marker positions
    against the central difference of J at +-0.01 degrees."""
import numpy as np
import pytest

from conftest import rand_halo
from ntustcfdlab_amd import configs
from ntustcfdlab_amd.airfoil import naca4_polygon
from ntustcfdlab_amd.grid import Grid, bc_codes, tanh_nodes
from ntustcfdlab_amd.hip import QUICK_CENTRE, SOR_MIXED, SOR_REDBLACK, NtustError, \
    ProjectionSolver
from ntustcfdlab_amd.markers import airfoil_marker_set
from oracle import OracleRun

pytestmark = pytest.mark.gpu


def same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


def small_airfoil():
    g = Grid(6, 56, 64, lx=0.375, ly=8.0, lz=8.0, Y=tanh_nodes(56, 8.0, 4.0, 3.0),
             Z=tanh_nodes(64, 8.0, 4.0, 3.0))
    return g, bc_codes(**configs.INFLOW_BC), naca4_polygon(poly=400)


@pytest.mark.parametrize("les", [0, 1])
def test_marker_steps_match_oracle(les, hip_lib):
    g, bc, (az0, ay0) = small_airfoil()
    mk = airfoil_marker_set(g, az0, ay0, 8.0, 4.0, 4.0, 0.05)
    kw = dict(dt=2e-3, nu=1.0 / 500.0, Cs=0.1, omega=1.5, zeta=1e-9, itmax=300.0, LES=les,
              scheme=0)
    ref = OracleRun(g, sor=1, bc=bc, lx=0.375, **kw)
    try:
        assert ref.set_markers(mk["zm"], mk["ym"], mk["wm"], mk["h"]) == 0
        f = ref.f
        f["w"][...] = 1.0
        ref.apply_bc()
        s = ProjectionSolver(g, bc)
        try:
            s.configure(sor=SOR_REDBLACK, lx=0.375, force_every=1, **kw)
            s.set_markers(mk["zm"], mk["ym"], mk["wm"], mk["h"])
            s.upload(f["u"], f["v"], f["w"], f["p"], f["ETA"])
            rows = [ref.step() for _ in range(3)]
            ik, pm, cd, cl, _ = s.step(3)
            out = {n: g.halo() for n in ("u", "v", "w", "p")}
            out.update({n: g.interior() for n in ("pre", "uc", "vc", "wc")})
            s.download(**out)
            U = s.marker_velocity()
        finally:
            s.close()
    finally:
        ref.o.clear_markers()
    print(f"markers {len(mk['zm'])}, sweeps {list(ik)}, C_D {list(cd)}, |U| at the markers "
          f"{np.abs(U[2]).mean():.3e}", flush=True)
    assert list(ik) == [r[0] for r in rows]
    assert same(cd, [r[2] for r in rows]) and same(cl, [r[3] for r in rows])
    assert cd[-1] != 0.0
    for n, a in out.items():
        assert same(a, f[n]), n


def test_marker_adjoint_dot_product(hip_lib):
    g, bc, (az0, ay0) = small_airfoil()
    mk = airfoil_marker_set(g, az0, ay0, 8.0, 4.0, 4.0, 0.05)
    rng = np.random.default_rng(5)
    state = [rand_halo(g, rng, 0.5) for _ in range(3)] + [rand_halo(g, rng, 1e-3)]
    dx = [rand_halo(g, rng, 1.0) for _ in range(3)] + [rand_halo(g, rng, 1e-3)]
    lam = [rand_halo(g, rng, 1.0) for _ in range(4)]
    eta = g.halo()
    s = ProjectionSolver(g, bc)
    try:
        s.configure(dt=2e-3, nu=1.0 / 500.0, Cs=0.1, LES=1, scheme=QUICK_CENTRE, omega=1.5,
                    zeta=0.0, itmax=30.0, sor=SOR_REDBLACK, lx=0.375, force_every=0)
        s.set_markers(mk["zm"], mk["ym"], mk["wm"], mk["h"])

        def one(st):
            s.upload(*[a.copy(order="F") for a in st], eta)
            s.step(1)
            o = [g.halo() for _ in range(4)]
            s.download(*o)
            return o
        h = 1e-7
        plus = one([x + h * d for x, d in zip(state, dx)])
        minus = one([x - h * d for x, d in zip(state, dx)])
        s.upload(*[a.copy(order="F") for a in state], eta)
        s.adjoint_set(*lam)
        s.adjoint_step()
        jtl = s.adjoint_get()
    finally:
        s.close()
    jdx = [(a - b) / (2 * h) for a, b in zip(plus, minus)]
    lhs = float(sum(np.vdot(a, b) for a, b in zip(lam, jdx)))
    rhs = float(sum(np.vdot(a, b) for a, b in zip(jtl, dx)))
    print(f"marker adjoint dot product {lhs:.12e} / {rhs:.12e}", flush=True)
    assert abs(lhs - rhs) <= 1e-6 * abs(rhs)


def test_set_markers_refuses_a_kernel_wider_than_its_window(hip_lib):
    """A kernel width with more than three grid points within 1.5 h of a
    marker would fall outside its window: refused."""
    g, bc, (az0, ay0) = small_airfoil()
    mk = airfoil_marker_set(g, az0, ay0, 8.0, 4.0, 4.0, 0.05)
    s = ProjectionSolver(g, bc)
    try:
        hx, hy, hz = mk["h"]
        with pytest.raises(NtustError, match="finer than the kernel"):
            s.set_markers(mk["zm"], mk["ym"], mk["wm"], (hx, 3.0 * hy, hz))
        s.set_markers()   # removal is always accepted
    finally:
        s.close()


def test_c5_marker_aoa_gradient_512x384x128(hip_lib):
    """J = mean C_D over 3 steps from the state two steps after the inflow
    (the airfoil forced by its markers, FP32-sweep solver); dJ/dAOA =
    sum_m dJ/dX_m . dX_m/dAOA against the central differences D(h) =
    (J(AOA+h) - J(AOA-h)) / 2h at h = 0.01 and 0.005 degrees, with the
    marker model (kernel widths, weights) of AOA.  D(h) carries a truncation
    error c h^2 of about 1e-4 of the derivative at 0.01 degrees, so the
    1e-4 bound is taken against the extrapolation (4 D(h/2) - D(h)) / 3,
    whose error is O(h^4), and D(0.01) itself is held to 5e-4."""
    g, bc, az0, ay0, aoa, zc, yc, prm = configs.airfoil_c5()
    steps, hdeg = 3, 0.01
    mk = airfoil_marker_set(g, az0, ay0, aoa, zc, yc, 0.01)
    s = ProjectionSolver(g, bc)
    try:
        configs.configure(s, prm, SOR_MIXED)
        eta = g.halo()
        u, v, w, p = g.halo(), g.halo(), g.halo(), g.halo()
        w[...] = 1.0
        s.boundary_conditions(u, v, w, p)
        s.set_markers(mk["zm"], mk["ym"], mk["wm"], mk["h"])
        s.upload(u, v, w, p, eta)
        s.step(2)
        state = [g.halo() for _ in range(4)]
        s.download(*state)

        def J_at(a):
            m = airfoil_marker_set(g, az0, ay0, a, zc, yc, 0.01, mk["h"], mk["hn"])
            s.set_markers(m["zm"], m["ym"], m["wm"], m["h"])
            s.upload(*[x.copy(order="F") for x in state], eta)
            ik, pm, cd, cl, _ = s.step(steps)
            tot = 0.0
            for x in cd:
                tot += float(x)
            return tot * (1.0 / steps), list(ik)
        s.set_markers(mk["zm"], mk["ym"], mk["wm"], mk["h"])
        s.adjoint_objective(0.0, 1.0, 0.0, eta_gradient=True)
        s.upload(*[x.copy(order="F") for x in state], eta)
        J, ik, ms = s.adjoint_run(steps, steps, steps)
        gz, gy = s.adjoint_marker_gradient()
        chain = float(np.dot(gz, mk["dzm"]) + np.dot(gy, mk["dym"]))
        fds, sweeps = [], [list(ik)]
        for hh in (hdeg, hdeg / 2):
            jp, ikp = J_at(aoa + hh)
            jm, ikm = J_at(aoa - hh)
            fds.append((jp - jm) / (2 * hh))
            sweeps += [ikp, ikm]
    finally:
        s.close()
    rich = (4.0 * fds[1] - fds[0]) / 3.0
    print(f"C5 markers ({len(mk['zm'])} x {g.nx}): J {J:.10f}, dJ/dAOA chain rule {chain:.8e} "
          f"per degree, central differences {fds[0]:.8e} (0.01 deg, rel "
          f"{abs(chain - fds[0]) / abs(fds[0]):.2e}), {fds[1]:.8e} (0.005 deg), extrapolated "
          f"{rich:.8e} (rel {abs(chain - rich) / abs(rich):.2e}); sweeps {sweeps}", flush=True)
    assert all(x == sweeps[0] for x in sweeps)
    assert abs(chain - rich) <= 1e-4 * abs(rich)
    assert abs(chain - fds[0]) <= 5e-4 * abs(fds[0])
