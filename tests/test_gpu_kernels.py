"""GPU parity of every drop-in entry point against the CPU restatement.

Each test feeds the same seeded inputs (random over the whole halo array, on
a non-uniform grid so every metric index matters) to the HIP library through
its C ABI and to the oracle, and requires bit-identical outputs: the
exact-mode kernels keep the reference's operation order without FMA.
"""
import numpy as np
import pytest

from conftest import rand_halo, rand_interior, stretched_grid
from ntustcfdlab_amd.grid import bc_codes
from ntustcfdlab_amd.hip import SOR_LEXICOGRAPHIC, SOR_REDBLACK, ProjectionSolver
from oracle import Oracle

pytestmark = pytest.mark.gpu

SHAPES = [(12, 10, 9), (17, 6, 11)]


def same(a, b):
    """bit-identical, NaN-aware"""
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.fixture(params=SHAPES, ids=lambda s: "x".join(map(str, s)))
def setup(request, hip_lib):
    g = stretched_grid(*request.param)
    s = ProjectionSolver(g)
    yield g, s, Oracle(g), np.random.default_rng(hash(request.param) & 0xFFFF)
    s.close()


@pytest.mark.parametrize("codes", [
    bc_codes(),
    bc_codes(WestWall_u="Neumann", EastWall_v="Dirichlet", SouthWall_w="Neumann",
             NorthWall_v="Neumann", BackWall_v="Neumann", FrontWall_w="Neumann",
             BackWall_w="Dirichlet"),
    np.array([3, 1, 2, 0, 1, 2, 2, 1, 0, 1, 0, 2, 0, 2, 1, 2, 2, 99], dtype=np.int32),
], ids=["cavity", "mixed", "all-rules"])
def test_boundary_conditions(setup, codes):
    g, s, o, rng = setup
    s.set_bc(codes)
    f = [rand_halo(g, rng) for _ in range(4)]
    ref = [a.copy(order="F") for a in f]
    outs = [g.halo() for _ in range(9)]
    routs = [g.halo() for _ in range(9)]
    s.boundary_conditions(*f, *outs)
    o.boundary_conditions(codes, *ref, *routs)
    for a, b in zip(f + outs, ref + routs):
        assert same(a, b)


@pytest.mark.parametrize("with_prebase", [False, True])
def test_gradient_phi_gauss(setup, with_prebase):
    g, s, o, rng = setup
    phi = rand_halo(g, rng)
    pre = rng.uniform(-1, 1, g.nx * g.ny) if with_prebase else None
    d = np.zeros((g.nx, g.ny, g.nz, 3), order="F")
    dr = d.copy(order="F")
    s.GradientPhiGauss(phi, d, g.Dxs, g.iDy, g.iDz, prebase=pre)
    o.gradient_phi_gauss(phi, dr, g.Dxs, g.iDy, g.iDz, prebase=pre)
    assert same(d, dr)


def test_smagorinsky(setup):
    g, s, o, rng = setup
    u, v, w = (rand_halo(g, rng) for _ in range(3))
    vis, vr = g.interior(), g.interior()
    grads = [np.zeros((g.nx, g.ny, g.nz, 3), order="F") for _ in range(3)]
    rgrads = [np.zeros((g.nx, g.ny, g.nz, 3), order="F") for _ in range(3)]
    Cs = float(np.float32(0.18))
    s.CalculateSmagorinskyViscosity(u, v, w, vis, Cs, *grads)
    o.smagorinsky(u, v, w, Cs, vr, *rgrads)
    assert same(vis, vr)
    for a, b in zip(grads, rgrads):
        assert same(a, b)


@pytest.mark.parametrize("LES", [0, 1])
def test_quick_centre(setup, LES):
    g, s, o, rng = setup
    u, v, w = (rand_halo(g, rng) for _ in range(3))
    vis = rand_interior(g, rng, 0.0, 0.05)
    outs = [rand_halo(g, rng) for _ in range(3)]
    refs = [a.copy(order="F") for a in outs]
    s.discretisation_QUICK_centre(u, v, w, vis, *outs, dt=2e-3, nu=1e-2, LES=LES)
    o.predictor_quick(u, v, w, vis, 2e-3, 1e-2, LES, *refs)
    for a, b in zip(outs, refs):
        assert same(a, b)


def test_quick_centre_slab(setup):
    """istart/iend restrict the k range exactly as the reference loops do"""
    g, s, o, rng = setup
    u, v, w = (rand_halo(g, rng) for _ in range(3))
    vis = rand_interior(g, rng, 0.0, 0.05)
    outs = [rand_halo(g, rng) for _ in range(3)]
    refs = [a.copy(order="F") for a in outs]
    s.discretisation_QUICK_centre(u, v, w, vis, *outs, dt=1e-3, nu=1e-3, LES=1, istart=3,
                                  iend=g.nz - 2)
    o.predictor_quick(u, v, w, vis, 1e-3, 1e-3, 1, *refs, kstart=3, kend=g.nz - 2)
    for a, b in zip(outs, refs):
        assert same(a, b)


def test_upwind(setup):
    g, s, o, rng = setup
    u, v, w = (rand_halo(g, rng) for _ in range(3))
    outs = [rand_halo(g, rng) for _ in range(3)]
    refs = [a.copy(order="F") for a in outs]
    s.discretisation_upwind(u, v, w, *outs, dt=2e-3, nu=1e-2)
    o.predictor_upwind(u, v, w, 2e-3, 1e-2, *refs)
    for a, b in zip(outs, refs):
        assert same(a, b)


@pytest.mark.parametrize("zeta,itmax", [(1e-6, 400.0), (1e-30, 37.0), (2.0, 50.0)])
def test_gauss_seidel_lexicographic(setup, zeta, itmax):
    g, s, o, rng = setup
    us, vs, ws = (rand_halo(g, rng) for _ in range(3))
    p = rand_halo(g, rng)
    pr = p.copy(order="F")
    pre, prer = g.interior(), g.interior()
    ik, pm, _ = s.gauss_seidel(us, vs, ws, p, pre, 1e-3, 1.5, zeta, itmax, SOR_LEXICOGRAPHIC)
    ikr, pmr = o.gauss_seidel(us, vs, ws, pr, prer, 1e-3, 1.5, zeta, itmax)
    assert ik == ikr
    assert pm == pmr
    assert same(p, pr)
    assert same(pre, prer)


@pytest.mark.parametrize("zeta,itmax", [(1e-6, 400.0), (1e-30, 37.0), (2.0, 50.0)])
def test_sor_redblack(setup, zeta, itmax):
    g, s, o, rng = setup
    us, vs, ws = (rand_halo(g, rng) for _ in range(3))
    p = rand_halo(g, rng)
    pr = p.copy(order="F")
    pre, prer = g.interior(), g.interior()
    ik, pm, _ = s.gauss_seidel(us, vs, ws, p, pre, 1e-3, 1.5, zeta, itmax, SOR_REDBLACK)
    b = g.interior()
    o.poisson_b(us, vs, ws, 1e-3, b)
    ikr, pmr = o.sor_redblack(b, pr, prer, 1.5, zeta, itmax)
    assert ik == ikr
    assert pm == pmr
    assert same(p, pr)
    assert same(pre, prer)


@pytest.mark.parametrize("eta_kind", ["zero", "one", "random"])
def test_calcul_new_velocity(setup, eta_kind):
    g, s, o, rng = setup
    us, vs, ws, p = (rand_halo(g, rng) for _ in range(4))
    eta = {"zero": g.halo(), "one": g.halo() + 1.0,
           "random": np.asfortranarray(rng.uniform(0, 1, g.halo_shape))}[eta_kind]
    outs = [rand_halo(g, rng) for _ in range(9)] + [rand_interior(g, rng) for _ in range(3)]
    refs = [a.copy(order="F") for a in outs]
    sol = (0.25, -0.5, 0.125)
    s.calcul_new_velocity(us, vs, ws, p, eta, *outs, dt=1e-3, solid=sol)
    o.calcul_new_velocity(us, vs, ws, p, eta, sol, 1e-3, *refs)
    for a, b in zip(outs, refs):
        assert same(a, b)


def test_updating_velocity(setup):
    g, s, o, rng = setup
    src = [rand_halo(g, rng) for _ in range(3)]
    dst = [rand_halo(g, rng) for _ in range(3)]
    ref = [a.copy(order="F") for a in dst]
    s.Updating_velocity(*src, *dst)
    o.updating_velocity(*src, *ref)
    for a, b in zip(dst, ref):
        assert same(a, b)


@pytest.mark.parametrize("layout", ["box", "touching", "random"])
def test_virtual_force_integrator(setup, layout):
    g, s, o, rng = setup
    FY, FZ = g.halo(), g.halo()
    if layout == "box":
        FY[4:8, 3:6, 3:7] = rng.uniform(-1, 1, (4, 3, 4))
        FZ[4:8, 3:6, 3:7] = rng.uniform(-1, 1, (4, 3, 4))
    elif layout == "touching":
        FY[2:g.nx + 2, 2:g.ny + 2, 3:6] = rng.uniform(-1, 1, (g.nx, g.ny, 3))
        FZ[2:g.nx + 2, 2:g.ny + 2, 3:6] = rng.uniform(-1, 1, (g.nx, g.ny, 3))
    else:
        FY[2:-2, 2:-2, 2:-2] = rng.uniform(-1, 1, g.shape) * (rng.uniform(0, 1, g.shape) > 0.5)
        FZ[2:-2, 2:-2, 2:-2] = rng.uniform(-1, 1, g.shape) * (rng.uniform(0, 1, g.shape) > 0.5)
    cd, cl = s.virtualForceIntegrator(FY, FZ, 1.0)
    cdr, clr = o.virtual_force_integrator(FY, FZ, 1.0)
    assert np.float64(cd).view(np.uint64) == np.float64(cdr).view(np.uint64)
    assert np.float64(cl).view(np.uint64) == np.float64(clr).view(np.uint64)


# ---- members of the Krylov pressure path ------------------------------------

def test_poisson_coeffs(setup):
    """solver_1st_member.f90: the 7-diagonal matrix and its offsets."""
    g, s, o, rng = setup
    coef, jc = s.solver_1st_member()
    rcoef, rjc = o.poisson_coeffs()
    assert jc == rjc == [0, 1, g.nx, g.nx * g.ny]
    assert same(coef, rcoef)


def test_poisson_rhs(setup):
    """b = mChange/dt in the face-area form of GS.f90:39-41,49."""
    g, s, o, rng = setup
    us, vs, ws = (rand_halo(g, rng) for _ in range(3))
    b = s.poisson_rhs(us, vs, ws, 1.00000004749745131e-03)
    rb = g.interior()
    o.poisson_b(us, vs, ws, 1.00000004749745131e-03, rb)
    assert same(b, rb)


def test_pressure_field(setup):
    """solver_pressure_field.f90: pre = x, p = pre on the interior only."""
    g, s, o, rng = setup
    x = rng.uniform(-1, 1, g.cells)
    p = rand_halo(g, rng)
    p0 = p.copy(order="F")
    pre = g.interior()
    s.solver_pressure_field(x, p, pre)
    xi = x.reshape(g.shape, order="F")
    assert same(pre, np.asfortranarray(xi))
    inner = (slice(2, g.nx + 2), slice(2, g.ny + 2), slice(2, g.nz + 2))
    assert same(np.ascontiguousarray(p[inner]), np.ascontiguousarray(xi))
    mask = np.ones(p.shape, bool)
    mask[inner] = False
    assert same(p[mask], p0[mask])
