"""GPU parity of ntust_ray_casting (RayCasting, airfoil.f90:580-892) and
ntust_dynamic_aoa against the CPU restatement, bit for bit: ETA on
stretched and uniform grids, several angles (including the pitching law's),
sub-grid resolutions and polygon sizes; the caller's ghost values kept; the
refusal of a polygon outside the grid; and resident steps on the device ETA
equal to the oracle's steps with the oracle's ETA."""
import numpy as np
import pytest

from ntustcfdlab_amd.airfoil import naca4_polygon
from ntustcfdlab_amd.grid import Grid, bc_codes, noise_ic, tanh_nodes
from ntustcfdlab_amd.hip import QUICK_CENTRE, SOR_REDBLACK, NtustError, ProjectionSolver, \
    dynamic_aoa
from oracle import Oracle, OracleRun
from oracle import dynamic_aoa as or_dynamic_aoa

pytestmark = pytest.mark.gpu


def same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


def airfoil_grid(nx, ny, nz, beta=2.5):
    return Grid(nx, ny, nz, ly=8.0, lz=8.0, Y=tanh_nodes(ny, 8.0, 4.0, beta),
                Z=tanh_nodes(nz, 8.0, 4.0, beta))


CASES = {
    "naca5000-aoa5": dict(grid=(6, 80, 96), poly=5000, AOA=5.0, nsub=50),
    "naca5000-aoa20-fine": dict(grid=(3, 128, 160), poly=5000, AOA=20.0, nsub=50),
    "naca400-aoa-7-nsub13": dict(grid=(5, 64, 72), poly=400, AOA=-7.0, nsub=13),
    "naca5000-pitching": dict(grid=(4, 96, 96), poly=5000,
                              AOA=or_dynamic_aoa(5.0, 0.137, 5e-4, 1000.0), nsub=50),
    "naca5000-uniform": dict(grid=(4, 96, 96), poly=5000, AOA=8.0, nsub=50, uniform=True),
}


@pytest.mark.parametrize("name", list(CASES))
def test_ray_casting_matches_oracle(name, hip_lib):
    c = CASES[name]
    nx, ny, nz = c["grid"]
    g = Grid(nx, ny, nz, ly=8.0, lz=8.0) if c.get("uniform") else airfoil_grid(nx, ny, nz)
    az0, ay0 = naca4_polygon(poly=c["poly"])
    o = Oracle(g)
    az, ay = o.airfoil_transform(az0, ay0, c["AOA"], 4.0, 4.0)
    want = g.halo()
    want[0], want[-1] = 7.0, -3.0                 # ghost planes: the caller's values
    got = want.copy(order="F")
    res = o.ray_casting(az, ay, g.Y, g.Z, c["nsub"], want)
    assert res is not None and res[1].sum() > 0
    s = ProjectionSolver(g)
    try:
        s.ray_casting(az0, ay0, c["AOA"], 4.0, 4.0, nsub=c["nsub"], ETA=got)
    finally:
        s.close()
    assert same(got, want)
    assert 0 < want[2:-2, 2:-2, 2:-2].sum()


def test_ray_casting_polygon_outside(hip_lib):
    """Wholly outside: ETA = 0 (the reference's unit cavity with the airfoil
    at (4, 4)); straddling the lower edge, or a bad nsub: refused."""
    az0, ay0 = naca4_polygon(poly=200)
    g = Grid(8, 8, 8)
    s = ProjectionSolver(g)
    try:
        eta = g.halo() + 0.5
        s.ray_casting(az0, ay0, 5.0, 4.0, 4.0, ETA=eta)
        assert not eta[2:-2, 2:-2, 2:-2].any() and np.all(eta[:2] == 0.5)
    finally:
        s.close()
    g = airfoil_grid(4, 24, 24)
    s = ProjectionSolver(g)
    try:
        with pytest.raises(NtustError, match="straddles"):
            s.ray_casting(az0, ay0, 5.0, -0.5, 4.0)
        with pytest.raises(NtustError, match="nsub"):
            s.ray_casting(az0, ay0, 5.0, 4.0, 4.0, nsub=0)
    finally:
        s.close()


def test_dynamic_aoa_matches_oracle(hip_lib):
    for t in (0.0, 0.4995, 0.5, 0.73, 3.0):
        assert same(dynamic_aoa(5.0, t, 5e-4, 1000.0), or_dynamic_aoa(5.0, t, 5e-4, 1000.0))


def test_steps_on_device_eta_match_oracle(hip_lib):
    """The resident ETA from ntust_ray_casting drives the direct forcing:
    steps equal the oracle's with the oracle's ETA (forces included)."""
    g = airfoil_grid(6, 40, 48, beta=1.5)
    bc = bc_codes()
    az0, ay0 = naca4_polygon(poly=1000)
    prm = dict(dt=5e-4, nu=1e-3, Cs=0.1, LES=1, scheme=QUICK_CENTRE, omega=1.5, zeta=1e-8,
               itmax=200.0, sor=SOR_REDBLACK, lx=1.0, force_every=1)
    ref = OracleRun(g, dt=prm["dt"], nu=prm["nu"], Cs=prm["Cs"], omega=1.5, zeta=prm["zeta"],
                    itmax=prm["itmax"], LES=1, scheme=0, sor=1, bc=bc)
    f = ref.f
    f["u"][...], f["v"][...], f["w"][...] = noise_ic(g, 1e-2, seed=21)
    ref.apply_bc()
    o = Oracle(g)
    az, ay = o.airfoil_transform(az0, ay0, 5.0, 4.0, 4.0)
    assert o.ray_casting(az, ay, g.Y, g.Z, 50, f["ETA"]) is not None
    s = ProjectionSolver(g, bc)
    try:
        s.configure(**prm)
        s.upload(f["u"], f["v"], f["w"], f["p"], g.halo())     # ETA = 0 uploaded
        s.ray_casting(az0, ay0, 5.0, 4.0, 4.0)                  # resident ETA only
        rows = [ref.step() for _ in range(3)]
        ik, pm, cd, cl, _ = s.step(3)
        assert list(ik) == [r[0] for r in rows]
        assert same(cd, [r[2] for r in rows]) and same(cl, [r[3] for r in rows])
        assert np.all(np.abs(cd) > 0)
        out = {n: g.halo() for n in ("u", "v", "w", "p")}
        s.download(**out)
        for n, a in out.items():
            assert same(a, f[n]), n
    finally:
        s.close()
