"""The workloads at their full sizes, checked through properties that do not
depend on the size (the oracle is too slow there):
  * the LES cavity on 128^3 and 256^3, one step from the seeded noise state:
    FP64 and FP32-sweep solvers run the same sweep count and agree on the
    velocity to 1e-6 relative; the projection leaves the interior cells
    with a small fraction of the divergence the noise state had;
  * the airfoil cast on 512x384x128: every span slice of ETA integrates to
    the area of the cast polygon (the sheared rotation scales the table's
    area by cos^2 AOA), and the slices are identical (extrusion along x);
  * the airfoil LES on that grid (stretched in y and z, inflow along z), 3
    steps from the inflow: FP64 and FP32-sweep solvers run the same sweep
    counts and agree on the velocity and on C_D, C_L to 1e-6 relative.
"""
import numpy as np
import pytest

from ntustcfdlab_amd.airfoil import naca4_polygon, polygon_area
from ntustcfdlab_amd.grid import Grid, bc_codes, noise_ic, tanh_nodes
from ntustcfdlab_amd.hip import QUICK_CENTRE, SOR_MIXED, SOR_REDBLACK, ProjectionSolver

pytestmark = pytest.mark.gpu


def interior_divergence(g, u, v, w):
    """max |D u| over the cells two or more cells from every wall (the
    wall-normal faces are reset by the boundary conditions after the
    projection)."""
    # cell i (1..n) has width iD[i+1] and faces u(i), u(i-1) at halo indices
    # i+1 and i
    iDx = g.iDx[2:g.nx + 2]
    iDy = g.iDy[2:g.ny + 2]
    iDz = g.iDz[2:g.nz + 2]
    du = (u[2:g.nx + 2, 2:-2, 2:-2] - u[1:g.nx + 1, 2:-2, 2:-2]) / iDx[:, None, None]
    dv = (v[2:-2, 2:g.ny + 2, 2:-2] - v[2:-2, 1:g.ny + 1, 2:-2]) / iDy[None, :, None]
    dw = (w[2:-2, 2:-2, 2:g.nz + 2] - w[2:-2, 2:-2, 1:g.nz + 1]) / iDz[None, None, :]
    d = du + dv + dw
    return float(np.abs(d[1:-1, 1:-1, 1:-1]).max())


@pytest.mark.parametrize("n", [128, 256])
def test_les_cavity_step_full_size(n, hip_lib):
    g = Grid(n, n, n)
    u0, v0, w0 = noise_ic(g, 1e-3, seed=2024)
    out = {}
    for sor in (SOR_REDBLACK, SOR_MIXED):
        s = ProjectionSolver(g, bc_codes())
        try:
            s.configure(dt=5e-4, nu=1.0 / 1000.0, Cs=0.1, LES=1, scheme=QUICK_CENTRE, omega=1.5,
                        zeta=1e-10, itmax=3000.0, sor=sor)
            u, v, w, p = u0.copy(order="F"), v0.copy(order="F"), w0.copy(order="F"), g.halo()
            s.boundary_conditions(u, v, w, p)
            if sor == SOR_REDBLACK:
                div0 = interior_divergence(g, u, v, w)
            s.upload(u, v, w, p, g.halo())
            ik, pm, cd, cl, rc = s.step(1)
            fin = [g.halo() for _ in range(4)]
            s.download(*fin)
            out[sor] = (int(ik[0]), fin)
        finally:
            s.close()
    (ik64, f64), (ik32, f32) = out[SOR_REDBLACK], out[SOR_MIXED]
    assert ik64 == ik32 and 1 <= ik64 <= 3000
    vel = max(np.abs(x).max() for x in f64[:3])
    err = max(np.abs(x - y).max() for x, y in zip(f32[:3], f64[:3])) / vel
    assert err < 1e-6, err
    div1 = interior_divergence(g, *f64[:3])
    assert div1 < 1e-2 * div0, (div1, div0)


def test_airfoil_cast_full_size(hip_lib):
    nx, ny, nz = 512, 384, 128
    g = Grid(nx, ny, nz, lx=1.0, ly=8.0, lz=8.0, Y=tanh_nodes(ny, 8.0, 4.0, 5.0),
             Z=tanh_nodes(nz, 8.0, 4.0, 5.0))
    az0, ay0 = naca4_polygon(poly=5000)
    aoa = 8.0
    s = ProjectionSolver(g)
    try:
        eta = g.halo()
        s.ray_casting(az0, ay0, aoa, 4.0, 4.0, ETA=eta)
    finally:
        s.close()
    e = eta[2:-2, 2:-2, 2:-2]
    assert np.array_equal(e, np.broadcast_to(e[:1], e.shape))
    dy = g.Y[3:ny + 3] - g.Y[2:ny + 2]
    dz = g.Z[3:nz + 3] - g.Z[2:nz + 2]
    area = float((e[0] * dy[:, None] * dz[None, :]).sum())
    expect = polygon_area(az0, ay0) * np.cos(np.radians(aoa)) ** 2
    assert abs(area - expect) <= 1e-2 * expect, (area, expect)
    assert ((e > 0.0) & (e < 1.0)).sum() > 0


def test_airfoil_les_steps_full_size(hip_lib):
    nx, ny, nz = 512, 384, 128
    g = Grid(nx, ny, nz, lx=1.0, ly=8.0, lz=8.0, Y=tanh_nodes(ny, 8.0, 4.0, 5.0),
             Z=tanh_nodes(nz, 8.0, 4.0, 5.0))
    bc = bc_codes(BackWall_w="Dirichlet", FrontWall_u="Neumann", FrontWall_v="Neumann",
                  FrontWall_w="Neumann", NorthWall_u="Neumann", NorthWall_w="Neumann",
                  SouthWall_u="Neumann", SouthWall_w="Neumann", WestWall_v="Neumann",
                  WestWall_w="Neumann", EastWall_v="Neumann", EastWall_w="Neumann")
    az0, ay0 = naca4_polygon(poly=5000)
    out = {}
    for sor in (SOR_REDBLACK, SOR_MIXED):
        s = ProjectionSolver(g, bc)
        try:
            s.configure(dt=1e-4, nu=1.0 / 5000.0, Cs=0.1, LES=1, scheme=QUICK_CENTRE, omega=1.5,
                        zeta=1e-10, itmax=3000.0, sor=sor, lx=1.0, force_every=1)
            u, v, w, p = g.halo(), g.halo(), g.halo(), g.halo()
            w[...] = 1.0
            s.boundary_conditions(u, v, w, p)
            eta = g.halo()
            s.ray_casting(az0, ay0, 8.0, 4.0, 4.0, ETA=eta)
            s.upload(u, v, w, p, eta)
            ik, pm, cd, cl, rc = s.step(3)
            fin = [g.halo() for _ in range(4)]
            s.download(*fin)
            out[sor] = (np.array(ik), np.array(cd), np.array(cl), fin)
        finally:
            s.close()
    a, b = out[SOR_REDBLACK], out[SOR_MIXED]
    assert np.array_equal(a[0], b[0]) and a[0][-1] > 1
    for x, y in ((a[1], b[1]), (a[2], b[2])):
        assert np.abs(x - y).max() <= 1e-6 * np.abs(x).max(), (x, y)
    vel = max(np.abs(x).max() for x in a[3][:3])
    err = max(np.abs(x - y).max() for x, y in zip(b[3][:3], a[3][:3])) / vel
    assert err < 1e-6, err
