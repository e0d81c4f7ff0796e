"""The workloads of the benchmark and the full-size tests, in one place.

Each builder returns the grid, the boundary-condition codes and the step
parameters of one configuration; bench.py times them and tests/ check them.
The reference ships one set-up, the Re = 10 lid-driven cavity
(variables_module.f90:9-23, main.f90:44-70); the others are the
configurations of the hot path's measurement plan, built from the same
subroutines: the LES cavity, the immersed-boundary airfoil at two sizes and
the adjoint cases.
"""
from __future__ import annotations

import numpy as np

from .airfoil import naca4_polygon
from .grid import Grid, bc_codes, tanh_nodes

# slip side walls with uniform inflow w = 1 through the back face and a
# Neumann outflow at the front: no-slip (zero) normal component, Neumann
# tangential components on the four side faces
INFLOW_BC = dict(BackWall_w="Dirichlet", FrontWall_u="Neumann", FrontWall_v="Neumann",
                 FrontWall_w="Neumann", NorthWall_u="Neumann", NorthWall_w="Neumann",
                 SouthWall_u="Neumann", SouthWall_w="Neumann", WestWall_v="Neumann",
                 WestWall_w="Neumann", EastWall_v="Neumann", EastWall_w="Neumann")


def cavity_c1() -> dict:
    """The reference's cavity as it ships: 40^3 uniform unit cube, Re = 10,
    dt = 1.0e-3 and zeta = 1.e-4 as single-precision literals widened to
    real*8 (variables_module.f90:11, main.f90:46), upwind, LES off, SOR
    omega 1.5, itmax 3000, lid w = 1 on the north wall, fluid at rest."""
    return dict(n=40, Re=10.0, dt=float(np.float32(1.0e-3)), zeta=float(np.float32(1.0e-4)),
                itmax=3000.0, omega=1.5, scheme=1, LES=0, Cs=float(np.float32(0.18)),
                steps=100)


def les_cavity(n: int, itmax: float = 3000.0) -> dict:
    """The LES cavity Re = 1000 on a uniform n^3 grid: QUICK, Smagorinsky
    Cs = 0.1 with filter width (iDx iDy iDz)^(1/3), dt = 5e-4, SOR omega
    1.5, zeta 1e-10, itmax 3000; rest plus seeded U(-1e-3, 1e-3) noise on
    u, v, w (seed 2024)."""
    return dict(nx=n, ny=n, nz=n, Re=1000.0, dt=5e-4, Cs=0.1, LES=1, scheme=0, omega=1.5,
                zeta=1e-10, itmax=itmax, noise=1e-3, seed=2024)


def airfoil_c3():
    """The immersed-boundary configuration at moderate size: NACA 0012 (5000
    vertices) at 5 degrees, Re 1000, QUICK, LES off, dt 2e-4; 384 cells
    streamwise (z, 12.5 long) x 192 normal (y, 3.95) x 64 span (x, 3.95),
    y and z clustered at the airfoil; uniform inflow w = 1, slip side walls.
    The reference places the airfoil at (z, y) = (4, 4)
    (variables_module.f90:144-146), outside a 3.95-high domain; here it sits
    at (4, 1.975).  Returns (grid, bc, az0, ay0, AOA, zc, yc, params)."""
    nx, ny, nz = 64, 192, 384
    zc, yc = 4.0, 1.975
    g = Grid(nx, ny, nz, lx=3.95, ly=3.95, lz=12.5, Y=tanh_nodes(ny, 3.95, yc, 4.0),
             Z=tanh_nodes(nz, 12.5, zc, 4.0))
    az0, ay0 = naca4_polygon(poly=5000)
    prm = dict(dt=2e-4, nu=1.0 / 1000.0, Cs=0.1, LES=0, scheme=0, omega=1.5, zeta=1e-10,
               itmax=3000.0, lx=3.95)
    return g, bc_codes(**INFLOW_BC), az0, ay0, 5.0, zc, yc, prm


def airfoil_c5():
    """The large airfoil configuration: NACA 0012 (5000 vertices) on 512 x 384
    x 128 cells (span x, normal y, chord z; y and z clustered at the airfoil
    with Roberts' stretching, finest spacing 0.0086 / 0.026 there, domain 1 x
    8 x 8), Re = 5000, QUICK, Smagorinsky LES Cs = 0.1, dt = 1e-4, SOR omega
    1.5, zeta 1e-10, itmax 3000, AOA 8 degrees with the quarter chord at
    (z, y) = (4, 4).  Uniform inflow w = 1 through the back face, Neumann
    outflow at the front, Neumann tangential components on the other faces
    (this configuration's own choice: the reference ships only the cavity's
    boundary conditions).  Returns (grid, bc, az0, ay0, AOA, zc, yc, params)."""
    nx, ny, nz = 512, 384, 128
    g = Grid(nx, ny, nz, lx=1.0, ly=8.0, lz=8.0, Y=tanh_nodes(ny, 8.0, 4.0, 5.0),
             Z=tanh_nodes(nz, 8.0, 4.0, 5.0))
    az0, ay0 = naca4_polygon(poly=5000)
    prm = dict(dt=1e-4, nu=1.0 / 5000.0, Cs=0.1, LES=1, scheme=0, omega=1.5, zeta=1e-10,
               itmax=3000.0, lx=1.0)
    return g, bc_codes(**INFLOW_BC), az0, ay0, 8.0, 4.0, 4.0, prm


def adjoint_c4() -> dict:
    """The adjoint case: the LES cavity at Re 400 on 96^3 (QUICK, Cs 0.1,
    dt 5e-4, FP64 red-black SOR omega 1.5, zeta 1e-10, itmax 3000) from
    rest; body force a Gaussian blob of amplitude 0.1 and width 0.1 at a
    seeded centre; objective the kinetic energy averaged over the last 32 of
    128 steps; checkpoints every 16 steps; 8 seeded central-difference
    directions with step 1e-6."""
    return dict(n=96, Re=400.0, dt=5e-4, Cs=0.1, omega=1.5, zeta=1e-10, itmax=3000.0,
                steps=128, every=16, navg=32, amp=0.1, width=0.1, seed=96, fd_dirs=8,
                fd_step=1e-6)


def actuator_strip(g, aoa, seed, zc=4.0, yc=4.0):
    """The actuator control of the large configuration: a seeded strip 0.05
    chord wide at x/c = 0.1, next to the surface (normal offset within 0.1
    chord of the chord line), uniform along the span like the extruded
    airfoil, pushing along the chord.  Chord frame of the cast airfoil
    (quarter chord at (z, y) = (zc, yc), rotated by -AOA): s = (z-zc) cos a -
    (y-yc) sin a + 0.25 from the leading edge.  Returns (fu, fv, fw, cells)
    with seeded U(0, 1) amplitudes.  Uniform along x keeps the flow
    x-invariant, so u stays exactly 0: an x-dependent field would make it
    non-zero with either sign, and the QUICK upwind choice at a zero face
    velocity (discretisation_QUICK_centre.f90:18-102, 0 takes the else
    branch) puts a kink in the map there that a central difference
    straddles."""
    _, ys, zs = g.centres()
    a = np.deg2rad(aoa)
    Z, Y = np.meshgrid(zs, ys, indexing="xy")          # (ny, nz)
    sc = (Z - zc) * np.cos(a) - (Y - yc) * np.sin(a) + 0.25
    nr = (Z - zc) * np.sin(a) + (Y - yc) * np.cos(a)
    mask = (sc >= 0.075) & (sc <= 0.125) & (np.abs(nr) <= 0.1)
    amp = np.random.default_rng(seed).uniform(0.0, 1.0, (g.ny, g.nz)) * mask
    amp = np.broadcast_to(amp[None, :, :], (g.nx, g.ny, g.nz))
    fu, fv, fw = g.halo(), g.halo(), g.halo()
    fv[2:-2, 2:-2, 2:-2] = -np.sin(a) * amp
    fw[2:-2, 2:-2, 2:-2] = np.cos(a) * amp
    return fu, fv, fw, int(mask.sum()) * g.nx


def configure(solver, prm: dict, sor: int, force_every: int = 1):
    """ProjectionSolver.configure with a configuration's step parameters."""
    solver.configure(dt=prm["dt"], nu=prm["nu"], Cs=prm["Cs"], LES=prm["LES"],
                     scheme=prm["scheme"], omega=prm["omega"], zeta=prm["zeta"],
                     itmax=prm["itmax"], sor=sor, lx=prm["lx"], force_every=force_every)
