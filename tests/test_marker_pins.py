"""Pins of the oracle's Lagrangian-marker forcing (oracle/ntust_markers.c)
and of the airfoil markers (ntustcfdlab_amd/markers.py).  The marker forcing
has no counterpart in the reference; these checks fix it by mathematics:
  * the kernel, the cubic B-spline: over unit-spaced points at arbitrary
    offsets its weights sum to 1 with central moments 0, 1/3, 0 (those of a
    sum of four uniform variables), its support, its derivative against a
    central difference, and continuity of the second derivative;
  * interpolation and spreading are transposes: <I u, F>_W = <u, S F>_dv
    on a stretched grid with random fields and markers;
  * interpolation reproduces fields linear in y and z on a uniform grid
    (away from walls), per staggered component;
  * the markers lie on the cast polygon (ntust_ray_casting's transform),
    their weights add up to perimeter * hn, and dX/dAOA is the derivative
    of the transform;
  * one forcing application leaves a small part of a uniform flow at the
    markers."""
import math

import numpy as np
import pytest

from conftest import stretched_grid
from ntustcfdlab_amd.airfoil import naca4_polygon
from ntustcfdlab_amd.grid import Grid
from ntustcfdlab_amd.markers import airfoil_markers, kernel_widths, transform, \
    transform_derivative
from oracle import Oracle


@pytest.fixture
def oracle_markers():
    made = []

    def make(g):
        o = Oracle(g)
        made.append(o)
        return o
    yield make
    for o in made:
        o.clear_markers()


def test_kernel_conditions():
    rng = np.random.default_rng(1)
    for x in rng.uniform(-0.5, 0.5, 50):
        k = np.arange(-3, 4)
        phi = np.array([Oracle.mk_phi(float(kk - x)) for kk in k])
        d = k - x
        assert abs(phi.sum() - 1.0) < 1e-15
        assert abs(np.dot(d, phi)) < 1e-15
        assert abs(np.dot(d * d, phi) - 1.0 / 3.0) < 1e-15
        assert abs(np.dot(d * d * d, phi)) < 1e-15
    assert Oracle.mk_phi(2.0) == 0.0 and Oracle.mk_phi(-2.3) == 0.0
    assert abs(Oracle.mk_phi(0.0) - 2.0 / 3.0) < 1e-16
    assert abs(Oracle.mk_phi(1.0) - 1.0 / 6.0) < 1e-16
    for r in (-1.7, -1.3, -0.8, -0.2, 0.1, 0.45, 0.7, 1.2, 1.9):
        h = 1e-6
        fd = (Oracle.mk_phi(r + h) - Oracle.mk_phi(r - h)) / (2 * h)
        assert abs(fd - Oracle.mk_dphi(r)) < 1e-8
    # twice continuously differentiable: phi' and phi'' continuous at the
    # breakpoints 0, 1 and at the edge of the support 2
    e = 1e-7
    for b in (0.0, 1.0, 2.0):
        assert abs(Oracle.mk_dphi(b - e) - Oracle.mk_dphi(b + e)) < 1e-6
        d2l = (Oracle.mk_dphi(b - e) - Oracle.mk_dphi(b - 2 * e)) / e
        d2r = (Oracle.mk_dphi(b + 2 * e) - Oracle.mk_dphi(b + e)) / e
        assert abs(d2l - d2r) < 1e-5, b


def random_markers(g, rng, n):
    y = rng.uniform(g.Y[2 + 3], g.Y[g.ny - 2], n)
    z = rng.uniform(g.Z[2 + 3], g.Z[g.nz - 2], n)
    wm = rng.uniform(0.5, 1.5, n) * 1e-2
    return z, y, wm


def test_interp_and_spread_are_transposes(oracle_markers):
    g = stretched_grid(9, 14, 12)
    o = oracle_markers(g)
    rng = np.random.default_rng(2)
    z, y, wm = random_markers(g, rng, 7)
    h = kernel_widths(g, z, y)
    assert o.set_markers(z, y, wm, h) == 0
    vel = [rng.standard_normal(g.halo_shape) for _ in range(3)]
    vel = [np.asfortranarray(a) for a in vel]
    U = o.marker_interp(*vel, len(z))
    F = rng.standard_normal(U.shape)
    f = o.marker_spread(F)
    W = wm[None, :, None] * np.asarray(g.iDx)[None, None, 2:g.nx + 2]
    lhs = float(np.sum(U * F * W))
    iDx, Dxs, iDy, Dys, iDz, Dzs = (np.asarray(m) for m in g.metrics())
    sl = (slice(2, g.nx + 2), slice(2, g.ny + 2), slice(2, g.nz + 2))
    dv = [Dxs[2:g.nx + 2, None, None] * iDy[None, 2:g.ny + 2, None] * iDz[None, None, 2:g.nz + 2],
          iDx[2:g.nx + 2, None, None] * Dys[None, 2:g.ny + 2, None] * iDz[None, None, 2:g.nz + 2],
          iDx[2:g.nx + 2, None, None] * iDy[None, 2:g.ny + 2, None] * Dzs[None, None, 2:g.nz + 2]]
    rhs = float(sum(np.sum(a[sl] * b[sl] * d) for a, b, d in zip(vel, f, dv)))
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs)
    assert abs(lhs) > 0


def test_interp_reproduces_linear_fields(oracle_markers):
    g = Grid(10, 16, 18, 1.0, 1.6, 1.8)   # uniform: cell 0.1
    o = oracle_markers(g)
    rng = np.random.default_rng(3)
    z, y, wm = random_markers(g, rng, 9)
    assert o.set_markers(z, y, wm, (0.1, 0.1, 0.1)) == 0
    xs, ys, zs = g.centres()
    X, Y, Z = (np.asarray(a) for a in (g.X, g.Y, g.Z))
    a0, ay, az = 0.7, -1.3, 2.1
    # component points: u at the east x face of its cell, X(i+1) (node i at
    # index i+1), v at the north y face, w at the front z face
    pos = [(X[3:g.nx + 3], ys, zs), (xs, Y[3:g.ny + 3], zs), (xs, ys, Z[3:g.nz + 3])]
    fields = []
    for px, py, pz in pos:
        f = g.halo()
        f[2:-2, 2:-2, 2:-2] = (a0 + ay * py[None, :, None] + az * pz[None, None, :]) * \
            np.ones((g.nx, 1, 1))
        fields.append(f)
    U = o.marker_interp(*fields, len(z))
    want = a0 + ay * y + az * z
    # markers away from the x walls (x windows of 5 interior points)
    for c in range(3):
        assert np.allclose(U[c][:, 2:-2], want[:, None], rtol=0, atol=1e-12), c


def test_airfoil_markers_on_the_cast_polygon():
    az0, ay0 = naca4_polygon(poly=400)
    aoa, zc, yc, hn = 7.0, 4.0, 4.0, 0.02
    zm, ym, wm, dz, dy = airfoil_markers(az0, ay0, aoa, zc, yc, 0.01, hn)
    azc, ayc = transform(az0, ay0, aoa, zc, yc)
    # each marker is on an edge of the cast polygon
    for zq, yq in zip(zm[::17], ym[::17]):
        pz, py = azc[:-1], ayc[:-1]
        ez, ey = np.diff(azc), np.diff(ayc)
        t = np.clip(((zq - pz) * ez + (yq - py) * ey) / (ez * ez + ey * ey), 0, 1)
        d = np.hypot(pz + t * ez - zq, py + t * ey - yq).min()
        assert d < 1e-12
    per = np.hypot(np.diff(az0), np.diff(ay0)).sum()
    assert abs(wm.sum() - per * hn) < 1e-12 * per * hn
    # the motion is the transform's derivative
    h = 1e-5
    zp, yp = transform(az0, ay0, aoa + h, zc, yc)
    zq, yq = transform(az0, ay0, aoa - h, zc, yc)
    gz, gy = transform_derivative(az0, ay0, aoa)
    assert np.allclose((zp - zq) / (2 * h), gz, atol=1e-9)
    assert np.allclose((yp - yq) / (2 * h), gy, atol=1e-9)
    # rotation about the quarter chord: its point does not move in z
    q = transform_derivative(np.array([0.25]), np.array([0.0]), aoa)
    assert abs(q[0][0]) < 1e-15 and abs(q[1][0]) < 1e-15


def test_marker_forcing_stops_a_uniform_flow(oracle_markers):
    """hn = 2 sqrt(hy hz): one application leaves a small part of a uniform
    flow at the markers of a circle on a uniform grid (the kernel's weights
    have sum of squares about 1/2 per axis, so a spread-then-interpolated
    force returns about hn / (2 h) of itself)."""
    g = Grid(6, 40, 40, 0.6, 4.0, 4.0)
    o = oracle_markers(g)
    n = 120
    th = (np.arange(n) + 0.5) * 2 * math.pi / n
    r = 1.0
    z, y = 2.0 + r * np.cos(th), 2.0 + r * np.sin(th)
    ds = 2 * math.pi * r / n
    h = (0.1, 0.1, 0.1)
    assert o.set_markers(z, y, np.full(n, ds * 2 * 0.1), h) == 0
    u, v, w = g.halo(), g.halo(), g.halo()
    w[...] = 1.0
    U = o.marker_interp(u, v, w, n)
    f = o.marker_spread(-U)
    U2 = o.marker_interp(u + f[0], v + f[1], w + f[2], n)
    # markers in the first and last x cell lose the point of their window
    # outside the domain (weight 1/6 of the kernel): the others
    assert np.abs(U[2][:, 1:-1] - 1.0).max() < 1e-12
    assert np.abs(U2[2][:, 1:-1]).mean() < 0.15
