"""Lagrangian markers of an airfoil for the marker direct forcing
(ntust_set_markers).

No counterpart in the reference, whose immersed boundary is the volume
fraction ETA cast from the airfoil table (airfoil.f90:580-892).  The markers
use the same table and the same placement: points at equal arc-length steps
along the table polygon are moved with the table's transform of
airfoil.f90:600-611 (shift to the quarter chord, rotation by -AOA with the
y rotation using the already rotated z -- a shear -- then the translation to
(zc, yc)).  The transform is linear in the table point, so a marker stays
at the same place of the cast polygon at every angle, and its velocity
dX/dAOA is the transform's derivative at its table point.
"""
from __future__ import annotations

import math

import numpy as np


def transform(az0, ay0, aoa, zc, yc):
    """airfoil.f90:600-611 on the table points (PI = 4 atan 1)."""
    t = -aoa * (4.0 * math.atan(1.0)) / 180.0
    c, s = math.cos(t), math.sin(t)
    z1 = np.asarray(az0, dtype=np.float64) - 0.25
    y0 = np.asarray(ay0, dtype=np.float64)
    z2 = c * z1 - s * y0
    y2 = s * z2 + c * y0
    return z2 + zc + 0.25, y2 + yc


def transform_derivative(az0, ay0, aoa):
    """d(z, y)/dAOA of transform(), per degree."""
    dt = -(4.0 * math.atan(1.0)) / 180.0
    t = -aoa * (4.0 * math.atan(1.0)) / 180.0
    c, s = math.cos(t), math.sin(t)
    z1 = np.asarray(az0, dtype=np.float64) - 0.25
    y0 = np.asarray(ay0, dtype=np.float64)
    z2 = c * z1 - s * y0
    dz2 = -s * z1 - c * y0
    dy2 = c * z2 + s * dz2 - s * y0
    return dz2 * dt, dy2 * dt


def airfoil_markers(az0, ay0, aoa, zc, yc, ds, hn):
    """Markers on the cast polygon: n = round(perimeter / ds) points at the
    midpoints of equal arc-length steps of the table polygon (its last
    vertex equal to the first), each of weight (step length) * hn per unit
    span, hn the thickness the forcing is spread over.  Returns (zm, ym, wm,
    dzm, dym): positions, weights and dX/dAOA per degree.  The weights do
    not depend on the angle."""
    az0 = np.asarray(az0, dtype=np.float64)
    ay0 = np.asarray(ay0, dtype=np.float64)
    seg = np.hypot(np.diff(az0), np.diff(ay0))
    s = np.concatenate([[0.0], np.cumsum(seg)])
    per = float(s[-1])
    n = max(8, int(round(per / ds)))
    step = per / n
    sm = (np.arange(n) + 0.5) * step
    z0 = np.interp(sm, s, az0)
    y0 = np.interp(sm, s, ay0)
    zm, ym = transform(z0, y0, aoa, zc, yc)
    dzm, dym = transform_derivative(z0, y0, aoa)
    wm = np.full(n, step * hn)
    return zm, ym, wm, dzm, dym


def kernel_widths(grid, zm, ym):
    """(hx, hy, hz) for markers at (zm, ym): the x spacing, and the smallest y
    and z node spacings within three cells of the markers, so that at most
    four points of every component fall within a kernel's support
    (|d| < 2 h); where the grid is coarser the volume weights of the
    interpolation keep its weights summing to about one."""
    hx = float(np.min(grid.iDx[2:grid.nx + 2]))

    def finest(nodes, pos):
        # centre spacings are means of node spacings: the node spacing bounds both
        lo = np.searchsorted(nodes, pos.min()) - 3
        hi = np.searchsorted(nodes, pos.max()) + 3
        lo, hi = max(lo, 0), min(hi, len(nodes) - 1)
        return float(np.min(np.diff(nodes[lo:hi + 1])))
    return hx, finest(np.asarray(grid.Y), np.asarray(ym)), finest(np.asarray(grid.Z),
                                                                    np.asarray(zm))


def airfoil_marker_set(grid, az0, ay0, aoa, zc, yc, ds, widths=None, hn=None):
    """Markers of the cast airfoil for ProjectionSolver.set_markers: a dict
    with zm, ym, wm, dzm, dym (as airfoil_markers), the kernel widths h
    (kernel_widths of the markers, unless given) and the thickness hn = 2
    sqrt(hy hz) (unless given: with the kernel's sum of squares about 1/2
    per axis one forcing step then returns about all of a force to its
    marker).
    Passing the widths and hn of another angle keeps the marker model the
    same, so J(AOA) differs only through the positions."""
    zm, ym, _, _, _ = airfoil_markers(az0, ay0, aoa, zc, yc, ds, 1.0)
    h = tuple(widths) if widths is not None else kernel_widths(grid, zm, ym)
    hn = hn if hn is not None else 2.0 * math.sqrt(h[1] * h[2])
    zm, ym, wm, dzm, dym = airfoil_markers(az0, ay0, aoa, zc, yc, ds, hn)
    return dict(zm=zm, ym=ym, wm=wm, dzm=dzm, dym=dym, h=h, hn=hn)
