"""Airfoil polygons for the immersed-boundary geometry (RayCasting,
airfoil.f90:580-892).

The reference reads a two-column table (z, y) of `poly` vertices from
NACA_filename with list-directed input (airfoil.f90:587-591): chord along z
from 0 to 1, the last vertex equal to the first.  `read_polygon` reads such
a table; `naca4_polygon` builds one of the same layout from the published
NACA four-digit thickness distribution, for runs without the table.
"""
from __future__ import annotations

import numpy as np


def read_polygon(path, poly: int | None = None):
    """(az, ay) from a two-column text table, the first `poly` rows."""
    tab = np.loadtxt(path, dtype=np.float64, max_rows=poly, ndmin=2)
    if poly is not None and len(tab) < poly:
        raise ValueError(f"{path}: {len(tab)} rows, {poly} expected")
    return np.ascontiguousarray(tab[:, 0]), np.ascontiguousarray(tab[:, 1])


def naca4_polygon(thickness: float = 0.12, poly: int = 5000):
    """Symmetric NACA 00tt section, unit chord along z: from the leading edge
    along the upper surface to the trailing edge, back along the lower
    surface, the last vertex repeating the first; cosine spacing in z.
    y_t = 5 t (0.2969 sqrt(z) - 0.1260 z - 0.3516 z^2 + 0.2843 z^3 - 0.1015 z^4)
    (NACA Report 824), so the trailing edge is open and closed by one edge."""
    if poly < 6 or poly % 2:
        raise ValueError("poly must be even and >= 6")
    h = (poly - 2) // 2                       # upper surface: h+1 points
    z = 0.5 * (1.0 - np.cos(np.linspace(0.0, np.pi, h + 1)))
    yt = 5.0 * thickness * (0.2969 * np.sqrt(z) - 0.1260 * z - 0.3516 * z**2
                            + 0.2843 * z**3 - 0.1015 * z**4)
    az = np.concatenate([z, z[:0:-1], z[:1]])
    ay = np.concatenate([yt, -yt[:0:-1], yt[:1]])
    return np.ascontiguousarray(az), np.ascontiguousarray(ay)


def polygon_area(az, ay) -> float:
    """Shoelace area of a closed polygon (last vertex = first)."""
    return 0.5 * abs(float(np.dot(az[:-1], ay[1:]) - np.dot(az[1:], ay[:-1])))
