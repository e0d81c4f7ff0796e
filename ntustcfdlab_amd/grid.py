"""Host-side set-up of a run: grid metrics, boundary-condition codes, initial
fields.  These are the driver-side steps of the reference (reading_data.f90,
gridder.f90, initial_conditions.f90, the BC strings of variables_module.f90)
restated for the Python host; the Fortran driver in fortran/ has its own.

Arrays use the reference's Fortran layout: halo arrays are numpy arrays of
shape (nx+4, ny+4, nz+4) in Fortran order, element (i,j,k) at
[i+1, j+1, k+1]; interior arrays (nx, ny, nz) at [i-1, j-1, k-1]; metric
arrays (n+4,) at [m+1].
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

# BC codes in the order the C ABI expects
FACES = ("West", "East", "South", "North", "Back", "Front")
BC_CODE = {"no-slip": 0, "Neumann": 1, "Dirichlet": 2}


def metrics_from_nodes(X: np.ndarray, n: int) -> tuple[np.ndarray, np.ndarray]:
    """Cell widths iD (X(i+1)-X(i)) and centre spacings Ds ((X(i+2)-X(i))/2)
    with the ghost rules of gridder.f90:267-327.  X holds nodes -1..n+3 at
    X[m+1].  Returns (iD, Ds), each (n+4,) indexed [m+1], m = -1..n+2."""
    x = lambda m: X[m + 1]  # noqa: E731
    iD = np.zeros(n + 4)
    Ds = np.zeros(n + 4)
    for m in range(1, n):
        iD[m + 1] = x(m + 1) - x(m)
        Ds[m + 1] = (x(m + 2) - x(m)) / 2.0
    iD[0 + 1] = iD[1 + 1]
    iD[-1 + 1] = iD[1 + 1]
    iD[n + 1] = x(n + 1) - x(n)
    iD[n + 1 + 1] = iD[n + 1]
    iD[n + 2 + 1] = iD[n + 1]
    Ds[0 + 1] = Ds[1 + 1]
    Ds[-1 + 1] = Ds[1 + 1]
    Ds[n + 1] = Ds[n - 1 + 1]
    Ds[n + 1 + 1] = Ds[n - 1 + 1]
    Ds[n + 2 + 1] = Ds[n - 1 + 1]
    return iD, Ds


def uniform_nodes(n: int, length: float) -> np.ndarray:
    """Nodes of gridder_equal (gridder.f90:235-263): X(1)=0, X(0)=-dx,
    X(-1)=-2dx, X(i)=X(i-1)+dx accumulated, dx = lx/nx (reading_data.f90:25)."""
    d = float(length) / n
    X = np.zeros(n + 5)
    X[1 + 1] = 0.0
    X[0 + 1] = X[1 + 1] - d
    X[-1 + 1] = X[0 + 1] - d
    for m in range(2, n + 4):
        X[m + 1] = X[m] + d
    return X


def tanh_nodes(n: int, length: float, centre: float, beta: float) -> np.ndarray:
    """Nodes clustered around the interior point `centre` (0 < centre <
    length, an absolute coordinate): Roberts' interior-clustering transform
    of a uniform parameter s in [0, 1],
        x(s) = c (1 + sinh(beta (s - A)) / sinh(beta A)),
        A = ln[(1 + (e^beta - 1) c/L) / (1 + (e^-beta - 1) c/L)] / (2 beta),
    which puts x(0) = 0, x(1) = L and the finest spacing at x = c (the
    spacing grows like cosh(beta (s - A)) away from it; beta > 0 sets the
    strength).  Uniform ghost spacing copied from the end cells.  The
    reference has only uniform and piecewise-uniform grids
    (gridder.f90:1-223); this generator serves the stretched airfoil grids."""
    if not (0.0 < centre < length) or not beta > 0.0:
        raise ValueError("tanh_nodes needs 0 < centre < length and beta > 0")
    s = np.linspace(0.0, 1.0, n + 1)
    r = centre / length
    A = np.log((1.0 + (np.exp(beta) - 1.0) * r) / (1.0 + (np.exp(-beta) - 1.0) * r)) / (2 * beta)
    xi = r * (1.0 + np.sinh(beta * (s - A)) / np.sinh(beta * A))
    xi[0], xi[-1] = 0.0, 1.0
    X = np.zeros(n + 5)
    X[2:n + 3] = xi * length           # nodes 1..n+1
    h0 = X[3] - X[2]
    h1 = X[n + 2] - X[n + 1]
    X[1] = X[2] - h0                   # node 0
    X[0] = X[1] - h0                   # node -1
    X[n + 3] = X[n + 2] + h1           # node n+2
    X[n + 4] = X[n + 3] + h1           # node n+3
    return X


@dataclass
class Grid:
    nx: int
    ny: int
    nz: int
    lx: float = 1.0
    ly: float = 1.0
    lz: float = 1.0
    X: np.ndarray = field(default=None, repr=False)
    Y: np.ndarray = field(default=None, repr=False)
    Z: np.ndarray = field(default=None, repr=False)

    def __post_init__(self):
        if self.X is None:
            self.X = uniform_nodes(self.nx, self.lx)
        if self.Y is None:
            self.Y = uniform_nodes(self.ny, self.ly)
        if self.Z is None:
            self.Z = uniform_nodes(self.nz, self.lz)
        self.iDx, self.Dxs = metrics_from_nodes(self.X, self.nx)
        self.iDy, self.Dys = metrics_from_nodes(self.Y, self.ny)
        self.iDz, self.Dzs = metrics_from_nodes(self.Z, self.nz)

    @property
    def halo_shape(self):
        return (self.nx + 4, self.ny + 4, self.nz + 4)

    @property
    def shape(self):
        return (self.nx, self.ny, self.nz)

    @property
    def cells(self) -> int:
        return self.nx * self.ny * self.nz

    def halo(self) -> np.ndarray:
        return np.zeros(self.halo_shape, order="F")

    def interior(self) -> np.ndarray:
        return np.zeros(self.shape, order="F")

    def metrics(self):
        return (self.iDx, self.Dxs, self.iDy, self.Dys, self.iDz, self.Dzs)

    def centres(self):
        """Cell centres Xs(1:nx), Ys, Zs: (X(i+1) + X(i)) / 2.0
        (gridder.f90:173-183, 343-353)."""
        def c(N, n):
            return np.ascontiguousarray((N[3:n + 3] + N[2:n + 2]) / 2.0)
        return c(self.X, self.nx), c(self.Y, self.ny), c(self.Z, self.nz)


def bc_codes(**over) -> np.ndarray:
    """The 18 BC codes, default = the reference's cavity
    (variables_module.f90:110-132: no-slip everywhere, NorthWall_w Dirichlet).
    Override with e.g. NorthWall_w='no-slip', BackWall_w='Dirichlet'."""
    names = {}
    for f in FACES:
        for c in "uvw":
            names[f"{f}Wall_{c}"] = "no-slip"
    names["NorthWall_w"] = "Dirichlet"
    for k, v in over.items():
        if k not in names:
            raise KeyError(k)
        names[k] = v
    return np.array([BC_CODE.get(names[f"{f}Wall_{c}"], 99) for f in FACES for c in "uvw"],
                    dtype=np.int32)


def noise_ic(grid: Grid, amp: float, seed: int):
    """u, v, w = seeded U(-amp, amp) on the interior faces 1..n of each
    axis (a rest state plus noise); ghosts zero.  Apply the boundary
    conditions afterwards.  numpy's generator is used (fast, seeded); the
    values are data, fixed by the seed."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(3):
        a = grid.halo()
        a[2:grid.nx + 2, 2:grid.ny + 2, 2:grid.nz + 2] = rng.uniform(-amp, amp, grid.shape)
        out.append(a)
    return out


def gaussian_blob(grid: Grid, amp: float, width: float, centre) -> list:
    """A body force f = amp*exp(-|x - centre|^2/width^2) on every interior
    face of each component (u faces at (X(i+1), yc(j), zc(k)), v and w
    alike; yc, zc cell centres), ghosts zero: three halo arrays."""
    def axes(N, n):
        node = N[3:n + 3]                      # X(i+1), i = 1..n
        cen = 0.5 * (N[2:n + 2] + N[3:n + 3])  # (X(i) + X(i+1))/2
        return node, cen
    (xn, xc), (yn, yc), (zn, zc) = (axes(grid.X, grid.nx), axes(grid.Y, grid.ny),
                                    axes(grid.Z, grid.nz))
    out = []
    for comp in range(3):
        ax = (xn if comp == 0 else xc, yn if comp == 1 else yc, zn if comp == 2 else zc)
        gx, gy, gz = np.meshgrid(*ax, indexing="ij")
        r2 = (gx - centre[0]) ** 2 + (gy - centre[1]) ** 2 + (gz - centre[2]) ** 2
        f = grid.halo()
        f[2:grid.nx + 2, 2:grid.ny + 2, 2:grid.nz + 2] = amp * np.exp(-r2 / width ** 2)
        out.append(f)
    return out
