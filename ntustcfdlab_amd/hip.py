"""Python host mirror of the reference's hot-path interface, over the C ABI
of libntust_hip.so (include/ntust.h).

`ProjectionSolver` exposes one method per reference subroutine, with the
reference's name and argument meaning, operating in place on numpy arrays in
the reference's Fortran layout (see grid.py) -- the same contract the
Fortran shims in fortran/hotpath.f90 give the Fortran driver.  The
device-resident stepping (configure / upload / step / download) keeps the
state in HBM across steps.

There is no fallback: if libntust_hip.so is missing or cannot be loaded,
constructing a solver raises.
"""
from __future__ import annotations

import ctypes as ct
from pathlib import Path

import numpy as np

from .grid import Grid, bc_codes

LIB = Path(__file__).resolve().parent / "libntust_hip.so"

SOR_LEXICOGRAPHIC, SOR_REDBLACK, SOR_MIXED = 0, 1, 2
QUICK_CENTRE, UPWIND = 0, 1
ARITH_EXACT, ARITH_FAST, ARITH_FP32 = 0, 1, 2
WARN_ITMAX = 1
E_STATE = -5
E_IO = -6
PROF_KERNELS = {0: ("k_rb_sweep<C>", "two-pass FP64 half-sweep"),
                2: ("k_rb_sweep32<C>", "two-pass FP32 half-sweep"),
                6: ("k_lex_sweep<C>", "lexicographic FP64 step: one colour, pipelined sweeps")}

_D = ct.POINTER(ct.c_double)
_F = ct.POINTER(ct.c_float)
_I = ct.POINTER(ct.c_int)
_lib = None


class NtustError(RuntimeError):
    pass


def load_library() -> ct.CDLL:
    """Load libntust_hip.so (built by ntustcfdlab_amd.build / __graft_entry__.build)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB.exists():
        raise NtustError(f"{LIB} is not built; run __graft_entry__.build()")
    lib = ct.CDLL(str(LIB))
    d, i, vp = ct.c_double, ct.c_int, ct.c_void_p
    sig = {
        "ntust_init": [i, i, i, _D, _D, _D, _D, _D, _D],
        "ntust_finalize": [],
        "ntust_set_bc": [_I],
        "ntust_gradient_phi_gauss": [_D, _D, _D, _D, _D, _D, i, i],
        "ntust_smagorinsky": [_D, _D, _D, _D, _D, _D, d, i, i, _D, _D, _D, _D],
        "ntust_discretisation_quick_centre": [_D, _D, _D, _D, d, d, i, i, i, _D, _D, _D],
        "ntust_discretisation_upwind": [_D, _D, _D, d, d, _D, _D, _D],
        "ntust_gauss_seidel": [_D, _D, _D, _D, _D, d, d, d, d, _I, _D],
        "ntust_calcul_new_velocity": [_D, _D, _D, _D, _D, d, d, d, d, i, i] + [_D] * 12,
        "ntust_updating_velocity": [_D, _D, _D, i, i, _D, _D, _D],
        "ntust_virtual_force_integrator": [_D, _D, d, _D, _D],
        "ntust_virtual_force_integrator_noalloc": [vp, vp, d, vp],
        "ntust_boundary_conditions": [_D] * 13,
        "ntust_configure": [d, d, d, i, i, d, d, d, i, d, i],
        "ntust_state_upload": [_D] * 5,
        "ntust_state_download": [_D] * 8,
        "ntust_step": [i, _I, _D, _D, _D],
        "ntust_prof_reset": [],
        "ntust_prof_read": [_D, ct.POINTER(ct.c_longlong)],
        "ntust_prof_copy": [ct.c_double, _D],
        "ntust_set_arithmetic": [i],
        "ntust_set_markers": [i, _D, _D, _D, d, d, d, _D, _D, _D],
        "ntust_marker_velocity": [_D],
        "ntust_adjoint_marker_gradient": [_D, _D],
        "ntust_set_body_force": [_D, _D, _D],
        "ntust_kinetic_energy": [_D],
        "ntust_checkpoint_reserve": [i],
        "ntust_checkpoint_save": [i],
        "ntust_checkpoint_load": [i],
        "ntust_adjoint_set": [_D] * 4,
        "ntust_adjoint_get": [_D] * 4,
        "ntust_adjoint_add_ke": [d],
        "ntust_adjoint_step": [_I],
        "ntust_adjoint_gradient": [_D] * 3,
        "ntust_adjoint_run": [i, i, i, _D, _I, _D],
        "ntust_adjoint_objective": [d, d, d, i],
        "ntust_adjoint_eta_gradient": [_D],
        "ntust_adjoint_checkpoint_plan": [i, d, _I, _D],
        "ntust_poisson_coeffs": [_D, _I],
        "ntust_poisson_rhs": [_D, _D, _D, d, _D],
        "ntust_pressure_field": [_D, _D, _D],
        "ntust_prof_kernel": [_D],
        "ntust_output_q": [ct.c_char_p, _F],
        "ntust_read_q": [ct.c_char_p, _F],
        "ntust_reading_variables": [_F],
        "ntust_write_mesh_x": [ct.c_char_p, _D, _D, _D],
        "ntust_check_steady": [i, i, _D],
        "ntust_restart_write": [ct.c_char_p, i, d],
        "ntust_restart_read": [ct.c_char_p, _I, _D],
        "ntust_ray_casting": [_D, _D, i, d, d, d, _D, _D, i, _D],
        "ntust_matmul_ell": [_D, _D, _I, _D, i, i],
        "ntust_solver_bicg3": [_D, _I, _D, _D, i, i, d, i, _I, _D],
        "ntust_dynamic_aoa": [d, d, d, d, _D],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = ct.c_int
    lib.ntust_last_error.restype = ct.c_char_p
    lib.ntust_last_error.argtypes = []
    lib.ntust_stream.restype = vp
    lib.ntust_stream.argtypes = []
    _lib = lib
    return lib


def dynamic_aoa(AOA1: float, time: float, dt: float, StartDynamic: float) -> float:
    """dynamic_AOA (airfoil.f90:311-321), evaluated by the library."""
    lib = load_library()
    out = ct.c_double()
    _check(lib.ntust_dynamic_aoa(float(AOA1), float(time), float(dt), float(StartDynamic),
                                 ct.byref(out)))
    return out.value


def _p(a):
    if a is None:
        return None
    if a.dtype != np.float64:
        raise TypeError("arrays must be float64")
    if not (a.flags.f_contiguous or a.ndim == 1):
        raise TypeError("arrays must be Fortran-contiguous")
    return a.ctypes.data_as(_D)


def _check(rc: int) -> int:
    if rc < 0:
        raise NtustError(f"ntust error {rc}: {_lib.ntust_last_error().decode()}")
    return rc


class ProjectionSolver:
    """One grid on the GPU.  Methods mirror the reference subroutines."""

    def __init__(self, grid: Grid, bc=None):
        self.lib = load_library()
        self.grid = grid
        self._m = [np.ascontiguousarray(m, dtype=np.float64) for m in grid.metrics()]
        _check(self.lib.ntust_init(grid.nx, grid.ny, grid.nz, *[_p(m) for m in self._m]))
        self.set_bc(bc_codes() if bc is None else bc)

    def close(self):
        _check(self.lib.ntust_finalize())

    def set_bc(self, codes):
        self.bc = np.ascontiguousarray(codes, dtype=np.int32)
        _check(self.lib.ntust_set_bc(self.bc.ctypes.data_as(_I)))

    # ---- drop-in subroutines ------------------------------------------

    def GradientPhiGauss(self, Phi, dPhidX, Dxm, Dym, Dzm, istart=1, iend=None, prebase=None):
        """smagorinsky.f90:42-87"""
        iend = self.grid.nz if iend is None else iend
        _check(self.lib.ntust_gradient_phi_gauss(_p(Phi), _p(prebase), _p(dPhidX), _p(Dxm),
                                                 _p(Dym), _p(Dzm), istart, iend))

    def CalculateSmagorinskyViscosity(self, u, v, w, Viseff, Cs=0.18, dudx=None, dvdx=None,
                                      dwdx=None, istart=1, iend=None, prebase=(None,) * 3):
        """smagorinsky.f90:1-37"""
        iend = self.grid.nz if iend is None else iend
        _check(self.lib.ntust_smagorinsky(_p(u), _p(v), _p(w), _p(prebase[0]), _p(prebase[1]),
                                          _p(prebase[2]), Cs, istart, iend, _p(dudx), _p(dvdx),
                                          _p(dwdx), _p(Viseff)))

    def discretisation_QUICK_centre(self, u, v, w, Viseff, u_star, v_star, w_star, dt, nu,
                                    LES, istart=1, iend=None):
        """discretisation_QUICK_centre.f90:1-345"""
        iend = self.grid.nz if iend is None else iend
        _check(self.lib.ntust_discretisation_quick_centre(_p(u), _p(v), _p(w), _p(Viseff), dt,
                                                          nu, int(LES), istart, iend,
                                                          _p(u_star), _p(v_star), _p(w_star)))

    def discretisation_upwind(self, u, v, w, u_star, v_star, w_star, dt, nu):
        """discretisation_upwind.f90:1-275 (Q1 = 0)"""
        _check(self.lib.ntust_discretisation_upwind(_p(u), _p(v), _p(w), dt, nu, _p(u_star),
                                                    _p(v_star), _p(w_star)))

    def gauss_seidel(self, u_star, v_star, w_star, p, pre, dt, omega, zeta, itmax,
                     ordering=SOR_LEXICOGRAPHIC):
        """GS.f90:1-106.  Returns (ik, pChangeMax, warn)."""
        self.configure(dt=dt, nu=0.0, omega=omega, zeta=zeta, itmax=itmax, sor=ordering)
        ik = ct.c_int()
        pm = ct.c_double()
        rc = _check(self.lib.ntust_gauss_seidel(_p(u_star), _p(v_star), _p(w_star), _p(p),
                                                _p(pre), dt, omega, zeta, float(itmax),
                                                ct.byref(ik), ct.byref(pm)))
        return ik.value, pm.value, rc

    def calcul_new_velocity(self, u_star, v_star, w_star, p, ETA, u1, v1, w1, u2, v2, w2, FX,
                            FY, FZ, uc, vc, wc, dt, solid=(0.0, 0.0, 0.0), istart=1,
                            iend=None):
        """calcul_new_velocity.f90:1-184"""
        iend = self.grid.nz if iend is None else iend
        _check(self.lib.ntust_calcul_new_velocity(
            _p(u_star), _p(v_star), _p(w_star), _p(p), _p(ETA), solid[0], solid[1], solid[2],
            dt, istart, iend, _p(u1), _p(v1), _p(w1), _p(u2), _p(v2), _p(w2), _p(FX), _p(FY),
            _p(FZ), _p(uc), _p(vc), _p(wc)))

    def Updating_velocity(self, u2, v2, w2, u, v, w, istart=1, iend=None):
        """calcul_new_velocity.f90:195-243"""
        iend = self.grid.nz if iend is None else iend
        _check(self.lib.ntust_updating_velocity(_p(u2), _p(v2), _p(w2), istart, iend, _p(u),
                                                _p(v), _p(w)))

    def virtualForceIntegrator(self, FY, FZ, lx):
        """virtualForceIntegrator.f90:3-85.  Returns (cDrag, cLift)."""
        cd, cl = ct.c_double(), ct.c_double()
        _check(self.lib.ntust_virtual_force_integrator(_p(FY), _p(FZ), lx, ct.byref(cd),
                                                       ct.byref(cl)))
        return cd.value, cl.value

    def virtualForceIntegrator_noalloc(self, FY_dev: int, FZ_dev: int, lx: float,
                                       coef_dev: int):
        """virtualForceIntegrator.f90:3-85 on device arrays, allocating nothing
        (ntust_virtual_force_integrator_noalloc): FY_dev, FZ_dev the device
        addresses of halo arrays, coef_dev of 2 doubles (cDrag, cLift)."""
        _check(self.lib.ntust_virtual_force_integrator_noalloc(
            ct.c_void_p(FY_dev), ct.c_void_p(FZ_dev), lx, ct.c_void_p(coef_dev)))

    def boundary_conditions(self, u, v, w, p=None, u1=None, v1=None, w1=None, u2=None,
                            v2=None, w2=None, u_star=None, v_star=None, w_star=None):
        """boundary_conditions.f90:1-376"""
        _check(self.lib.ntust_boundary_conditions(_p(u), _p(v), _p(w), _p(p), _p(u1), _p(v1),
                                                  _p(w1), _p(u2), _p(v2), _p(w2), _p(u_star),
                                                  _p(v_star), _p(w_star)))

    # ---- members of the Krylov pressure path ----------------------------

    def solver_1st_member(self):
        """solver_1st_member.f90: (coef (ndim, 7) Fortran order, jcoef[4])."""
        coef = np.zeros((self.grid.cells, 7), order="F")
        jc = (ct.c_int * 4)()
        _check(self.lib.ntust_poisson_coeffs(_p(coef), jc))
        return coef, list(jc)

    def poisson_rhs(self, u_star, v_star, w_star, dt):
        """b = mChange/dt (GS.f90:39-41,49), interior array."""
        b = self.grid.interior()
        _check(self.lib.ntust_poisson_rhs(_p(u_star), _p(v_star), _p(w_star), dt, _p(b)))
        return b

    def solver_pressure_field(self, x, p, pre):
        """solver_pressure_field.f90: pre = x, p = pre on the interior."""
        _check(self.lib.ntust_pressure_field(_p(np.ascontiguousarray(x, dtype=np.float64)),
                                             _p(p), _p(pre)))

    # ---- device-resident stepping --------------------------------------

    def configure(self, *, dt, nu, Cs=0.18, LES=0, scheme=QUICK_CENTRE, omega=1.5, zeta=1e-4,
                  itmax=3000.0, sor=SOR_LEXICOGRAPHIC, lx=1.0, force_every=1):
        _check(self.lib.ntust_configure(dt, nu, Cs, int(LES), int(scheme), omega, zeta,
                                        float(itmax), int(sor), lx, int(force_every)))

    def upload(self, u, v, w, p, ETA):
        _check(self.lib.ntust_state_upload(_p(u), _p(v), _p(w), _p(p), _p(ETA)))

    def download(self, u=None, v=None, w=None, p=None, pre=None, uc=None, vc=None, wc=None):
        _check(self.lib.ntust_state_download(_p(u), _p(v), _p(w), _p(p), _p(pre), _p(uc),
                                             _p(vc), _p(wc)))

    def step(self, nsteps: int):
        """Returns per-step (ik, pChangeMax, cDrag, cLift) arrays and the status."""
        ik = np.zeros(nsteps, dtype=np.int32)
        pm = np.zeros(nsteps)
        cd = np.zeros(nsteps)
        cl = np.zeros(nsteps)
        rc = _check(self.lib.ntust_step(nsteps, ik.ctypes.data_as(_I), _p(pm), _p(cd), _p(cl)))
        return ik, pm, cd, cl, rc

    def stream(self) -> int:
        return self.lib.ntust_stream()

    def prof_reset(self):
        self.lib.ntust_prof_reset()

    def prof_kernel(self):
        """(kernel id, algorithmic bytes per grid cell of one timed launch)."""
        b = ct.c_double()
        k = self.lib.ntust_prof_kernel(ct.byref(b))
        return k, b.value

    def set_markers(self, zm=None, ym=None, wm=None, h=(1.0, 1.0, 1.0)):
        """Lagrangian-marker direct forcing (ntust_set_markers): marker
        positions zm, ym and weights wm per unit span in the (z, y) plane,
        kernel widths h = (hx, hy, hz); no arguments remove the markers."""
        if zm is None:
            _check(self.lib.ntust_set_markers(0, None, None, None, 1.0, 1.0, 1.0, None, None,
                                              None))
            self.n_markers = 0
            return
        zm, ym, wm = (np.ascontiguousarray(a, dtype=np.float64) for a in (zm, ym, wm))
        X, Y, Z = (np.ascontiguousarray(a, dtype=np.float64)
                   for a in (self.grid.X, self.grid.Y, self.grid.Z))
        _check(self.lib.ntust_set_markers(len(zm), _p(zm), _p(ym), _p(wm), float(h[0]),
                                          float(h[1]), float(h[2]), _p(X), _p(Y), _p(Z)))
        self.n_markers = len(zm)

    def marker_velocity(self):
        """U[a, m, ix] of the last step."""
        U = np.zeros(3 * self.n_markers * self.grid.nx)
        _check(self.lib.ntust_marker_velocity(_p(U)))
        return U.reshape(3, self.n_markers, self.grid.nx)

    def adjoint_marker_gradient(self):
        """(dJ/dZ_m, dJ/dY_m) of the markers."""
        gz, gy = np.zeros(self.n_markers), np.zeros(self.n_markers)
        _check(self.lib.ntust_adjoint_marker_gradient(_p(gz), _p(gy)))
        return gz, gy

    def set_arithmetic(self, mode: int):
        """ARITH_EXACT (default), ARITH_FAST or ARITH_FP32 for the resident
        step's QUICK predictor and Smagorinsky viscosity."""
        _check(self.lib.ntust_set_arithmetic(int(mode)))

    def prof_copy_calibration(self, sizes=(64 << 20, 1 << 30)):
        """Plain device copies of the given sizes (ntust_prof_copy), for the
        counter profiles.  Returns [(bytes, ms)]."""
        out = []
        for n in sizes:
            ms = ct.c_double()
            _check(self.lib.ntust_prof_copy(float(n), ct.byref(ms)))
            out.append((n, ms.value))
        return out

    def prof_read(self):
        ms = ct.c_double()
        n = ct.c_longlong()
        self.lib.ntust_prof_read(ct.byref(ms), ct.byref(n))
        return ms.value, n.value

    # ---- discrete adjoint (no counterpart in the reference) --------------

    def set_body_force(self, fu=None, fv=None, fw=None):
        """Body force added to the predictor, u* += dt*f (None removes it)."""
        _check(self.lib.ntust_set_body_force(_p(fu), _p(fv), _p(fw)))

    def kinetic_energy(self) -> float:
        ke = ct.c_double()
        _check(self.lib.ntust_kinetic_energy(ct.byref(ke)))
        return ke.value

    # ---- output, steady check, restart (resident state) ----

    def _q(self):
        g = self.grid
        return np.zeros((g.nx, g.ny, g.nz, 5), np.float32, order="F")

    def output_q(self, path=None) -> np.ndarray:
        """filereachtime (filer.f90:161-187): Qout from the resident state;
        with a path, the PLOT3D records are appended to it."""
        q = self._q()
        _check(self.lib.ntust_output_q(None if path is None else str(path).encode(),
                                       q.ctypes.data_as(_F)))
        return q

    def read_q(self, path) -> np.ndarray:
        q = self._q()
        _check(self.lib.ntust_read_q(str(path).encode(), q.ctypes.data_as(_F)))
        return q

    def reading_variables(self, q: np.ndarray):
        """reading_data.f90:47-83 on the resident state, then the boundary
        conditions (main.f90:96-100)."""
        if q.dtype != np.float32 or not q.flags.f_contiguous or q.size != 5 * self.grid.cells:
            raise TypeError("Qout must be float32 (nx,ny,nz,5) in Fortran order")
        _check(self.lib.ntust_reading_variables(q.ctypes.data_as(_F)))

    def write_mesh_x(self, path, Xs, Ys, Zs):
        _check(self.lib.ntust_write_mesh_x(str(path).encode(), _p(Xs), _p(Ys), _p(Zs)))

    def check_steady(self, istep: int, isto: int, velocity_difference: float) -> float:
        """check_steady.f90:1-30; returns the new VelocityDifference (the
        given one when istep == isto)."""
        d = ct.c_double(velocity_difference)
        _check(self.lib.ntust_check_steady(int(istep), int(isto), ct.byref(d)))
        return d.value

    def restart_write(self, path, istep: int, time: float):
        _check(self.lib.ntust_restart_write(str(path).encode(), int(istep), float(time)))

    def restart_read(self, path):
        st, t = ct.c_int(), ct.c_double()
        _check(self.lib.ntust_restart_read(str(path).encode(), ct.byref(st), ct.byref(t)))
        return st.value, t.value

    # ---- Krylov path (solver.f90) ----

    def matmul_ell(self, coef, jcoef, y):
        """matmul_ell (solver.f90:146-190): returns A y."""
        y = np.ascontiguousarray(y, np.float64)
        x = np.zeros_like(y)
        jc = np.ascontiguousarray(jcoef, np.int32)
        _check(self.lib.ntust_matmul_ell(_p(x), _p(np.asfortranarray(coef)),
                                         jc.ctypes.data_as(_I), _p(y), len(y), len(jc)))
        return x

    def solver_bicg3(self, coef, jcoef, b, x, eps: float, itmax: int):
        """solver_BICG3 (solver.f90:1-144): x (initial guess) is overwritten by
        the solution; returns (iterations, norm)."""
        jc = np.ascontiguousarray(jcoef, np.int32)
        it, nrm = ct.c_int(), ct.c_double()
        _check(self.lib.ntust_solver_bicg3(_p(np.asfortranarray(coef)), jc.ctypes.data_as(_I),
                                           _p(np.ascontiguousarray(b, np.float64)), _p(x),
                                           len(x), len(jc), float(eps), int(itmax),
                                           ct.byref(it), ct.byref(nrm)))
        return it.value, nrm.value

    # ---- airfoil geometry ----

    def ray_casting(self, az0, ay0, AOA: float, zc: float, yc: float, nsub: int = 50,
                    ETA: np.ndarray | None = None):
        """RayCasting (airfoil.f90:580-892) on the device: the resident ETA,
        and ETA (a halo array) in place when given.  az0, ay0: the polygon
        table as read (see airfoil.py)."""
        az0 = np.ascontiguousarray(az0, np.float64)
        ay0 = np.ascontiguousarray(ay0, np.float64)
        if az0.shape != ay0.shape:
            raise ValueError("az0 and ay0 differ in length")
        g = self.grid
        _check(self.lib.ntust_ray_casting(_p(az0), _p(ay0), len(az0), float(AOA), float(zc),
                                          float(yc), _p(np.ascontiguousarray(g.Y)),
                                          _p(np.ascontiguousarray(g.Z)), int(nsub), _p(ETA)))

    def checkpoint_reserve(self, n: int):
        _check(self.lib.ntust_checkpoint_reserve(int(n)))

    def checkpoint_save(self, slot: int):
        _check(self.lib.ntust_checkpoint_save(int(slot)))

    def checkpoint_load(self, slot: int):
        _check(self.lib.ntust_checkpoint_load(int(slot)))

    def adjoint_set(self, lu=None, lv=None, lw=None, lp=None):
        _check(self.lib.ntust_adjoint_set(_p(lu), _p(lv), _p(lw), _p(lp)))

    def adjoint_get(self):
        out = [self.grid.halo() for _ in range(4)]
        _check(self.lib.ntust_adjoint_get(*[_p(a) for a in out]))
        return out

    def adjoint_add_ke(self, scale: float):
        _check(self.lib.ntust_adjoint_add_ke(scale))

    def adjoint_step(self) -> int:
        ik = ct.c_int()
        _check(self.lib.ntust_adjoint_step(ct.byref(ik)))
        return ik.value

    def adjoint_gradient(self):
        out = [self.grid.halo() for _ in range(3)]
        _check(self.lib.ntust_adjoint_gradient(*[_p(a) for a in out]))
        return out

    def adjoint_objective(self, w_ke: float = 1.0, w_cd: float = 0.0, w_cl: float = 0.0,
                          eta_gradient: bool = False):
        """J = mean over the averaging window of w_ke*KE + w_cd*C_D + w_cl*C_L;
        eta_gradient also accumulates dJ/dETA (adjoint_eta_gradient)."""
        _check(self.lib.ntust_adjoint_objective(float(w_ke), float(w_cd), float(w_cl),
                                                1 if eta_gradient else 0))

    def adjoint_eta_gradient(self):
        out = self.grid.halo()
        _check(self.lib.ntust_adjoint_eta_gradient(_p(out)))
        return out

    def adjoint_checkpoint_plan(self, nsteps: int, reserve_bytes: float):
        """(checkpoint spacing, bytes of the run's stores): nsteps when every
        step's state fits in free device memory with reserve_bytes left,
        else ceil(sqrt(nsteps))."""
        k = ct.c_int()
        b = ct.c_double()
        _check(self.lib.ntust_adjoint_checkpoint_plan(int(nsteps), float(reserve_bytes),
                                                      ct.byref(k), ct.byref(b)))
        return k.value, b.value

    def adjoint_run(self, nsteps: int, checkpoint_every: int, navg: int):
        """Returns (J, per-step sweep counts, ms[forward, reverse, recompute, adjoint])."""
        J = ct.c_double()
        ik = np.zeros(nsteps, dtype=np.int32)
        ms = np.zeros(4)
        _check(self.lib.ntust_adjoint_run(int(nsteps), int(checkpoint_every), int(navg),
                                          ct.byref(J), ik.ctypes.data_as(_I), _p(ms)))
        return J.value, ik, ms
