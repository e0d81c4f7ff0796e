"""The Fortran drop-in layer: the driver (fortran/driver.f90) calls the
hot-path shims with the reference's subroutine names (fortran/hotpath.f90)
through iso_c_binding, or the device-resident step, and its final state
must equal the CPU restatement stepped in main.f90 order, bit for bit."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from ntustcfdlab_amd.grid import Grid, bc_codes, noise_ic
from oracle import OracleRun

pytestmark = pytest.mark.gpu

DRIVER = Path(__file__).resolve().parents[1] / "ntustcfdlab_amd" / "fortran" / "ntust_driver"


def read_out(path, g, nstep):
    raw = path.read_bytes()
    off = 4
    assert int(np.frombuffer(raw[:4], np.int32)[0]) == nstep
    ik = np.frombuffer(raw[off:off + 4 * nstep], np.int32)
    off += 4 * nstep
    cd = np.frombuffer(raw[off:off + 8 * nstep], np.float64); off += 8 * nstep
    cl = np.frombuffer(raw[off:off + 8 * nstep], np.float64); off += 8 * nstep
    out = {}
    nh, ni = int(np.prod(g.halo_shape)), g.cells
    for n in ("u", "v", "w", "p"):
        out[n] = np.frombuffer(raw[off:off + 8 * nh], np.float64).reshape(g.halo_shape, order="F")
        off += 8 * nh
    for n in ("pre", "uc", "vc", "wc"):
        out[n] = np.frombuffer(raw[off:off + 8 * ni], np.float64).reshape(g.shape, order="F")
        off += 8 * ni
    assert off == len(raw)
    return ik, cd, cl, out


CASES = {
    # the reference's cavity as it ships: Re = 10, dt = 1.0e-3 (single
    # literal widened), zeta = 1.e-4, upwind, LES off, 40 cells reduced to 16
    "cavity-upwind": dict(n=16, Re=10.0, dt=float(np.float32(1e-3)), steps=5, scheme=1, LES=0,
                          Cs=float(np.float32(0.18)), zeta=float(np.float32(1e-4)),
                          itmax=3000.0, sor=0, noise=0.0),
    "les-quick": dict(n=12, Re=1000.0, dt=5e-4, steps=4, scheme=0, LES=1, Cs=0.1, zeta=1e-9,
                      itmax=500.0, sor=0, noise=1e-3),
    "les-quick-rb": dict(n=12, Re=1000.0, dt=5e-4, steps=4, scheme=0, LES=1, Cs=0.1, zeta=1e-9,
                         itmax=500.0, sor=1, noise=1e-3),
}


@pytest.mark.parametrize("mode", [0, 1], ids=["per-call-shims", "resident"])
@pytest.mark.parametrize("name", list(CASES))
def test_fortran_driver_matches_oracle(name, mode, tmp_path, hip_lib):
    assert DRIVER.exists(), "Fortran driver not built"
    c = CASES[name]
    g = Grid(c["n"], c["n"], c["n"])
    ref = OracleRun(g, dt=c["dt"], nu=1.0 / c["Re"], Cs=c["Cs"], omega=1.5, zeta=c["zeta"],
                    itmax=c["itmax"], LES=c["LES"], scheme=c["scheme"], sor=c["sor"],
                    bc=bc_codes())
    args = [str(DRIVER), str(tmp_path / "out.bin"), *(str(c["n"]),) * 3, repr(c["Re"]),
            repr(c["dt"]), str(c["steps"]), str(c["scheme"]), str(c["LES"]), repr(c["Cs"]),
            "1.5", repr(c["zeta"]), repr(c["itmax"]), str(c["sor"]), str(mode)]
    if c["noise"]:
        u, v, w = noise_ic(g, c["noise"], seed=11)
        ref.f["u"][...], ref.f["v"][...], ref.f["w"][...] = u, v, w
        ic = tmp_path / "ic.bin"
        ic.write_bytes(b"".join(a.tobytes(order="F") for a in (u, v, w)))
        args.append(str(ic))
    ref.apply_bc()
    rows = [ref.step() for _ in range(c["steps"])]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    ik, cd, cl, out = read_out(tmp_path / "out.bin", g, c["steps"])
    assert list(ik) == [x[0] for x in rows]
    assert np.array_equal(cd.view(np.uint64), np.array([x[2] for x in rows]).view(np.uint64))
    for n in ("u", "v", "w", "p", "pre", "uc", "vc", "wc"):
        assert np.array_equal(out[n].view(np.uint64), ref.f[n].view(np.uint64)), n


@pytest.mark.parametrize("mode", [0, 1], ids=["host-fortran-io", "resident-library"])
def test_fortran_driver_output_files(mode, tmp_path, hip_lib):
    """The output of main.f90:113-118,315-328 (a solution file at the start
    and every isto steps, check_steady after each): mode 0 writes the files
    by Fortran I/O from the host arrays, mode 1 through ntust_output_q and
    ntust_check_steady from the resident state.  Both must equal the
    oracle's files byte for byte and its VelocityDifference sequence."""
    assert DRIVER.exists(), "Fortran driver not built"
    c = CASES["les-quick-rb"]
    n, steps, isto = c["n"], 6, 2
    g = Grid(n, n, n)
    ref = OracleRun(g, dt=c["dt"], nu=1.0 / c["Re"], Cs=c["Cs"], omega=1.5, zeta=c["zeta"],
                    itmax=c["itmax"], LES=c["LES"], scheme=c["scheme"], sor=c["sor"],
                    bc=bc_codes())
    u, v, w = noise_ic(g, c["noise"], seed=11)
    ref.f["u"][...], ref.f["v"][...], ref.f["w"][...] = u, v, w
    ic = tmp_path / "ic.bin"
    ic.write_bytes(b"".join(a.tobytes(order="F") for a in (u, v, w)))
    ref.apply_bc()
    run = tmp_path / "run"
    run.mkdir()
    args = [str(DRIVER), str(tmp_path / "out.bin"), *(str(n),) * 3, repr(c["Re"]), repr(c["dt"]),
            str(steps), str(c["scheme"]), str(c["LES"]), repr(c["Cs"]), "1.5", repr(c["zeta"]),
            repr(c["itmax"]), str(c["sor"]), str(mode), str(ic), str(isto)]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120, cwd=run)
    assert r.returncode == 0, r.stdout + r.stderr

    from oracle import Oracle
    o = Oracle(g)
    f = ref.f
    want = run.parent / "want"
    want.mkdir()
    last = g.interior()
    vds = []

    def out(num):
        o.write_q(want / f"{num:04d}.q", o.pack_q(f["pre"], f["uc"], f["vc"], f["wc"], f["ETA"]))

    out(0)
    for istep in range(1, steps + 1):
        ref.step()
        if istep % isto == 0:
            out(istep // isto)
            vds.append(o.check_steady(istep, isto, f["wc"], last, 1024.0))
    names = sorted(p.name for p in want.iterdir())
    assert names == [f"{k:04d}.q" for k in range(steps // isto + 1)]
    for name in names:
        assert (run / name).read_bytes() == (want / name).read_bytes(), name
    raw = (run / "vdiff.bin").read_bytes()
    assert int(np.frombuffer(raw[:4], np.int32)[0]) == len(vds)
    got = np.frombuffer(raw[4:], np.float64)
    assert np.array_equal(got.view(np.uint64), np.array(vds).view(np.uint64))
    assert vds[0] == 1024.0 and vds[1] != 1024.0


@pytest.mark.parametrize("mode", [0, 1], ids=["per-call-shims", "resident"])
def test_fortran_driver_airfoil_pitching(mode, tmp_path, hip_lib):
    """RayCasting at start-up and, from StartDynamic on, dynamic_AOA +
    RayCasting before every step (main.f90:109-110, 144-155): in mode 0 the
    airfoil_gpu.f90 shims with the reference's names, in mode 1 the
    resident ETA.  The state, sweep counts and force coefficients must
    equal the oracle's, whose ETA comes from its own RayCasting."""
    from ntustcfdlab_amd.airfoil import naca4_polygon
    from oracle import Oracle, dynamic_aoa
    assert DRIVER.exists(), "Fortran driver not built"
    nx, n, L, steps, start, isto, poly = 6, 24, 8.0, 6, 4.0, 3, 1000
    az0, ay0 = naca4_polygon(poly=poly)
    table = tmp_path / "naca.dat"
    table.write_text("".join(f"{float(z)!r} {float(y)!r}\n" for z, y in zip(az0, ay0)))
    dt, Re = 5e-4, 1000.0
    g = Grid(nx, n, n, ly=L, lz=L)
    ref = OracleRun(g, dt=dt, nu=1.0 / Re, Cs=0.1, omega=1.5, zeta=1e-9, itmax=300.0, LES=1,
                    scheme=0, sor=1, bc=bc_codes())
    u, v, w = noise_ic(g, 1e-2, seed=17)
    ref.f["u"][...], ref.f["v"][...], ref.f["w"][...] = u, v, w
    ref.apply_bc()
    o = Oracle(g)

    def cast(aoa):
        az, ay = o.airfoil_transform(az0, ay0, aoa, 4.0, 4.0)
        assert o.ray_casting(az, ay, g.Y, g.Z, 50, ref.f["ETA"]) is not None

    cast(5.0)
    assert ref.f["ETA"].any()
    rows, time = [], 0.0
    for istep in range(1, steps + 1):
        if istep >= start:
            cast(dynamic_aoa(5.0, time, dt, start))
        rows.append(ref.step())
        time = time + dt
    ic = tmp_path / "ic.bin"
    ic.write_bytes(b"".join(a.tobytes(order="F") for a in (u, v, w)))
    run = tmp_path / "run"
    run.mkdir()
    args = [str(DRIVER), str(tmp_path / "out.bin"), str(nx), str(n), str(n), repr(Re), repr(dt),
            str(steps), "0", "1", "0.1", "1.5", "1e-9", "300.0", "1", str(mode), str(ic),
            str(isto), str(table), "5.0", repr(start), repr(L), str(poly)]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120, cwd=run)
    assert r.returncode == 0, r.stdout + r.stderr
    ik, cd, cl, out = read_out(tmp_path / "out.bin", g, steps)
    assert list(ik) == [x[0] for x in rows]
    assert np.array_equal(cd.view(np.uint64), np.array([x[2] for x in rows]).view(np.uint64))
    assert np.array_equal(cl.view(np.uint64), np.array([x[3] for x in rows]).view(np.uint64))
    assert np.all(cd != 0)
    for name in ("u", "v", "w", "p", "pre", "uc", "vc", "wc"):
        assert np.array_equal(out[name].view(np.uint64), ref.f[name].view(np.uint64)), name


def test_fortran_driver_krylov_path(tmp_path, hip_lib):
    """solver_1st_member, solver_BICG3 and matmul_ell through the shims of
    hotpath.f90 with the reference's argument lists: the solution equals the
    library's own call bit for bit, the printed count and norm are the
    reference's messages, and A x is the restatement's product of x."""
    from ntustcfdlab_amd.hip import ProjectionSolver
    from oracle import Oracle
    assert DRIVER.exists(), "Fortran driver not built"
    n, eps, itmax = 14, 1e-9, 2000
    g = Grid(n, n, n)
    b = np.random.default_rng(4).standard_normal(g.cells)
    ic = tmp_path / "b.bin"
    ic.write_bytes(b.tobytes())
    args = [str(DRIVER), str(tmp_path / "out.bin"), *(str(n),) * 3, "1000.0", "5e-4", "1",
            "0", "0", "0.1", "1.5", repr(eps), repr(float(itmax)), "1", "2", str(ic)]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.frombuffer((tmp_path / "out.bin").read_bytes(), np.float64)
    x, ax = raw[:g.cells], raw[g.cells:]
    line = [ln for ln in r.stdout.splitlines() if "Iterations BICG3" in ln]
    assert len(line) == 1, r.stdout
    it, nrm = line[0].split("=")[1].split()
    s = ProjectionSolver(g)
    try:
        coef, jcoef = s.solver_1st_member()
        xl = np.zeros(g.cells)
        itl, nrml = s.solver_bicg3(coef, jcoef, b, xl, eps, itmax)
    finally:
        s.close()
    assert int(it) == itl and float(nrm) == pytest.approx(nrml, rel=1e-12) and nrml <= eps
    assert np.array_equal(x.view(np.uint64), xl.view(np.uint64))
    assert np.array_equal(ax.view(np.uint64), Oracle.matmul_ell(coef, jcoef, x).view(np.uint64))
