"""GPU parity of the output, steady-check and restart entry points on the
resident state, against the CPU restatement (whose file layout and restart
arithmetic are pinned against the Fortran runtime in test_output_format.py):
  - ntust_output_q: Qout bit for bit, the appended .q file byte for byte;
  - ntust_read_q / ntust_reading_variables: the state rebuilt from a .q
    file, then the boundary conditions, bit for bit; refusals;
  - ntust_write_mesh_x: byte for byte;
  - ntust_check_steady: the VelocityDifference sequence, NaN skipped;
  - ntust_restart_write / _read: a run continued from the file equals the
    uninterrupted run bit for bit; other grids and damaged files refused
    with the state left as it was."""
import numpy as np
import pytest

from conftest import stretched_grid
from ntustcfdlab_amd.grid import bc_codes, noise_ic
from ntustcfdlab_amd.hip import E_IO, QUICK_CENTRE, SOR_REDBLACK, NtustError, ProjectionSolver
from oracle import Oracle, OracleRun

pytestmark = pytest.mark.gpu

PRM = dict(dt=5e-4, nu=1e-3, Cs=0.1, LES=1, scheme=QUICK_CENTRE, omega=1.5, zeta=1e-8,
           itmax=300.0, sor=SOR_REDBLACK, lx=1.0, force_every=1)


def same(a, b):
    """Bit-for-bit equality (float64 or float32, any memory order)."""
    a, b = np.asarray(a), np.asarray(b)
    if a.shape != b.shape or a.dtype != b.dtype:
        return False
    u = {4: np.uint32, 8: np.uint64}[a.dtype.itemsize]
    return np.array_equal(a.ravel(order="F").view(u), b.ravel(order="F").view(u))


def _case(shape=(12, 10, 9), seed=7):
    g = stretched_grid(*shape)
    bc = bc_codes()
    ref = OracleRun(g, dt=PRM["dt"], nu=PRM["nu"], Cs=PRM["Cs"], omega=1.5, zeta=PRM["zeta"],
                    itmax=PRM["itmax"], LES=1, scheme=0, sor=1, bc=bc)
    f = ref.f
    f["u"][...], f["v"][...], f["w"][...] = noise_ic(g, 1e-3, seed=seed)
    f["ETA"][4:8, 4:7, 5:8] = 1.0
    f["ETA"][3, 4:7, 5:8] = 0.5
    ref.apply_bc()
    s = ProjectionSolver(g, bc)
    s.configure(**PRM)
    s.upload(f["u"], f["v"], f["w"], f["p"], f["ETA"])
    return g, ref, s


def _download(s, g):
    out = {n: g.halo() for n in ("u", "v", "w", "p")}
    out.update({n: g.interior() for n in ("pre", "uc", "vc", "wc")})
    s.download(**out)
    return out


def test_output_q_matches_oracle(tmp_path, hip_lib):
    g, ref, s = _case()
    try:
        o = Oracle(g)
        f = ref.f
        # before any step: pre, uc, vc, wc are the zero module arrays
        q0 = s.output_q(tmp_path / "0000.q")
        assert same(q0, o.pack_q(f["pre"], f["uc"], f["vc"], f["wc"], f["ETA"]))
        for _ in range(3):
            ref.step()
        s.step(3)
        q1 = s.output_q(tmp_path / "0001.q")
        want = o.pack_q(f["pre"], f["uc"], f["vc"], f["wc"], f["ETA"])
        assert same(q1, want)
        assert np.any(q1[..., 1] != 0)
        o.write_q(tmp_path / "ref.q", want)
        assert (tmp_path / "0001.q").read_bytes() == (tmp_path / "ref.q").read_bytes()
        # position='append': a second write to the same name appends
        s.output_q(tmp_path / "0001.q")
        o.write_q(tmp_path / "ref.q", want)
        assert (tmp_path / "0001.q").read_bytes() == (tmp_path / "ref.q").read_bytes()
        assert same(s.read_q(tmp_path / "0000.q"), q0)
    finally:
        s.close()


def test_reading_variables_matches_oracle(tmp_path, hip_lib):
    g, ref, s = _case(seed=9)
    try:
        o = Oracle(g)
        f = ref.f
        rng = np.random.default_rng(4)
        q = np.asfortranarray(rng.standard_normal((g.nx, g.ny, g.nz, 5)).astype(np.float32))
        o.write_q(tmp_path / "in.q", q)
        qr = s.read_q(tmp_path / "in.q")
        assert same(qr, q)
        s.reading_variables(qr)
        o.reading_variables(q, f["u"], f["v"], f["w"], f["p"], f["pre"], f["uc"], f["vc"],
                            f["wc"])
        ref.apply_bc()
        out = _download(s, g)
        for n, a in out.items():
            assert same(a, f[n]), n
        # and the steps that follow
        rows = [ref.step() for _ in range(2)]
        ik, pm, cd, cl, _ = s.step(2)
        assert list(ik) == [r[0] for r in rows]
        out = _download(s, g)
        for n, a in out.items():
            assert same(a, f[n]), n
    finally:
        s.close()


def test_read_q_refusals(tmp_path, hip_lib):
    g, ref, s = _case()
    try:
        other = stretched_grid(8, 8, 8)
        o = Oracle(other)
        o.write_q(tmp_path / "other.q", np.zeros((8, 8, 8, 5), np.float32, order="F"))
        with pytest.raises(NtustError, match="grid"):
            s.read_q(tmp_path / "other.q")
        Oracle(g).write_q(tmp_path / "ok.q", np.zeros((g.nx, g.ny, g.nz, 5), np.float32,
                                                      order="F"))
        raw = (tmp_path / "ok.q").read_bytes()
        (tmp_path / "cut.q").write_bytes(raw[:-100])
        with pytest.raises(NtustError, match=str(E_IO)):
            s.read_q(tmp_path / "cut.q")
        with pytest.raises(NtustError, match=str(E_IO)):
            s.read_q(tmp_path / "missing.q")
    finally:
        s.close()


def test_mesh_x_matches_oracle(tmp_path, hip_lib):
    g, ref, s = _case()
    try:
        Xs, Ys, Zs = g.centres()
        s.write_mesh_x(tmp_path / "mesh.x", Xs, Ys, Zs)
        Oracle(g).write_mesh_x(tmp_path / "ref.x", Xs, Ys, Zs)
        assert (tmp_path / "mesh.x").read_bytes() == (tmp_path / "ref.x").read_bytes()
    finally:
        s.close()


def test_check_steady_sequence(tmp_path, hip_lib):
    g, ref, s = _case()
    try:
        o = Oracle(g)
        last = g.interior()
        isto = 2
        for istep in range(isto, 4 * isto + 1, isto):
            for _ in range(isto):
                ref.step()
            s.step(isto)
            want = o.check_steady(istep, isto, ref.f["wc"], last, 1024.0)
            got = s.check_steady(istep, isto, 1024.0)
            assert same(got, want), istep
        assert want != 1024.0 and want > 0.0
        # a NaN in wc (through reading_variables) never wins the maximum
        rng = np.random.default_rng(2)
        q = np.asfortranarray(rng.standard_normal((g.nx, g.ny, g.nz, 5)).astype(np.float32))
        q[3, 4, 5, 3] = np.nan
        s.reading_variables(q)
        f = ref.f
        o.reading_variables(q, f["u"], f["v"], f["w"], f["p"], f["pre"], f["uc"], f["vc"],
                            f["wc"])
        ref.apply_bc()
        assert np.isnan(f["wc"]).any()
        want = o.check_steady(99, isto, f["wc"], last, 1024.0)
        got = s.check_steady(99, isto, 1024.0)
        assert np.isfinite(want) and same(got, want)
    finally:
        s.close()


def test_restart_continues_bit_for_bit(tmp_path, hip_lib):
    g, ref, s = _case(shape=(14, 12, 10), seed=5)
    try:
        s.step(2)
        s.check_steady(2, 2, 1024.0)
        s.restart_write(tmp_path / "r.bin", 2, 2 * PRM["dt"])
        ik_a, pm_a, cd_a, cl_a, _ = s.step(3)
        vd_a = s.check_steady(5, 2, 1024.0)
        out_a = _download(s, g)
        q_a = s.output_q()
    finally:
        s.close()
    s = ProjectionSolver(g, bc_codes())
    try:
        s.configure(**PRM)
        istep, t = s.restart_read(tmp_path / "r.bin")
        assert istep == 2 and t == 2 * PRM["dt"]
        ik_b, pm_b, cd_b, cl_b, _ = s.step(3)
        vd_b = s.check_steady(5, 2, 1024.0)
        out_b = _download(s, g)
        assert list(ik_a) == list(ik_b)
        assert same(pm_a, pm_b) and same(cd_a, cd_b) and same(cl_a, cl_b)
        assert same(vd_a, vd_b)
        for n in out_a:
            assert same(out_a[n], out_b[n]), n
        assert same(q_a, s.output_q())
    finally:
        s.close()


def test_restart_refusals_keep_state(tmp_path, hip_lib):
    g, ref, s = _case(seed=3)
    try:
        s.step(1)
        s.restart_write(tmp_path / "r.bin", 1, 0.5)
        before = _download(s, g)
        raw = bytearray((tmp_path / "r.bin").read_bytes())
        raw[64 + 8 * 1000] ^= 0x10                 # one payload bit
        (tmp_path / "bad.bin").write_bytes(bytes(raw))
        with pytest.raises(NtustError, match="hash"):
            s.restart_read(tmp_path / "bad.bin")
        (tmp_path / "cut.bin").write_bytes(bytes(raw[:len(raw) // 2]))
        with pytest.raises(NtustError, match=str(E_IO)):
            s.restart_read(tmp_path / "cut.bin")
        after = _download(s, g)
        for n in before:
            assert same(before[n], after[n]), n
    finally:
        s.close()
    s = ProjectionSolver(stretched_grid(10, 12, 14), bc_codes())
    try:
        with pytest.raises(NtustError, match="grid"):
            s.restart_read(tmp_path / "r.bin")
    finally:
        s.close()
