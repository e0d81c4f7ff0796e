"""GPU parity of the device-resident time step (ntust_step) against the CPU
restatement stepped in main.f90 order (or_step), bit for bit, over several
steps: state u, v, w, p (whole halo arrays), pre and the centred velocity of
the last step, SOR sweep counts, force coefficients."""
import numpy as np
import pytest

from conftest import stretched_grid
from ntustcfdlab_amd.grid import Grid, bc_codes, noise_ic
from ntustcfdlab_amd.hip import QUICK_CENTRE, SOR_LEXICOGRAPHIC, SOR_REDBLACK, UPWIND, \
    ProjectionSolver
from oracle import OracleRun

pytestmark = pytest.mark.gpu


def same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


CASES = {
    # the reference's own cavity (Re=10, dt = 1.0e-3 widened from single,
    # upwind, LES off) on a small uniform grid
    "cavity-upwind-lex": dict(grid=lambda: Grid(16, 16, 16), scheme=UPWIND, LES=0, sor=0,
                              dt=float(np.float32(1.0e-3)), nu=1.0 / 10.0, Cs=0.18,
                              zeta=float(np.float32(1.0e-4)), itmax=3000.0, steps=6,
                              noise=0.0, body=False),
    "stretched-quick-les-lex": dict(grid=lambda: stretched_grid(12, 10, 9), scheme=QUICK_CENTRE,
                                    LES=1, sor=0, dt=5e-4, nu=1e-3, Cs=0.1, zeta=1e-8,
                                    itmax=400.0, steps=4, noise=1e-3, body=False),
    "stretched-quick-les-rb": dict(grid=lambda: stretched_grid(12, 10, 9), scheme=QUICK_CENTRE,
                                   LES=1, sor=1, dt=5e-4, nu=1e-3, Cs=0.1, zeta=1e-8,
                                   itmax=400.0, steps=4, noise=1e-3, body=False),
    "body-quick-rb": dict(grid=lambda: stretched_grid(14, 12, 16), scheme=QUICK_CENTRE, LES=0,
                          sor=1, dt=1e-3, nu=1e-2, Cs=0.18, zeta=1e-7, itmax=300.0, steps=4,
                          noise=1e-2, body=True),
    "body-upwind-lex": dict(grid=lambda: stretched_grid(10, 9, 12), scheme=UPWIND, LES=0,
                            sor=0, dt=1e-3, nu=1e-2, Cs=0.18, zeta=1e-7, itmax=300.0, steps=3,
                            noise=1e-2, body=True),
}


@pytest.mark.parametrize("name", list(CASES))
def test_resident_steps_match_oracle(name, hip_lib):
    c = CASES[name]
    g = c["grid"]()
    bc = bc_codes()
    ref = OracleRun(g, dt=c["dt"], nu=c["nu"], Cs=c["Cs"], omega=1.5, zeta=c["zeta"],
                    itmax=c["itmax"], LES=c["LES"], scheme=c["scheme"], sor=c["sor"], bc=bc)
    f = ref.f
    if c["noise"]:
        u, v, w = noise_ic(g, c["noise"], seed=7)
        f["u"][...], f["v"][...], f["w"][...] = u, v, w
    if c["body"]:
        eta = f["ETA"]
        eta[4:8, 4:7, 5:9] = 1.0
        eta[3, 4:7, 5:9] = 0.5
    ref.apply_bc()

    s = ProjectionSolver(g, bc)
    try:
        s.configure(dt=c["dt"], nu=c["nu"], Cs=c["Cs"], LES=c["LES"], scheme=c["scheme"],
                    omega=1.5, zeta=c["zeta"], itmax=c["itmax"],
                    sor=SOR_REDBLACK if c["sor"] else SOR_LEXICOGRAPHIC, lx=1.0, force_every=1)
        s.upload(f["u"], f["v"], f["w"], f["p"], f["ETA"])
        rows = [ref.step() for _ in range(c["steps"])]
        ik, pm, cd, cl, _ = s.step(c["steps"])
        assert list(ik) == [r[0] for r in rows]
        assert same(pm, [r[1] for r in rows])
        assert same(cd, [r[2] for r in rows])
        assert same(cl, [r[3] for r in rows])
        out = {n: g.halo() for n in ("u", "v", "w", "p")}
        out.update({n: g.interior() for n in ("pre", "uc", "vc", "wc")})
        s.download(**out)
        for n, a in out.items():
            assert same(a, f[n]), n
    finally:
        s.close()
