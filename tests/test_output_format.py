"""Pins of the oracle's output and restart-reader restatements against the
Fortran runtime itself: a small Fortran program (tests/fortran/qio_probe.f90)
writes the solution and grid files with unformatted sequential I/O, as
filer.f90:173-182 and gridder.f90:368-378 do, and evaluates the
reading_variables expressions (reading_data.f90:60-80) with flang.  The
oracle's files must equal the runtime's byte for byte, and its restart
arithmetic must equal flang's bit for bit (a double-precision face average,
the plausible mistake, differs in most cells).  Also pins check_steady's
NaN rule.  CPU only; skipped where flang is absent."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from ntustcfdlab_amd.build import _flang
from ntustcfdlab_amd.grid import Grid, tanh_nodes
from oracle import Oracle

PROBE = Path(__file__).resolve().parent / "fortran" / "qio_probe.f90"


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    fc = _flang()
    if fc is None:
        pytest.skip("flang not available")
    d = tmp_path_factory.mktemp("qio")
    exe = d / "qio_probe"
    r = subprocess.run([fc, "-O2", "-module-dir", str(d), str(PROBE), "-o", str(exe)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return exe


def _grid():
    n = (9, 7, 6)
    return Grid(*n, X=tanh_nodes(n[0], 1.0, 0.5, 1.3), Y=tanh_nodes(n[1], 1.0, 0.4, 1.1))


def _inputs(g, seed):
    rng = np.random.default_rng(seed)
    q = (rng.standard_normal((g.nx, g.ny, g.nz, 5)) * 10.0 ** rng.integers(-6, 3, (g.nx, g.ny, g.nz, 5)))
    q = np.asfortranarray(q.astype(np.float32))
    u, v, w = (np.asfortranarray(rng.standard_normal(g.halo_shape)) for _ in range(3))
    return q, u, v, w


def test_q_and_mesh_files_match_fortran_runtime(probe, tmp_path):
    g = _grid()
    q, u, v, w = _inputs(g, 5)
    Xs, Ys, Zs = g.centres()
    fin = tmp_path / "in.bin"
    fin.write_bytes(b"".join(a.tobytes(order="F") for a in (q, Xs, Ys, Zs, u, v, w)))
    r = subprocess.run([str(probe), str(g.nx), str(g.ny), str(g.nz), str(fin),
                        str(tmp_path / "f.q"), str(tmp_path / "f.x"), str(tmp_path / "f.r")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    o = Oracle(g)
    o.write_q(tmp_path / "o.q", q)
    o.write_q(tmp_path / "o.q", q)       # appended, as the probe does
    o.write_mesh_x(tmp_path / "o.x", Xs, Ys, Zs)
    assert (tmp_path / "o.q").read_bytes() == (tmp_path / "f.q").read_bytes()
    assert (tmp_path / "o.x").read_bytes() == (tmp_path / "f.x").read_bytes()
    # the record layout itself: 4-byte lengths around each record
    raw = (tmp_path / "o.x").read_bytes()
    assert np.frombuffer(raw[:12], np.int32).tolist() == [4, 1, 4]
    assert len(raw) == 12 + 20 + 8 + 12 * g.cells


def test_pack_q_rounds_to_nearest(tmp_path):
    g = _grid()
    rng = np.random.default_rng(8)
    f = [np.asfortranarray(rng.standard_normal(g.shape)) for _ in range(4)]
    eta = np.asfortranarray(rng.random(g.halo_shape))
    q = Oracle(g).pack_q(*f, eta)
    for h in range(4):
        assert np.array_equal(q[..., h], f[h].astype(np.float32))
    assert np.array_equal(q[..., 4], eta[2:-2, 2:-2, 2:-2].astype(np.float32))


def test_reading_variables_matches_flang(probe, tmp_path):
    g = _grid()
    q, u, v, w = _inputs(g, 6)
    Xs, Ys, Zs = g.centres()
    fin = tmp_path / "in.bin"
    fin.write_bytes(b"".join(a.tobytes(order="F") for a in (q, Xs, Ys, Zs, u, v, w)))
    r = subprocess.run([str(probe), str(g.nx), str(g.ny), str(g.nz), str(fin),
                        str(tmp_path / "f.q"), str(tmp_path / "f.x"), str(tmp_path / "f.r")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    raw = (tmp_path / "f.r").read_bytes()
    nh, ni = int(np.prod(g.halo_shape)), g.cells
    sizes = [nh] * 3 + [ni] * 4
    parts, off = [], 0
    for n in sizes:
        parts.append(np.frombuffer(raw[off:off + 8 * n], np.float64))
        off += 8 * n
    assert off == len(raw)
    ou, ov, ow = (a.copy(order="F") for a in (u, v, w))
    p = g.halo()
    pre, uc, vc, wc = (g.interior() for _ in range(4))
    Oracle(g).reading_variables(q, ou, ov, ow, p, pre, uc, vc, wc)
    for mine, ref in zip((ou, ov, ow), parts[:3]):
        assert np.array_equal(mine.ravel(order="F").view(np.uint64), ref.view(np.uint64))
    assert np.array_equal(p[2:-2, 2:-2, 2:-2].ravel(order="F").view(np.uint64),
                          parts[3].view(np.uint64))
    assert np.array_equal(pre, p[2:-2, 2:-2, 2:-2])
    for mine, ref in zip((uc, vc, wc), parts[4:]):
        assert np.array_equal(mine.ravel(order="F").view(np.uint64), ref.view(np.uint64))
    # the face average is a single-precision one: the double-precision
    # average of the same two values differs in most cells
    dbl = 0.5 * (q[:-1, :-1, :-1, 1].astype(np.float64) + q[1:, :-1, :-1, 1])
    assert np.mean(dbl != ou[2:g.nx + 1, 2:g.ny + 1, 2:g.nz + 1]) > 0.3


def test_check_steady_rule():
    g = _grid()
    rng = np.random.default_rng(3)
    o = Oracle(g)
    wc = np.asfortranarray(rng.standard_normal(g.shape))
    last = g.interior()
    # istep == isto: no comparison, VelocityDifference untouched, last = wc
    assert o.check_steady(100, 100, wc, last, 1024.0) == 1024.0
    assert np.array_equal(last, wc)
    wc2 = wc + np.asfortranarray(rng.standard_normal(g.shape)) * 1e-3
    want = np.max(np.abs(wc2 - wc))
    assert o.check_steady(200, 100, wc2, last, 1024.0) == want
    # a NaN difference never compares greater: it is skipped
    wc3 = wc2.copy(order="F")
    wc3[1, 2, 3] = np.nan
    d = o.check_steady(300, 100, wc3, last, 1024.0)
    mask = np.ones(g.shape, bool)
    mask[1, 2, 3] = False
    assert d == np.max(np.abs(wc3 - wc2)[mask])
