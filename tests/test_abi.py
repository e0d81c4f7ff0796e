"""The C ABI: the library loads without a GPU and exports every function
include/ntust.h declares; the Python mirror binds them; calls fail loudly
when no device is present instead of falling back to anything."""
import ctypes as ct
import re
from pathlib import Path

import pytest

from conftest import gpu_available

ROOT = Path(__file__).resolve().parents[1]


def declared():
    text = (ROOT / "include" / "ntust.h").read_text()
    return sorted(set(re.findall(r"\b(ntust_\w+)\s*\(", text)))


def test_header_declares_the_hot_path():
    names = declared()
    for must in ("ntust_init", "ntust_finalize", "ntust_gradient_phi_gauss", "ntust_smagorinsky",
                 "ntust_discretisation_quick_centre", "ntust_discretisation_upwind",
                 "ntust_gauss_seidel", "ntust_calcul_new_velocity", "ntust_updating_velocity",
                 "ntust_virtual_force_integrator", "ntust_boundary_conditions", "ntust_step"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from ntustcfdlab_amd import hip
    lib = hip.load_library()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(gpu_available(), reason="checks the no-device error path")
def test_no_device_raises():
    from ntustcfdlab_amd import hip
    from ntustcfdlab_amd.grid import Grid
    with pytest.raises(hip.NtustError):
        hip.ProjectionSolver(Grid(4, 4, 4))


def test_calls_before_init_report_not_init():
    from ntustcfdlab_amd import hip
    lib = hip.load_library()
    lib.ntust_finalize()
    rc = lib.ntust_step(1, None, None, None, None)
    assert rc == -1
    assert b"ntust_init" in lib.ntust_last_error()


def test_init_rejects_bad_arguments_before_touching_the_device():
    """Argument errors are reported as NTUST_E_ARG with a message, before any
    HIP call (so they behave the same with and without a GPU)."""
    import numpy as np
    from ntustcfdlab_amd import hip
    lib = hip.load_library()
    m = np.ones(16)
    p = m.ctypes.data_as(hip.ct.POINTER(hip.ct.c_double))
    for n in ((1, 4, 4), (4, 0, 4), (4, 4, -3)):
        assert lib.ntust_init(*n, p, p, p, p, p, p) == -2
        assert b"must be >= 2" in lib.ntust_last_error()
    assert lib.ntust_init(4, 4, 4, p, p, None, p, p, p) == -2
    assert b"NULL" in lib.ntust_last_error()


def test_new_entry_points_report_not_init():
    from ntustcfdlab_amd import hip
    lib = hip.load_library()
    lib.ntust_finalize()
    assert lib.ntust_adjoint_objective(1.0, 0.0, 0.0, 0) == -1
    assert lib.ntust_adjoint_eta_gradient(None) == -1
    assert lib.ntust_adjoint_run(2, 1, 1, None, None, None) == -1
    assert b"ntust_init" in lib.ntust_last_error()
    k = hip.ct.c_int()
    assert lib.ntust_adjoint_checkpoint_plan(4, 0.0, hip.ct.byref(k), None) == -1
    assert b"ntust_init" in lib.ntust_last_error()
