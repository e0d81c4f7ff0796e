"""Build recipes: the HIP library (gfx950), the Fortran drop-in layer and the
test oracle.  Every artefact is written in-tree, next to its sources, so it
travels with the repository snapshot.

    python -m ntustcfdlab_amd.build            # all three
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
FSRC = PKG / "fortran"
LIB = PKG / "libntust_hip.so"
ORACLE_DIR = ROOT / "oracle"

HIP_SOURCES = ["k_les.hip", "k_predict.hip", "k_sor.hip", "k_sor_lex.hip", "k_sor_mixed.hip",
               "k_project.hip", "k_fast.hip", "k_marker.hip", "k_output.hip", "k_ray.hip", "k_krylov.hip", "k_bc.hip",
               "k_adjoint.hip", "ntust_api.hip"]

# -ffp-contract=off: no fused multiply-add unless the source writes fma(),
# which keeps the exact-mode kernels bit-identical to the CPU restatement.
# -packed-fp32-ops: no packed FP32 instructions in generated code.
HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-Wall", "-Wno-unused-function"]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def _flang() -> str | None:
    for cand in ("/opt/rocm/lib/llvm/bin/flang", shutil.which("flang"), shutil.which("amdflang")):
        if cand and Path(cand).exists():
            return cand
    return None


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed: %s\n%s" % (" ".join(map(str, cmd)), r.stdout))
    return r.stdout


def build_hip(verbose: bool = False) -> Path:
    """Compile every kernel file and the C ABI into libntust_hip.so."""
    hipcc = _hipcc()
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    objs = [str(objdir / (Path(src).stem + ".o")) for src in HIP_SOURCES]
    # one hipcc per translation unit, a few at a time
    jobs = max(1, min(len(HIP_SOURCES), os.cpu_count() or 1, 8))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        outs = list(ex.map(lambda so: _run([hipcc, *HIPCC_FLAGS, "-c", str(CSRC / so[0]),
                                            "-o", so[1]]),
                           zip(HIP_SOURCES, objs)))
    if verbose:
        for out in outs:
            if out.strip():
                print(out)
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *objs])
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> Path:
    """The CPU restatement (test infrastructure only)."""
    _run(["make", "-s", "-C", str(ORACLE_DIR)])
    return ORACLE_DIR / "libntust_oracle.so"


def build_fortran() -> Path | None:
    """The iso_c_binding layer and the Fortran driver, linked to the HIP library."""
    flang = _flang()
    if flang is None:
        return None
    out = FSRC / "ntust_driver"
    moddir = PKG / "_build" / "fmod"
    moddir.mkdir(parents=True, exist_ok=True)
    srcs = [FSRC / f for f in ("ntust_hip.f90", "variables.f90", "hotpath.f90",
                                "airfoil_gpu.f90", "driver.f90")]
    if not all(s.exists() for s in srcs):
        return None
    _run([flang, "-O2", "-module-dir", str(moddir), *map(str, srcs), "-o", str(out),
          "-L" + str(PKG), "-lntust_hip", "-Wl,-rpath,$ORIGIN/.."])
    return out


def build_all(verbose: bool = False) -> None:
    build_hip(verbose)
    build_oracle()
    build_fortran()


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print("built", LIB)
