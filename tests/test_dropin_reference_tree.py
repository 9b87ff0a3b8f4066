"""The drop-in linked into the reference's own source tree (the maintainer's
route): the reference's MODULE global is replaced by
ib3d-gpu_amd/fortran/ib3d_global.f90, the reference files whose routines the
drop-in provides are left out, and every remaining reference file
(main.f90, subReadInput.f90, subWriteOutput.f90, subCalculateuc.f90) is
compiled unchanged against the drop-in's module with amdflang and linked
with the drop-in and libib3d_hip.

Compiled where the reference's sources lie; skipped where they are not
mounted.  Without a GPU the linked program must run the reference's own
readers and then stop in allocateArrays with the library's error: the HIP
path has no CPU fallback."""
import os
import shutil
import subprocess

import pytest

from ib3d_amd import cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ib3d-gpu_amd")
REF = "/root/reference"
FC = "/opt/rocm/bin/amdflang"
FFLAGS = ["-O2", "-fdefault-real-8", "-fdefault-double-8", "-ffp-contract=off"]

# reference files replaced as a whole by the drop-in (every routine they hold
# is one of its routines, or, for nsMomentumAB in subNavierStokes.f90, never
# called by main.f90)
REPLACED = ["modGlobal.f90", "subAllocateArrays.f90", "subBoundaryConditions.f90", "subCoefficientMatrix.f90",
            "subForcing.f90", "subInitialConditions.f90", "subLastConditions.f90", "subNavierStokes.f90",
            "subPcorVcor.f90", "subSearch.f90", "subStressCal1.f90", "subStressCal2.f90"]
KEPT = ["subReadInput.f90", "subWriteOutput.f90", "subCalculateuc.f90", "main.f90"]
DROPIN = ["ib3d_global.f90", "ib3d_cabi.f90", "ib3d_dropin.f90"]

pytestmark = pytest.mark.skipif(not (os.path.isdir(REF) and os.path.exists(FC)),
                                reason="the reference's sources or amdflang are not available here")


def test_reference_files_cover_the_tree():
    """Every .f90 file of the reference is either replaced or kept."""
    ref = sorted(f for f in os.listdir(REF) if f.endswith(".f90"))
    assert sorted(REPLACED + KEPT) == ref


def test_remaining_reference_files_compile_and_link_with_the_dropin(tmp_path):
    lib = os.path.join(PKG, "lib", "libib3d_hip.so")
    assert os.path.exists(lib), "build the library first"
    mod = tmp_path / "mod"
    mod.mkdir()
    objs = []

    def fc(src, name):
        obj = str(mod / (name + ".o"))
        r = subprocess.run([FC] + FFLAGS + ["-module-dir", str(mod), "-I", str(mod), "-c", src, "-o", obj],
                           capture_output=True, text=True)
        assert r.returncode == 0, f"{src} does not compile against the drop-in module:\n{r.stderr[-3000:]}"
        objs.append(obj)

    for f in DROPIN:
        fc(os.path.join(PKG, "fortran", f), f[:-4])
    for f in KEPT:
        fc(os.path.join(REF, f), "ref_" + f[:-4])
    exe = str(tmp_path / "ib3d_ref_tree")
    r = subprocess.run([FC] + FFLAGS + ["-o", exe] + objs +
                       ["-L" + os.path.dirname(lib), "-lib3d_hip", "-Wl,-rpath," + os.path.dirname(lib)],
                       capture_output=True, text=True)
    assert r.returncode == 0, f"link failed:\n{r.stderr[-3000:]}"

    # no GPU: the reference's readers run, then allocateArrays stops with the
    # library's message
    case = tmp_path / "case"
    cases.write_case(cases.CONFIGS["tiny"], str(case))
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1")
    r = subprocess.run([exe], cwd=str(case), capture_output=True, text=True, env=env, timeout=120)
    out = r.stdout + r.stderr
    assert r.returncode != 0
    assert "SURFACE MESH READING COMPLETE" in out
    assert "ib3d: ib3d_init" in out, out[-2000:]
