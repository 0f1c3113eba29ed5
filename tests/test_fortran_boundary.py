"""The drop-in boundary as compiled code: the binding module
spral_amd/fortran/spral_ssids_gpu_subtree.f90 replaces the reference's
src/ssids/gpu/subtree.f90, the two patches (fkeep.F90 enquire/alter branches
for gpu_numeric_subtree; anal.F90 options%gpu_only) apply, and the
reference's whole Fortran host (src/ssids/ssids.f90 and every module it
uses) compiles against them with amdflang (oracle/Makefile target
`boundary`).  Runs where the reference tree is mounted."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
REF = Path("/root/reference")
BDIR = ROOT / "oracle" / "_ref" / "boundary"

needs_ref = pytest.mark.skipif(
    not (REF / "src" / "ssids" / "subtree.f90").exists() or shutil.which("amdflang") is None,
    reason="reference tree or amdflang absent")


@needs_ref
@pytest.mark.parametrize("patch,target", [("fkeep_gpu_branches.patch", "src/ssids/fkeep.F90"),
                                          ("anal_gpu_only.patch", "src/ssids/anal.F90")])
def test_patches_apply(patch, target, tmp_path):
    out = tmp_path / "out"
    r = subprocess.run(["patch", "-s", "-o", str(out), str(REF / target),
                        str(ROOT / "spral_amd" / "fortran" / "patches" / patch)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    text = out.read_text()
    if "fkeep" in patch:
        assert text.count("type is (gpu_numeric_subtree)") == 3
        assert "use spral_ssids_gpu_subtree, only : gpu_numeric_subtree" in text
    else:
        assert "options%gpu_only" in text


@needs_ref
def test_reference_host_compiles_with_binding(lib):
    r = subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "boundary"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert (BDIR / "host_compiled.ok").exists()
    for obj in ["ssids_ssids.f90.o", "PATCHED_fkeep.F90.o", "PATCHED_anal.F90.o",
                "ssids_contrib_free.f90.o", "gpu_subtree.o"]:
        assert (BDIR / obj).exists(), obj
    assert (BDIR / "hip_subtree_driver").exists()
    # the driver binds nothing but the C-ABI and the reference's modules
    nm = subprocess.run(["nm", "-u", str(BDIR / "hip_subtree_driver")],
                        capture_output=True, text=True).stdout
    hip = sorted({ln.split()[-1] for ln in nm.splitlines() if "spral_ssids_hip_" in ln})
    assert "spral_ssids_hip_factor" in hip and "spral_ssids_hip_solve_diag_bwd" in hip
