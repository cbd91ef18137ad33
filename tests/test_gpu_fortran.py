"""The Fortran host path: a Fortran 2008 program calls maketree ->
densityiterate -> cons2prim -> force through module phsph (iso_c_binding)
and checks the reference's lattice answers (57 neighbours, two density
evaluations per particle, gradh = 1.01948); a second one runs whole
leapfrog steps on the reference's step test (test_step.F90:63-160)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "phantom_amd", "fortran", "shim_driver")

pytestmark = pytest.mark.gpu


def test_fortran_driver(gpu):
    if not os.path.exists(EXE):
        pytest.skip("Fortran driver not built (amdflang absent at build time)")
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "shim_driver: ok" in r.stdout


STEP_EXE = os.path.join(ROOT, "phantom_amd", "fortran", "step_driver")


def test_fortran_step_driver(gpu):
    """Whole leapfrog steps from Fortran (the control flow of step,
    step_leapfrog.f90:91-760) on the reference's step test lattice."""
    if not os.path.exists(STEP_EXE):
        pytest.skip("Fortran driver not built (amdflang absent at build time)")
    r = subprocess.run([STEP_EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "step_driver: ok" in r.stdout
