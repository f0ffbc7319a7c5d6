"""The reference-named drop-in bodies (nrg-mpi_amd/fortran/dropin/
nrg_dropin.F90) compile against the reference's own modules: the module
files that oracle/Makefile builds from the reference sources into
oracle/_ref/mod, with the library's interface module nrg_gpu_api.  Compile
only (nothing is linked or run); skipped where that build or amdflang is
absent.  CPU only."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MOD = os.path.join(ROOT, "oracle", "_ref", "mod")
FC = "/opt/rocm/bin/amdflang"


@pytest.mark.skipif(not os.path.isdir(MOD) or not os.path.exists(os.path.join(MOD, "hamiltonian.mod")),
                    reason="reference modules not built (oracle/_ref/mod)")
@pytest.mark.skipif(not os.path.exists(FC), reason="amdflang not found")
def test_dropin_compiles_against_reference_modules(tmp_path):
    api = os.path.join(ROOT, "nrg-mpi_amd", "fortran", "nrg_gpu_api.F90")
    dropin = os.path.join(ROOT, "nrg-mpi_amd", "fortran", "dropin", "nrg_dropin.F90")
    d = str(tmp_path)
    subprocess.run([FC, "-c", "-module-dir", d, api, "-o", os.path.join(d, "api.o")], check=True,
                   capture_output=True, text=True)
    r = subprocess.run([FC, "-c", "-I", MOD, "-I", d, "-module-dir", d, dropin, "-o", os.path.join(d, "dropin.o")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr + r.stdout
    # every replaced routine is defined with the reference's name
    nm = shutil.which("nm")
    if nm:
        syms = subprocess.run([nm, os.path.join(d, "dropin.o")], capture_output=True, text=True, check=True).stdout
        for name in ("diagonalization_", "invariantmatrixdgemm_", "invariantmatrixforspectrumdgemm_",
                     "initializereduceddensity_", "reduceddensitymatrixcalculation_", "spikezerotemperature_",
                     "spikecfs_", "gpusetup_", "gpustock_", "gpuspectrumbins_"):
            assert f" T {name}" in syms, name
