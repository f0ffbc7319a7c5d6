"""The C-ABI libraries load and export every entry point their headers and
bindings declare, and the error path works without a device (no compute
call is made).  CPU only."""
import ctypes
import os
import re

import pytest

from nrgamd import gpu, host

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nrg_gpu.h")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(nrg_gpu_\w+)\s*\(", txt)))


def test_header_declares_the_boundary():
    names = declared()
    for must in ("nrg_gpu_init", "nrg_gpu_diagonalize", "nrg_gpu_invariant", "nrg_gpu_invariant_spectrum",
                 "nrg_gpu_stock", "nrg_gpu_rho_init", "nrg_gpu_rho_shell", "nrg_gpu_spike_cfs",
                 "nrg_gpu_spike_zero_temperature", "nrg_gpu_spectrum_bins", "nrg_gpu_finalize"):
        assert must in names


def test_gpu_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(gpu.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers the whole header
    assert set(declared()) == set(gpu.SIGNATURES)


def test_host_library_exports_its_driver_surface():
    lib = host.load()
    for n in host.SIGNATURES:
        assert hasattr(lib, n), n


def test_calls_before_init_fail_cleanly():
    lib = gpu.load()
    rc = lib.nrg_gpu_invariant(None, None)
    assert rc == 3
    assert b"nrg_gpu_init" in lib.nrg_gpu_last_error()
    rc = lib.nrg_gpu_stock(0, None, 0)
    assert rc != 0
    # a clean call clears the message
    assert lib.nrg_gpu_stage_times((ctypes.c_double * 10)(), 10) == 0
    assert lib.nrg_gpu_last_error() == b""


@pytest.mark.gpu
def test_injected_fault_is_reported(tmp_path):
    """An error inside the device path reaches the Fortran host and the
    caller as a status with the library's message."""
    from nrgamd import models
    d = str(tmp_path / "siam")
    models.write_inputs(models.workload("siam_nk256_it40", niter=6, spectrum=False), d)
    h = host.Host(d, device=0)
    try:
        h.iterate(-1)
        assert gpu.load().nrg_gpu_inject_fault(0) == 0
        with pytest.raises(host.HostError, match="injected fault"):
            h.iterate(0)
    finally:
        h.close()
