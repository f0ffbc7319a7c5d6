"""Restart of the CFS spectrum phase (saveRestart / loadRestart,
module.F90:1723-1890; main.F90:165-221): a run that saves the restart file
before every CFS step and stops after the save at an intermediate
iteration, resumed by a fresh host that skips the NRG loop and the
reduced-density sweep, ends with the bins of an uninterrupted run and of
the reference run (1e-8 of the peak bin).  The uninterrupted run made its
own forward pass, and two forward passes on the GPU agree to rounding only
(the SYMV's LDS atomics add in no fixed order: ~3e-12 of the peak bin after
40 iterations), so that comparison is held to 1e-10 -- a history entry, a
group or a reduced-density block missing from the file would shift bins by
whole spike weights.  Two resumes from the same file agree to the rounding
of the atomic bin sums (1e-14)."""
import os
import struct

import numpy as np
import pytest

from nrgamd import models

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_cfs_restart_resumes_to_the_same_spectrum(tmp_path):
    from nrgamd.host import Host
    fx = np.load(os.path.join(GOLD, "ref_siam_nk256_it40.npz"))
    spec = models.workload("siam_nk256_it40")
    d = str(tmp_path / "in")
    models.write_inputs(spec, d)
    h = Host(d, device=0)
    try:
        h.run_forward()
        h.spectrum()
        neg0, pos0 = h.bins()
        st, N = h.info()["start_trunc"], h.info()["n_iter"]
    finally:
        h.close()
    K = st + (N - st) // 2
    print(f"truncation from {st}, last iteration {N}, interrupted after the save at {K}", flush=True)
    h = Host(d, device=0)
    try:
        h.set_restart(save=True, stop_at=K)
        h.run_forward()
        h.spectrum()
    finally:
        h.close()
    path = os.path.join(d, "restart.dat")
    with open(path, "rb") as f:
        mode, it = struct.unpack("<ii", f.read(8))
    assert (mode, it) == (1002, K)
    import shutil
    keep = os.path.join(d, "restart_K.dat")
    shutil.copyfile(path, keep)
    resumed = []
    for save in (True, False):
        h = Host(d, device=0)
        try:
            h.set_restart(load=True, save=save, path=None if save else keep)
            h.spectrum()                     # no forward pass: everything from the file
            resumed.append(h.bins())
            assert h.info()["start_trunc"] == st
        finally:
            h.close()
    (neg1, pos1), (neg2, pos2) = resumed
    peak = max(np.abs(neg0).max(), np.abs(pos0).max())
    d_self = max(np.abs(neg1 - neg0).max(), np.abs(pos1 - pos0).max()) / peak
    d_twice = max(np.abs(neg1 - neg2).max(), np.abs(pos1 - pos2).max()) / peak
    d_ref = max(np.abs(neg1 - fx["bins_neg"]).max(), np.abs(pos1 - fx["bins_pos"]).max()) / peak
    print(f"resumed vs uninterrupted {d_self:.2e}, two resumes {d_twice:.2e}, vs reference {d_ref:.2e} "
          f"of the peak bin")
    assert d_self <= 1e-10
    assert d_twice <= 1e-14
    assert d_ref <= 1e-8
    with open(path, "rb") as f:          # rewritten after the restart step
        assert struct.unpack("<ii", f.read(8)) == (1002, N)


def test_broken_restart_file_is_rejected(tmp_path):
    from nrgamd.host import Host, HostError
    spec = models.workload("siam_nk256_it40")
    d = str(tmp_path / "in")
    models.write_inputs(spec, d)
    with open(os.path.join(d, "restart.dat"), "wb") as f:
        f.write(struct.pack("<ii", 1001, 5))    # MODE_RESTART_NRG: not a CFS restart
    h = Host(d, device=0)
    try:
        h.set_restart(load=True)
        with pytest.raises(HostError, match="broken restart file"):
            h.spectrum()
    finally:
        h.close()
