"""Reader for the binary dumps written by oracle/ref_driver.F90.

Test infrastructure: used only by make_fixtures.py to turn a reference run
into compact golden fixtures.
"""
import glob
import os

import numpy as np


class _Stream:
    def __init__(self, path):
        with open(path, "rb") as f:
            self.b = f.read()
        self.p = 0

    def take(self, dtype, n):
        dt = np.dtype(dtype)
        a = np.frombuffer(self.b, dtype=dt, count=n, offset=self.p)
        self.p += dt.itemsize * n
        return a.copy()

    def i4(self, n=1):
        a = self.take("<i4", n)
        return int(a[0]) if n == 1 else a

    def eof(self):
        return self.p >= len(self.b)


def read_iteration(path):
    s = _Stream(path)
    it, nv, nsub, nfull, nkept, start_trunc, ncm = (int(x) for x in s.take("<i4", 7))
    d = dict(it=it, nv=nv, nsub=nsub, nfull=nfull, nkept=nkept, start_trunc=start_trunc, ncm=ncm)
    d["subspace"] = s.take("<i4", nsub * (nv + 3)).reshape(nsub, nv + 3)
    d["basis_input"] = s.take("<i4", nfull * (nv + 3)).reshape(nfull, nv + 3)
    d["basis_output"] = s.take("<i4", nfull * (nv + 3)).reshape(nfull, nv + 3)
    d["eigenvalue"] = s.take("<f8", nfull)
    full = s.i4()
    if full:
        sz = int(s.take("<i8", 1)[0])
        d["eigenvector"] = s.take("<f8", sz)
        k = s.i4()
        if k:
            d["invariant_matrix"] = s.take("<f8", k * k * ncm).reshape(ncm, k, k)  # [t, l, r] (Fortran (r,l,t))
            k2 = s.i4()
            if k2:
                rest = (len(s.b) - s.p) // 8
                nmat = rest // (k2 * k2)
                d["invariant_matrix_spectrum"] = s.take("<f8", k2 * k2 * nmat).reshape(nmat, k2, k2)
        else:
            s.i4()
    return d


def read_spectrum(path, nv):
    s = _Stream(path)
    nbin, nmat, start_trunc, nfullsub, size_rho, nv2 = (int(x) for x in s.take("<i4", 6))
    d = dict(nbin=nbin, nmat=nmat, start_trunc=start_trunc, nfullsub=nfullsub, size_rho=size_rho)
    d["bins_pos"] = s.take("<f8", nbin * nmat).reshape(nmat, nbin)
    d["bins_neg"] = s.take("<f8", nbin * nmat).reshape(nmat, nbin)
    return d, s


def read_run(rundir):
    its = [read_iteration(p) for p in sorted(glob.glob(os.path.join(rundir, "dump_it*.bin")))]
    spec = None
    sp = os.path.join(rundir, "dump_spectrum.bin")
    if os.path.exists(sp):
        nv = its[0]["nv"]
        spec, s = read_spectrum(sp, nv)
        N = its[-1]["it"]
        spec["map_full"] = s.take("<i4", N - spec["start_trunc"] + 1)
        spec["full_subspace"] = s.take("<i4", spec["nfullsub"] * (nv + 6)).reshape(spec["nfullsub"], nv + 6)
        spec["rho"] = s.take("<f8", spec["size_rho"])
    return its, spec
