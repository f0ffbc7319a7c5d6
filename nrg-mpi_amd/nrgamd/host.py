"""ctypes driver of libnrg_host.so (the Fortran host of the NRG loop).

The Fortran host owns the Wilson-chain loop, the (Q,Sz) block bookkeeping and
the truncation; it calls libnrg_gpu.so for the numerical bodies.  This module
only forwards calls and copies results into numpy arrays.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

from . import gpu as _gpu

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG, "lib", "libnrg_host.so")

_lib = None
_i32p = ctypes.POINTER(ctypes.c_int32)
_f64p = ctypes.POINTER(ctypes.c_double)

SIGNATURES = {
    "nrg_host_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int]),
    "nrg_host_open": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "nrg_host_iterate": (ctypes.c_int, [ctypes.c_int]),
    "nrg_host_prepare": (ctypes.c_int, [ctypes.c_int]),
    "nrg_host_set_basis_output": (ctypes.c_int, [ctypes.c_int, _i32p, _f64p]),
    "nrg_host_truncate": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _f64p]),
    "nrg_host_info": (ctypes.c_int, [_i32p]),
    "nrg_host_get_eigenvalues": (ctypes.c_int, [_f64p, ctypes.c_int]),
    "nrg_host_get_ints": (ctypes.c_int, [ctypes.c_int, _i32p, ctypes.c_int]),
    "nrg_host_get_eigenvectors": (ctypes.c_int, [_f64p, ctypes.c_int64]),
    "nrg_host_get_times": (ctypes.c_int, [_f64p]),
    "nrg_host_close": (ctypes.c_int, []),
    "nrg_host_set_output": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int]),
    "nrg_host_set_restart": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p,
                                            ctypes.c_int]),
    "nrg_host_spectrum": (ctypes.c_int, [ctypes.c_int]),
    "nrg_host_get_bins": (ctypes.c_int, [_f64p, _f64p, ctypes.c_int]),
    "nrg_host_get_rho_traces": (ctypes.c_int, [_f64p, ctypes.c_int]),
    "nrg_host_get_spectrum_times": (ctypes.c_int, [_f64p]),
    "nrg_host_sort_keys": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_int64), _i32p, ctypes.c_int]),
    "nrg_host_sort_values": (ctypes.c_int, [ctypes.c_int, _f64p, _i32p]),
}


def sort_keys(keys, mode=0):
    """The host's basis sort on int64 keys: (sorted keys, 0-based source
    positions).  mode 0: O(N log N) replay, 1: swap-by-swap exchange sort."""
    k = np.ascontiguousarray(keys, dtype=np.int64).copy()
    p = np.zeros(len(k), dtype=np.int32)
    load().nrg_host_sort_keys(len(k), k.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), p.ctypes.data_as(_i32p),
                              mode)
    return k, p.astype(np.int64) - 1


def sort_values(values):
    """The truncation's eigenvalue sort on float64 values: (sorted values,
    0-based source positions), ties in input order."""
    v = np.ascontiguousarray(values, dtype=np.float64).copy()
    p = np.zeros(len(v), dtype=np.int32)
    load().nrg_host_sort_values(len(v), v.ctypes.data_as(_f64p), p.ctypes.data_as(_i32p))
    return v, p.astype(np.int64) - 1


class HostError(RuntimeError):
    pass


def load(path: Optional[str] = None):
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError(f"{p} not built (run __graft_entry__.build() or make -C nrg-mpi_amd)")
    # the host library links libnrg_gpu.so from its own directory
    _gpu.load()
    lib = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _check(rc, what):
    if rc != 0:
        msg = ""
        try:
            msg = _gpu.load().nrg_gpu_last_error().decode()
        except Exception:  # pragma: no cover
            pass
        raise HostError(f"{what} failed with status {rc}: {msg}")


INFO_KEYS = ["iteration", "nsub", "nfull", "nkept", "start_trunc", "nold", "i_first", "n_iter",
             "spectrum", "nv", "nops", "ncm", "nmat", "nkeep", "flag_trunc", "nchain"]


class Host:
    """One NRG run driven by the Fortran host."""

    def __init__(self, inputs_dir: str, device: Optional[int] = 0, sort_mode: int = 0, keep_vectors: bool = False):
        self.lib = load()
        d = inputs_dir.encode()
        if device is None:
            _check(self.lib.nrg_host_read(d, len(d)), "nrg_host_read")
        else:
            _check(self.lib.nrg_host_open(d, len(d), device, sort_mode, int(keep_vectors)), "nrg_host_open")
        self.info()

    def info(self):
        a = np.zeros(16, dtype=np.int32)
        self.lib.nrg_host_info(a.ctypes.data_as(_i32p))
        self.i = dict(zip(INFO_KEYS, a.tolist()))
        return self.i

    def iterate(self, it: int):
        _check(self.lib.nrg_host_iterate(it), f"nrg_host_iterate({it})")
        return self.info()

    def iterations(self):
        return range(self.i["i_first"], self.i["n_iter"] + 1)

    def eigenvalues(self):
        n = self.info()["nfull"]
        e = np.zeros(n)
        self.lib.nrg_host_get_eigenvalues(e.ctypes.data_as(_f64p), n)
        return e

    def _ints(self, which, rows, cols):
        a = np.zeros(rows * cols, dtype=np.int32)
        self.lib.nrg_host_get_ints(which, a.ctypes.data_as(_i32p), a.size)
        return a.reshape(cols, rows) if rows > 1 else a

    def subspace(self):
        i = self.info()
        return self._ints(1, i["nv"] + 3, i["nsub"])

    def basis_input(self):
        i = self.info()
        return self._ints(2, i["nv"] + 3, i["nfull"])

    def basis_output(self):
        i = self.info()
        return self._ints(3, i["nv"] + 3, i["nfull"])

    def kept_per_block(self):
        i = self.info()
        return self._ints(4, 1, i["nsub"])

    def set_basis_output(self, basis, e):
        basis = np.ascontiguousarray(basis, dtype=np.int32)
        e = np.ascontiguousarray(e, dtype=np.float64)
        self.lib.nrg_host_set_basis_output(len(e), basis.ctypes.data_as(_i32p), e.ctypes.data_as(_f64p))

    def prepare(self, sort_mode=0):
        _check(self.lib.nrg_host_prepare(sort_mode), "nrg_host_prepare")
        return self.info()

    def truncate(self, it, e_block_order):
        e = np.ascontiguousarray(e_block_order, dtype=np.float64)
        return self.lib.nrg_host_truncate(it, len(e), e.ctypes.data_as(_f64p))

    def times(self):
        t = np.zeros(9)
        self.lib.nrg_host_get_times(t.ctypes.data_as(_f64p))
        keys = ["prepare", "diagonalization", "postprocess", "invariant", "iteration", "Xn", "Sn", "Cn", "trunc_margin"]
        return dict(zip(keys, t.tolist()))

    def run_forward(self):
        for it in self.iterations():
            self.iterate(it)

    def spectrum(self, keep_history: bool = False):
        """CFS spectrum phase after the last iteration: rho sweep + spikes.
        keep_history leaves rho and the products on the device (for
        nrgamd.gpu.rho_get); otherwise the history is freed."""
        _check(self.lib.nrg_host_spectrum(int(keep_history)), "nrg_host_spectrum")

    def bins(self):
        """(negative, positive) bin weights, each [nmat, numberOfBin]."""
        i = self.info()
        cap = 64 * 1024 * max(i["nmat"], 1)
        neg, pos = np.zeros(cap), np.zeros(cap)
        nbin = self.lib.nrg_host_get_bins(neg.ctypes.data_as(_f64p), pos.ctypes.data_as(_f64p), cap)
        n = nbin * i["nmat"]
        return neg[:n].reshape(i["nmat"], nbin), pos[:n].reshape(i["nmat"], nbin)

    def rho_traces(self):
        """{iteration: Tr rho} for the swept shells and the last iteration."""
        i = self.info()
        n = i["n_iter"] - i["i_first"] + 1
        t = np.zeros(n)
        self.lib.nrg_host_get_rho_traces(t.ctypes.data_as(_f64p), n)
        return {i["i_first"] + j: t[j] for j in range(n) if i["start_trunc"] <= i["i_first"] + j}

    def spectrum_times(self):
        t = np.zeros(3)
        self.lib.nrg_host_get_spectrum_times(t.ctypes.data_as(_f64p))
        return dict(zip(["stock", "rho", "cfs"], t.tolist()))

    def set_output(self, directory: Optional[str]):
        """Write the reference's output files (Xn, Sn, Cn, eigenodd.txt,
        eigeneven.txt, keepingState.txt, error.log, and after the spectrum
        phase CFS-spectrum-* and the zero-temperature step_* files) into
        `directory` from now on; None stops writing them."""
        d = (directory or "").encode()
        _check(self.lib.nrg_host_set_output(d, len(d)), "nrg_host_set_output")

    def set_restart(self, save: bool = False, load: bool = False, stop_at: Optional[int] = None,
                    path: Optional[str] = None):
        """Restart of the CFS phase (saveRestart / loadRestart): save the
        file before every CFS step, or resume from it; stop_at ends the
        spectrum phase right after the save at that iteration."""
        p = (path or "").encode()
        _check(self.lib.nrg_host_set_restart(int(save), int(load), -1 if stop_at is None else int(stop_at), p,
                                             len(p)), "nrg_host_set_restart")

    def close(self):
        return self.lib.nrg_host_close()
