"""ctypes binding of libnrg_gpu.so (the C ABI declared in include/nrg_gpu.h).

The library is loaded from the package's in-tree ``lib/`` directory.  There is
no fallback: if the shared library is missing or the device is absent the
calls raise.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG, "lib", "libnrg_gpu.so")

_lib = None

_i32p = ctypes.POINTER(ctypes.c_int32)
_f64p = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)

SIGNATURES = {
    "nrg_gpu_init": (ctypes.c_int, [ctypes.c_int]),
    "nrg_gpu_finalize": (ctypes.c_int, []),
    "nrg_gpu_last_error": (ctypes.c_char_p, []),
    "nrg_gpu_set_model": (ctypes.c_int, [ctypes.c_int] * 5 + [_i32p, _i32p, _f64p, _i32p, _f64p, _i32p,
                                                              ctypes.c_int, _f64p, _i32p]),
    "nrg_gpu_set_bcs": (ctypes.c_int, [ctypes.c_int, _i32p, ctypes.c_int, _f64p]),
    "nrg_gpu_set_invariant": (ctypes.c_int, [ctypes.c_int, _f64p]),
    "nrg_gpu_set_invariant_spectrum": (ctypes.c_int, [ctypes.c_int, _f64p]),
    "nrg_gpu_diagonalize": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _i32p, ctypes.c_int, _i32p, ctypes.c_int,
                                           _f64p, _f64p, _f64p, ctypes.c_double, _f64p, _f64p, ctypes.c_int64]),
    "nrg_gpu_invariant": (ctypes.c_int, [_i32p, _f64p]),
    "nrg_gpu_invariant_spectrum": (ctypes.c_int, [_i32p, _f64p]),
    "nrg_gpu_invariant_get": (ctypes.c_int, [_f64p, ctypes.c_int64]),
    "nrg_gpu_stage_times": (ctypes.c_int, [_f64p, ctypes.c_int]),
    "nrg_gpu_syev_batched": (ctypes.c_int, [ctypes.c_int, _i32p, _f64p, _f64p, _f64p]),
    "nrg_gpu_dgemm": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_double, _f64p, ctypes.c_int, _f64p, ctypes.c_int, ctypes.c_double,
                                     _f64p, ctypes.c_int]),
    "nrg_gpu_inject_fault": (ctypes.c_int, [ctypes.c_int]),
    "nrg_gpu_profile": (ctypes.c_int, [ctypes.c_int]),
    "nrg_gpu_kernel_stats": (ctypes.c_int, [_f64p, ctypes.c_int]),
    "nrg_gpu_profile_sample": (ctypes.c_int, [ctypes.c_int]),
    "nrg_gpu_profile_exclusive": (ctypes.c_int, [ctypes.c_int]),
    "nrg_gpu_set_tridiag_finish": (ctypes.c_int, [ctypes.c_int]),
    "nrg_gpu_launch_log": (ctypes.c_int, [ctypes.c_int]),
    "nrg_gpu_work_counters": (ctypes.c_int, [_f64p, ctypes.c_int, ctypes.c_int]),
    "nrg_gpu_set_tridiag_method": (ctypes.c_int, [ctypes.c_int]),
    "nrg_gpu_set_tridiag_groups": (ctypes.c_int, [ctypes.c_int]),
    "nrg_gpu_set_tridiag_symv": (ctypes.c_int, [ctypes.c_int]),
    "nrg_gpu_set_gemm_split_families": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_double]),
    "nrg_gpu_copy_calibration": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "nrg_gpu_launch_log_get": (ctypes.c_int, [_f64p, ctypes.c_int64, _i64p]),
    "nrg_gpu_launch_log_kernel": (ctypes.c_int, [ctypes.c_int, ctypes.c_char_p, ctypes.c_int]),
    "nrg_gpu_set_spectrum_operators": (ctypes.c_int, [ctypes.c_int, _i32p, _i32p]),
    "nrg_gpu_stock": (ctypes.c_int, [ctypes.c_int, _i32p, ctypes.c_int]),
    "nrg_gpu_rho_init": (ctypes.c_int, [ctypes.c_int]),
    "nrg_gpu_rho_shell": (ctypes.c_int, [ctypes.c_int, _f64p]),
    "nrg_gpu_rho_get": (ctypes.c_int, [ctypes.c_int, _f64p, ctypes.c_int64]),
    "nrg_gpu_spectrum_setup": (ctypes.c_int, [ctypes.c_int]),
    "nrg_gpu_spike_cfs": (ctypes.c_int, [ctypes.c_int, ctypes.c_double]),
    "nrg_gpu_spike_zero_temperature": (ctypes.c_int, [ctypes.c_int, ctypes.c_double]),
    "nrg_gpu_spectrum_bins": (ctypes.c_int, [_f64p, _f64p, ctypes.c_int64]),
    "nrg_gpu_zero_temperature_weights": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _f64p, ctypes.c_int,
                                                        _i32p, _i32p]),
    "nrg_gpu_history_clear": (ctypes.c_int, []),
    "nrg_gpu_save_restart": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _i32p]),
    "nrg_gpu_load_restart": (ctypes.c_int, [ctypes.c_char_p, _i32p, ctypes.c_int, _i32p]),
    "nrg_gpu_set_batch_limit": (ctypes.c_int, [ctypes.c_int]),
    "nrg_gpu_set_tridiag_schedule": (ctypes.c_int, [ctypes.c_int]),
    "nrg_gpu_set_graphs": (ctypes.c_int, [ctypes.c_int]),
    "nrg_gpu_set_tuning": (ctypes.c_int, [ctypes.c_int, ctypes.c_char_p]),
    "nrg_gpu_set_gemm_split": (ctypes.c_int, [ctypes.c_int]),
    "nrg_gpu_dgemm_batched": (ctypes.c_int, [ctypes.c_int, _i32p, ctypes.c_int, ctypes.c_int, _f64p, _i64p, _f64p,
                                             _i64p, _f64p, _i64p, ctypes.c_int, ctypes.c_int, _f64p]),
    "nrg_gpu_dgemm_batched_noalloc": (ctypes.c_int, [ctypes.c_int, _i32p, ctypes.c_int, ctypes.c_int, _f64p, _i64p,
                                                     _f64p, _i64p, _f64p, _i64p]),
}

KERNEL_FAMILIES = ["assemble", "tridiag_col", "tridiag_symv", "tridiag_gemm", "dc_leaf", "dc_prep",
                   "dc_secular", "dc_vectors", "dc_gemm", "dc_finish", "back_small", "back_gemm",
                   "invariant_gemm", "spectral_gemm", "cfs_products", "cfs_rho_weight", "cfs_binning",
                   "rho_gemm", "rho_gather", "tridiag_band", "back_band", "other"]


class NrgGpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def load(path: Optional[str] = None):
    """Load the shared library (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FileNotFoundError(f"{p} not built (run __graft_entry__.build() or make -C nrg-mpi_amd)")
    lib = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _p(a, t):
    if a is None:
        return None
    return a.ctypes.data_as(t)


def _check(rc):
    if rc != 0:
        raise NrgGpuError(rc, _lib.nrg_gpu_last_error().decode())


def init(device: int = 0):
    _check(load().nrg_gpu_init(device))


def finalize():
    _check(load().nrg_gpu_finalize())


def dgemm(A, B, C=None, transa=False, transb=False, alpha=1.0, beta=0.0):
    A = np.asfortranarray(A, dtype=np.float64)
    B = np.asfortranarray(B, dtype=np.float64)
    m = A.shape[1] if transa else A.shape[0]
    k = A.shape[0] if transa else A.shape[1]
    n = B.shape[0] if transb else B.shape[1]
    C = np.zeros((m, n), order="F") if C is None else np.asfortranarray(C, dtype=np.float64).copy(order="F")
    _check(load().nrg_gpu_dgemm(int(transa), int(transb), m, n, k, alpha, _p(A, _f64p), max(A.shape[0], 1),
                                _p(B, _f64p), max(B.shape[0], 1), beta, _p(C, _f64p), max(m, 1)))
    return C


def dgemm_batched(As, Bs, transa=False, transb=False, slices=0, repeat=0):
    """C_i = op(A_i) op(B_i) for every pair in one batched launch; slices 0:
    FP64 matrix cores, 3..8: split precision.  Returns (list of C_i, mean ms
    per batch over `repeat` device-timed runs or None)."""
    mnk, oa, ob, oc = [], [], [], []
    a_parts, b_parts = [], []
    pa = pb = pc = 0
    for A, B in zip(As, Bs):
        A = np.asfortranarray(A, dtype=np.float64)
        B = np.asfortranarray(B, dtype=np.float64)
        m = A.shape[1] if transa else A.shape[0]
        k = A.shape[0] if transa else A.shape[1]
        n = B.shape[0] if transb else B.shape[1]
        mnk += [m, n, k]
        oa.append(pa); ob.append(pb); oc.append(pc)
        a_parts.append(A.ravel(order="F")); b_parts.append(B.ravel(order="F"))
        pa += A.size; pb += B.size; pc += m * n
    Ah = np.concatenate(a_parts)
    Bh = np.concatenate(b_parts)
    Ch = np.zeros(pc)
    mnk = np.array(mnk, dtype=np.int32)
    oa, ob, oc = (np.array(x, dtype=np.int64) for x in (oa, ob, oc))
    ms = np.zeros(1)
    _check(load().nrg_gpu_dgemm_batched(len(As), _p(mnk, _i32p), int(transa), int(transb), _p(Ah, _f64p),
                                        _p(oa, _i64p), _p(Bh, _f64p), _p(ob, _i64p), _p(Ch, _f64p),
                                        _p(oc, _i64p), slices, repeat, _p(ms, _f64p)))
    out = []
    for i in range(len(As)):
        m, n = int(mnk[3 * i]), int(mnk[3 * i + 1])
        out.append(Ch[oc[i]:oc[i] + m * n].reshape(m, n, order="F"))
    return out, (float(ms[0]) if repeat > 0 else None)


def syev_batched(blocks):
    """Eigen-decompose a list of symmetric matrices in one batched call.

    Returns lists (w_b ascending, U_b) like numpy.linalg.eigh per block.
    """
    n = np.array([b.shape[0] for b in blocks], dtype=np.int32)
    H = np.concatenate([np.asfortranarray(b, dtype=np.float64).ravel(order="F") for b in blocks])
    w = np.zeros(int(n.sum()))
    U = np.zeros_like(H)
    _check(load().nrg_gpu_syev_batched(len(blocks), _p(n, _i32p), _p(H, _f64p), _p(w, _f64p), _p(U, _f64p)))
    ws, Us = [], []
    o1 = o2 = 0
    for nb in n:
        nb = int(nb)
        ws.append(w[o1:o1 + nb])
        Us.append(U[o2:o2 + nb * nb].reshape(nb, nb, order="F"))
        o1 += nb
        o2 += nb * nb
    return ws, Us


def invariant_get(k: int, ncm: int) -> np.ndarray:
    """The device-resident invariant matrix of the current kept basis as
    [t, l, r] = invariant_matrix(r, l, t) (k x k per conduction type)."""
    a = np.zeros(k * k * ncm)
    _check(load().nrg_gpu_invariant_get(_p(a, _f64p), a.size))
    return a.reshape(ncm, k, k)


def rho_get(iteration: int, kept_per_block):
    """Reduced density matrix of a swept iteration: list of k_b x k_b blocks
    (stocked blocks in block order; row-major as the reference stores it)."""
    kb = [int(k) for k in kept_per_block]
    tot = sum(k * k for k in kb)
    a = np.zeros(max(tot, 1))
    _check(load().nrg_gpu_rho_get(iteration, _p(a, _f64p), tot))
    out, o = [], 0
    for k in kb:
        out.append(a[o:o + k * k].reshape(k, k))
        o += k * k
    return out


def profile(enable=True, families=None, every=1, exclusive=True):
    """Start (clearing) or stop the per-family kernel timing; `families`
    restricts it to the named families (see KERNEL_FAMILIES); `every` times
    only every n-th launch of a family; `exclusive` fences a timed SYMV
    launch of the tridiagonalisation against the other stream groups
    (False: timed in place)."""
    _check(load().nrg_gpu_profile_sample(int(every)))
    _check(load().nrg_gpu_profile_exclusive(1 if exclusive else 0))
    if not enable:
        code = 0
    elif families is None:
        code = -1
    else:
        code = 0
        for f in families:
            code |= 1 << KERNEL_FAMILIES.index(f)
    _check(load().nrg_gpu_profile(code))


def launch_log(families=None):
    """Start (families: names) or stop (None) the launch log."""
    code = 0
    for f in families or []:
        code |= 1 << KERNEL_FAMILIES.index(f)
    _check(load().nrg_gpu_launch_log(code))


def launch_log_get():
    """Records (family, threads, kernel id, ordinal, bytes, flops, timed) as
    an (n, 7) array; timed is 1 for launches that carried timing events."""
    n = np.zeros(1, dtype=np.int64)
    _check(load().nrg_gpu_launch_log_get(None, 0, _p(n, _i64p)))
    a = np.zeros(7 * max(int(n[0]), 1))
    _check(load().nrg_gpu_launch_log_get(_p(a, _f64p), int(n[0]), _p(n, _i64p)))
    return a[:7 * int(n[0])].reshape(-1, 7)


def launch_log_kernel(kid):
    buf = ctypes.create_string_buffer(512)
    _check(load().nrg_gpu_launch_log_kernel(int(kid), buf, 512))
    return buf.value.decode()


def work_counters(reset=False):
    """{family: {launches, bytes, flops}} issued since the last reset."""
    n = len(KERNEL_FAMILIES)
    a = np.zeros(3 * n)
    _check(load().nrg_gpu_work_counters(_p(a, _f64p), n, int(reset)))
    a = a.reshape(n, 3)
    return {f: dict(launches=int(a[i, 0]), bytes=a[i, 1], flops=a[i, 2]) for i, f in enumerate(KERNEL_FAMILIES)}


def kernel_stats():
    n = len(KERNEL_FAMILIES)
    a = np.zeros(4 * n)
    _check(load().nrg_gpu_kernel_stats(_p(a, _f64p), n))
    a = a.reshape(n, 4)
    return {f: dict(ms=a[i, 0], launches=int(a[i, 1]), bytes=a[i, 2], flops=a[i, 3])
            for i, f in enumerate(KERNEL_FAMILIES)}


def stage_times():
    t = np.zeros(13)
    _check(load().nrg_gpu_stage_times(_p(t, _f64p), 13))
    keys = ["assemble", "eig", "tridiag", "dc", "back", "invariant", "spectral", "rho", "cfs", "eig_flops",
            "graph_captures", "graph_replays", "tridiag_issue"]
    return dict(zip(keys, t.tolist()))
