"""The no-allocation batched FP64 product (nrg_gpu_dgemm_batched_noalloc)
against nrg_gpu_dgemm_batched on the same products: identical results (the
same task decomposition and kernels), on two calls in a row, the second
with fewer and smaller products (the scratch array keeps the first call's
tables), including a long-K product that takes the split-K path; and the
limit on the number of products per call is enforced with a clear error.
The operands are device arrays owned by the test (hipMalloc through the
HIP runtime the library links)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


def to_device(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    p = ctypes.c_void_p()
    assert hip().hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
    assert hip().hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0   # host to device
    return p


def from_device(p, n):
    out = np.zeros(n)
    assert hip().hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0   # device to host
    return out


def pack(rng, shapes, transa, transb):
    """Operands of products C_i = op(A_i) op(B_i) packed column-major."""
    A, B, oa, ob, oc, mnk = [], [], [], [], [], []
    pa = pb = pc = 0
    for m, n, k in shapes:
        a = rng.standard_normal((k, m) if transa else (m, k))
        b = rng.standard_normal((n, k) if transb else (k, n))
        A.append(a.ravel(order="F")); B.append(b.ravel(order="F"))
        oa.append(pa); ob.append(pb); oc.append(pc)
        pa += a.size; pb += b.size; pc += m * n
        mnk += [m, n, k]
    i64 = lambda x: np.array(x, dtype=np.int64)
    return (np.concatenate(A), np.concatenate(B), i64(oa), i64(ob), i64(oc), np.array(mnk, dtype=np.int32), pc)


def run_both(gpu, shapes, transa, transb, seed):
    rng = np.random.default_rng(seed)
    A, B, oa, ob, oc, mnk, nc = pack(rng, shapes, transa, transb)
    lib = gpu.load()
    P = gpu._p
    C_ref = np.zeros(nc)
    ms = np.zeros(1)
    gpu._check(lib.nrg_gpu_dgemm_batched(len(shapes), P(mnk, gpu._i32p), transa, transb, P(A, gpu._f64p),
                                         P(oa, gpu._i64p), P(B, gpu._f64p), P(ob, gpu._i64p), P(C_ref, gpu._f64p),
                                         P(oc, gpu._i64p), 0, 0, P(ms, gpu._f64p)))
    dA, dB, dC = to_device(A), to_device(B), to_device(np.full(nc, np.nan))
    try:
        gpu._check(lib.nrg_gpu_dgemm_batched_noalloc(len(shapes), P(mnk, gpu._i32p), transa, transb,
                                                     ctypes.cast(dA, gpu._f64p), P(oa, gpu._i64p),
                                                     ctypes.cast(dB, gpu._f64p), P(ob, gpu._i64p),
                                                     ctypes.cast(dC, gpu._f64p), P(oc, gpu._i64p)))
        C = from_device(dC, nc)
    finally:
        for p in (dA, dB, dC):
            hip().hipFree(p)
    # against numpy, for the products themselves
    for i, (m, n, k) in enumerate(shapes):
        a = A[oa[i]:oa[i] + m * k].reshape((k, m) if transa else (m, k), order="F")
        b = B[ob[i]:ob[i] + k * n].reshape((n, k) if transb else (k, n), order="F")
        ref = (a.T if transa else a) @ (b.T if transb else b)
        got = C[oc[i]:oc[i] + m * n].reshape(m, n, order="F")
        assert np.max(np.abs(got - ref)) <= 1e-13 * k, (m, n, k)
    return C, C_ref


def test_noalloc_matches_the_allocating_entry_point(gpu):
    # first call: seven products, one of them long-K with few output tiles
    # (split along K into partial products in the scratch array)
    big = [(100, 100, 4475), (130, 77, 200), (64, 64, 64), (17, 5, 3), (1, 1, 1), (256, 300, 129), (33, 7, 1030)]
    C, C_ref = run_both(gpu, big, 1, 0, seed=1)
    np.testing.assert_array_equal(C, C_ref)
    # second call: fewer, smaller products with other transposes
    small = [(40, 30, 2048), (9, 11, 13)]
    C, C_ref = run_both(gpu, small, 0, 1, seed=2)
    np.testing.assert_array_equal(C, C_ref)


def test_noalloc_rejects_too_many_products(gpu):
    lib = gpu.load()
    n = 4097
    mnk = np.ones(3 * n, dtype=np.int32)
    off = np.zeros(n, dtype=np.int64)
    P = gpu._p
    rc = lib.nrg_gpu_dgemm_batched_noalloc(n, P(mnk, gpu._i32p), 0, 0, None, P(off, gpu._i64p), None,
                                           P(off, gpu._i64p), None, P(off, gpu._i64p))
    assert rc == 2
    assert b"at most 4096" in lib.nrg_gpu_last_error()
