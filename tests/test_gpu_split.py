"""Split-precision (Ozaki slicing) GEMM on int8 matrix cores against FP64.

Pinned by: exact products of small-integer matrices (two digits hold each
entry exactly, every digit product and diagonal sum is an exact int32 and
the weighted sum fits 53 bits, so the split result must equal the exact
product bit for bit); normwise agreement with numpy's FP64 product on
random and transposed operands of uneven sizes; the synthetic block batch of
U^T O U contractions (Haar-rotated NRG-like spectra, four dense operators
per block); the NRG iterations with every contraction product in split form
against the reference-run fixtures."""
import os

import numpy as np
import pytest

from nrgamd import models

pytestmark = pytest.mark.gpu


@pytest.fixture
def gpu():
    from nrgamd import gpu as g
    g.init(0)
    return g


def test_integer_products_exact(gpu):
    rng = np.random.default_rng(1)
    As, Bs = [], []
    for m, n, k in [(64, 64, 64), (70, 33, 1000), (1, 5, 300), (130, 129, 515)]:
        As.append(rng.integers(-128, 129, size=(m, k)).astype(np.float64))
        Bs.append(rng.integers(-128, 129, size=(k, n)).astype(np.float64))
    for s in (3, 8):
        Cs, _ = gpu.dgemm_batched(As, Bs, slices=s)
        for A, B, C in zip(As, Bs, Cs):
            np.testing.assert_array_equal(C, A @ B)


@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
def test_random_products(gpu, ta, tb):
    rng = np.random.default_rng(2)
    As, Bs, ref = [], [], []
    for m, n, k in [(200, 150, 300), (17, 400, 65), (513, 64, 1024), (3, 3, 3)]:
        A = rng.standard_normal((m, k)) * np.exp(rng.uniform(-3, 3, size=(m, 1)))
        B = rng.standard_normal((k, n)) * np.exp(rng.uniform(-3, 3, size=(1, n)))
        ref.append(A @ B)
        As.append(np.asfortranarray(A.T) if ta else A)
        Bs.append(np.asfortranarray(B.T) if tb else B)
    C8, _ = gpu.dgemm_batched(As, Bs, transa=ta, transb=tb, slices=8)
    C0, _ = gpu.dgemm_batched(As, Bs, transa=ta, transb=tb, slices=0)
    for c8, c0, r, A, B in zip(C8, C0, ref, As, Bs):
        scale = np.abs(A).max() * np.abs(B).max() * max(A.shape) + 1e-300
        e8 = np.abs(c8 - r).max() / scale
        e0 = np.abs(c0 - r).max() / scale
        assert e8 < 1e-14, (e8, e0)
        assert e0 < 1e-14


def synthetic_batch(nblk, median, seed):
    """Symmetric blocks with eigenvalues on Lambda^{-k/2} ladders rotated by
    Haar orthogonal matrices (QR of a Gaussian with sign fix) and four dense
    N(0,1) operators each; returns (U list, O list)."""
    rng = np.random.default_rng(seed)
    sizes = np.clip(np.exp(rng.normal(np.log(median), 1.0, size=nblk)), 16, 4 * median).astype(int)
    Us, Os = [], []
    for n in sizes:
        q, r = np.linalg.qr(rng.standard_normal((n, n)))
        q = q * np.sign(np.diag(r))
        ev = 2.0 ** (-rng.integers(0, 61, size=n) / 2.0)
        H = (q * ev) @ q.T
        _, U = np.linalg.eigh(H)
        Us.append(U)
        Os.append([rng.standard_normal((n, n)) for _ in range(4)])
    return Us, Os


def test_synthetic_block_batch(gpu):
    Us, Os = synthetic_batch(24, 120, seed=3)
    A1, B1 = [], []
    for U, ops in zip(Us, Os):
        for O in ops:
            A1.append(O); B1.append(U)
    for s in (8, 6):
        X, _ = gpu.dgemm_batched(A1, B1, slices=s)               # O U
        A2 = [B1[i] for i in range(len(X))]
        M, _ = gpu.dgemm_batched(A2, X, transa=True, slices=s)   # U^T (O U)
        worst = 0.0
        for i, m in enumerate(M):
            r = B1[i].T @ (A1[i] @ B1[i])
            worst = max(worst, np.abs(m - r).max() / np.abs(r).max())
        print("slices", s, "worst relative error of U^T O U", worst)
        assert worst < (1e-13 if s == 8 else 1e-9)


def test_nrg_iterations_split(tmp_path, gpu):
    from test_gpu_nrg import run_and_compare
    gpu.load().nrg_gpu_set_gemm_split(8)
    try:
        run_and_compare(tmp_path, "siam_nk256_it40", {})
        run_and_compare(tmp_path, "two_orbital_nk4000_it45", {}, last=2)
    finally:
        gpu.load().nrg_gpu_set_gemm_split(0)
