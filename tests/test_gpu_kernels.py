"""Kernel-level parity of the HIP library against numpy/LAPACK on shared inputs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (17, 5, 3), (64, 64, 64), (130, 77, 200), (5, 300, 0),
                                   (128, 100, 4475), (33, 7, 1030)])   # the last two split along K
def test_dgemm(gpu, ta, tb, m, n, k):
    rng = np.random.default_rng(m * 1000 + n * 10 + k)
    A = rng.standard_normal((k, m) if ta else (m, k))
    B = rng.standard_normal((n, k) if tb else (k, n))
    C0 = rng.standard_normal((m, n))
    C = gpu.dgemm(A, B, C0, transa=ta, transb=tb, alpha=-0.7, beta=0.3)
    ref = -0.7 * ((A.T if ta else A) @ (B.T if tb else B)) + 0.3 * C0
    # FP64 accumulate: error ~ k * eps * |A||B|
    assert np.max(np.abs(C - ref)) <= 1e-13 * max(1, k)


def _haar(rng, n):
    q, r = np.linalg.qr(rng.standard_normal((n, n)))
    return q * np.sign(np.diag(r))


def _check_eig(H, w, U, tol=1e-12):
    n = H.shape[0]
    nrm = max(np.linalg.norm(H, 2), 1e-300)
    wr = np.linalg.eigvalsh(H)
    assert np.all(np.diff(w) >= -1e-14 * nrm), "eigenvalues not ascending"
    assert np.max(np.abs(w - wr)) <= tol * nrm * max(1, n / 100)
    res = np.linalg.norm(H @ U - U * w) / nrm
    orth = np.linalg.norm(U.T @ U - np.eye(n))
    assert res <= 1e-13 * max(1, n / 10), res
    assert orth <= 1e-13 * max(1, n / 10), orth


def test_syev_random_sizes(gpu):
    rng = np.random.default_rng(7)
    sizes = [1, 2, 3, 5, 16, 31, 32, 33, 47, 64, 65, 100, 129, 257, 400]
    blocks = []
    for n in sizes:
        a = rng.standard_normal((n, n))
        blocks.append(a + a.T)
    ws, Us = gpu.syev_batched(blocks)
    for H, w, U in zip(blocks, ws, Us):
        _check_eig(H, w, U)


def test_syev_degenerate_nrg_ladder(gpu):
    """Exact degeneracies (deflation) and an NRG-like geometric ladder."""
    rng = np.random.default_rng(11)
    blocks = []
    for n in [40, 90, 333]:
        k = rng.integers(0, 61, size=n)
        lam = 2.0 ** (-k / 2.0)
        Q = _haar(rng, n)
        blocks.append((Q * lam) @ Q.T)
    # exactly degenerate multiplets
    for n in [48, 200]:
        lam = np.repeat(np.arange(n // 4, dtype=float), 4)[:n]
        Q = _haar(rng, n)
        blocks.append((Q * lam) @ Q.T)
    # already diagonal / tridiagonal with zero couplings
    blocks.append(np.diag(np.arange(70, dtype=float)[::-1]))
    ws, Us = gpu.syev_batched(blocks)
    for H, w, U in zip(blocks, ws, Us):
        _check_eig(H, w, U)


def test_syev_large_block(gpu):
    rng = np.random.default_rng(3)
    n = 1500
    a = rng.standard_normal((n, n))
    H = (a + a.T) / np.sqrt(n)
    ws, Us = gpu.syev_batched([H])
    _check_eig(H, ws[0], Us[0])
