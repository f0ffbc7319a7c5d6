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


def _fast_rotation(rng, n):
    """A dense random orthogonal matrix applied in O(n^2 log n): random
    signs, an orthonormal DCT-II and a random permutation.  Returns
    apply(X) = Q X and apply_t(X) = Q^T X for column blocks X."""
    import scipy.fft as sf
    s1 = rng.choice([-1.0, 1.0], size=n)
    s2 = rng.choice([-1.0, 1.0], size=n)
    p = rng.permutation(n)

    def apply(X):   # Q = S2 C S1 P
        return s2[:, None] * sf.dct(s1[:, None] * X[p], type=2, norm="ortho", axis=0, workers=-1)

    def apply_t(X):  # Q^T = P^T S1 C^T S2
        Y = s1[:, None] * sf.idct(s2[:, None] * X, type=2, norm="ortho", axis=0, workers=-1)
        out = np.empty_like(Y)
        out[p] = Y
        return out
    return apply, apply_t


def _planted(rng, n, lam):
    """H = Q diag(lam) Q^T with a dense random orthogonal Q."""
    apply, apply_t = _fast_rotation(rng, n)
    Y = apply_t(np.eye(n)) * lam[:, None]   # diag(lam) Q^T
    H = apply(Y)
    return 0.5 * (H + H.T)


def _check_planted(H, w, U, lam, rng, tol):
    """Eigenvalues against the planted spectrum (a closed form), residual
    and orthogonality by random probes (matrix-vector products only)."""
    n = H.shape[0]
    nrm = np.abs(lam).max()
    assert np.all(np.diff(w) >= -1e-14 * nrm), "eigenvalues not ascending"
    assert np.max(np.abs(w - np.sort(lam))) <= tol * nrm
    X = rng.standard_normal((n, 4))
    UX = U @ X
    res = np.linalg.norm(H @ UX - U @ (w[:, None] * X)) / (nrm * np.linalg.norm(X))
    orth = np.linalg.norm(U.T @ UX - X) / np.linalg.norm(X)
    assert res <= tol, res
    assert orth <= tol, orth
    return res, orth


@pytest.mark.parametrize("n,ladder", [(8500, "integer"), (10000, "continuous")])
def test_syev_blocks_above_lds_limits(gpu, n, ladder):
    """Blocks above the LDS-resident paths: the top divide-and-conquer merge
    of m > 8192 runs its deflation scan from global memory, the continuous
    ladder leaves > 4000 secular roots (poles from global memory), and at
    n = 10000 the first ~280 SYMV launches keep v in global memory (their two
    LDS vectors would exceed 160 KB).  Spectra Lambda^(-k/2), Lambda = 2,
    k uniform on 0..60 (integer: exact degeneracies, heavy deflation)."""
    rng = np.random.default_rng(n)
    k = rng.integers(0, 61, size=n).astype(float) if ladder == "integer" else rng.uniform(0, 60, size=n)
    lam = 2.0 ** (-k / 2.0)
    H = _planted(rng, n, lam)
    print(f"syev n={n} {ladder} ladder", flush=True)
    ws, Us = gpu.syev_batched([H])
    # eigenvalue error ~ n eps |H|, residual and orthogonality ~ n eps
    res, orth = _check_planted(H, ws[0], Us[0], lam, rng, tol=2e-16 * n * 8)
    print("residual", res, "orthogonality", orth)


@pytest.mark.parametrize("finish", [192, 96, 0])
def test_syev_tail_finisher(gpu, finish):
    """The last columns of every block reduced in LDS by one launch (trailing
    blocks of at most `finish` rows) against the column steps run to the end
    (finish 0) and against LAPACK: random blocks around the 192-row boundary
    and larger ones, with degenerate (deflating) and geometric spectra."""
    rng = np.random.default_rng(29)
    blocks = []
    for n in [1, 2, 31, 96, 97, 150, 191, 192, 193, 226, 450, 700]:
        a = rng.standard_normal((n, n))
        blocks.append(a + a.T)
    for n in [180, 500]:
        lam = np.repeat(np.arange(n // 4, dtype=float), 4)[:n]
        Q = _haar(rng, n)
        blocks.append((Q * lam) @ Q.T)
    k = rng.integers(0, 61, size=260)
    Q = _haar(rng, 260)
    blocks.append((Q * 2.0 ** (-k / 2.0)) @ Q.T)
    lib = gpu.load()
    try:
        gpu._check(lib.nrg_gpu_set_tridiag_finish(finish))
        ws, Us = gpu.syev_batched(blocks)
    finally:
        gpu._check(lib.nrg_gpu_set_tridiag_finish(192))
    for H, w, U in zip(blocks, ws, Us):
        _check_eig(H, w, U)
