"""The two-stage tridiagonalisation (dense -> band of 32 by panel QRs and
matrix-core updates -> tridiagonal by bulge chasing, back-transform through
both stages): batched eigensolves against LAPACK (eigenvalues, residual,
orthogonality) on sizes around the band width and the 64-row tiles, exact
degeneracies, an NRG ladder and mixed batches (end-aligned panel schedule
over two stream groups); then NRG forward passes through the Fortran host
against the reference fixtures (block tables, basis records and kept counts
bit-exact, eigenvalues to 1e-12)."""
import numpy as np
import pytest

from test_gpu_kernels import _check_eig, _haar
from test_gpu_nrg import run_and_compare

pytestmark = pytest.mark.gpu


@pytest.fixture
def two_stage(gpu):
    gpu.load().nrg_gpu_set_tridiag_method(1)
    yield gpu
    gpu.init(0)
    gpu.load().nrg_gpu_set_tridiag_method(0)


def test_two_stage_sizes_around_the_band(two_stage):
    rng = np.random.default_rng(71)
    sizes = [1, 2, 3, 4, 5, 16, 31, 32, 33, 34, 35, 36, 47, 63, 64, 65, 66, 67, 96, 97, 98, 100, 129, 130, 161, 257]
    blocks = []
    for n in sizes:
        a = rng.standard_normal((n, n))
        blocks.append(a + a.T)
    print("two-stage syev, sizes", sizes, flush=True)
    ws, Us = two_stage.syev_batched(blocks)
    for H, w, U in zip(blocks, ws, Us):
        _check_eig(H, w, U)


def test_two_stage_degenerate_and_ladder(two_stage):
    rng = np.random.default_rng(13)
    blocks = []
    for n in [40, 90, 333, 700]:
        k = rng.integers(0, 61, size=n)
        lam = 2.0 ** (-k / 2.0)
        Q = _haar(rng, n)
        blocks.append((Q * lam) @ Q.T)
    for n in [48, 200]:
        lam = np.repeat(np.arange(n // 4, dtype=float), 4)[:n]
        Q = _haar(rng, n)
        blocks.append((Q * lam) @ Q.T)
    blocks.append(np.diag(np.arange(70, dtype=float)[::-1]))
    print("two-stage syev, degenerate and ladder blocks", flush=True)
    ws, Us = two_stage.syev_batched(blocks)
    for H, w, U in zip(blocks, ws, Us):
        _check_eig(H, w, U)


def test_two_stage_large_block(two_stage):
    rng = np.random.default_rng(5)
    n = 1500
    a = rng.standard_normal((n, n))
    H = (a + a.T) / np.sqrt(n)
    print("two-stage syev n=1500", flush=True)
    ws, Us = two_stage.syev_batched([H])
    _check_eig(H, ws[0], Us[0])


def test_two_stage_nrg_siam(tmp_path, two_stage):
    w = run_and_compare(tmp_path, "siam_nk256_it40", {})
    print("SIAM Nkeep=256, two-stage: worst eigenvalue error", w)


def test_two_stage_nrg_two_orbital_window(tmp_path, two_stage):
    w = run_and_compare(tmp_path, "two_orbital_nk4000_it45", {}, last=8)
    print("two-orbital Nkeep=4000 iterations -1..8, two-stage: worst eigenvalue error", w)
