"""The row-chunk SYMV of the one-stage tridiagonalisation
(nrg_gpu_set_tridiag_symv(1): row parts summed in registers and LDS chunks
in a fixed order instead of LDS atomics): batched eigensolves against LAPACK
and the two-orbital Nkeep=4000 window against the reference fixture, and
bit-identical results from two runs of the same solve."""
import numpy as np
import pytest

from test_gpu_kernels import _check_eig
from test_gpu_nrg import run_and_compare

pytestmark = pytest.mark.gpu


@pytest.fixture
def symv_rows(gpu):
    gpu.load().nrg_gpu_set_tridiag_symv(1)
    yield gpu
    gpu.init(0)
    gpu.load().nrg_gpu_set_tridiag_symv(0)


def test_symv_rows_batched_and_reproducible(symv_rows):
    rng = np.random.default_rng(17)
    sizes = [1, 2, 3, 17, 33, 64, 65, 255, 256, 257, 300, 513, 1500, 2600]
    blocks = []
    for n in sizes:
        a = rng.standard_normal((n, n))
        blocks.append((a + a.T) / np.sqrt(n))
    print("row-chunk SYMV syev, sizes", sizes, flush=True)
    ws, Us = symv_rows.syev_batched(blocks)
    for H, w, U in zip(blocks, ws, Us):
        _check_eig(H, w, U)
    ws2, _ = symv_rows.syev_batched(blocks)
    for w, w2 in zip(ws, ws2):
        assert np.array_equal(w, w2)


def test_symv_rows_nrg_window(tmp_path, symv_rows):
    w = run_and_compare(tmp_path, "two_orbital_nk4000_it45", {}, last=12)
    print("two-orbital Nkeep=4000 iterations -1..12, row-chunk SYMV: worst eigenvalue error", w)
