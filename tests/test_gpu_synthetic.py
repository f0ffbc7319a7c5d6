"""Synthetic block batch at its stated size: 256 dense symmetric FP64 blocks,
sizes log-normal (median 600, sigma 1.0 in ln-space, seed 20240) clipped to
16..4096, eigenvalues on the ladder 2^(-k/2) with k uniform on the integers
0..60 (exact degeneracies), Haar-rotated; four N(0,1) operators per block.

Checked: eigenvalues against the planted ladder (a closed form), residual
and orthogonality by random probes, all within c n eps; U^T O U in FP64
against numpy on a sample of blocks (the largest included); the
split-precision products against the FP64 products on every block."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# backward-stable bounds per unit of block size (the checks divide by n)
EIG_TOL_OVER_N = 8 * 2.0 ** -52


# the numpy checks of the split products cover every block (minutes of CPU)
@pytest.mark.timeout(600)
def test_synthetic_batch_256_blocks(gpu):
    from nrgamd import synthetic as S
    print("generating the 256-block batch", flush=True)
    batch = S.make_batch(seed=20240, nblk=256, median=600.0, sigma=1.0)
    assert len(batch.n) == 256 and batch.n.min() >= 16 and batch.n.max() <= 4096
    r = S.run(batch, nops=4, split=(8, 6), repeat=0, sample=12, log=lambda m: print(m, flush=True))
    print({k: v for k, v in r.items() if k != "contraction"})
    print(r["contraction"])
    assert r["eig_err_over_n"] <= EIG_TOL_OVER_N
    assert r["residual_over_n"] <= EIG_TOL_OVER_N
    assert r["orthogonality_over_n"] <= EIG_TOL_OVER_N
    c = r["contraction"]
    # FP64 products: k eps |U||O||U| with k <= 4096, relative to max |C|
    assert c["fp64"]["err_vs_numpy"] <= 1e-12
    # 8 slices carry ~53 bits, 6 slices ~40 bits of each operand
    assert c["split8"]["err_vs_fp64"] <= 1e-12
    assert c["split6"]["err_vs_fp64"] <= 1e-10
