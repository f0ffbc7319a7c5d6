"""The Fortran host's bookkeeping against the reference run fixtures (CPU).

prepareBasis (prepareBasis.F90:41-114) with the exact replay of the
reference's exchange sort (module.F90:531-602) must reproduce the reference's
basis_input records bit for bit (SHA-1 of the int32 records) and its
subspace table; truncation (postprocess.F90:79-160) must reproduce the kept
counts and the kept-state layout.
"""
import hashlib
import os

import numpy as np
import pytest

from nrgamd import models

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def sha(a):
    return hashlib.sha1(np.ascontiguousarray(a, dtype="<i4").tobytes()).hexdigest()


def deck(tmp_path, name, **ov):
    d = str(tmp_path / name)
    models.write_inputs(models.workload(name, **ov), d)
    return d


def kept_basis(sub, kept, nv):
    rows = []
    for row, k in zip(sub, kept):
        for j in range(k):
            rows.append(list(row[:nv]) + [j + 1, 0, 0])
    return np.array(rows, dtype=np.int32)


CASES = [("siam_nk256_it40", {}, None), ("siam_nk2048_it50", {}, 12),
         ("two_orbital_nk4000_it45", {}, None), ("siam_bcs_nk128_it12", {}, None),
         ("siam_h3_nk128_it12", {}, None)]


@pytest.mark.parametrize("name,ov,limit", CASES)
def test_prepare_basis_exact(tmp_path, name, ov, limit):
    from nrgamd.host import Host
    fx = np.load(os.path.join(GOLD, f"ref_{name}.npz"))
    h = Host(deck(tmp_path, name, **ov), device=None)
    nv = h.i["nv"]
    its = list(fx["iterations"])
    if limit:
        its = its[:limit]
    for it in its:
        if it != its[0]:
            prev = it - 1
            sub = fx[f"it{prev}_subspace"]
            kept = fx[f"it{prev}_kept_blk"]
            bo = kept_basis(sub, kept, nv)
            h.set_basis_output(bo, np.zeros(len(bo)))
        h.prepare(0)
        np.testing.assert_array_equal(h.subspace(), fx[f"it{it}_subspace"])
        assert sha(h.basis_input()) == str(fx[f"it{it}_sha_basis_input"]), f"iteration {it}"


def test_canonical_sort_same_blocks(tmp_path):
    """Canonical mode: same block table, records equal as a multiset per block."""
    from nrgamd.host import Host
    fx = np.load(os.path.join(GOLD, "ref_two_orbital_nk4000_it45.npz"))
    h = Host(deck(tmp_path, "two_orbital_nk4000_it45"), device=None)
    nv = h.i["nv"]
    sub = fx["it1_subspace"]
    bo = kept_basis(sub, fx["it1_kept_blk"], nv)
    h.set_basis_output(bo, np.zeros(len(bo)))
    h.prepare(0)
    exact = h.basis_input().copy()
    h.set_basis_output(bo, np.zeros(len(bo)))
    h.prepare(1)
    canon = h.basis_input()
    np.testing.assert_array_equal(h.subspace(), fx["it2_subspace"])
    # same (labels, reference) sequence; ties may differ only in operation order
    np.testing.assert_array_equal(exact[:, :nv + 1], canon[:, :nv + 1])
    key = lambda a: np.lexsort(a.T[::-1])
    np.testing.assert_array_equal(exact[key(exact)], canon[key(canon)])


def test_truncation_matches_reference(tmp_path):
    """Kept counts, kept layout and eigenvalue order after postprocess."""
    from nrgamd.host import Host
    name = "siam_nk256_it40"
    fx = np.load(os.path.join(GOLD, f"ref_{name}.npz"))
    h = Host(deck(tmp_path, name), device=None)
    nv = h.i["nv"]
    its = list(fx["iterations"])
    for it in its[1:]:
        sub = fx[f"it{it - 1}_subspace"]
        bo = kept_basis(sub, fx[f"it{it - 1}_kept_blk"], nv)
        h.set_basis_output(bo, np.zeros(len(bo)))
        h.prepare(0)
        # rebuild the block-order eigenvalues from the post-truncation arrays
        E = fx[f"it{it}_eigenvalue"]
        bout = fx[f"it{it}_basis_output"].astype(np.int64)
        sub_now = fx[f"it{it}_subspace"]
        start = {tuple(r[:nv]): int(r[nv + 2]) for r in sub_now}
        eb = np.zeros_like(E)
        for e, row in zip(E, bout):
            eb[start[tuple(row[:nv])] - 1 + int(row[nv]) - 1] = e
        nk = h.truncate(it, eb)
        assert nk == int(fx["nkept"][its.index(it)]), f"iteration {it}"
        np.testing.assert_array_equal(h.kept_per_block(), fx[f"it{it}_kept_blk"])
        np.testing.assert_array_equal(h.eigenvalues(), E)
        assert sha(h.basis_output()) == str(fx[f"it{it}_sha_basis_output"])


@pytest.mark.parametrize("seed", range(6))
def test_lazy_sort_replays_exchange_sort(seed):
    """The O(N log N) replay of the reference's exchange sort gives the same
    permutation, tie order included, as the swap-by-swap loop and as the
    oracle's independent restatement, on keys with many ties."""
    import nrg_oracle as O
    from nrgamd import host
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 400))
    keys = rng.integers(0, max(2, n // 3), n).astype(np.int64)
    k0, p0 = host.sort_keys(keys, mode=0)
    k1, p1 = host.sort_keys(keys, mode=1)
    np.testing.assert_array_equal(k0, np.sort(keys))
    np.testing.assert_array_equal(p0, p1)
    np.testing.assert_array_equal(p0, O.exchange_sort_permutation(keys))


@pytest.mark.parametrize("seed", range(6))
def test_eigenvalue_sort_is_stable(seed):
    """The truncation's eigenvalue sort merges the ascending runs of its input
    (one per block); its permutation must be the stable one (ties in input
    order, as the reference's insertion sort with strict comparisons), which
    numpy's stable argsort fixes.  Inputs: ascending runs of random length
    with many exact ties, plus single-element, descending and all-equal cases."""
    from nrgamd import host
    rng = np.random.default_rng(seed)
    runs = []
    for _ in range(int(rng.integers(1, 40))):
        r = np.sort(rng.integers(0, 50, int(rng.integers(1, 60))).astype(np.float64) * 0.25)
        runs.append(r)
    cases = [np.concatenate(runs), np.array([3.0]), np.arange(30.0)[::-1].copy(), np.full(17, 2.5),
             rng.integers(0, 5, 200).astype(np.float64)]
    for v in cases:
        vs, p = host.sort_values(v)
        ref = np.argsort(v, kind="stable")
        np.testing.assert_array_equal(p, ref)
        np.testing.assert_array_equal(vs, v[ref])


@pytest.mark.parametrize("grid", [8.0, 1.0e6])
def test_truncation_with_ties_matches_oracle(tmp_path, grid):
    """The host's truncation (k-way merge of the per-block runs as far as the
    cutoff window reaches, discarded part formed on request) against the
    oracle's truncation (a full stable sort, pinned by the reference
    fixtures) on synthetic block eigenvalues: on a coarse grid (many exact
    ties within and across blocks) and a fine one.  Kept count, the whole
    eigenvalue array, the whole basis_output and the kept count per block."""
    from types import SimpleNamespace
    from nrgamd.host import Host
    import nrg_oracle as orc
    name = "siam_nk256_it40"
    fx = np.load(os.path.join(GOLD, f"ref_{name}.npz"))
    h = Host(deck(tmp_path, name), device=None)
    nv = h.i["nv"]
    it = 12
    bo = kept_basis(fx[f"it{it - 1}_subspace"], fx[f"it{it - 1}_kept_blk"], nv)
    h.set_basis_output(bo, np.zeros(len(bo)))
    h.prepare(0)
    sub = h.subspace()
    rng = np.random.default_rng(int(grid))
    eb = np.concatenate([np.sort(np.round(rng.random(int(r[nv])) * 3.0 * grid) / grid) for r in sub])
    nfull = len(eb)
    assert nfull > 256
    nk = h.truncate(it, eb)
    # the oracle's state for the same iteration: eigenvalues in block order,
    # basis_output (block labels, column index, 0, 0) in index order
    rows = []
    for r in sub:
        for j in range(int(r[nv])):
            rows.append(list(r[:nv]) + [j + 1, 0, 0])
    o = orc.NRG.__new__(orc.NRG)
    o.m = SimpleNamespace(nkeep=256, i0=int(h.i["i_first"]), N=int(h.i["n_iter"]), nv=nv)
    o.eigenvalue = eb.copy()
    o.basis_output = np.array(rows, dtype=np.int64)
    o.subspace = sub.astype(np.int64)
    o.start_trunc = 10 ** 9
    o.flag_trunc = False
    nb = o.truncate(it)
    assert nk == nb and nb < nfull
    np.testing.assert_array_equal(h.eigenvalues(), o.eigenvalue)
    np.testing.assert_array_equal(h.basis_output(), o.basis_output)
    np.testing.assert_array_equal(h.kept_per_block(), o.kept_per_block())
