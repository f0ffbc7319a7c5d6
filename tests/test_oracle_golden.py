"""The oracle (numpy restatement, oracle/nrg_oracle.py) against the golden
fixtures written from reference runs: block tables and basis-record checksums
bit-exact, kept counts exact, eigenvalues to 1e-12 relative, binned spectra
to 1e-8 of the peak, Tr rho = 1 per shell, reduced-density block traces and
spectra to 1e-10.  CPU only."""
import hashlib
import os

import numpy as np
import pytest

import nrg_oracle as O
from nrgamd import models

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def sha(a):
    return hashlib.sha1(np.ascontiguousarray(a, dtype="<i4").tobytes()).hexdigest()


def kept_blk(rec, nv):
    lab = rec.basis_output[:rec.nkept, :nv]
    return np.array([int(np.all(lab == row[:nv], axis=1).sum()) for row in rec.subspace])


def compare_forward(nrg, fx, upto=None):
    nv = nrg.m.nv
    worst = 0.0
    for k, rec in enumerate(nrg.records):
        it = rec.it
        if upto is not None and it > upto:
            break
        np.testing.assert_array_equal(rec.subspace, fx[f"it{it}_subspace"])
        assert sha(rec.basis_input) == str(fx[f"it{it}_sha_basis_input"]), it
        assert rec.nkept == int(fx["nkept"][k])
        if f"it{it}_basis_output" in fx:
            # kept prefix in index order: exact; discarded states are in
            # energy order, where rounding-level differences between two
            # eigensolvers may permute (near-)degenerate states: same set
            ref = fx[f"it{it}_basis_output"].astype(np.int64)
            nk = rec.nkept
            np.testing.assert_array_equal(rec.basis_output[:nk], ref[:nk])
            a = rec.basis_output[nk:]
            b = ref[nk:]
            np.testing.assert_array_equal(a[np.lexsort(a.T[::-1])], b[np.lexsort(b.T[::-1])])
        np.testing.assert_array_equal(kept_blk(rec, nv), fx[f"it{it}_kept_blk"])
        E = rec.eigenvalue
        sc = max(1.0, float(np.abs(E).max()))
        Es = np.sort(E)
        if f"it{it}_eigenvalue" in fx:
            err = np.abs(E - fx[f"it{it}_eigenvalue"]).max()
        else:
            lo = int(fx[f"it{it}_E_cut_window_start"])
            w = fx[f"it{it}_E_cut_window"]
            err = max(np.abs(Es[:64] - fx[f"it{it}_E_low"][:len(Es[:64])]).max(),
                      np.abs(Es[lo:lo + len(w)] - w).max())
        worst = max(worst, err / sc)
    assert worst <= 1e-12, worst
    return worst


def compare_spectrum(nrg, fx):
    nv = nrg.m.nv
    peak = max(np.abs(fx["bins_pos"]).max(), np.abs(fx["bins_neg"]).max())
    err = max(np.abs(nrg.bins_pos - fx["bins_pos"]).max(), np.abs(nrg.bins_neg - fx["bins_neg"]).max())
    assert err <= 1e-8 * peak, err
    np.testing.assert_allclose(nrg.bins_pos.sum(1) + nrg.bins_neg.sum(1), 1.0, atol=1e-10)
    for it, tr in nrg.rho_trace.items():
        assert abs(tr - 1.0) < 1e-12, (it, tr)
    fs, ref_tr, ref_eig = fx["full_subspace"], fx["rho_block_trace"], fx["rho_block_eigs"]
    e0 = np.concatenate(([0], np.cumsum(fs[:, nv + 1])))
    mp = fx["map_full"]
    st = nrg.start_trunc
    worst = 0.0
    for row in range(len(fs)):
        it = st + int(np.searchsorted(mp, row + 1, side="right")) - 1
        r = nrg.rho[(it, tuple(int(x) for x in fs[row, :nv]))]
        worst = max(worst, abs(np.trace(r) - ref_tr[row]),
                    np.abs(np.linalg.eigvalsh(0.5 * (r + r.T)) - ref_eig[e0[row]:e0[row + 1]]).max())
    assert worst <= 1e-10, worst


def run_oracle(tmp_path, name, base=None, upto=None, **ov):
    d = str(tmp_path / name)
    models.write_inputs(models.workload(base or name, **ov), d)
    nrg = O.NRG(O.read_inputs(d))
    nrg.run_forward(upto)
    return nrg


def test_siam_nk256_forward_and_spectrum(tmp_path):
    fx = np.load(os.path.join(GOLD, "ref_siam_nk256_it40.npz"))
    nrg = run_oracle(tmp_path, "siam_nk256_it40")
    assert nrg.start_trunc == int(fx["start_trunc"])
    compare_forward(nrg, fx)
    nrg.run_spectrum()
    compare_spectrum(nrg, fx)


def test_two_orbital_spectrum(tmp_path):
    fx = np.load(os.path.join(GOLD, "ref_two_orbital_nk300_it20_spec.npz"))
    nrg = run_oracle(tmp_path, "two_orbital_nk300_it20_spec", base="two_orbital_nk4000_it45",
                     nkeep=300, niter=20, spectrum=True)
    compare_forward(nrg, fx)
    nrg.run_spectrum()
    compare_spectrum(nrg, fx)


def test_siam_nk2048_first_iterations(tmp_path):
    fx = np.load(os.path.join(GOLD, "ref_siam_nk2048_it50.npz"))
    nrg = run_oracle(tmp_path, "siam_nk2048_it50", upto=5)
    compare_forward(nrg, fx)


def test_two_orbital_nk4000_first_iterations(tmp_path):
    fx = np.load(os.path.join(GOLD, "ref_two_orbital_nk4000_it45.npz"))
    nrg = run_oracle(tmp_path, "two_orbital_nk4000_it45", upto=0)
    compare_forward(nrg, fx)


def test_per_state_and_grouped_cfs_agree(tmp_path):
    """The oracle's grouped evaluation of spikeCFS equals the per-state walk
    of the reference order (one iteration of the SIAM run)."""
    nrg = run_oracle(tmp_path, "siam_nk256_it40")
    nrg.setup_bins()
    nrg.init_reduced_density()
    nrg.reduced_density()
    rec = {r.it: r for r in nrg.records}[nrg.start_trunc + 3]
    nrg.spike_cfs(rec, per_state=True)
    a = (nrg.bins_pos.copy(), nrg.bins_neg.copy())
    nrg.setup_bins()
    nrg.spike_cfs(rec)
    np.testing.assert_allclose(nrg.bins_pos, a[0], atol=1e-15, rtol=1e-12)
    np.testing.assert_allclose(nrg.bins_neg, a[1], atol=1e-15, rtol=1e-12)
    assert np.abs(a[0]).sum() > 0


@pytest.mark.parametrize("mutation", ["drop_hopping_sign", "swap_spectral_sides"])
def test_fixture_rejects_planted_mistakes(tmp_path, mutation, monkeypatch):
    """A plausible mistake planted in the oracle must fail a fixture check."""
    fx = np.load(os.path.join(GOLD, "ref_siam_nk256_it40.npz"))
    if mutation == "drop_hopping_sign":
        orig = O.NRG.hopping_lookup

        def bad(self):
            T = orig(self)
            T = dict(T)
            T["coef"] = np.abs(T["coef"])
            self._hop = T
            return T
        monkeypatch.setattr(O.NRG, "hopping_lookup", bad)
        nrg = run_oracle(tmp_path, "siam_nk256_it40", upto=6)
        with pytest.raises(AssertionError):
            compare_forward(nrg, fx)
    else:
        nrg = run_oracle(tmp_path, "siam_nk256_it40")
        nrg.m.pos_op, nrg.m.neg_op = nrg.m.neg_op, nrg.m.pos_op
        nrg.run_spectrum()
        with pytest.raises(AssertionError):
            compare_spectrum(nrg, fx)


def _siam_eigenvalues(tmp_path, nkeep, eigh, tag):
    spec = models.siam(nkeep=nkeep, niter=8, spectrum=False)
    d = str(tmp_path / f"siam{nkeep}_{tag}")
    models.write_inputs(spec, d)
    nrg = O.NRG(O.read_inputs(d), eigh=eigh)
    return [np.sort(nrg.iteration(it).eigenvalue) for it in range(spec.first_iteration, spec.last_iteration + 1)]


def test_smoke_workload_is_well_conditioned(tmp_path):
    """The GPU smoke check compares SIAM Nkeep=128 (8 iterations) with the
    oracle at 1e-12 of the largest eigenvalue.  That bound is meaningful only
    if the workload's eigenvalues do not depend on the eigensolver beyond
    rounding: two LAPACK drivers (numpy's eigh and scipy's MRRR driver) must
    agree to it.  At Nkeep=64 they do not (the truncation falls inside
    near-degenerate levels), which this test also records."""
    import scipy.linalg as sl
    alt = lambda a: sl.eigh(a, driver="evr")
    ok = zip(_siam_eigenvalues(tmp_path, 128, np.linalg.eigh, "a"), _siam_eigenvalues(tmp_path, 128, alt, "b"))
    for k, (e0, e1) in enumerate(ok):
        assert len(e0) == len(e1)
        assert np.max(np.abs(e0 - e1)) <= 1e-12 * max(1.0, np.max(np.abs(e0))), k
    bad = list(zip(_siam_eigenvalues(tmp_path, 64, np.linalg.eigh, "c"), _siam_eigenvalues(tmp_path, 64, alt, "d")))
    worst = max(np.max(np.abs(e0 - e1)) / max(1.0, np.max(np.abs(e0))) for e0, e1 in bad if len(e0) == len(e1))
    assert worst > 1e-11


@pytest.mark.parametrize("name", ["siam_h3_nk128_it12", "siam_bcs_nk128_it12"])
def test_input_feature_decks(tmp_path, name):
    """Decks that exercise reference input features the physical workloads
    leave unused, against reference runs of the same decks:
    siam_h3 lists a reversed hopping row before its forward row (the
    reversed match swaps the references for good, diagonalization.F90:195-211);
    siam_bcs is a flag_BCS deck with 2Sz-only blocks whose pairing rows
    replace the hopping term (:215-245)."""
    fx = np.load(os.path.join(GOLD, f"ref_{name}.npz"))
    nrg = run_oracle(tmp_path, name)
    assert nrg.start_trunc == int(fx["start_trunc"])
    compare_forward(nrg, fx)


def test_input_feature_decks_reject_planted_mistakes(tmp_path, monkeypatch):
    """The two decks catch the mistakes they are made for: a swap flag taken
    from the last matching row only, and BCS rows that are ignored."""
    fx = np.load(os.path.join(GOLD, "ref_siam_h3_nk128_it12.npz"))
    orig = O.NRG.hopping_lookup

    def last_match(self):
        T = dict(orig(self))
        sw = np.zeros_like(T["swap"])
        for oi in range(self.m.nops):
            for oj in range(self.m.nops):
                for a, b, t, ch in self.m.coef_diag_type:
                    if (oi, oj) == (a - 1, b - 1):
                        sw[oi, oj] = False
                    elif (oi, oj) == (b - 1, a - 1):
                        sw[oi, oj] = True
        T["swap"] = sw
        self._hop = T
        return T
    monkeypatch.setattr(O.NRG, "hopping_lookup", last_match)
    nrg = run_oracle(tmp_path, "siam_h3_nk128_it12")
    with pytest.raises(AssertionError):
        compare_forward(nrg, fx)
    monkeypatch.setattr(O.NRG, "hopping_lookup", orig)

    fx = np.load(os.path.join(GOLD, "ref_siam_bcs_nk128_it12.npz"))
    d = str(tmp_path / "bcs_off")
    models.write_inputs(models.workload("siam_bcs_nk128_it12"), d)
    m = O.read_inputs(d)
    m.coef_bcs = m.coef_bcs[:0]
    nrg = O.NRG(m)
    nrg.run_forward(None)
    with pytest.raises(AssertionError):
        compare_forward(nrg, fx)


@pytest.mark.parametrize("name", ["siam_nk256_it40", "siam_bcs_nk128_it12"])
def test_thermodynamics_against_reference_files(tmp_path, name):
    """calcThermo (module.F90:1931-1982) of every iteration against the
    Xn, Sn, Cn files the reference run wrote (postprocess.F90:60-67)."""
    fx = np.load(os.path.join(GOLD, f"ref_{name}.npz"))
    nrg = run_oracle(tmp_path, name)
    ref = np.array([[float(x) for x in str(fx["file:" + f]).split()] for f in ("Xn", "Sn", "Cn")]).T
    got = np.array(nrg.thermo)
    assert got.shape == ref.shape
    np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-13)
