"""The 200-bin log view of the CFS spectrum (nrgamd.spectrum.log_view).

Pinned by what the construction fixes: a spike binned by the oracle's
binningSpike (module.F90:1210-1254) lands whole in the coarse bin whose
cells hold its two mesh points; the coarse weights plus the weight outside
the window equal the native total; on the reference's own SIAM Nkeep=256
bins (particle-hole symmetric deck) the view keeps A(ω) = A(−ω) and the
spectral sum rule.  CPU only."""
import math
import os

import numpy as np
import pytest

import nrg_oracle as O
from nrgamd.spectrum import log_view

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


class _Bins:
    nbin = 16001
    bin_spike = O.NRG.bin_spike
    bin_spikes = O.NRG.bin_spikes


def test_spike_lands_in_its_coarse_bin():
    b = _Bins()
    c = np.zeros((1, b.nbin))
    omega, weight = 3.0e-5, 0.7
    b.bin_spike(c, omega, weight, 0)
    w, A, cw, out = log_view(c)
    # the two mesh points around 3e-5 are cells 10477 and 10478 (0-based:
    # log10 3e-5 = -4.5229, position (-4.5229 + 15) * 1000 = 10477.1);
    # coarse bin j holds cells 7000 + 40 j .. 7039 + 40 j
    assert np.nonzero(c[0])[0].tolist() == [10477, 10478]
    j = (10477 - 7000) // 40
    assert (10478 - 7000) // 40 == j
    np.testing.assert_allclose(cw[0, j], weight, rtol=0, atol=1e-15)
    assert np.count_nonzero(cw[0]) == 1
    assert abs(out[0]) < 1e-15
    edges_lo = 10.0 ** (-15 + (7000 + 40 * j - 0.5) / 1000)
    edges_hi = 10.0 ** (-15 + (7040 + 40 * j - 0.5) / 1000)
    assert edges_lo <= omega < edges_hi
    np.testing.assert_allclose(A[0, j] * (edges_hi - edges_lo), weight, rtol=1e-14)
    np.testing.assert_allclose(w[j], math.sqrt(edges_lo * edges_hi), rtol=1e-14)


def test_grid_is_25_per_decade():
    w, A, cw, out = log_view(np.zeros(16001))
    assert len(w) == 200
    np.testing.assert_allclose(np.diff(np.log10(w)), 0.04, rtol=0, atol=1e-12)
    assert 1e-8 < w[0] < 1.1e-8 and 0.9 < w[-1] < 1.0


def test_rejects_windows_that_cut_native_cells():
    with pytest.raises(ValueError):
        log_view(np.zeros(16001), lo=-8, hi=0, nbins=300)
    with pytest.raises(ValueError):
        log_view(np.zeros(16001), lo=-16, hi=0, nbins=200)
    with pytest.raises(ValueError):
        log_view(np.zeros(9001), lo=-8, hi=0, nbins=200)


def test_reference_siam_view_conserves_weight_and_symmetry():
    z = np.load(os.path.join(GOLD, "ref_siam_nk256_it40.npz"))
    pos, neg = z["bins_pos"], z["bins_neg"]
    for bins in (pos, neg):
        w, A, cw, out = log_view(bins)
        np.testing.assert_allclose(cw.sum(axis=1) + out, bins.sum(axis=1), rtol=0, atol=1e-14)
        # the window 1e-8 .. 1 holds the spectrum of a 40-iteration Lambda=2 run
        assert np.all(np.abs(out) < 1e-3 * bins.sum(axis=1))
    # sum rule: total weight of both signs is 1 per operator
    np.testing.assert_allclose(pos.sum(axis=1) + neg.sum(axis=1), 1.0, rtol=0, atol=1e-10)
    # particle-hole symmetric deck (eps_d = -U/2): A(w) = A(-w) on the view
    _, Ap, _, _ = log_view(pos)
    _, An, _, _ = log_view(neg)
    peak = float(np.max(np.abs(Ap)))
    assert float(np.max(np.abs(Ap - An))) < 1e-6 * peak
