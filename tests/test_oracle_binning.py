"""binningSpike (module.F90:1210-1254) in the oracle against hand-computed
cases: the log-mesh position pos = (log10|w| + 15) * 1000 + 1, the linear
split over bins floor(pos) and floor(pos)+1, the clamps to bins 1 and
numberOfBin, omega = 0, the sign-blind |omega|, and the vectorised form
against the scalar one.  CPU only."""
import math

import numpy as np
import pytest

import nrg_oracle as O


class _Bins:
    nbin = 16001
    bin_spike = O.NRG.bin_spike
    bin_spikes = O.NRG.bin_spikes


def fresh():
    b = _Bins()
    return b, np.zeros((2, b.nbin))


@pytest.mark.parametrize("omega,expect", [
    (1e-15, {0: 1.0}),                 # pos = 1 exactly -> bin 1 (index 0), weight split 1/0
    (1e-14, {1000: 1.0}),              # pos = 1001
    (10 ** -14.5, {500: 1.0}),         # pos = 501
    (1e-16, {0: 1.0}),                 # pos = -999 -> clamp to bin 1
    (0.0, {0: 1.0}),                   # log10(0) = -inf -> bin 1
    (100.0, {16000: 1.0}),             # pos = 17001 > numberOfBin -> clamp to the last bin
    (10.0, {16000: 1.0}),              # pos = 16001 = numberOfBin -> last bin
    (-1e-14, {1000: 1.0}),             # |omega|
])
def test_exact_positions(omega, expect):
    b, c = fresh()
    b.bin_spike(c, omega, 1.0, 0)
    got = {int(i): float(c[0, i]) for i in np.nonzero(c[0])[0]}
    assert set(got) <= set(expect) | {i + 1 for i in expect}
    for i, w in expect.items():
        assert got.get(i, 0.0) == pytest.approx(w, abs=1e-9)
    assert c[1].sum() == 0.0


def test_linear_split():
    b, c = fresh()
    om = 10 ** (-14.5 + 0.25e-3)      # pos = 501.25
    b.bin_spike(c, om, 2.0, 1)
    pos = (math.log10(om) + 15.0) * 1000.0 + 1.0
    w = pos - int(pos)
    assert int(pos) == 501
    assert c[1, 500] == pytest.approx(2.0 * (1.0 - w), rel=1e-12)
    assert c[1, 501] == pytest.approx(2.0 * w, rel=1e-12)
    assert w == pytest.approx(0.25, abs=1e-9)


def test_below_first_bin_quirk():
    # 0 < pos < 1: positionReference = int(pos) = 0 is reset to 1, the split
    # weight pos - 1 is negative (the reference keeps this)
    b, c = fresh()
    om = 10 ** (-15 - 0.5e-3)          # pos = 0.5
    b.bin_spike(c, om, 1.0, 0)
    assert c[0, 0] == pytest.approx(1.5, rel=1e-9)
    assert c[0, 1] == pytest.approx(-0.5, rel=1e-9)


def test_vectorised_equals_scalar():
    rng = np.random.default_rng(3)
    om = np.concatenate([10 ** rng.uniform(-17, 1.5, 3000), [0.0, 1e-15, 5.0, -2e-3]])
    wt = rng.standard_normal(len(om))
    b, c1 = fresh()
    for o, w in zip(om, wt):
        b.bin_spike(c1, o, w, 1)
    _, c2 = fresh()
    b.bin_spikes(c2, om, wt, 1)
    np.testing.assert_allclose(c2, c1, rtol=1e-12, atol=1e-14)
    # total weight is conserved by the split
    assert c2.sum() == pytest.approx(wt.sum(), rel=1e-12)
