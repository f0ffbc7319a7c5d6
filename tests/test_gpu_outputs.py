"""The reference's output files written by the Fortran host (Xn, Sn, Cn,
eigenodd.txt, eigeneven.txt, keepingState.txt, CFS-spectrum-*, the
zero-temperature spike files step_ 02_itr_NNN.*) against the files the
reference run wrote for the same decks (postprocess.F90:60-67, 128-137,
163-167; main.F90:228-252; calculateSpectrumDGEMM.F90:80-90,157-159,203-205).

Both are written by the same list-directed Fortran statements from the same
compiler, so every record has the same layout: the lines are compared with
their digit runs masked (signs, decimal points and exponent fields must
agree; the shortest round-trip digits may differ in count),
integers exactly, single-precision energies to a few units of their last
place, thermodynamic values to 1e-10 relative, binned weights to 1e-8 of the
peak bin.  The spectrum files list the bins above 1e-16 scale(N); a bin
within rounding of that threshold may appear in only one of them."""
import os
import re

import numpy as np
import pytest

from nrgamd import models

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def shape(line):
    # digit runs collapsed: list-directed output prints the shortest digits
    # that round-trip, so a last-bit difference can change their count
    return re.sub(r"\d+", "0", line)


def tokens(line):
    return line.split()


def compare_records(name, mine, ref, rtol_real):
    a, b = mine.splitlines(), ref.splitlines()
    assert len(a) == len(b), (name, len(a), len(b))
    worst = 0.0
    for la, lb in zip(a, b):
        assert shape(la) == shape(lb), (name, la, lb)
        for x, y in zip(tokens(la), tokens(lb)):
            if re.fullmatch(r"-?\d+", y):
                assert x == y, (name, la, lb)
            else:
                fx, fy = float(x), float(y)
                err = abs(fx - fy) / max(abs(fy), 1e-300)
                worst = max(worst, err)
                assert err <= rtol_real or abs(fx - fy) <= 1e-12, (name, la, lb)
    return worst


def compare_step_file(name, mine, ref):
    """Zero-temperature spike records: same records in the same order,
    frequencies to 1e-12 of the file's largest (the eigenvalue parity is
    relative to the largest eigenvalue); weights summed over each run of equal
    frequencies (a degenerate multiplet of the excited block, whose
    individual eigenvectors are any rotation of each other, so only the
    multiplet's total is solver independent) to 1e-10 of the file's largest."""
    a = np.array([[float(t) for t in l.split()] for l in mine.splitlines()]).reshape(-1, 2)
    b = np.array([[float(t) for t in l.split()] for l in ref.splitlines()]).reshape(-1, 2)
    assert a.shape == b.shape, (name, a.shape, b.shape)
    if not len(b):
        return 0.0
    np.testing.assert_allclose(a[:, 0], b[:, 0], rtol=0, atol=1e-12 * np.abs(b[:, 0]).max())
    brk = np.ones(len(b), dtype=bool)
    brk[1:] = np.abs(np.diff(b[:, 0])) > 1e-9 * np.abs(b[1:, 0])
    run = np.cumsum(brk) - 1
    sa = np.bincount(run, weights=a[:, 1])
    sb = np.bincount(run, weights=b[:, 1])
    wmax = max(np.abs(sb).max(), 1e-300)
    err = float(np.abs(sa - sb).max() / wmax)
    assert err <= 1e-10, (name, err)
    return err


def compare_spectrum_file(name, mine, ref):
    pa = np.array([[float(t) for t in l.split()] for l in mine.splitlines()]).reshape(-1, 2)
    pb = np.array([[float(t) for t in l.split()] for l in ref.splitlines()]).reshape(-1, 2)
    for la, lb in zip(mine.splitlines()[:3], ref.splitlines()[:3]):
        assert shape(la) == shape(lb), (name, la, lb)
    peak = max(np.abs(pb[:, 1]).max(), 1e-300)
    keys_a = {round(x, 12): w for x, w in pa}
    keys_b = {round(x, 12): w for x, w in pb}
    only = set(keys_a) ^ set(keys_b)
    for k in only:   # bins at the listing threshold only
        w = keys_a.get(k, keys_b.get(k))
        assert abs(w) <= 1e-12 * peak, (name, k, w)
    common = sorted(set(keys_a) & set(keys_b))
    err = max(abs(keys_a[k] - keys_b[k]) for k in common)
    assert err <= 1e-8 * peak, (name, err, peak)
    return err / peak


@pytest.mark.parametrize("name", ["siam_nk256_it40", "siam_bcs_nk128_it12", "two_orbital_nk300_it20_spec"])
def test_reference_output_files(tmp_path, name):
    from nrgamd.host import Host
    fx = np.load(os.path.join(GOLD, f"ref_{name}.npz"))
    if name == "two_orbital_nk300_it20_spec":
        spec = models.workload("two_orbital_nk4000_it45", nkeep=300, niter=20, spectrum=True)
    else:
        spec = models.workload(name)
    d = str(tmp_path / "in")
    out = str(tmp_path / "out")
    os.makedirs(out)
    models.write_inputs(spec, d)
    h = Host(d, device=0)
    try:
        h.set_output(out)
        print(f"[{name}] forward", flush=True)
        h.run_forward()
        if spec.spectrum:
            print(f"[{name}] spectrum", flush=True)
            h.spectrum()
        h.set_output(None)
    finally:
        h.close()
    files = [str(x) for x in fx["output_files"]]
    assert any(f.startswith("CFS-spectrum") for f in files) == bool(spec.spectrum)
    assert any(f.startswith("step_") for f in files) == bool(spec.spectrum)
    for f in files:
        path = os.path.join(out, f)
        assert os.path.exists(path), f
        mine = open(path).read()
        ref = str(fx["file:" + f])
        if f.startswith("CFS-spectrum"):
            e = compare_spectrum_file(f, mine, ref)
        elif f.startswith("step_"):
            e = compare_step_file(f, mine, ref)
        elif f in ("Xn", "Sn", "Cn"):
            e = compare_records(f, mine, ref, 1e-10)
        else:
            # REAL(4) energies: a GPU/LAPACK difference at 1e-13 can move
            # the rounding of the last printed digit
            e = compare_records(f, mine, ref, 3e-7)
        print(f, "worst relative difference", e)
