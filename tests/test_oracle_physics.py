"""Physics pins of the input decks and the oracle, independent of the
reference: before the first truncation an NRG iteration diagonalises the
Wilson-chain Hamiltonian exactly, so its spectrum must equal

* exact diagonalisation of the chain Hamiltonian built here from
  Jordan-Wigner fermions (impurity + Wilson sites 0..N+1, in units of
  scale(N), shifted to a zero ground state);
* at U = 0 (and J = 0), the sums of occupied single-particle levels of the
  tridiagonal chain matrix (a closed form of the free-fermion spectrum).

Iteration N of the oracle assembles H from the previous eigenvalues times
sqrt(Lambda) and chain_nondiagonal(N) = t_N / scale(N) (diagonalization.F90:
166-237; read_initial_basis.F90:222-261), so these checks pin the deck
generator's operators, coefficient signs and hybridisation as well as the
oracle's assembly.  CPU only."""
import math

import numpy as np
import pytest

import nrg_oracle as O
from nrgamd import models


def jw_creators(nmodes):
    """c_m^dag on 2^nmodes states (scipy sparse), Jordan-Wigner strings to
    the left."""
    import scipy.sparse as sp
    I = sp.identity(2, format="csr")
    Z = sp.diags([1.0, -1.0], format="csr")
    up = sp.csr_matrix(np.array([[0.0, 0.0], [1.0, 0.0]]))   # |1><0|
    out = []
    for m in range(nmodes):
        a = sp.identity(1, format="csr")
        for x in [Z] * m + [up] + [I] * (nmodes - m - 1):
            a = sp.kron(a, x, format="csr")
        out.append(a)
    return out


def chain_hamiltonian(spec, N):
    """Impurity + Wilson sites 0..N+1 in units of scale(N).  Modes are
    (site, orbital, spin) with the impurity as site -1."""
    norb = spec.nchain
    nsite = N + 3                     # impurity and sites 0 .. N+1
    nm = 2 * norb * nsite
    cd = jw_creators(nm)
    c = [x.T for x in cd]

    def mode(site, orb, spin):        # site -1 = impurity
        return ((site + 1) * norb + orb) * 2 + spin

    n = {m: cd[m] @ c[m] for m in range(nm)}
    H = 0 * n[0]
    imp = [mode(-1, o, s) for o in range(norb) for s in range(2)]
    H += spec.eps_d * sum(n[m] for m in imp)
    if norb == 1:
        H += spec.U * n[imp[0]] @ n[imp[1]]
    else:
        U, J = spec.U, spec.J
        Up = U - 2.0 * J
        a_u, a_d, b_u, b_d = imp
        H += U * (n[a_u] @ n[a_d] + n[b_u] @ n[b_d])
        H += Up * (n[a_u] @ n[b_d] + n[a_d] @ n[b_u])
        H += (Up - J) * (n[a_u] @ n[b_u] + n[a_d] @ n[b_d])
        sf = cd[a_u] @ c[a_d] @ cd[b_d] @ c[b_u]
        H += -J * (sf + sf.T)
        ph = cd[a_u] @ cd[a_d] @ c[b_d] @ c[b_u]
        H += J * (ph + ph.T)
    V = math.sqrt(2.0 * spec.D * spec.Gamma / math.pi)
    for o in range(norb):
        for s in range(2):
            hop = [(mode(-1, o, s), mode(0, o, s), V)]
            hop += [(mode(k, o, s), mode(k + 1, o, s), models.wilson_hopping(k, spec.Lam, spec.D))
                    for k in range(0, N + 1)]
            for a, b, t in hop:
                h = cd[a] @ c[b]
                H += t * (h + h.T)
    return H.toarray() / models.scale_wilson(N, spec.Lam, spec.D, spec.z)


def single_particle_levels(spec, N):
    """Levels of the U = J = 0 chain (impurity level eps_d, hybridisation V,
    Wilson hoppings t_0..t_N) per orbital and spin, in units of scale(N)."""
    V = math.sqrt(2.0 * spec.D * spec.Gamma / math.pi)
    t = [V] + [models.wilson_hopping(k, spec.Lam, spec.D) for k in range(0, N + 1)]
    T = np.diag(np.r_[spec.eps_d, np.zeros(N + 2)]) + np.diag(t, 1) + np.diag(t, -1)
    e = np.linalg.eigvalsh(T) / models.scale_wilson(N, spec.Lam, spec.D, spec.z)
    return np.tile(e, 2 * spec.nchain)


def oracle_spectra(tmp_path, spec, upto):
    d = str(tmp_path / spec.name)
    models.write_inputs(spec, d)
    nrg = O.NRG(O.read_inputs(d))
    out = {}
    for it in range(spec.first_iteration, upto + 1):
        rec = nrg.iteration(it)
        out[it] = np.sort(rec.eigenvalue)
    return out


def shifted(e):
    e = np.sort(e)
    return e - e[0]


@pytest.mark.parametrize("U,eps", [(0.2, -0.1), (0.5, -0.15)])
def test_siam_equals_exact_diagonalisation(tmp_path, U, eps):
    # Nkeep=256: iterations -1..2 (16, 64, 256, 1024 states) are solved
    # before any truncation reaches them
    spec = models.siam(nkeep=256, niter=6, spectrum=False, U=U, eps_d=eps)
    nrg = oracle_spectra(tmp_path, spec, 2)
    for it, e in nrg.items():
        ed = shifted(np.linalg.eigvalsh(chain_hamiltonian(spec, it)))
        assert len(ed) == len(e), it
        err = np.max(np.abs(e - ed)) / max(1.0, ed.max())
        assert err <= 1e-11, (it, err)


def test_two_orbital_equals_exact_diagonalisation(tmp_path):
    # Kanamori interaction with Hund's coupling: iterations -1 and 0 (256
    # and 4096 states), untruncated with Nkeep = 5000
    spec = models.two_orbital(nkeep=5000, niter=4, spectrum=False)
    nrg = oracle_spectra(tmp_path, spec, 0)
    for it, e in nrg.items():
        ed = shifted(np.linalg.eigvalsh(chain_hamiltonian(spec, it)))
        assert len(ed) == len(e), it
        err = np.max(np.abs(e - ed)) / max(1.0, ed.max())
        assert err <= 1e-11, (it, err)


@pytest.mark.parametrize("nchain", [1, 2])
def test_free_fermion_levels_at_u0(tmp_path, nchain):
    """At U = J = 0 the many-body levels are sums of occupied single-particle
    levels of the chain: all 2^M subsets of the M levels."""
    if nchain == 1:
        spec = models.siam(nkeep=256, niter=6, spectrum=False, U=0.0, eps_d=0.0)
        upto = 2
    else:
        spec = models.two_orbital(nkeep=5000, niter=4, spectrum=False, U=0.0, J=0.0, eps_d=0.0)
        upto = 0
    nrg = oracle_spectra(tmp_path, spec, upto)
    for it, e in nrg.items():
        lev = single_particle_levels(spec, it)
        M = len(lev)
        # all subset sums, by doubling: sums over the first k levels
        sums = np.zeros(1)
        for x in lev:
            sums = np.concatenate((sums, sums + x))
        ref = shifted(sums)
        assert len(ref) == len(e) == 2 ** M, it
        err = np.max(np.abs(e - ref)) / max(1.0, ref.max())
        assert err <= 1e-11, (it, err)


def test_exact_diagonalisation_detects_a_sign_error(tmp_path):
    """The check is sharp: the oracle with the sign of one hopping row
    flipped (a plausible generator mistake) no longer matches ED."""
    spec = models.siam(nkeep=256, niter=6, spectrum=False)
    d = str(tmp_path / "flip")
    models.write_inputs(spec, d)
    m = O.read_inputs(d)
    # one hopping row with the wrong sign
    m.coef_diag = m.coef_diag.copy()
    m.coef_diag[0] = -m.coef_diag[0]
    nrg = O.NRG(m)
    errs = []
    for it in range(-1, 2):
        e = np.sort(nrg.iteration(it).eigenvalue)
        ed = shifted(np.linalg.eigvalsh(chain_hamiltonian(spec, it)))
        errs.append(np.max(np.abs(e - ed)))
    assert max(errs) > 1e-6
