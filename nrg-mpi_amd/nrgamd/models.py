"""Input-deck generator for the NRG hot path.

The reference program reads three files whose generator is not part of the
reference repository (``read_initial_basis.F90:113`` says the eigenvalues were
"scaled already in Java code").  This module writes them from first
principles for the two models the workloads use:

* the single-impurity Anderson model (SIAM): one Wilson chain, local site of
  dimension 4, conserved labels (Q, 2Sz);
* the two-orbital Anderson model with Kanamori interaction (U, U'=U-2J, J):
  two Wilson chains (one per orbital), local site of dimension 16, labels
  (Q, 2Sz).

File formats follow the readers they feed:

* ``hamiltonianInfo.dat`` -- namelist ``/hamiltonianInfo/`` (``module.F90:846-861``)
* ``parameter.txt``       -- namelist ``/nrg/`` (``module.F90:665-675``)
* ``initial.dat``         -- list-directed records (``read_initial_basis.F90:146-302``)

Conventions (all derived here, none copied):

* A product state of iteration N+1 is ``|r, s> = c_s^dag |r>`` where ``c_s^dag``
  creates the local configuration ``s`` of the new site(s) to the left of all
  older operators.  Labels add, and because ``prepareBasis.F90:47-48`` forms
  ``basis_input = basis_output - operation``, ``operation(s, 1:nv)`` stores the
  *negated* label shift of ``s``; columns ``nv+c`` hold the electron count of
  channel ``c`` (used with the on-site energies, ``diagonalization.F90:167-171``).
* ``invariant_matrix(r, l, t) = <l| f_t^dag |r>`` (``invariantMatrixDGEMM.F90:330``).
* Hopping table rows ``(op_l, op_r, t, channel, coef)`` carry
  ``coef = <s_l| f_t |s_r> (-1)^{n(s_l)}``: the matrix element of
  ``f_{N,t}^dag f_{N+1,t}`` between product states is ``coef * <r_l|f_t^dag|r_r>``
  (the reversed match at ``diagonalization.F90:202-211`` supplies the h.c.).
* Invariant-matrix table rows ``(op_a, op_b, t, coef)`` carry
  ``coef = <s_a| f_t^dag |s_b>`` (``invariantMatrixDGEMM.F90:248-253``).
* Spectral coefficients are the fermionic parity ``(-1)^{n(s)}`` picked up when
  the impurity ``d^dag`` passes the new-site operators
  (``invariantMatrixForSpectrumDGEMM.F90:251-252``).

Energy scaling (``scalingType = 1``, ``module.F90:699-703`` and
``read_initial_basis.F90:246-258``): iteration ``it`` works in units of
``scale(it)``.  The reader rescales ``chain_*(n)`` for ``n >= 0`` only, so the
impurity-to-site-0 hybridisation ``chain_nondiagonal(-1)`` is written already
divided by ``scale(-1)``, and the impurity eigenvalues are written divided by
``scale(-2)`` because the first H-assembly multiplies them by ``sqrt(Lambda)``
(``diagonalization.F90:166``).
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import List, Sequence, Tuple

import numpy as np


# --------------------------------------------------------------------------
# Fock-space helpers
# --------------------------------------------------------------------------

def _popcount(x: int) -> int:
    return bin(x).count("1")


def _create(mask: int, mode: int) -> Tuple[int, int]:
    """Apply c_mode^dag to the occupation-number state |mask>.

    Ordering convention: |mask> = prod_{set bits, ascending} c^dag |0>, the
    lowest mode leftmost.  Returns (sign, new_mask); sign 0 if occupied.
    """
    if mask >> mode & 1:
        return 0, mask
    below = _popcount(mask & ((1 << mode) - 1))
    return (-1 if below & 1 else 1), mask | (1 << mode)


def _annihilate(mask: int, mode: int) -> Tuple[int, int]:
    if not (mask >> mode & 1):
        return 0, mask
    below = _popcount(mask & ((1 << mode) - 1))
    return (-1 if below & 1 else 1), mask & ~(1 << mode)


def _fock_ops(nmodes: int) -> List[np.ndarray]:
    """Dense creation matrices c_m^dag on the 2^nmodes Fock space."""
    dim = 1 << nmodes
    ops = []
    for m in range(nmodes):
        c = np.zeros((dim, dim))
        for s in range(dim):
            sg, t = _create(s, m)
            if sg:
                c[t, s] = sg
        ops.append(c)
    return ops


@dataclasses.dataclass
class LocalSite:
    """The site added at every NRG step (all channels together)."""

    nchain: int                     # channels (orbitals)
    modes: List[Tuple[int, int]]    # (channel 0-based, spin +1/-1) per mode

    @property
    def nmodes(self) -> int:
        return len(self.modes)

    @property
    def nops(self) -> int:
        return 1 << self.nmodes

    def q(self, s: int) -> int:
        # charge measured from half filling of the site
        return _popcount(s) - self.nmodes // 2

    def twosz(self, s: int) -> int:
        return sum(sp for m, (_, sp) in enumerate(self.modes) if s >> m & 1)

    def nchan(self, s: int, c: int) -> int:
        return sum(1 for m, (ch, _) in enumerate(self.modes) if ch == c and s >> m & 1)

    def operation(self, sz_only: bool = False) -> np.ndarray:
        """operation(nops, nv+nchain), op index = mask+1."""
        rows = []
        for s in range(self.nops):
            lab = [-self.twosz(s)] if sz_only else [-self.q(s), -self.twosz(s)]
            rows.append(lab + [self.nchan(s, c) for c in range(self.nchain)])
        return np.array(rows, dtype=np.int64)

    def hopping_table(self):
        """Rows (op_l, op_r, type, channel, coef), coef=<s_l|f_t|s_r>(-1)^{n(s_l)}."""
        rows = []
        for t, (ch, _) in enumerate(self.modes):
            for sl in range(self.nops):
                sg, sr = _create(sl, t)      # f_t |sr> = sg' |sl>  <=>  f_t^dag |sl> = sg |sr>
                if not sg:
                    continue
                # <sl|f_t|sr> = <sr|f_t^dag|sl> = sg (real)
                coef = sg * (-1 if _popcount(sl) & 1 else 1)
                rows.append((sl + 1, sr + 1, t + 1, ch + 1, float(coef)))
        return rows

    def invariant_table(self):
        """Rows (op_a, op_b, type, coef), coef = <s_a|f_t^dag|s_b>."""
        rows = []
        for t in range(self.nmodes):
            for sb in range(self.nops):
                sg, sa = _create(sb, t)
                if not sg:
                    continue
                rows.append((sa + 1, sb + 1, t + 1, float(sg)))
        return rows

    def spectral_coefficients(self, nmat: int) -> np.ndarray:
        """coefficient_invariant_matrix_spectrum(nMatrix, nops) = (-1)^{n(s)}."""
        par = np.array([(-1.0 if _popcount(s) & 1 else 1.0) for s in range(self.nops)])
        return np.tile(par, (nmat, 1))


# --------------------------------------------------------------------------
# Wilson chain
# --------------------------------------------------------------------------

def wilson_hopping(n: int, Lam: float, D: float = 1.0) -> float:
    """Flat-band Wilson chain hopping t_n (Krishna-murthy, Wilkins, Wilson 1980)."""
    num = (1.0 - Lam ** (-n - 1))
    den = math.sqrt((1.0 - Lam ** (-2 * n - 1)) * (1.0 - Lam ** (-2 * n - 3)))
    return 0.5 * D * (1.0 + 1.0 / Lam) * num / den * Lam ** (-0.5 * n)


def scale_wilson(N: int, Lam: float, D: float, z: float) -> float:
    """Energy scale of iteration N for scalingType 1 (``module.F90:699-703``)."""
    return D * 0.5 * (1.0 + 1.0 / Lam) * Lam ** (-0.5 * N - z + 1.0)


# --------------------------------------------------------------------------
# Model specification
# --------------------------------------------------------------------------

@dataclasses.dataclass
class ModelSpec:
    name: str
    nchain: int
    U: float
    Gamma: float
    eps_d: float
    J: float = 0.0
    Lam: float = 2.0
    D: float = 1.0
    z: float = 1.0
    nkeep: int = 256
    niter: int = 40                 # total number of NRG iterations (loop i0..N)
    first_iteration: int = -1
    spectrum: bool = False
    # on-site pairing Delta of a superconducting (BCS) Wilson chain; nonzero
    # writes a flag_BCS deck whose blocks carry 2Sz only (charge is not
    # conserved), chain_bcs(n) = Delta / scale(n-1) and BCS coefficient rows
    # joining the empty and doubly occupied configurations of the new site
    bcs_delta: float = 0.0
    # coefficient-table probe: a reversed copy (half the coefficient) of the
    # first hopping row is written after it (and a duplicate of the first
    # invariant row, the tables share numberOfCoefficient), so that for the
    # pair that occurs a reversed match precedes a forward one
    h3_reversed_row: bool = False

    @property
    def last_iteration(self) -> int:
        return self.first_iteration + self.niter - 1


def siam(nkeep=256, niter=40, spectrum=True, U=0.2, Gamma=0.01, eps_d=None) -> ModelSpec:
    if eps_d is None:
        eps_d = -U / 2
    return ModelSpec("siam", 1, U, Gamma, eps_d, nkeep=nkeep, niter=niter, spectrum=spectrum)


def two_orbital(nkeep=4000, niter=45, spectrum=False, U=0.3, J=0.05, Gamma=0.01, eps_d=None) -> ModelSpec:
    if eps_d is None:
        # particle-hole symmetric point of the two-orbital Kanamori model
        eps_d = -(3.0 * U - 5.0 * J) / 2.0
    return ModelSpec("two_orbital", 2, U, Gamma, eps_d, J=J, nkeep=nkeep, niter=niter, spectrum=spectrum)


def siam_bcs(nkeep=128, niter=12, U=0.2, Gamma=0.01, delta=0.02) -> ModelSpec:
    """SIAM coupled to a superconducting Wilson chain (on-site pairing delta)."""
    return ModelSpec("siam", 1, U, Gamma, -U / 2, nkeep=nkeep, niter=niter, spectrum=False, bcs_delta=delta)


def siam_h3(nkeep=128, niter=12) -> ModelSpec:
    """SIAM deck in which a pair's reversed hopping row precedes a forward row."""
    return ModelSpec("siam", 1, 0.2, 0.01, -0.1, nkeep=nkeep, niter=niter, spectrum=False, h3_reversed_row=True)


WORKLOADS = {
    # name: (factory, kwargs) -- sizes are the ones the workloads are quoted on
    "siam_nk256_it40": (siam, dict(nkeep=256, niter=40, spectrum=True)),
    "siam_nk2048_it50": (siam, dict(nkeep=2048, niter=50, spectrum=False)),
    "two_orbital_nk4000_it45": (two_orbital, dict(nkeep=4000, niter=45, spectrum=False)),
    "two_orbital_nk8000_it45": (two_orbital, dict(nkeep=8000, niter=45, spectrum=True)),
    # decks exercising input features of the reference the workloads above
    # leave unused: the BCS pairing override and the coefficient-search order
    "siam_bcs_nk128_it12": (siam_bcs, dict(nkeep=128, niter=12)),
    "siam_h3_nk128_it12": (siam_h3, dict(nkeep=128, niter=12)),
}


def workload(name: str, **override) -> ModelSpec:
    fac, kw = WORKLOADS[name]
    kw = dict(kw)
    kw.update(override)
    return fac(**kw)


# --------------------------------------------------------------------------
# Impurity
# --------------------------------------------------------------------------

def _impurity_modes(spec: ModelSpec):
    modes = []
    for c in range(spec.nchain):
        modes += [(c, +1), (c, -1)]
    return modes


def impurity_hamiltonian(spec: ModelSpec) -> np.ndarray:
    """Many-body impurity Hamiltonian on the 2^(2*nchain) Fock space."""
    modes = _impurity_modes(spec)
    cd = _fock_ops(len(modes))
    c = [m.T for m in cd]
    n = [cd[m] @ c[m] for m in range(len(modes))]
    dim = cd[0].shape[0]
    H = spec.eps_d * sum(n)
    if spec.nchain == 1:
        H = H + spec.U * n[0] @ n[1]
        return H
    # Kanamori: orbitals a,b; modes (a up, a dn, b up, b dn) = 0,1,2,3
    U, J = spec.U, spec.J
    Up = U - 2.0 * J
    H = H + U * (n[0] @ n[1] + n[2] @ n[3])
    H = H + Up * (n[0] @ n[3] + n[1] @ n[2])          # opposite spins, different orbitals
    H = H + (Up - J) * (n[0] @ n[2] + n[1] @ n[3])    # same spin, different orbitals
    # spin flip  -J (a_up^dag a_dn b_dn^dag b_up + h.c.)
    sf = cd[0] @ c[1] @ cd[3] @ c[2]
    H = H - J * (sf + sf.T)
    # pair hopping  +J (a_up^dag a_dn^dag b_dn b_up + h.c.)
    ph = cd[0] @ cd[1] @ c[3] @ c[2]
    H = H + J * (ph + ph.T)
    assert np.allclose(H, H.T)
    return H


def impurity_basis(spec: ModelSpec):
    """Diagonalise the impurity per (Q, 2Sz) block.

    Returns a list of (labels, reference, energy) sorted by block label then
    energy, the global eigenvector matrix (Fock x states) and the creation
    operators in the eigenbasis, ``ops[t][l, r] = <l|d_t^dag|r>``.
    """
    modes = _impurity_modes(spec)
    nm = len(modes)
    site = LocalSite(spec.nchain, modes)
    H = impurity_hamiltonian(spec)
    dim = 1 << nm
    labels = {}
    for s in range(dim):
        lab = (site.twosz(s),) if spec.bcs_delta else (site.q(s), site.twosz(s))
        labels.setdefault(lab, []).append(s)
    states = []
    vecs = []
    for lab in sorted(labels):
        idx = labels[lab]
        hb = H[np.ix_(idx, idx)]
        w, v = np.linalg.eigh(hb)
        for k in range(len(idx)):
            col = np.zeros(dim)
            col[idx] = v[:, k]
            # deterministic sign: largest component positive
            j = int(np.argmax(np.abs(col)))
            if col[j] < 0:
                col = -col
            states.append((lab, k + 1, float(w[k])))
            vecs.append(col)
    V = np.array(vecs).T
    cd = _fock_ops(nm)
    ops = [V.T @ cdm @ V for cdm in cd]
    return states, ops, site


# --------------------------------------------------------------------------
# Writer
# --------------------------------------------------------------------------

def _fmt(x: float) -> str:
    return repr(float(x))


def write_inputs(spec: ModelSpec, outdir: str) -> dict:
    """Write hamiltonianInfo.dat, parameter.txt and initial.dat into outdir."""
    os.makedirs(outdir, exist_ok=True)
    states, ops, imp_site = impurity_basis(spec)
    modes = _impurity_modes(spec)
    site = LocalSite(spec.nchain, modes)          # Wilson sites have the same modes
    nb = len(states)
    nops = site.nops
    ncm = site.nmodes
    nmat = site.nmodes
    hop = site.hopping_table()
    inv = site.invariant_table()
    assert len(hop) == len(inv)
    if spec.h3_reversed_row:
        # inside a block the product state made by op a (one electron less
        # on the new site than op b) always has the larger reference, so the
        # pair that occurs is (op_i, op_j) = (b, a): row (a, b) is a reversed
        # match for it; the copy (b, a) written after it is a forward match
        # (last match: coef/2) and must leave the references swapped
        a, b, t, ch, c = hop[0]
        hop = [hop[0], (b, a, t, ch, 0.5 * c)] + hop[1:]
        inv = [inv[0], inv[0]] + inv[1:]
    ncoef = len(hop)
    i0, N = spec.first_iteration, spec.last_iteration
    bcs = bool(spec.bcs_delta)
    nv = 1 if bcs else 2
    if bcs:
        # pair (empty, doubly occupied) of the new site, both orders: the
        # reference tests the forward order only (diagonalization.F90:217)
        full = nops
        bcs_rows = [(1, full, 1), (full, 1, 1)]

    with open(os.path.join(outdir, "hamiltonianInfo.dat"), "w") as f:
        f.write("&hamiltonianInfo\n")
        f.write(f" numberOfIteration = {N}\n")
        f.write(f" numberOfChain = {spec.nchain}\n")
        f.write(f" numberOfVariation = {nv}\n")
        # charge range 0..0 selects the 2Sz-only pair rule (invariantMatrixDGEMM.F90:168-176)
        f.write(" conservationBlock = 0, 1, 0, 1\n" if bcs else " conservationBlock = 1, 2, 1, 2\n")
        f.write(f" numberOfMatrix = {nmat}\n")
        f.write(f" numberOfConductionMatrix = {ncm}\n")
        f.write(f" numberOfBasis = {nb}\n")
        f.write(f" numberOfCoefficient = {ncoef}\n")
        f.write(f" numberOfBCSCoefficient = {len(bcs_rows) if bcs else 0}\n")
        f.write(" numberOfCoefficientImpurity = 0\n")
        f.write(f" numberOfOperation = {nops}\n")
        f.write(f" indexOfFirstIteration = {i0}\n")
        f.write(f" flag_spectrum = {'.true.' if spec.spectrum else '.false.'}\n")
        f.write(" flag_local = .false.\n")
        f.write(f" flag_BCS = {'.true.' if bcs else '.false.'}\n")
        f.write("/\n")

    with open(os.path.join(outdir, "parameter.txt"), "w") as f:
        f.write("&nrg\n")
        f.write(f" bandWidth = {_fmt(spec.D)}\n")
        f.write(f" ztrick = {_fmt(spec.z)}\n")
        f.write(f" Lambda = {_fmt(spec.Lam)}\n")
        f.write(" scalingType = 1\n")
        f.write(f" truncation = {spec.nkeep}\n")
        f.write(" cfsnrg = .false.\n")
        f.write(" noimpurity = .false.\n")
        f.write(" loadRestart = .false.\n")
        f.write(" saveRestart = .false.\n")
        f.write("/\n")

    s_m2 = scale_wilson(i0 - 1, spec.Lam, spec.D, spec.z)
    s_m1 = scale_wilson(i0, spec.Lam, spec.D, spec.z)
    V = math.sqrt(2.0 * spec.D * spec.Gamma / math.pi)
    L = []
    L.append("basis")
    for (lab, ref, e) in states:
        L.append(" ".join(str(x) for x in lab) + f" {ref} {_fmt(e / s_m2)}")
    L.append("invariant_matrix")
    for t in range(ncm):
        for j in range(nb):
            for k in range(nb):
                L.append(_fmt(ops[t][j, k]))       # invariant_matrix(k,j,t) = <j|d_t^dag|k>
    L.append("coefficient_invariant_matrix")
    for (a, b, t, c) in inv:
        L.append(f"{a} {b} {t} {_fmt(c)}")
    L.append("coefficient_diagonalization")
    for (a, b, t, ch, c) in hop:
        L.append(f"{a} {b} {t} {ch} {_fmt(c)}")
    L.append("operation")
    for row in site.operation(sz_only=bcs):
        L.append(" ".join(str(int(x)) for x in row))
    L.append("chain")
    for c in range(spec.nchain):
        L.append(f"diagonal{c + 1}")
        for n in range(0, N + 2):
            L.append(_fmt(0.0))
    for c in range(spec.nchain):
        L.append(f"nondiagonal{c + 1}")
        for n in range(i0, N + 2):
            if n == i0:
                L.append(_fmt(V / s_m1))
            else:
                L.append(_fmt(wilson_hopping(n, spec.Lam, spec.D)))
    if spec.spectrum:
        L.append("coefficient_spectrum")
        cs = site.spectral_coefficients(nmat)
        for im in range(nmat):
            for s in range(nops):
                L.append(_fmt(cs[im, s]))
        L.append("invariant_matrix_spectrum")
        for im in range(nmat):
            for j in range(nb):
                for i in range(nb):
                    L.append(_fmt(ops[im][j, i]))  # (i,j,imat) = <j|d^dag|i>
        L.append("conservation_difference_spectrum")
        for (ch, sp) in modes:
            L.append(f"1 {sp}")
        L.append("positive_spectrum_operation")
        for (ch, sp) in modes:
            L.append(f"1 {sp}")
        L.append("negative_spectrum_operation")
        for (ch, sp) in modes:
            L.append(f"-1 {-sp}")
    if bcs:
        assert not spec.spectrum, "spectra of BCS decks are not generated"
        # read raw, without the chain rescaling (read_initial_basis.F90:185-201)
        for c in range(spec.nchain):
            L.append(f"chain_bcs{c + 1}")
            for n in range(0, N + 2):
                L.append(_fmt(spec.bcs_delta / scale_wilson(n - 1, spec.Lam, spec.D, spec.z)))
        L.append("coefficient_BCS")
        for (a, b, ch) in bcs_rows:
            L.append(f"{a} {b} {ch}")
    with open(os.path.join(outdir, "initial.dat"), "w") as f:
        f.write("\n".join(L) + "\n")
    return dict(nb=nb, nops=nops, ncm=ncm, nmat=nmat, ncoef=ncoef, i0=i0, N=N)


if __name__ == "__main__":  # pragma: no cover
    import argparse
    ap = argparse.ArgumentParser(description="write an NRG input deck")
    ap.add_argument("workload", choices=sorted(WORKLOADS))
    ap.add_argument("outdir")
    ap.add_argument("--niter", type=int)
    ap.add_argument("--nkeep", type=int)
    ap.add_argument("--spectrum", choices=["on", "off"])
    a = ap.parse_args()
    ov = {}
    if a.niter:
        ov["niter"] = a.niter
    if a.nkeep:
        ov["nkeep"] = a.nkeep
    if a.spectrum:
        ov["spectrum"] = a.spectrum == "on"
    print(write_inputs(workload(a.workload, **ov), a.outdir))
