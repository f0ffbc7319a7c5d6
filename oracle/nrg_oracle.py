"""CPU restatement of the reference NRG hot path (numpy) -- TEST INFRASTRUCTURE.

This module is the oracle.  Only tests, the smoke check and the CPU-baseline
leg of the benchmark may import it, and only as the checker.  The product
path (Fortran host + HIP library) never calls it.

It restates, step for step, what the reference Fortran program computes
(file:line citations refer to the reference tree):

* ``read_inputs``            -- ``module.F90:656-688, 826-944``; ``read_initial_basis.F90:108-302``
* ``prepare_basis``          -- ``prepareBasis.F90:41-114``; exchange sort ``module.F90:531-602``
* ``hamiltonian_block``      -- ``diagonalization.F90:153-263``
* ``diagonalize``            -- ``diagonalization.F90:66-407`` (``dsyev`` -> LAPACK via numpy)
* ``postprocess``            -- ``postprocess.F90:71-217``; sorts ``module.F90:16-70``
* ``invariant_matrix``       -- ``invariantMatrixDGEMM.F90:67-332``
* ``invariant_matrix_spec``  -- ``invariantMatrixForSpectrumDGEMM.F90:71-370``
* ``stock``                  -- ``postprocess.F90:221-374``
* ``init_reduced_density``   -- ``initializeReducedDensity.F90:30-78``
* ``reduced_density``        -- ``reducedDensityMatrixDGEMM.F90:62-321``
* ``spike_zero_temperature`` -- ``calculateSpectrumDGEMM.F90:55-212``
* ``spike_cfs``              -- ``calculateSpectrumDGEMM.F90:296-643``
* ``spectrum_element``       -- ``module.F90:1014-1125``
* ``bin_spike``              -- ``module.F90:1210-1254``
* ``calc_thermo``            -- ``module.F90:1931-1982``

Arrays use 0-based numpy indexing; Fortran's 1-based integers stored in the
records (reference, variation, operation, start offsets) are kept 1-based.

Parity pins (see tests/test_oracle_*.py): the reference run fixtures under
tests/golden/ (block tables, kept counts, basis-record checksums, all
eigenvalues, binned spectra, reduced-density traces/spectra); brute-force
exact diagonalisation of the first iterations; U=0 free-fermion levels; the
spectral sum rule; Tr rho = 1; the binning formula on hand cases.
"""
from __future__ import annotations

import dataclasses
import math
import os
import re
from typing import Dict, List, Optional

import numpy as np

NUMERICAL_THRESHOLD = 1.0e-12       # module.F90:5
EIGENVECTOR_THRESHOLD = 1.0e-14     # module.F90:6
DEGENERACY_THRESHOLD = 1.0e-8       # module.F90:7
TRUNCATION_THRESHOLD = 1.0e-6       # module.F90:629
TRUNCATION_LIMIT_SCALE = 1.2        # module.F90:630
BETABAR = float(np.float32(0.8))    # module.F90:652 (a default-real literal)
MESH_OF_BIN = 1000                  # module.F90:964
LOG10_EMIN = -15.0                  # module.F90:970


# ---------------------------------------------------------------------------
# inputs
# ---------------------------------------------------------------------------

def _namelist(path: str) -> Dict[str, str]:
    txt = open(path).read()
    body = txt[txt.index("&"):]
    body = body[body.index("\n"):]
    body = body[:body.rindex("/")]
    out = {}
    for m in re.finditer(r"(\w+)\s*=\s*([^=]*?)(?=\s*\w+\s*=|\Z)", body, re.S):
        out[m.group(1).lower()] = m.group(2).strip().rstrip(",")
    return out


def _flt(s):
    return float(s.replace("d", "e").replace("D", "e"))


def _bool(s):
    return s.strip().lower() in (".true.", "t", ".t.", "true")


@dataclasses.dataclass
class Model:
    N: int
    i0: int
    nchain: int
    nv: int
    cons: np.ndarray            # conservationBlock(2,2) as [[chargeStart, chargeEnd],[spinStart, spinEnd]]
    nmat: int
    ncm: int
    nb0: int
    ncoef: int
    nops: int
    spectrum: bool
    D: float
    z: float
    Lam: float
    scaling: int
    nkeep: int
    basis_output0: np.ndarray   # [nb0, nv+3]
    eigenvalue0: np.ndarray
    invariant0: np.ndarray      # [ncm, l, r]
    coef_inv_type: np.ndarray   # [ncoef, 3]
    coef_inv: np.ndarray
    coef_diag_type: np.ndarray  # [ncoef, 4]
    coef_diag: np.ndarray
    operation: np.ndarray       # [nops, nv+nchain]
    chain_diag: np.ndarray      # [N+2, nchain], index n
    chain_nondiag: np.ndarray   # [N+2-i0, nchain], index n-i0
    coef_spec: Optional[np.ndarray] = None   # [nmat, nops]
    spec0: Optional[np.ndarray] = None       # [nmat, l, r]
    cons_diff_spec: Optional[np.ndarray] = None
    pos_op: Optional[np.ndarray] = None
    neg_op: Optional[np.ndarray] = None
    coef_bcs: Optional[np.ndarray] = None    # [nbcs, 3] (op_i, op_j, channel), 1-based
    chain_bcs: Optional[np.ndarray] = None   # [N+2, nchain], index n, raw

    def scale(self, n: int) -> float:
        """module.F90:690-711"""
        if self.scaling == 0:
            return self.Lam ** (-n * 0.5)
        if self.scaling == 1:
            return self.D * 0.5 * (1.0 + 1.0 / self.Lam) * self.Lam ** (-n * 0.5 - self.z + 1.0)
        return self.Lam ** (-n * 0.5 - self.z) * self.D * (self.Lam - 1.0) / math.log(self.Lam)

    def cnd(self, n: int, c: int) -> float:
        return self.chain_nondiag[n - self.i0, c]


def read_inputs(d: str) -> Model:
    hi = _namelist(os.path.join(d, "hamiltonianInfo.dat"))
    pa = _namelist(os.path.join(d, "parameter.txt"))
    N = int(hi["numberofiteration"])
    nchain = int(hi["numberofchain"])
    nv = int(hi["numberofvariation"])
    cb = [int(x) for x in re.split(r"[,\s]+", hi["conservationblock"].strip()) if x]
    cons = np.array(cb).reshape(2, 2, order="F")
    nmat = int(hi["numberofmatrix"])
    ncm = int(hi["numberofconductionmatrix"])
    nb = int(hi["numberofbasis"])
    ncoef = int(hi["numberofcoefficient"])
    nops = int(hi["numberofoperation"])
    i0 = int(hi["indexoffirstiteration"])
    spectrum = _bool(hi["flag_spectrum"])
    bcs = _bool(hi.get("flag_bcs", ".false."))
    nbcs = int(hi.get("numberofbcscoefficient", "0"))
    D, z, Lam = _flt(pa["bandwidth"]), _flt(pa["ztrick"]), _flt(pa["lambda"])
    scaling, nkeep = int(pa["scalingtype"]), int(pa["truncation"])

    tok = open(os.path.join(d, "initial.dat")).read().split("\n")
    lines = [l for l in tok if l.strip() != ""]
    p = 0

    def hdr():
        nonlocal p
        p += 1

    def row():
        nonlocal p
        r = lines[p].split()
        p += 1
        return r

    hdr()
    bo = np.zeros((nb, nv + 3), dtype=np.int64)
    e0 = np.zeros(nb)
    for i in range(nb):
        r = row()
        bo[i, :nv] = [int(x) for x in r[:nv]]
        bo[i, nv] = int(r[nv])
        e0[i] = _flt(r[nv + 1])
    hdr()
    inv = np.zeros((ncm, nb, nb))
    for t in range(ncm):
        for j in range(nb):
            for k in range(nb):
                inv[t, j, k] = _flt(row()[0])      # invariant_matrix(k,j,t)
    hdr()
    cit = np.zeros((ncoef, 3), dtype=np.int64)
    ci = np.zeros(ncoef)
    for i in range(ncoef):
        r = row()
        cit[i] = [int(x) for x in r[:3]]
        ci[i] = _flt(r[3])
    hdr()
    cdt = np.zeros((ncoef, 4), dtype=np.int64)
    cd = np.zeros(ncoef)
    for i in range(ncoef):
        r = row()
        cdt[i] = [int(x) for x in r[:4]]
        cd[i] = _flt(r[4])
    hdr()
    op = np.zeros((nops, nv + nchain), dtype=np.int64)
    for i in range(nops):
        op[i] = [int(x) for x in row()[:nv + nchain]]
    hdr()
    chd = np.zeros((N + 2, nchain))
    for c in range(nchain):
        hdr()
        for n in range(0, N + 2):
            chd[n, c] = _flt(row()[0])
    chn = np.zeros((N + 2 - i0, nchain))
    for c in range(nchain):
        hdr()
        for n in range(i0, N + 2):
            chn[n - i0, c] = _flt(row()[0])
    # chain rescaling, read_initial_basis.F90:222-261 (indices >= 0 only)
    for c in range(nchain):
        for n in range(0, N + 2):
            if scaling == 0:
                chd[n, c] *= Lam ** ((n - 1) * 0.5)
                chn[n - i0, c] *= Lam ** (n * 0.5)
            elif scaling == 2:
                f = math.log(Lam) / D / (Lam - 1.0)
                chd[n, c] *= Lam ** ((n - 1) * 0.5 + z) * f
                chn[n - i0, c] *= Lam ** (n * 0.5 + z) * f
            else:
                f = 1.0 / D * 2.0 / (1.0 + 1.0 / Lam)
                chd[n, c] *= Lam ** ((n - 1) * 0.5 + z - 1.0) * f
                chn[n - i0, c] *= Lam ** (n * 0.5 + z - 1.0) * f
    m = Model(N, i0, nchain, nv, cons, nmat, ncm, nb, ncoef, nops, spectrum, D, z, Lam,
              scaling, nkeep, bo, e0, inv, cit, ci, cdt, cd, op, chd, chn)
    if spectrum:
        hdr()
        cs = np.zeros((nmat, nops))
        for im in range(nmat):
            for i in range(nops):
                cs[im, i] = _flt(row()[0])
        hdr()
        sp = np.zeros((nmat, nb, nb))
        for im in range(nmat):
            for j in range(nb):
                for i in range(nb):
                    sp[im, j, i] = _flt(row()[0])   # (i,j,imat)
        hdr()
        cds = np.array([[int(x) for x in row()[:nv]] for _ in range(nmat)])
        hdr()
        po = np.array([[int(x) for x in row()[:nv]] for _ in range(nmat)])
        hdr()
        no = np.array([[int(x) for x in row()[:nv]] for _ in range(nmat)])
        m.coef_spec, m.spec0, m.cons_diff_spec, m.pos_op, m.neg_op = cs, sp, cds, po, no
    if bcs:
        # read_initial_basis.F90:185-201: chain_bcs(0:N+1) per chain, not
        # rescaled, then the coefficient_BCS rows
        cb = np.zeros((N + 2, nchain))
        for c in range(nchain):
            hdr()
            for n in range(0, N + 2):
                cb[n, c] = _flt(row()[0])
        hdr()
        m.coef_bcs = np.array([[int(x) for x in row()[:3]] for _ in range(nbcs)], dtype=np.int64).reshape(nbcs, 3)
        m.chain_bcs = cb
    return m


# ---------------------------------------------------------------------------
# sorting helpers with the reference's exact permutations
# ---------------------------------------------------------------------------

def exchange_sort_permutation(keys: np.ndarray) -> np.ndarray:
    """Permutation produced by the exchange sort of ``module.F90:537-543``.

    ``for n: for m>n: if key(A(n)) > key(A(m)): swap``.  Equal keys are never
    swapped, so the order of ties depends on the swap history; this function
    replays it exactly (vectorised per outer step: the swaps of step n are a
    rotation along the strict prefix minima of A(n+1:)).
    Returns ``perm`` with ``sorted[i] = original[perm[i]]``.
    """
    k = np.asarray(keys, dtype=np.int64).copy()
    perm = np.arange(len(k))
    n = len(k)
    for i in range(n - 1):
        tail = k[i + 1:]
        # running minimum seen before position m, starting with k[i]
        run = np.minimum.accumulate(np.concatenate(([k[i]], tail)))[:-1]
        hits = np.nonzero(tail < run)[0]
        if hits.size == 0:
            continue
        pos = hits + i + 1
        # rotation: A(i) <- A(pos[-1]); A(pos[j]) <- A(pos[j-1]); A(pos[0]) <- A(i)
        src = np.concatenate(([i], pos[:-1]))
        dst = np.concatenate((pos, [i]))
        ksrc = np.concatenate((k[src], [k[pos[-1]]]))
        psrc = np.concatenate((perm[src], [perm[pos[-1]]]))
        k[dst[:-1]] = ksrc[:-1]
        perm[dst[:-1]] = psrc[:-1]
        k[i] = ksrc[-1]
        perm[i] = psrc[-1]
    return perm


def basis_keys(labels: np.ndarray, reference: np.ndarray) -> np.ndarray:
    """Lexicographic (labels..., reference) key as one int64 (``module.F90:549-577``)."""
    lab = np.asarray(labels, dtype=np.int64)
    key = np.zeros(len(lab), dtype=np.int64)
    for c in range(lab.shape[1]):
        key = key * 1024 + (lab[:, c] + 512)
    return key * (1 << 24) + np.asarray(reference, dtype=np.int64)


def stable_argsort(x):
    """sortWithIndex (``module.F90:16-44``) is a stable insertion sort."""
    return np.argsort(x, kind="stable")


# ---------------------------------------------------------------------------
# one NRG run
# ---------------------------------------------------------------------------

@dataclasses.dataclass
class IterationRecord:
    it: int
    subspace: np.ndarray        # [nsub, nv+3] (labels, dim, count_eigenvector, start_input)
    basis_input: np.ndarray     # [nfull, nv+3]
    basis_output: np.ndarray    # [nfull, nv+3] after postprocess
    eigenvalue: np.ndarray      # [nfull] after postprocess
    nkept: int
    U: List[np.ndarray]         # per block eigenvectors (n_b x n_b)
    spec_old: Optional[np.ndarray] = None    # spectral invariant of it-1 (restore)


class NRG:
    """Forward NRG + CFS in the reference's semantics."""

    def __init__(self, model: Model, keep_history: Optional[bool] = None, eigh=None, workers: int = 1):
        self.m = model
        self.keep = model.spectrum if keep_history is None else keep_history
        self.eigh = eigh or np.linalg.eigh
        # blocks solved by a pool of `workers` threads, one single-threaded
        # LAPACK call each (the reference's OpenMP loop over blocks); the
        # results do not depend on the count
        self.workers = max(1, int(workers))
        self.basis_output = model.basis_output0.copy()
        self.eigenvalue = model.eigenvalue0.copy()
        self.inv = model.invariant0.copy()
        self.spec = model.spec0.copy() if model.spectrum else None
        self.nold = model.nb0
        self.start_trunc = model.N + 1
        self.flag_trunc = False
        self.records: List[IterationRecord] = []
        self.thermo = []
        self.sort_swaps = 0

    # -- prepareBasis.F90:41-114 ------------------------------------------
    def prepare_basis(self):
        m = self.m
        nv, nops = m.nv, m.nops
        nfull = self.nold * nops
        bi = np.zeros((nfull, nv + 3), dtype=np.int64)
        i = np.repeat(np.arange(self.nold), nops)
        j = np.tile(np.arange(nops), self.nold)
        bi[:, :nv] = self.basis_output[i, :nv] - m.operation[j, :nv]
        bi[:, nv] = i + 1
        bi[:, nv + 1] = self.basis_output[i, nv]
        bi[:, nv + 2] = j + 1
        perm = exchange_sort_permutation(basis_keys(bi[:, :nv], bi[:, nv]))
        bi = bi[perm]
        # subspaceInfo (prepareBasis.F90:63-114)
        lab = bi[:, :nv]
        brk = np.ones(nfull, dtype=bool)
        brk[1:] = np.any(lab[1:] != lab[:-1], axis=1)
        starts = np.nonzero(brk)[0]
        dims = np.diff(np.append(starts, nfull))
        sub = np.zeros((len(starts), nv + 3), dtype=np.int64)
        sub[:, :nv] = lab[starts]
        sub[:, nv] = dims
        sub[:, nv + 1] = 1 + np.concatenate(([0], np.cumsum(dims[:-1] ** 2)))
        sub[:, nv + 2] = starts + 1
        self.basis_input = bi
        self.subspace = sub

    # -- diagonalization.F90:153-263 ---------------------------------------
    def hamiltonian_block(self, b: int, it: int) -> np.ndarray:
        m = self.m
        nv = m.nv
        dim = int(self.subspace[b, nv])
        s0 = int(self.subspace[b, nv + 2]) - 1
        rows = self.basis_input[s0:s0 + dim]
        ops = rows[:, nv + 2] - 1
        refs = rows[:, nv] - 1
        H = np.zeros((dim, dim))
        diag = self.eigenvalue[refs] * math.sqrt(m.Lam)
        for c in range(m.nchain):
            diag = diag + m.operation[ops, nv + c] * m.chain_diag[it + 1, c]
        H[np.arange(dim), np.arange(dim)] = diag
        T = self.hopping_lookup()
        iu, ju = np.triu_indices(dim, 1)
        oi, oj = ops[iu], ops[ju]
        ok = T["valid"][oi, oj]
        iu, ju, oi, oj = iu[ok], ju[ok], oi[ok], oj[ok]
        sw = T["swap"][oi, oj]
        ir = np.where(sw, refs[ju], refs[iu])
        jr = np.where(sw, refs[iu], refs[ju])
        t = T["type"][oi, oj]
        ch = T["chan"][oi, oj]
        cnd = np.array([m.cnd(it, c) for c in range(m.nchain)])
        # invariant_matrix(jref, iref, t) = inv[t, iref, jref]
        v = cnd[ch] * T["coef"][oi, oj] * self.inv[t, ir, jr]
        H[iu, ju] = v
        H[ju, iu] = v
        if m.coef_bcs is not None and len(m.coef_bcs):
            # BCS pairing (diagonalization.F90:215-245): a forward row of
            # (op_i, op_j) with equal references; the last match sets the
            # channel and the value replaces the hopping term
            iu, ju = np.triu_indices(dim, 1)
            bch = np.full((m.nops, m.nops), -1, dtype=np.int64)
            for a, b, c in m.coef_bcs:
                bch[a - 1, b - 1] = c - 1
            cb = bch[ops[iu], ops[ju]]
            hit = (cb >= 0) & (refs[iu] == refs[ju])
            val = m.chain_bcs[it + 1, np.maximum(cb[hit], 0)]
            H[iu[hit], ju[hit]] = val
            H[ju[hit], iu[hit]] = val
        return H

    def hopping_lookup(self):
        """(op_i, op_j) -> (type, channel, coef, swap) replaying the coefficient
        search of ``diagonalization.F90:195-213``: every row is tested, the last
        match sets (type, channel, coef), and a reversed match swaps the
        references for good (``:208-209`` reassign ireference/jreference and a
        later forward match does not restore them)."""
        if getattr(self, "_hop", None) is not None:
            return self._hop
        m = self.m
        n = m.nops
        T = dict(valid=np.zeros((n, n), bool), swap=np.zeros((n, n), bool),
                 type=np.zeros((n, n), np.int64), chan=np.zeros((n, n), np.int64), coef=np.zeros((n, n)))
        for oi in range(n):
            for oj in range(n):
                for ic in range(m.ncoef):
                    a, b, t, ch = (int(x) - 1 for x in m.coef_diag_type[ic])
                    if oi == a and oj == b:
                        hit = (False,)
                    elif oi == b and oj == a:
                        hit = (True,)
                    else:
                        continue
                    T["valid"][oi, oj] = True
                    T["swap"][oi, oj] |= hit[0]
                    T["type"][oi, oj] = t
                    T["chan"][oi, oj] = ch
                    T["coef"][oi, oj] = m.coef_diag[ic]
        self._hop = T
        return T

    # -- diagonalization.F90:66-407 ----------------------------------------
    def diagonalize(self, it: int):
        nv = self.m.nv
        nsub = len(self.subspace)
        nfull = len(self.basis_input)
        E = np.zeros(nfull)

        def solve(b):
            w, U = self.eigh(self.hamiltonian_block(b, it))
            return w, np.where(np.abs(U) < EIGENVECTOR_THRESHOLD, 0.0, U)
        if self.workers > 1:
            from concurrent.futures import ThreadPoolExecutor
            from threadpoolctl import threadpool_limits
            order = sorted(range(nsub), key=lambda b: -int(self.subspace[b, nv]))   # largest first
            with threadpool_limits(limits=1), ThreadPoolExecutor(self.workers) as ex:
                done = dict(zip(order, ex.map(solve, order)))
            res = [done[b] for b in range(nsub)]
        else:
            res = [solve(b) for b in range(nsub)]
        Us = []
        for b, (w, U) in enumerate(res):
            s0 = int(self.subspace[b, nv + 2]) - 1
            E[s0:s0 + len(w)] = w
            Us.append(U)
        E = E - E.min()
        E[E <= NUMERICAL_THRESHOLD] = 0.0
        bo = np.zeros((nfull, nv + 3), dtype=np.int64)
        for b in range(nsub):
            s0 = int(self.subspace[b, nv + 2]) - 1
            d = int(self.subspace[b, nv])
            bo[s0:s0 + d, :nv] = self.subspace[b, :nv]
            bo[s0:s0 + d, nv] = np.arange(1, d + 1)
        self.eigenvalue = E
        self.U = Us
        self.basis_output = bo

    # -- module.F90:1931-1982 -------------------------------------------------
    def calc_thermo(self):
        m = self.m
        E = self.eigenvalue
        bE = BETABAR * E
        ex = np.exp(-bE)
        Z = ex.sum()
        En = (bE * ex).sum()
        cs, ce = m.cons[1, 0] - 1, m.cons[1, 1]
        sz = self.basis_output[:, cs:ce].sum(axis=1) * 0.5
        Sz2 = (ex * sz * sz).sum()
        Sz = (ex * sz).sum()
        Cn = (((bE - En / Z) ** 2) * ex).sum() / Z
        Sn = En / Z + math.log(Z)
        Xn = Sz2 / Z - (Sz / Z) ** 2
        self.thermo.append((Xn, Sn, Cn))

    # -- postprocess.F90:71-160 -----------------------------------------------
    def truncate(self, it: int):
        m = self.m
        nfull = len(self.eigenvalue)
        idx = stable_argsort(self.eigenvalue)
        Es = self.eigenvalue[idx]
        nb = nfull
        self.trunc_margin = None
        if it != m.i0 and it != m.N and nfull > m.nkeep:
            if self.start_trunc > it:
                self.start_trunc = it
                self.flag_trunc = True
            limit = int(m.nkeep * TRUNCATION_LIMIT_SCALE)
            ecut = Es[m.nkeep - 1] + TRUNCATION_THRESHOLD
            cutoff = m.nkeep
            i = m.nkeep
            while i < nfull:
                if Es[i] <= ecut:
                    cutoff += 1
                else:
                    break
                i += 1
            while cutoff > limit:
                ecut = ecut - 0.001
                for i in range(cutoff, 0, -1):
                    if Es[i - 1] > ecut:
                        cutoff -= 1
                    else:
                        break
            nb = cutoff
            self.trunc_margin = float(np.min(np.abs(Es - ecut)))
        # kept prefix back to index order (sortByIndex on 1..nb)
        kept = np.sort(idx[:nb])
        order = np.concatenate((kept, idx[nb:]))
        self.eigenvalue = self.eigenvalue[order]
        if nb < nfull:
            self.basis_output = self.basis_output[order]
        self.nkept = nb
        return nb

    # -- invariantMatrixDGEMM.F90:84-100 --------------------------------------
    def kept_per_block(self) -> np.ndarray:
        nv = self.m.nv
        lab = self.basis_output[:self.nkept, :nv]
        out = np.zeros(len(self.subspace), dtype=np.int64)
        for b in range(len(self.subspace)):
            out[b] = int(np.all(lab == self.subspace[b, :nv], axis=1).sum())
        return out

    def _block_rows(self, b):
        nv = self.m.nv
        d = int(self.subspace[b, nv])
        s0 = int(self.subspace[b, nv + 2]) - 1
        return self.basis_input[s0:s0 + d]

    # -- invariantMatrixDGEMM.F90:154-332 -------------------------------------
    def invariant_matrix(self):
        m = self.m
        nv = m.nv
        kb = self.kept_per_block()
        starts = np.concatenate(([0], np.cumsum(kb)[:-1]))
        k = int(kb.sum())
        inv = np.zeros((m.ncm, k, k))
        cs0, cs1 = m.cons[0, 0] - 1, m.cons[0, 1]
        ss0, ss1 = m.cons[1, 0] - 1, m.cons[1, 1]
        nsub = len(self.subspace)
        for L in range(nsub):
            for R in range(nsub):
                if kb[L] == 0 or kb[R] == 0:
                    continue
                lab_l, lab_r = self.subspace[L, :nv], self.subspace[R, :nv]
                dsz = int((lab_l[ss0:ss1] - lab_r[ss0:ss1]).sum())
                if m.cons[0, 0] == m.cons[0, 1] == 0:
                    if abs(dsz) != 1:
                        continue
                else:
                    dq = int((lab_l[cs0:cs1] - lab_r[cs0:cs1]).sum())
                    if dq != 1 or abs(dsz) != 1:
                        continue
                C = self._coef_matrix(L, R)
                UL = self.U[L][:, :kb[L]]
                UR = self.U[R][:, :kb[R]]
                for t in range(m.ncm):
                    M = UL.T @ C[t] @ UR          # (kL x kR)
                    # invariant_matrix(startr+isubr-1, startl+isubl-1, t) = M(isubl, isubr)
                    inv[t, starts[L]:starts[L] + kb[L], starts[R]:starts[R] + kb[R]] = M
        self.inv = inv
        self.kb = kb

    def _coef_matrix(self, L, R):
        """C_t(a,b) (``invariantMatrixDGEMM.F90:236-256``): for every table row
        with (op_a, op_b) matching and equal variations, slot t of that row."""
        m = self.m
        nv = m.nv
        rl, rr = self._block_rows(L), self._block_rows(R)
        C = np.zeros((m.ncm, len(rl), len(rr)))
        va, oa = rl[:, nv + 1], rl[:, nv + 2]
        vb, ob = rr[:, nv + 1], rr[:, nv + 2]
        same_var = va[:, None] == vb[None, :]
        for ic in range(m.ncoef):
            a, b, t = (int(x) for x in m.coef_inv_type[ic])
            mask = same_var & (oa[:, None] == a) & (ob[None, :] == b)
            C[t - 1][mask] = m.coef_inv[ic]
        return C

    # -- invariantMatrixForSpectrumDGEMM.F90:148-370 --------------------------
    def invariant_matrix_spec(self):
        m = self.m
        nv = m.nv
        kb = self.kb
        starts = np.concatenate(([0], np.cumsum(kb)[:-1]))
        k = int(kb.sum())
        new = np.zeros((m.nmat, k, k))
        nsub = len(self.subspace)
        for L in range(nsub):
            for R in range(nsub):
                if kb[L] == 0 or kb[R] == 0:
                    continue
                diff = self.subspace[L, :nv] - self.subspace[R, :nv]
                mats = [im for im in range(m.nmat) if np.all(diff == m.cons_diff_spec[im])]
                if not mats:
                    continue
                rl, rr = self._block_rows(L), self._block_rows(R)
                oa, ra = rl[:, nv + 2] - 1, rl[:, nv] - 1
                ob, rb = rr[:, nv + 2] - 1, rr[:, nv] - 1
                same_op = oa[:, None] == ob[None, :]
                UL = self.U[L][:, :kb[L]]
                UR = self.U[R][:, :kb[R]]
                for im in mats:
                    # coefMatrix(a,b,im) = cspec(im, op_a) * spec(ref_b, ref_a, im)
                    C = np.where(same_op, m.coef_spec[im, oa][:, None] * self.spec[im][ra[:, None], rb[None, :]], 0.0)
                    new[im, starts[L]:starts[L] + kb[L], starts[R]:starts[R] + kb[R]] = UL.T @ C @ UR
        self.spec = new

    # -- postprocess.F90:221-374 (stock of the kept columns) -------------------
    def stock(self, it):
        nv = self.m.nv
        if it == self.start_trunc:
            self.full = []          # one entry per stocked (it, block)
            self.full_map = {}
        self.full_map[it] = len(self.full)
        lab_k = self.basis_output[:self.nkept, :nv]
        for b in range(len(self.subspace)):
            match = np.all(lab_k == self.subspace[b, :nv], axis=1)
            kept = int(match.sum())
            if kept == 0:
                continue
            first = int(np.argmax(match))
            self.full.append(dict(
                it=it, labels=self.subspace[b, :nv].copy(), dim=int(self.subspace[b, nv]), kept=kept,
                U=self.U[b][:, :kept].copy(), rows=self._block_rows(b).copy(),
                out_refs=self.basis_output[first:first + kept, nv].copy(),
                E=self.eigenvalue[first:first + kept].copy()))

    def iteration(self, it):
        self.prepare_basis()
        self.diagonalize(it)
        self.calc_thermo()
        nfull = len(self.eigenvalue)
        self.truncate(it)
        spec_old = self.spec.copy() if (self.m.spectrum and self.keep) else None
        if it < self.m.N:
            self.invariant_matrix()
            if self.m.spectrum:
                self.invariant_matrix_spec()
        self.nold = self.nkept
        rec = IterationRecord(it, self.subspace.copy(), self.basis_input.copy(), self.basis_output.copy(),
                              self.eigenvalue.copy(), self.nkept, self.U if self.keep else [], spec_old)
        self.records.append(rec)
        if self.m.spectrum and it >= self.start_trunc:
            self.stock(it)
        return rec

    def seed(self, basis_output, eigenvalue, invariant, nkept: int, start_trunc: int):
        """Continue a run from the state another implementation reached:
        the kept basis records (first nkept rows of basis_output), their
        energies and the invariant matrix [t, l, r] of the kept basis."""
        self.basis_output = np.asarray(basis_output, dtype=np.int64).copy()
        self.eigenvalue = np.asarray(eigenvalue, dtype=np.float64).copy()
        self.inv = np.asarray(invariant, dtype=np.float64)
        self.nold = int(nkept)
        self.start_trunc = int(start_trunc)
        self.flag_trunc = self.start_trunc <= self.m.N

    def run_forward(self, upto: Optional[int] = None):
        last = self.m.N if upto is None else upto
        for it in range(self.m.i0, last + 1):
            self.iteration(it)

    # -- initializeReducedDensity.F90:30-78 ------------------------------------
    def init_reduced_density(self):
        last = self.records[-1]
        g = int((last.eigenvalue < DEGENERACY_THRESHOLD).sum())
        self.rho = {}
        base = self.full_map[self.m.N]
        for e in self.full[base:]:
            r = np.zeros((e["kept"], e["kept"]))
            d = np.where(e["E"] < DEGENERACY_THRESHOLD, 1.0 / g, 0.0)
            r[np.arange(e["kept"]), np.arange(e["kept"])] = d
            self.rho[(self.m.N, tuple(e["labels"]))] = r
        self.gs_degeneracy = g

    # -- reducedDensityMatrixDGEMM.F90:62-321 ----------------------------------
    def reduced_density(self):
        m = self.m
        nv = m.nv
        self.rho_trace = {m.N: sum(np.trace(r) for r in self.rho.values())}
        by_it = {}
        for e in self.full:
            by_it.setdefault(e["it"], {})[tuple(e["labels"])] = e
        for n in range(m.N - 1, self.start_trunc - 1, -1):
            tr = 0.0
            for lab, e in by_it[n].items():
                kB = e["kept"]
                acc = np.zeros((kB, kB))
                for o in range(m.nops):
                    plab = tuple(np.array(lab) - m.operation[o, :nv])
                    P = by_it[n + 1].get(plab)
                    if P is None:
                        continue
                    rows = P["rows"]
                    # V(il, k) = U_P(row, k), row = first input of P with variation = out_ref(il), op = o+1
                    V = np.zeros((kB, P["kept"]))
                    for il in range(kB):
                        hit = np.nonzero((rows[:, nv + 1] == e["out_refs"][il]) & (rows[:, nv + 2] == o + 1))[0]
                        V[il] = P["U"][hit[0]]
                    acc += V @ self.rho[(n + 1, plab)] @ V.T
                self.rho[(n, lab)] = acc
                tr += np.trace(acc)
            self.rho_trace[n] = tr

    # -- module.F90:1210-1254 --------------------------------------------------
    def bin_spike(self, container, omega, weight, im):
        nbin = self.nbin
        pos = (math.log10(abs(omega)) - LOG10_EMIN) * MESH_OF_BIN + 1.0 if omega != 0.0 else -math.inf
        i1 = i2 = 0
        if pos <= 0.0:
            i1 = i2 = 1
        elif pos >= nbin:
            i1 = i2 = nbin
        pref = int(pos) if math.isfinite(pos) else -1
        if pref <= 0:
            pref = 1
        if i1 == i2 and i1 > 0:
            container[im, i1 - 1] += weight
        else:
            w = pos - pref
            container[im, pref - 1] += weight * (1.0 - w)
            container[im, pref] += weight * w

    def bin_spikes(self, container, omega, weight, im):
        """Vectorised ``bin_spike`` for many spikes (same arithmetic)."""
        omega = np.asarray(omega, dtype=np.float64)
        weight = np.asarray(weight, dtype=np.float64)
        nbin = self.nbin
        with np.errstate(divide="ignore"):
            pos = (np.log10(np.abs(omega)) - LOG10_EMIN) * MESH_OF_BIN + 1.0
        edge_lo = pos <= 0.0
        edge_hi = (~edge_lo) & (pos >= nbin)
        mid = ~(edge_lo | edge_hi)
        np.add.at(container[im], np.zeros(int(edge_lo.sum()), np.int64), weight[edge_lo])
        np.add.at(container[im], np.full(int(edge_hi.sum()), nbin - 1, np.int64), weight[edge_hi])
        pm = pos[mid]
        pref = np.trunc(pm).astype(np.int64)
        pref[pref <= 0] = 1
        w = pm - pref
        np.add.at(container[im], pref - 1, weight[mid] * (1.0 - w))
        np.add.at(container[im], pref, weight[mid] * w)

    def setup_bins(self):
        emax = self.m.scale(-1)
        log10emax = int(math.log10(emax)) + 1.0
        self.nbin = int(log10emax - LOG10_EMIN) * MESH_OF_BIN + 1
        self.bins_pos = np.zeros((self.m.nmat, self.nbin))
        self.bins_neg = np.zeros((self.m.nmat, self.nbin))

    def _block_of(self, rec, lab):
        nv = self.m.nv
        hit = np.nonzero(np.all(rec.subspace[:, :nv] == lab, axis=1))[0]
        return int(hit[0]) if len(hit) else None

    def _spectral_element(self, rec, bl, br, im, UL, UR):
        """module.F90:1014-1125: UL^T C UR with C(a,b) = cspec(im,op_a) spec_old(ref_b, ref_a, im), op_a == op_b."""
        nv = self.m.nv
        sl, sr = rec.subspace[bl], rec.subspace[br]
        rl = rec.basis_input[int(sl[nv + 2]) - 1:int(sl[nv + 2]) - 1 + int(sl[nv])]
        rr = rec.basis_input[int(sr[nv + 2]) - 1:int(sr[nv + 2]) - 1 + int(sr[nv])]
        oa, ra = rl[:, nv + 2] - 1, rl[:, nv] - 1
        ob, rb = rr[:, nv + 2] - 1, rr[:, nv] - 1
        C = np.where(oa[:, None] == ob[None, :],
                     self.m.coef_spec[im, oa][:, None] * rec.spec_old[im][ra[:, None], rb[None, :]], 0.0)
        return UL.T @ C @ UR

    # -- calculateSpectrumDGEMM.F90:55-212 -------------------------------------
    def spike_zero_temperature(self, rec):
        m = self.m
        nv = m.nv
        E = rec.eigenvalue
        gs = np.nonzero(E < DEGENERACY_THRESHOLD)[0]
        pf = float(len(gs))
        sc = m.scale(rec.it)
        for g in gs:
            lab = rec.basis_output[g, :nv]
            bg = self._block_of(rec, lab)
            ug = rec.U[bg][:, :1]                 # column 1 of the ground-state block
            for sgn, ops, cont in ((+1, m.pos_op, self.bins_pos), (-1, m.neg_op, self.bins_neg)):
                for im in range(m.nmat):
                    be = self._block_of(rec, lab + ops[im])
                    if be is None:
                        continue
                    ue = rec.U[be]
                    if sgn > 0:
                        w = self._spectral_element(rec, be, bg, im, ue, ug)[:, 0]
                    else:
                        w = self._spectral_element(rec, bg, be, im, ug, ue)[0, :]
                    s0 = int(rec.subspace[be, nv + 2]) - 1
                    self.bin_spikes(cont, sgn * E[s0:s0 + len(w)] * sc, w * w / pf, im)

    # -- calculateSpectrumDGEMM.F90:296-643 -------------------------------------
    def spike_cfs(self, rec, per_state: bool = False):
        """Spikes of every discarded state l (positions nkept+1..nfull).

        per_state=True walks the discarded states one at a time in the
        reference's order (``:296-302, :363``).  The default groups the
        discarded states of one block L and evaluates the same products for
        all of them at once (the columns of U_L[:, lvar]); the two orders
        agree to rounding and are checked against each other."""
        m = self.m
        nv = m.nv
        E = rec.eigenvalue
        sc = m.scale(rec.it)
        full_it = {tuple(e["labels"]): e for e in self.full if e["it"] == rec.it}
        kept_lab = rec.basis_output[:rec.nkept, :nv]
        disc = np.arange(rec.nkept, len(E))
        if per_state:
            groups = [(int(il),) for il in disc]
        else:
            labs = rec.basis_output[disc, :nv]
            keys = [tuple(x) for x in labs]
            order = {}
            for j, k in enumerate(keys):
                order.setdefault(k, []).append(int(disc[j]))
            groups = [tuple(v) for v in order.values()]
        for grp in groups:
            il = np.array(grp)
            lvar = rec.basis_output[il, nv].astype(np.int64)
            El = E[il]
            lab = rec.basis_output[il[0], :nv]
            bl = self._block_of(rec, lab)
            ul = rec.U[bl][:, lvar - 1]
            for sgn, ops, cont in ((-1, m.pos_op, self.bins_neg), (+1, m.neg_op, self.bins_pos)):
                for im in range(m.nmat):
                    klab = lab + ops[im]
                    fe = full_it.get(tuple(klab))
                    bk = self._block_of(rec, klab)
                    if fe is None or bk is None:
                        continue
                    kk = fe["kept"]
                    start = int(np.argmax(np.all(kept_lab == klab, axis=1)))
                    uk = rec.U[bk][:, :kk]
                    if sgn < 0:
                        w = self._spectral_element(rec, bk, bl, im, uk, ul)          # (kk, nl)
                    else:
                        w = self._spectral_element(rec, bl, bk, im, ul, uk).T        # (kk, nl)
                    rho = self.rho[(rec.it, tuple(klab))]
                    refs = rec.basis_output[start:start + kk, nv] - 1
                    # weight(k1) = w(k1) * sum_k2 w(k2) rho(k2, k1)
                    x = rho[np.ix_(refs, refs)].T @ w
                    Ek = E[start:start + kk]
                    om = (Ek[:, None] - El[None, :]) * sc if sgn < 0 else (El[None, :] - Ek[:, None]) * sc
                    self.bin_spikes(cont, om.ravel(), (w * x).ravel(), im)

    def run_spectrum(self):
        self.setup_bins()
        if self.flag_trunc:
            self.init_reduced_density()
        self.reduced_density()
        recs = {r.it: r for r in self.records}
        for it in range(self.start_trunc, self.m.N + 1):
            if it == self.m.N:
                self.spike_zero_temperature(recs[it])
            else:
                self.spike_cfs(recs[it])


def run(inputs_dir: str, upto: Optional[int] = None) -> NRG:
    nrg = NRG(read_inputs(inputs_dir))
    nrg.run_forward(upto)
    if nrg.m.spectrum and upto is None:
        nrg.run_spectrum()
    return nrg
