"""Synthetic block batch: the eigensolver and contraction products on blocks
of NRG-like shape, without an NRG run.

A batch is `nblk` dense symmetric FP64 blocks whose sizes are drawn
log-normal (median `median`, `sigma` in ln-space, rounded, clipped to
[lo, hi]), with eigenvalues on the NRG-like geometric ladder
Lambda^(-k/2), Lambda = 2 (k uniform on the integers 0..60, which plants
exact degeneracies, or uniform on [0, 60] for distinct values), rotated by
Haar-random orthogonal matrices (QR of a Gaussian matrix with the signs of
R's diagonal moved into Q, Mezzadri 2007).  Each block carries `nops` dense
operator matrices with N(0, 1) entries; the contraction under test is
U^T O U with U the eigenvectors the library returns, the product shape of
the reference's invariant-matrix and spectral-weight call sites
(invariantMatrixDGEMM.F90:300,316; module.F90:1100,1113).

Arrays are packed column-major at offsets sum n^2, the layout of the
library's batched entry points.  The planted spectrum is a closed form the
eigenvalues are checked against; residual and orthogonality are checked
with random probe vectors (matrix-vector products only).
"""
from __future__ import annotations

import dataclasses
import time
from typing import Dict, List, Optional

import numpy as np

from . import gpu


@dataclasses.dataclass
class Batch:
    seed: int
    n: np.ndarray          # block sizes (int32)
    off: np.ndarray        # offsets sum n^2 (int64), len nblk + 1
    lam: List[np.ndarray]  # planted eigenvalues per block (unsorted)
    H: np.ndarray          # packed blocks


def sizes(seed: int, nblk: int = 256, median: float = 600.0, sigma: float = 1.0, lo: int = 16,
          hi: int = 4096) -> np.ndarray:
    rng = np.random.default_rng(seed)
    n = np.round(median * np.exp(sigma * rng.standard_normal(nblk)))
    return np.clip(n, lo, hi).astype(np.int32)


def _threads():
    try:
        from threadpoolctl import threadpool_limits
        import os
        return threadpool_limits(limits=min(16, os.cpu_count() or 1))
    except Exception:   # pragma: no cover - threadpoolctl is optional
        import contextlib
        return contextlib.nullcontext()


def make_batch(seed: int = 20240, nblk: int = 256, median: float = 600.0, sigma: float = 1.0,
               ladder: str = "integer", lo: int = 16, hi: int = 4096) -> Batch:
    n = sizes(seed, nblk, median, sigma, lo, hi)
    off = np.zeros(nblk + 1, dtype=np.int64)
    off[1:] = np.cumsum(n.astype(np.int64) ** 2)
    H = np.empty(int(off[-1]))
    rng = np.random.default_rng(seed + 1)
    lams = []
    with _threads():
        for b, nb in enumerate(n):
            nb = int(nb)
            if ladder == "integer":
                k = rng.integers(0, 61, size=nb).astype(float)
            else:
                k = rng.uniform(0.0, 60.0, size=nb)
            lam = 2.0 ** (-k / 2.0)
            Q, R = np.linalg.qr(rng.standard_normal((nb, nb)))
            Q *= np.sign(np.diag(R))
            Hb = (Q * lam) @ Q.T
            Hb = 0.5 * (Hb + Hb.T)
            H[off[b]:off[b + 1]] = Hb.ravel(order="F")
            lams.append(lam)
    return Batch(seed, n, off, lams, H)


def block(a: np.ndarray, off: np.ndarray, n: np.ndarray, b: int) -> np.ndarray:
    nb = int(n[b])
    return a[off[b]:off[b + 1]].reshape(nb, nb, order="F")


def operators(batch: Batch, t: int) -> np.ndarray:
    """Operator t of every block (N(0,1) entries), packed like the blocks."""
    rng = np.random.default_rng(batch.seed * 7919 + 17 + t)
    return rng.standard_normal(int(batch.off[-1]))


def check_eigen(batch: Batch, w: np.ndarray, U: np.ndarray, seed: int = 5) -> Dict[str, float]:
    """Worst eigenvalue error against the planted ladder, probe residual
    ||H U x - U (w x)|| / (|lambda|max ||x||) and orthogonality
    ||U^T U x - x|| / ||x||, over all blocks."""
    rng = np.random.default_rng(seed)
    ev = res = orth = 0.0
    wo = 0
    with _threads():
        for b, nb in enumerate(batch.n):
            nb = int(nb)
            lam = batch.lam[b]
            nrm = float(np.abs(lam).max())
            wb = w[wo:wo + nb]
            wo += nb
            ev = max(ev, float(np.max(np.abs(wb - np.sort(lam)))) / nrm / nb)
            Hb = block(batch.H, batch.off, batch.n, b)
            Ub = block(U, batch.off, batch.n, b)
            X = rng.standard_normal((nb, 2))
            UX = Ub @ X
            xn = float(np.linalg.norm(X))
            res = max(res, float(np.linalg.norm(Hb @ UX - Ub @ (wb[:, None] * X))) / (nrm * xn) / nb)
            orth = max(orth, float(np.linalg.norm(Ub.T @ UX - X)) / xn / nb)
    # errors per unit of block size: the backward-stable bounds grow like n eps
    return dict(eig_err_over_n=ev, residual_over_n=res, orthogonality_over_n=orth)


def contraction_tasks(batch: Batch):
    """Task arrays of the two batched products Y = O U and C = U^T Y."""
    nb = len(batch.n)
    mnk = np.empty(3 * nb, dtype=np.int32)
    mnk[0::3] = batch.n
    mnk[1::3] = batch.n
    mnk[2::3] = batch.n
    return mnk, batch.off[:-1].copy()


def contract(batch: Batch, U: np.ndarray, O: np.ndarray, slices: int, repeat: int):
    """C_b = U_b^T O_b U_b for all blocks: two batched products on the
    device.  Returns (C packed, mean device ms of the two products)."""
    mnk, off = contraction_tasks(batch)
    lib = gpu.load()
    nb = len(batch.n)
    Y = np.zeros_like(O)
    C = np.zeros_like(O)
    ms1 = np.zeros(1)
    ms2 = np.zeros(1)
    P = gpu._p
    gpu._check(lib.nrg_gpu_dgemm_batched(nb, P(mnk, gpu._i32p), 0, 0, P(O, gpu._f64p), P(off, gpu._i64p),
                                         P(U, gpu._f64p), P(off, gpu._i64p), P(Y, gpu._f64p), P(off, gpu._i64p),
                                         slices, repeat, P(ms1, gpu._f64p)))
    gpu._check(lib.nrg_gpu_dgemm_batched(nb, P(mnk, gpu._i32p), 1, 0, P(U, gpu._f64p), P(off, gpu._i64p),
                                         P(Y, gpu._f64p), P(off, gpu._i64p), P(C, gpu._f64p), P(off, gpu._i64p),
                                         slices, repeat, P(ms2, gpu._f64p)))
    return C, float(ms1[0] + ms2[0])


def run(batch: Batch, nops: int = 4, split: tuple = (8, 6), repeat: int = 3, sample: int = 12,
        log=print) -> Dict[str, object]:
    """Eigensolve the batch, check it, then form U^T O U for `nops` operators
    in FP64 and in each split-precision slice count.  FP64 products are
    checked against numpy on `sample` blocks (the largest among them);
    split products against the FP64 products on every block."""
    out: Dict[str, object] = {}
    n = batch.n
    n3 = float(np.sum(n.astype(np.float64) ** 3))
    out["blocks"] = int(len(n))
    out["sizes"] = dict(min=int(n.min()), median=float(np.median(n)), max=int(n.max()),
                        sum_n2=int(batch.off[-1]), sum_n3=n3)
    w = np.zeros(int(n.sum()))
    U = np.zeros_like(batch.H)
    lib = gpu.load()
    P = gpu._p
    t0 = time.perf_counter()
    gpu._check(lib.nrg_gpu_syev_batched(len(n), P(n, gpu._i32p), P(batch.H, gpu._f64p), P(w, gpu._f64p),
                                        P(U, gpu._f64p)))
    out["eigensolve_s_incl_transfers"] = time.perf_counter() - t0
    st = gpu.stage_times()
    out["eigensolve_device_ms"] = dict(tridiag=st["tridiag"], dc=st["dc"], back=st["back"])
    log(f"eigensolve of {len(n)} blocks done")
    out.update(check_eigen(batch, w, U))
    log(f"eigen checks {out['eig_err_over_n']:.2e} {out['residual_over_n']:.2e} {out['orthogonality_over_n']:.2e}")
    flops = 4.0 * n3 * nops          # two n^3 products (2 n^3 flops each) per operator
    rng = np.random.default_rng(batch.seed + 99)
    pick = sorted(set([int(np.argmax(n))] + list(rng.choice(len(n), size=min(sample, len(n)) - 1,
                                                            replace=False))))
    forms = {"fp64": 0}
    forms.update({f"split{s}": s for s in split})
    res = {f: dict(ms=0.0, err_vs_fp64=0.0, err_vs_numpy=0.0) for f in forms}
    for t in range(nops):
        O = operators(batch, t)
        C64, ms = contract(batch, U, O, 0, repeat)
        res["fp64"]["ms"] += ms
        with _threads():
            for b in pick:
                Ub = block(U, batch.off, n, b)
                ref = Ub.T @ block(O, batch.off, n, b) @ Ub
                e = float(np.abs(block(C64, batch.off, n, b) - ref).max() / np.abs(ref).max())
                res["fp64"]["err_vs_numpy"] = max(res["fp64"]["err_vs_numpy"], e)
        for f, s in forms.items():
            if s == 0:
                continue
            Cs, ms = contract(batch, U, O, s, repeat)
            res[f]["ms"] += ms
            worst = 0.0
            for b in range(len(n)):
                r = block(C64, batch.off, n, b)
                worst = max(worst, float(np.abs(block(Cs, batch.off, n, b) - r).max() / np.abs(r).max()))
            res[f]["err_vs_fp64"] = max(res[f]["err_vs_fp64"], worst)
            del Cs
        del C64, O
        log(f"operator {t}: products done")
    for f, r in res.items():
        r["fp64_equiv_tflops"] = flops / (r["ms"] * 1e-3) / 1e12 if r["ms"] > 0 else None
        s = forms[f]
        if s and r["ms"] > 0:
            # S(S+1)/2 digit products of 2mnk int8 operations each
            pairs = s * (s + 1) // 2
            r["int8_mfma_tops"] = pairs * flops / (r["ms"] * 1e-3) / 1e12
            r["frac_int8_peak"] = r["int8_mfma_tops"] / 5000.0
    out["contraction"] = res
    out["contraction_flops"] = flops
    return out


def main(argv=None):
    """python -m nrgamd.synthetic [--nblk N] [--median M] [--nops K] [--split 8,6]:
    one batch on device 0, results as one JSON line (the kernel-profiling
    workload of the split-precision products)."""
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser()
    ap.add_argument("--seed", type=int, default=20240)
    ap.add_argument("--nblk", type=int, default=64)
    ap.add_argument("--median", type=float, default=600.0)
    ap.add_argument("--sigma", type=float, default=1.0)
    ap.add_argument("--nops", type=int, default=1)
    ap.add_argument("--split", default="8,6")
    ap.add_argument("--repeat", type=int, default=3)
    a = ap.parse_args(argv)
    gpu.init(0)
    b = make_batch(seed=a.seed, nblk=a.nblk, median=a.median, sigma=a.sigma)
    split = tuple(int(x) for x in a.split.split(",") if x.strip())

    def log(m):
        print(m, file=sys.stderr, flush=True)
    r = run(b, nops=a.nops, split=split, repeat=a.repeat, sample=4, log=log)
    print(json.dumps(r))


if __name__ == "__main__":
    main()
