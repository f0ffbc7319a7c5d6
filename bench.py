#!/usr/bin/env python3
"""Benchmark: NRG iterations per second on one MI355X.

Workload: two-orbital Anderson model with Kanamori interaction (U=0.3,
J=0.05, Gamma=0.01 per orbital, two Wilson chains, Lambda=2, local site of
dimension 16), Nkeep=4000.  One step is one NRG iteration of the forward
pass: prepareBasis + block-Hamiltonian assembly + batched FP64 eigensolve of
all (Q,Sz) blocks + truncation + invariant-matrix contractions, driven by the
Fortran host (nrg-mpi_amd/lib/libnrg_host.so) through the HIP library.
Warmup iterations run the growth phase (-1, 0, 1, ...); the timed steps are
steady-state truncating iterations (~64k states, ~90 blocks up to ~4700).

Prints one JSON line: metric, value (iter/s), ms_per_step, config, the
per-stage seconds, `roofline` of the dominant kernel family (algorithmic
bytes or flops over its HIP-event time, against 8 TB/s HBM or 78.6 TFLOP/s
FP64 MFMA) and `cpu_baseline` (LAPACK dsyev on one host core over a sample
of one iteration's blocks).
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "nrg-mpi_amd"))

METRIC = "NRG iter/s at Nkeep=4000, batched+split-prec vs plain FP64; % MFMA peak"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="two_orbital_nk4000_it45")
    ap.add_argument("--sort", choices=["exact", "canonical"], default="exact")
    ap.add_argument("--no-profile", action="store_true", help="skip the HIP-event kernel timing")
    ap.add_argument("--profile-every", type=int, default=61,
                    help="time every n-th launch of the profiled families (61: prime, so coprime to the "
                         "32-column panel; a timed tridiagonal SYMV launch runs with the device to itself, "
                         "so sampling more often also slows the multi-stream schedule)")
    ap.add_argument("--profile-families", default="tridiag_symv",
                    help="comma-separated kernel families whose launches carry HIP timing events in the "
                         "timed steps (default: the dominant one, tridiag_symv; 'all' for every family)")
    ap.add_argument("--launch-log", default="",
                    help="write a CSV of every launch of the profiled families (ordinal per kernel function "
                         "from the first iteration, algorithmic bytes and flops; phase warmup/timed)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a plain device-to-device copy of BYTES bytes five times (the reference "
                         "kernel that scales the FETCH_SIZE/WRITE_SIZE counters to bytes) and exit")
    ap.add_argument("--tridiag-align", choices=["end", "start"], default="end",
                    help="column schedule of the batched tridiagonalisation (blocks end or start together)")
    ap.add_argument("--split", type=int, default=0,
                    help="slices of the split-precision contraction GEMMs in the timed steps (0: FP64 matrix "
                         "cores, the faster form on MI355X; the split form is measured in its own iteration)")
    ap.add_argument("--split-leg", type=int, default=8,
                    help="slices of the one-iteration split-precision comparison (0: skip)")
    ap.add_argument("--no-spectrum-leg", dest="spectrum_leg", action="store_false",
                    help="skip the full-pipeline comparison with spectra (forward + rho sweep + CFS)")
    ap.add_argument("--no-synthetic", dest="synthetic", action="store_false",
                    help="skip the synthetic block-batch GEMM comparison (FP64 vs split precision)")
    ap.add_argument("--no-plain", dest="plain", action="store_false",
                    help="skip the one-iteration plain (one block per pass) comparison")
    return ap.parse_args()


def cpu_baseline(block_sizes, budget_s):
    """Reference-equivalent CPU eigensolve (LAPACK dsyev, the routine the
    reference calls per block, diagonalization.F90:267) on one host core, on a
    bounded sample of the iteration's blocks; the remaining blocks are
    extrapolated with the n^3 cost of the largest measured block."""
    os.environ.setdefault("OPENBLAS_NUM_THREADS", "1")
    import numpy as np
    from scipy.linalg import eigh
    try:
        from threadpoolctl import threadpool_limits
        lim = threadpool_limits(1)
    except Exception:  # pragma: no cover
        lim = None
    rng = np.random.default_rng(0)
    sizes = sorted(int(n) for n in block_sizes)
    t_meas, n_meas, last_rate = 0.0, [], None
    for n in sizes:
        if t_meas > budget_s:
            break
        a = rng.standard_normal((n, n))
        a = a + a.T
        t0 = time.perf_counter()
        eigh(a, driver="ev", overwrite_a=True, check_finite=False)
        dt = time.perf_counter() - t0
        t_meas += dt
        n_meas.append(n)
        if n >= 64:
            last_rate = dt / float(n) ** 3
    rest = sizes[len(n_meas):]
    t_rest = sum(last_rate * float(n) ** 3 for n in rest) if rest else 0.0
    total = t_meas + t_rest
    if lim is not None:
        lim.unregister() if hasattr(lim, "unregister") else None
    sample = (f"dsyev (scipy/OpenBLAS, 1 thread) on {len(n_meas)} of {len(sizes)} blocks of one steady-state "
              f"iteration (n <= {max(n_meas) if n_meas else 0}, {t_meas:.1f} s measured), the other "
              f"{len(rest)} blocks extrapolated by n^3; eigensolve stage only, host bookkeeping excluded")
    return dict(value=1.0 / total, unit="iter/s", cores=1, kind="port", sample=sample,
                seconds_per_iteration=total)


def profiled_traffic(family, workload):
    """HBM bytes per launch of `family` on `workload` from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (a sample of the family's
    dispatches in a timed step, with their algorithmic bytes)."""
    name = f"{family}_traffic_{workload.rsplit('_it', 1)[0]}.md"
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        k, _, v = line.partition("=")
        if _ and k.strip() in ("traffic_bytes_per_launch", "alg_bytes_per_launch", "traffic_over_algorithmic"):
            vals[k.strip()] = float(v)
    if "traffic_bytes_per_launch" not in vals:
        return None
    vals["source"] = f"profiles/{name}"
    return vals


def synthetic_gemm(gpu, slices, nblk=32, median=600, seed=0):
    """FP64 vs split-precision matrix-core products on a synthetic block batch:
    block sizes log-normal (median `median`, sigma 1 in ln) clipped to
    [16, 4096], one dense N(0,1) operator O and a dense N(0,1) U per block,
    the product O U of every block in one batched launch, device-timed."""
    import numpy as np
    gpu.init(0)
    rng = np.random.default_rng(seed)
    sizes = np.clip(np.exp(rng.normal(np.log(median), 1.0, size=nblk)), 16, 4096).astype(int)
    As = [rng.standard_normal((n, n)) for n in sizes]
    Bs = [rng.standard_normal((n, n)) for n in sizes]
    flops = float(sum(2.0 * n ** 3 for n in sizes))
    C0, t0 = gpu.dgemm_batched(As, Bs, slices=0, repeat=3)
    out = {"config": f"{nblk} blocks, n log-normal(median {median}, sigma 1) in [16, 4096], seed {seed}, "
                     f"one product O U per block (n <= {int(sizes.max())}, sum 2n^3 = {flops:.3e})",
           "fp64": {"ms": t0, "tflops": flops / (t0 * 1e-3) / 1e12}}
    for s in slices:
        Cs, ts = gpu.dgemm_batched(As, Bs, slices=s, repeat=3)
        err = max(float(np.abs(c - r).max() / np.abs(r).max()) for c, r in zip(Cs, C0))
        pairs = s * (s + 1) // 2
        out[f"split{s}"] = {"ms": ts, "tflops_fp64_equiv": flops / (ts * 1e-3) / 1e12,
                            "fp16_mfma_tflops": pairs * flops / (ts * 1e-3) / 1e12,
                            "max_rel_err_vs_fp64": err, "digit_products": pairs}
    return out


def spectrum_pipeline(gpu, models, Host, nkeep=1000, niter=25):
    """Full pipeline with spectra (forward with stocks, rho sweep, CFS
    spikes and binning) on the two-orbital Kanamori model at a reduced
    Nkeep, once with FP64 contraction products and once split (8 slices)."""
    import numpy as np
    out = {"config": f"two-orbital Kanamori, Nkeep={nkeep}, {niter} iterations, 4 spectral types"}
    bins = {}
    for split in (0, 8):
        tmp = tempfile.mkdtemp(prefix="nrgspec_")
        models.write_inputs(models.workload("two_orbital_nk4000_it45", nkeep=nkeep, niter=niter, spectrum=True), tmp)
        h = Host(tmp, device=0)
        gpu.load().nrg_gpu_set_gemm_split(split)
        try:
            t0 = time.perf_counter()
            h.run_forward()
            t1 = time.perf_counter()
            h.spectrum()
            t2 = time.perf_counter()
            st = h.spectrum_times()
            neg, pos = h.bins()
            i = h.info()
            nit = i["n_iter"] - i["i_first"] + 1
            out["split8" if split else "fp64"] = {
                "forward_s": t1 - t0, "forward_iter_per_s": nit / (t1 - t0),
                "stock_total_s": st["stock"], "rho_sweep_s": st["rho"], "cfs_s": st["cfs"],
                "spectrum_phase_s": t2 - t1, "pipeline_s": t2 - t0,
                "sum_rule_max_dev": float(np.max(np.abs(neg.sum(1) + pos.sum(1) - 1.0)))}
            bins[split] = (neg, pos)
            if split:
                peak = max(np.abs(bins[0][0]).max(), np.abs(bins[0][1]).max())
                diff = max(np.abs(neg - bins[0][0]).max(), np.abs(pos - bins[0][1]).max())
                out["split8"]["max_bin_diff_vs_fp64_rel_peak"] = float(diff / peak)
        finally:
            gpu.load().nrg_gpu_set_gemm_split(0)
            h.close()
    return out


def copy_calibration(nbytes):
    """A plain HBM copy of known size: under `rocprofv3 --pmc FETCH_SIZE` (or
    WRITE_SIZE) its counter reading per dispatch, against nbytes, gives the
    factor that turns the counters of the NRG kernels into bytes."""
    import torch
    n = nbytes // 8
    src = torch.arange(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    ok = bool(torch.equal(dst[-4:], src[-4:]))
    print(json.dumps({"copy_bytes": 8 * n, "dispatches": 5, "ok": ok}))


def main():
    a = parse()
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print(f"bench.py: runs on one GPU only; --gpus {a.gpus} (WORLD_SIZE={world}) is not supported: "
              f"the NRG hot path is measured on a single MI355X and block sharding across GPUs is not "
              f"implemented", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    from nrgamd import gpu, models
    from nrgamd.host import Host

    spec = models.workload(a.workload)
    need = a.warmup + a.steps
    if need > spec.niter:
        raise SystemExit(f"warmup+steps={need} exceeds the {spec.niter} iterations of {a.workload}")
    tmp = tempfile.mkdtemp(prefix="nrgbench_")
    models.write_inputs(spec, tmp)
    h = Host(tmp, device=0, sort_mode=0 if a.sort == "exact" else 1)
    gpu.load().nrg_gpu_set_tridiag_schedule(1 if a.tridiag_align == "end" else 0)
    gpu.load().nrg_gpu_set_gemm_split(a.split)
    log_fams = [f for f in a.profile_families.split(",") if f and f != "all"] or list(gpu.KERNEL_FAMILIES)
    if a.launch_log:
        gpu.launch_log(log_fams)
    its = list(range(spec.first_iteration, spec.first_iteration + need))
    for it in its[:a.warmup]:
        h.iterate(it)
        print(f"warmup iteration {it}: nfull={h.i['nfull']} nsub={h.i['nsub']} kept={h.i['nkept']}",
              file=sys.stderr, flush=True)
    if not a.no_profile:
        fams = None if a.profile_families == "all" else [f for f in a.profile_families.split(",") if f]
        gpu.profile(True, families=fams, every=a.profile_every)
    stages = {}
    flops_eig = 0.0
    n_warm_log = len(gpu.launch_log_get()) if a.launch_log else 0
    t0 = time.perf_counter()
    for it in its[a.warmup:]:
        h.iterate(it)
        t = h.times()
        st = gpu.stage_times()
        flops_eig += st["eig_flops"]
        for k in ("prepare", "diagonalization", "postprocess", "invariant"):
            stages[k] = stages.get(k, 0.0) + t[k]
        for k in ("assemble", "eig", "tridiag", "dc", "back"):
            stages["gpu_" + k] = stages.get("gpu_" + k, 0.0) + st[k] * 1e-3
        print(f"timed iteration {it}: {t['iteration']:.3f} s  (prepare {t['prepare']:.3f}, diag "
              f"{t['diagonalization']:.3f} [tri {st['tridiag']:.0f} ms, dc {st['dc']:.0f} ms, back "
              f"{st['back']:.0f} ms], post {t['postprocess']:.3f}) nfull={h.i['nfull']} nsub={h.i['nsub']}",
              file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t0
    if a.launch_log:
        L = gpu.launch_log_get()
        names = {int(k): gpu.launch_log_kernel(int(k)) for k in set(L[:, 2].tolist())}
        gpu.launch_log(None)
        with open(a.launch_log, "w") as f:
            f.write("family,kernel,threads,ordinal,bytes,flops,phase,events\n")
            for i, r in enumerate(L):
                f.write(f"{gpu.KERNEL_FAMILIES[int(r[0])]},{names[int(r[2])]},{int(r[1])},{int(r[3])},"
                        f"{r[4]:.0f},{r[5]:.0f},{'warmup' if i < n_warm_log else 'timed'},{int(r[6])}\n")
    ks = gpu.kernel_stats() if not a.no_profile else {}
    if not a.no_profile:
        gpu.profile(False)
    plain = batched64 = batched_split = None
    nxt = its[-1] + 1
    if a.plain and a.split > 0 and nxt <= spec.last_iteration:
        # the batched path with FP64 contraction products, one iteration
        gpu.load().nrg_gpu_set_gemm_split(0)
        tb = time.perf_counter()
        h.iterate(nxt)
        tb = time.perf_counter() - tb
        batched64 = {"value": 1.0 / tb, "unit": "iter/s", "iteration": nxt,
                     "form": "batched FP64: all blocks per eigensolver pass, FP64 contraction products"}
        print(f"batched FP64 iteration {nxt}: {tb:.3f} s", file=sys.stderr, flush=True)
        nxt += 1
    if a.plain and a.split == 0 and a.split_leg > 0 and nxt <= spec.last_iteration:
        # the batched path with split-precision contraction products, one iteration
        gpu.load().nrg_gpu_set_gemm_split(a.split_leg)
        tb = time.perf_counter()
        h.iterate(nxt)
        tb = time.perf_counter() - tb
        batched_split = {"value": 1.0 / tb, "unit": "iter/s", "iteration": nxt,
                         "form": f"batched eigensolve + split-precision contraction products ({a.split_leg} "
                                 f"slices, FP16 matrix cores)"}
        print(f"batched split iteration {nxt}: {tb:.3f} s", file=sys.stderr, flush=True)
        nxt += 1
    gpu.load().nrg_gpu_set_gemm_split(0)
    if a.plain and nxt <= spec.last_iteration:
        # one block per eigensolver pass (the per-block schedule of the
        # reference's dsyev loop) and FP64 products, one iteration
        gpu.load().nrg_gpu_set_batch_limit(1)
        tp = time.perf_counter()
        h.iterate(nxt)
        tp = time.perf_counter() - tp
        gpu.load().nrg_gpu_set_batch_limit(0)
        st = gpu.stage_times()
        plain = {"value": 1.0 / tp, "unit": "iter/s", "iteration": nxt,
                 "eig_s": st["eig"] * 1e-3, "form": "plain FP64: one block per eigensolver pass, FP64 products"}
        print(f"plain iteration {nxt}: {tp:.3f} s", file=sys.stderr, flush=True)
    sizes = h.subspace()[:, h.i["nv"]].tolist()
    nfull, nsub, nmax = h.i["nfull"], h.i["nsub"], int(max(sizes))
    h.close()

    value = a.steps / elapsed
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "iter/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * elapsed / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": a.workload,
                   "model": "two-orbital Anderson, Kanamori U=0.3 J=0.05, Gamma=0.01/orbital, 2 Wilson chains, "
                            "Lambda=2, D=1, local dim 16, labels (Q,2Sz)",
                   "nkeep": spec.nkeep, "timed_iterations": [its[a.warmup], its[-1]],
                   "states_per_iteration": nfull, "blocks_per_iteration": nsub, "largest_block": nmax,
                   "sort": a.sort, "tridiag_schedule": a.tridiag_align,
                   "form": ("batched eigensolve (FP64) + split-precision contraction GEMMs "
                            f"({a.split} slices on FP16 matrix cores)") if a.split > 0 else
                           "batched eigensolve (FP64) + FP64 contraction GEMMs"},
        "stages_s_per_step": {k: v / a.steps for k, v in stages.items()},
        "eig_alg_tflops": flops_eig / max(stages.get("gpu_eig", 1e-9), 1e-9) / 1e12,
        "plain_fp64": plain,
        "batched_fp64": batched64,
        "batched_split": batched_split,
        "batched_over_plain": (value / plain["value"]) if plain else None,
    }
    if a.spectrum_leg:
        out["spectrum_pipeline"] = spectrum_pipeline(gpu, models, Host)
    if a.synthetic:
        out["synthetic_split_gemm"] = synthetic_gemm(gpu, (8, 6))
    if ks:
        fam = max(ks, key=lambda f: ks[f]["ms"])
        k = ks[fam]
        if fam in ("tridiag_symv", "tridiag_col", "assemble"):
            ach = k["bytes"] / (k["ms"] * 1e-3) / 1e9
            roof = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": ach / HBM_PEAK_GBS, "traffic": None}
        else:
            ach = k["flops"] / (k["ms"] * 1e-3) / 1e12
            roof = {"bound": "mfma", "achieved": ach, "peak": FP64_PEAK_TFS, "unit": "TFLOP/s",
                    "frac": ach / FP64_PEAK_TFS, "traffic": None}
        roof["kernel_family"] = fam
        roof["launches_timed"] = k["launches"]
        roof["timing"] = (f"HIP events on every {a.profile_every}-th launch of the family over the timed steps, "
                          f"each such tridiagonal SYMV launch with the device to itself (achieved = algorithmic "
                          f"bytes or flops / event time of those launches)")
        roof["avg_launch_us"] = 1e3 * k["ms"] / max(k["launches"], 1)
        roof["alg_bytes_per_launch"] = k["bytes"] / max(k["launches"], 1)
        roof["alg_flops_per_launch"] = k["flops"] / max(k["launches"], 1)
        tr = profiled_traffic(fam, a.workload)
        if tr:
            roof["traffic"] = tr["traffic_bytes_per_launch"]
            roof["traffic_sample"] = {"alg_bytes_per_launch": tr.get("alg_bytes_per_launch"),
                                      "traffic_over_algorithmic": tr.get("traffic_over_algorithmic"),
                                      "source": tr["source"]}
        out["roofline"] = roof
        # sampled launches scaled to all launches of the family: kernel time
        # summed over launches (launches of different streams overlap in the
        # wall time of a step)
        out["kernel_ms_per_step_est"] = {f: v["ms"] * a.profile_every / a.steps for f, v in ks.items()
                                         if v["launches"]}
    else:
        out["roofline"] = None
    if not a.no_cpu_baseline:
        cb = cpu_baseline(sizes, a.cpu_seconds)
        out["cpu_baseline"] = cb
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
