#!/usr/bin/env python3
"""Benchmark: NRG iterations per second on one MI355X.

Workload: two-orbital Anderson model with Kanamori interaction (U=0.3,
J=0.05, Gamma=0.01 per orbital, two Wilson chains, Lambda=2, local site of
dimension 16), Nkeep=4000, 45 iterations.  One step is one NRG iteration of
the forward pass: prepareBasis + block-Hamiltonian assembly + batched FP64
eigensolve of all (Q,Sz) blocks + truncation + invariant-matrix
contractions, driven by the Fortran host (nrg-mpi_amd/lib/libnrg_host.so)
through the HIP library.  Warmup iterations run the growth phase (-1, 0, 1,
...); the timed steps are steady-state truncating iterations (~64k states,
~99 blocks up to ~4100).

Prints one JSON line: metric, value (iter/s), ms_per_step, config, the
per-stage seconds, the executed FP64 flops and the percentage of the FP64
matrix-core peak they reach over the step, `roofline` of the dominant
kernel family (algorithmic bytes over its HIP-event time against 8 TB/s),
three-iteration comparison legs (plain FP64: one block per eigensolver
pass; batched with split-precision contraction products), `cpu_baseline`
(one whole steady-state iteration of the oracle measured on the host
cores, next to the reference's own timing from its fixture run), the
synthetic 256-block batch, and the Nkeep=8000 pipeline with spectra in FP64
and split precision.
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
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="two_orbital_nk4000_it45")
    ap.add_argument("--sort", choices=["exact", "canonical"], default="exact")
    ap.add_argument("--no-profile", action="store_true", help="skip the HIP-event kernel timing")
    ap.add_argument("--profile-every", type=int, default=307,
                    help="time every n-th launch of the profiled families in the timed steps (307: prime, so "
                         "coprime to the 32-column panel); the launches are timed in place, sharing the device "
                         "with the other stream groups")
    ap.add_argument("--isolated-every", type=int, default=61,
                    help="sampling stride of the isolated-launch leg after the timed steps (two iterations whose "
                         "timed tridiagonal SYMV launches are fenced against the other stream groups, so that "
                         "their event interval is the kernel alone; 0 skips the leg)")
    ap.add_argument("--profile-families", default="tridiag_symv",
                    help="comma-separated kernel families whose launches carry HIP timing events in the "
                         "timed steps (default: the dominant one, tridiag_symv; 'all' for every family)")
    ap.add_argument("--launch-log", default="",
                    help="write a CSV of every launch of the profiled families (ordinal per kernel function "
                         "from the first iteration, algorithmic bytes and flops; phase warmup/timed)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--graphs", choices=["on", "off"], default="off",
                    help="capture the tridiagonalisation's launch sequence in a HIP graph per block set and "
                         "replay it when the same block set returns (default off: the block tables change "
                         "from one NRG iteration to the next, so graphs are captured and never replayed, and "
                         "capture costs more than launching; the bench measures the graph form in a leg of "
                         "its own)")
    ap.add_argument("--tuning", choices=["cache", "off"], default="cache",
                    help="launch configuration of the tridiagonalisation from the tuner's cache (per block-size "
                         "histogram) or the defaults")
    ap.add_argument("--tuning-cache", default=os.path.join(ROOT, "nrg-mpi_amd", "tuning", "tridiag_mi355x.txt"))
    ap.add_argument("--tune", action="store_true",
                    help="only run the tuner over the workload (--tune-passes passes, exploring and updating the "
                         "cache file) and exit")
    ap.add_argument("--tune-passes", type=int, default=2,
                    help="passes over the workload with --tune; a block set seen once per pass needs 15 passes "
                         "for every candidate to be timed three times")
    ap.add_argument("--leg-budget", type=float, default=240.0,
                    help="start the Nkeep=8000 leg only if the run so far took less than this many seconds")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a plain device-to-device copy of BYTES bytes five times (the reference "
                         "kernel that scales the FETCH_SIZE/WRITE_SIZE counters to bytes) and exit")
    ap.add_argument("--tridiag", choices=["one-stage", "two-stage"], default="one-stage",
                    help="tridiagonalisation: one-stage Householder columns with a trailing SYMV, or two-stage "
                         "(band of 32 by panel QRs and matrix-core updates, then bulge chasing)")
    ap.add_argument("--symv", choices=["atomic", "rows"], default="atomic",
                    help="SYMV of the one-stage reduction: row parts through LDS atomics, or summed in registers and "
                         "LDS row chunks in a fixed order")
    ap.add_argument("--tridiag-align", choices=["end", "start"], default="end",
                    help="column schedule of the batched tridiagonalisation (blocks end or start together)")
    ap.add_argument("--split", type=int, default=0,
                    help="slices of the split-precision contraction GEMMs in the timed steps (0: FP64 matrix "
                         "cores, the faster form on MI355X; the split form is measured in its own iteration)")
    ap.add_argument("--split-leg", type=int, default=8,
                    help="slices of the invariant-matrix products in the three-iteration split-precision "
                         "comparison and in the split form of the Nkeep=8000 leg (0: skip)")
    ap.add_argument("--split-spectral", type=int, default=6,
                    help="slices of the spectral-side products (spectral invariant, CFS weights, rho) in the "
                         "split forms: they feed only the binned spectrum, whose 1e-8 budget 6 slices meet")
    ap.add_argument("--split-min-flops", type=float, default=4e9,
                    help="batches below this many FP64-equivalent flops stay on FP64 matrix cores in the split "
                         "forms")
    ap.add_argument("--no-spectrum-leg", dest="spectrum_leg", action="store_false",
                    help="skip the Nkeep=8000 full pipeline with spectra (forward + rho sweep + CFS, FP64 and "
                         "split contraction products)")
    ap.add_argument("--no-synthetic", dest="synthetic", action="store_false",
                    help="skip the synthetic 256-block batch (eigensolve, FP64 vs split-precision U^T O U)")
    ap.add_argument("--no-graph-leg", dest="graph_leg", action="store_false",
                    help="skip the three-iteration leg with HIP graphs (profiled runs skip it: under "
                         "rocprofv3 --kernel-trace its first graph launch failed)")
    ap.add_argument("--no-two-stage-leg", dest="two_stage_leg", action="store_false",
                    help="skip the three-iteration leg with the two-stage tridiagonalisation")
    ap.add_argument("--no-plain", dest="plain", action="store_false",
                    help="skip the three-iteration plain (one block per pass) comparison")
    return ap.parse_args()


MFMA_FAMILIES = ("tridiag_gemm", "dc_gemm", "back_gemm", "invariant_gemm", "spectral_gemm", "cfs_products",
                 "cfs_rho_weight", "rho_gemm")
EIG_FAMILIES = ("tridiag_symv", "tridiag_gemm", "dc_gemm", "back_gemm")


def reference_count_csv(workload):
    """The reference's own per-iteration cost in this build's fixture run
    (count.f90's nrgIteration row: wall seconds over the run's iterations)."""
    import numpy as np
    path = os.path.join(ROOT, "tests", "golden", f"ref_{workload}.npz")
    if not os.path.exists(path):
        return None
    fx = np.load(path)
    if "count_csv" not in fx:
        return None
    rows = [l.split(",") for l in str(fx["count_csv"]).splitlines()]
    for r in rows:
        if r and r[0].strip() == "nrgIteration":
            tot = float(r[1])
            n = int(r[2])
            out = {"seconds_total": tot, "iterations": n, "mpi_ranks": int(fx["mpi_ranks"])}
            if "iteration_wall_s" in fx:
                w = fx["iteration_wall_s"]
                steady = [float(x) for it, x in zip(fx["iterations"], w) if it >= 4]
                out["steady_state_s_per_iteration"] = float(np.mean(steady)) if steady else None
            out.update({
                    "note": "reference build (amdflang -O2, MKL sequential) in the build container, "
                            "iterations -1..%d of the 45-iteration deck, count.csv nrgIteration row" %
                            int(fx["iterations"][-1])})
            return out
    return None


def cpu_baseline(h, gpu, deck, it, workers):
    """One whole steady-state NRG iteration of the oracle (the numpy
    restatement of the reference path) on the host cores, measured: it
    continues from the state the GPU run reached before iteration `it`
    (kept basis records, energies, invariant matrix), runs prepareBasis with
    the reference's exchange-sort permutation, the block assembly, one LAPACK
    eigensolve per block on a pool of `workers` threads (the reference's
    OpenMP loop over blocks), the truncation and the invariant matrices
    (dense coefficient matrices and two GEMMs per block pair, the
    reference's form).  Returns the baseline and the oracle's record of the
    iteration for comparison with the GPU's."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import nrg_oracle as O
    m = O.read_inputs(deck)
    nrg = O.NRG(m, workers=workers)
    k = h.i["nkept"]
    nrg.seed(h.basis_output(), h.eigenvalues(), gpu.invariant_get(k, h.i["ncm"]), k, h.i["start_trunc"])
    t = [time.perf_counter()]
    nrg.prepare_basis()
    t.append(time.perf_counter())
    nrg.diagonalize(it)
    t.append(time.perf_counter())
    nrg.calc_thermo()
    nrg.truncate(it)
    t.append(time.perf_counter())
    nrg.invariant_matrix()
    t.append(time.perf_counter())
    total = t[-1] - t[0]
    stages = dict(prepare=t[1] - t[0], diagonalize=t[2] - t[1], truncate=t[3] - t[2], invariant=t[4] - t[3])
    rec = dict(nkept=nrg.nkept, kept_blk=nrg.kept_per_block(), E=np.sort(nrg.eigenvalue))
    return dict(value=1.0 / total, unit="iter/s", cores=workers, kind="port",
                sample=f"one whole steady-state iteration ({it}) of the two-orbital Nkeep=4000 workload, "
                       f"measured, nothing extrapolated: the oracle's prepareBasis (exact exchange-sort "
                       f"permutation), block assembly, LAPACK eigensolve per block on {workers} threads, "
                       f"truncation and invariant matrices, continued from the GPU run's state after "
                       f"iteration {it - 1}",
                seconds_per_iteration=total, stages_s=stages, nproc=os.cpu_count()), rec


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


def synthetic_batch(log):
    """The synthetic block batch at its stated size: 256 blocks, sizes
    log-normal (median 600, sigma 1.0 in ln-space, seed 20240) clipped to
    16..4096, eigenvalues 2^(-k/2), k uniform on 0..60, Haar rotations, four
    N(0,1) operators per block.  Batched FP64 eigensolve checked against the
    planted ladder; U^T O U in FP64 and in split precision (8 and 6 slices),
    device-timed (mean of one repeat), errors against numpy (FP64, sampled
    blocks) and against the FP64 products (split, every block)."""
    from nrgamd import gpu, synthetic as S
    gpu.init(0)
    t0 = time.perf_counter()
    b = S.make_batch(seed=20240, nblk=256, median=600.0, sigma=1.0)
    gen = time.perf_counter() - t0
    r = S.run(b, nops=4, split=(8, 6), repeat=1, sample=8, log=log)
    r["generation_s"] = gen
    r["config"] = ("256 blocks, n ~ log-normal(median 600, sigma 1.0) clipped to [16, 4096], seed 20240, "
                   "eigenvalues 2^(-k/2) k in 0..60 (integers), Haar rotations, 4 N(0,1) operators per block")
    return r


def workload_nk8000(gpu, models, Host, log, split_cfg):
    """Two-orbital Kanamori model at Nkeep=8000 with spectra (45 iterations,
    4 spectral types): forward pass with stocks, rho sweep and CFS spikes,
    with FP64 contraction products and with split-precision products
    (split_cfg = invariant slices, spectral slices, minimum batch flops);
    the FP64 form is the oracle of the split one."""
    import numpy as np
    out = {"config": "two-orbital Kanamori U=0.3 J=0.05 Gamma=0.01/orbital, Nkeep=8000, 45 iterations, "
                     "4 spectral types"}
    bins = {}
    for split in (0, split_cfg[0]):
        tmp = tempfile.mkdtemp(prefix="nrg8000_")
        models.write_inputs(models.workload("two_orbital_nk8000_it45"), tmp)
        h = Host(tmp, device=0)
        if split:
            gpu.load().nrg_gpu_set_gemm_split_families(split_cfg[0], split_cfg[1], split_cfg[2])
        else:
            gpu.load().nrg_gpu_set_gemm_split(0)
        try:
            t0 = time.perf_counter()
            steady = []
            for it in h.iterations():
                ti = time.perf_counter()
                h.iterate(it)
                if 4 <= it < h.i["n_iter"]:
                    steady.append(time.perf_counter() - ti)
            t1 = time.perf_counter()
            h.spectrum()
            t2 = time.perf_counter()
            st = h.spectrum_times()
            neg, pos = h.bins()
            tr = h.rho_traces()
            key = "split" if split else "fp64"
            out[key] = {"forward_s": t1 - t0, "steady_state_iter_per_s": len(steady) / sum(steady),
                        "stock_total_s": st["stock"], "rho_sweep_s": st["rho"], "cfs_s": st["cfs"],
                        "pipeline_s": t2 - t0,
                        "sum_rule_max_dev": float(np.max(np.abs(neg.sum(1) + pos.sum(1) - 1.0))),
                        "trace_rho_max_dev": float(max(abs(v - 1.0) for v in tr.values()))}
            if split:
                out[key]["slices"] = {"invariant": split_cfg[0], "spectral": split_cfg[1],
                                      "min_batch_flops": split_cfg[2]}
            bins[split] = (neg, pos)
            if split:
                peak = max(np.abs(bins[0][0]).max(), np.abs(bins[0][1]).max())
                diff = max(np.abs(neg - bins[0][0]).max(), np.abs(pos - bins[0][1]).max())
                out[key]["max_bin_diff_vs_fp64_rel_peak"] = float(diff / peak)
            log(f"Nkeep=8000 {key}: pipeline {t2 - t0:.1f} s")
        finally:
            gpu.load().nrg_gpu_set_gemm_split(0)
            h.close()
    return out


def copy_calibration(nbytes):
    """A plain HBM copy of known size (the library's k_stream_copy): under
    `rocprofv3 --pmc FETCH_SIZE` (or WRITE_SIZE) its counter reading per
    dispatch, against nbytes, gives the factor that turns the counters of the
    NRG kernels into bytes."""
    import ctypes
    from nrgamd import gpu
    gpu.init(0)
    ok = ctypes.c_int(0)
    gpu._check(gpu.load().nrg_gpu_copy_calibration(int(nbytes) // 16 * 16, 5, ctypes.byref(ok)))
    print(json.dumps({"copy_bytes": int(nbytes) // 16 * 16, "dispatches": 5, "kernel": "k_stream_copy",
                      "ok": bool(ok.value)}))


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
    t_start = time.perf_counter()

    def log(msg):
        print(f"[{time.perf_counter() - t_start:7.1f} s] {msg}", file=sys.stderr, flush=True)

    spec = models.workload(a.workload)
    if a.tune:
        os.makedirs(os.path.dirname(a.tuning_cache), exist_ok=True)
        for p in range(a.tune_passes):
            tmp = tempfile.mkdtemp(prefix="nrgtune_")
            models.write_inputs(spec, tmp)
            h = Host(tmp, device=0)
            gpu.load().nrg_gpu_set_tuning(1, a.tuning_cache.encode())
            for it in h.iterations():
                h.iterate(it)
            h.close()
            log(f"tuning pass {p}: cache {a.tuning_cache}")
        print(open(a.tuning_cache).read())
        return
    need = a.warmup + a.steps
    legs = 3 * (int(a.plain) + int(a.split_leg > 0 or a.split > 0) + 1 + int(a.graph_leg) +
                int(a.two_stage_leg and a.tridiag == "one-stage")) + (2 if a.isolated_every > 0 else 0)
    if need + legs > spec.niter:
        raise SystemExit(f"warmup+steps+legs={need + legs} exceeds the {spec.niter} iterations of {a.workload}")
    tmp = tempfile.mkdtemp(prefix="nrgbench_")
    models.write_inputs(spec, tmp)
    h = Host(tmp, device=0, sort_mode=0 if a.sort == "exact" else 1)
    gpu.load().nrg_gpu_set_tridiag_schedule(1 if a.tridiag_align == "end" else 0)
    gpu.load().nrg_gpu_set_tridiag_method(1 if a.tridiag == "two-stage" else 0)
    gpu.load().nrg_gpu_set_tridiag_symv(1 if a.symv == "rows" else 0)
    gpu.load().nrg_gpu_set_gemm_split(a.split)
    tuned = a.tuning == "cache" and os.path.exists(a.tuning_cache)
    gpu.load().nrg_gpu_set_tuning(2 if tuned else 0, a.tuning_cache.encode() if tuned else None)
    gpu.load().nrg_gpu_set_graphs(1 if a.graphs == "on" else 0)
    log_fams = [f for f in a.profile_families.split(",") if f and f != "all"] or list(gpu.KERNEL_FAMILIES)
    if a.launch_log:
        gpu.launch_log(log_fams)
    its = list(range(spec.first_iteration, spec.first_iteration + need))
    for it in its[:a.warmup]:
        h.iterate(it)
        log(f"warmup iteration {it}: nfull={h.i['nfull']} nsub={h.i['nsub']} kept={h.i['nkept']}")
    if not a.no_profile:
        fams = None if a.profile_families == "all" else [f for f in a.profile_families.split(",") if f]
        gpu.profile(True, families=fams, every=a.profile_every, exclusive=False)
    stages = {}
    n_warm_log = len(gpu.launch_log_get()) if a.launch_log else 0
    gpu.work_counters(reset=True)
    t0 = time.perf_counter()
    for it in its[a.warmup:]:
        h.iterate(it)
        t = h.times()
        st = gpu.stage_times()
        for k in ("prepare", "diagonalization", "postprocess", "invariant"):
            stages[k] = stages.get(k, 0.0) + t[k]
        for k in ("assemble", "eig", "tridiag", "tridiag_issue", "dc", "back"):
            stages["gpu_" + k] = stages.get("gpu_" + k, 0.0) + st[k] * 1e-3
        log(f"timed iteration {it}: {t['iteration']:.3f} s  (prepare {t['prepare']:.3f}, diag "
            f"{t['diagonalization']:.3f} [tri {st['tridiag']:.0f} ms (issued in {st['tridiag_issue']:.0f} ms), "
            f"dc {st['dc']:.0f} ms, back "
            f"{st['back']:.0f} ms], post {t['postprocess']:.3f}) nfull={h.i['nfull']} nsub={h.i['nsub']}")
    elapsed = time.perf_counter() - t0
    work = gpu.work_counters(reset=True)
    stg = gpu.stage_times()
    gpu_graphs = (int(stg["graph_captures"]), int(stg["graph_replays"]))
    # the launch log runs on through the first leg and the isolated leg (a
    # launch's ordinal counts from the first iteration, as the profiler's
    # per-kernel dispatch counter does); records are labelled by phase
    n_timed_log = len(gpu.launch_log_get()) if a.launch_log else 0
    n_iso_log = None

    def write_launch_log():
        L = gpu.launch_log_get()
        names = {int(k): gpu.launch_log_kernel(int(k)) for k in set(L[:, 2].tolist())}
        gpu.launch_log(None)

        def phase(i):
            if i < n_warm_log:
                return "warmup"
            if i < n_timed_log:
                return "timed"
            return "isolated" if (n_iso_log is not None and i >= n_iso_log) else "leg"
        with open(a.launch_log, "w") as f:
            f.write("family,kernel,threads,ordinal,bytes,flops,phase,events\n")
            for i, r in enumerate(L):
                f.write(f"{gpu.KERNEL_FAMILIES[int(r[0])]},{names[int(r[2])]},{int(r[1])},{int(r[3])},"
                        f"{r[4]:.0f},{r[5]:.0f},{phase(i)},{int(r[6])}\n")
    ks = gpu.kernel_stats() if not a.no_profile else {}
    if not a.no_profile:
        gpu.profile(False)
    sizes = h.subspace()[:, h.i["nv"]].tolist()
    nfull, nsub, nmax = h.i["nfull"], h.i["nsub"], int(max(sizes))
    value = a.steps / elapsed
    ms_step = 1e3 * elapsed / a.steps

    # ---- CPU baseline: the oracle continues from the GPU run's state and
    # runs the next iteration on the host cores; the first comparison leg
    # below computes the same iteration on the GPU
    nxt = its[-1] + 1
    cb = None
    cpu_rec = None
    if not a.no_cpu_baseline:
        workers = min(16, os.cpu_count() or 1)
        log(f"CPU baseline: oracle iteration {nxt} on {workers} threads")
        cb, cpu_rec = cpu_baseline(h, gpu, tmp, nxt, workers)
        cb["reference_run"] = reference_count_csv(a.workload)
        log(f"CPU baseline: {cb['seconds_per_iteration']:.1f} s per iteration")

    def leg(split, batch_limit, label, graphs=False, use_tuning=False):
        nonlocal nxt
        if split:
            gpu.load().nrg_gpu_set_gemm_split_families(split, a.split_spectral, a.split_min_flops)
        else:
            gpu.load().nrg_gpu_set_gemm_split(0)
        gpu.load().nrg_gpu_set_batch_limit(batch_limit)
        gpu.load().nrg_gpu_set_graphs(1 if graphs else 0)
        if use_tuning and tuned:
            gpu.load().nrg_gpu_set_tuning(2, a.tuning_cache.encode())
        else:
            gpu.load().nrg_gpu_set_tuning(0, None)
        dts, itl = [], []
        try:
            for _ in range(3):
                tb = time.perf_counter()
                h.iterate(nxt)
                dts.append(time.perf_counter() - tb)
                itl.append(nxt)
                if cpu_rec is not None and nxt == its[-1] + 1:
                    E = np.sort(h.eigenvalues())
                    sc = max(1.0, float(E.max()))
                    n = min(len(E), len(cpu_rec["E"]))
                    cb["agrees_with_gpu"] = {
                        "iteration": nxt, "kept_equal": int(h.i["nkept"]) == int(cpu_rec["nkept"]),
                        "kept_per_block_equal": bool(np.array_equal(h.kept_per_block(), cpu_rec["kept_blk"])),
                        "max_eigenvalue_diff_rel": float(np.max(np.abs(E[:n] - cpu_rec["E"][:n])) / sc)}
                nxt += 1
        finally:
            gpu.load().nrg_gpu_set_batch_limit(0)
            gpu.load().nrg_gpu_set_gemm_split(0)
        log(f"{label} iterations {itl}: {[round(x, 3) for x in dts]} s")
        return {"value": len(dts) / sum(dts), "unit": "iter/s", "iterations": itl, "seconds": dts}

    plain = batched_split = None
    # the timed form with the default launch configuration instead of the
    # tuned one, and with HIP graphs
    untuned = leg(0, 0, "batched FP64, no graphs, default launch configuration")
    untuned["form"] = "batched FP64 eigensolve and contractions, launches issued one by one, default configuration"
    # the dominant kernel with the device to itself: two more iterations in
    # the timed form whose sampled SYMV launches are fenced against the
    # other stream groups (not part of the timed value)
    isolated = None
    if a.isolated_every > 0 and not a.no_profile:
        if a.launch_log:
            n_iso_log = len(gpu.launch_log_get())
        gpu.load().nrg_gpu_set_tuning(2 if tuned else 0, a.tuning_cache.encode() if tuned else None)
        gpu.profile(True, families=["tridiag_symv"], every=a.isolated_every, exclusive=True)
        itl = []
        for _ in range(2):
            h.iterate(nxt)
            itl.append(nxt)
            nxt += 1
        ki = gpu.kernel_stats().get("tridiag_symv", {})
        gpu.profile(False)
        if ki.get("launches") and ki.get("ms", 0) > 0:
            ach = ki["bytes"] / (ki["ms"] * 1e-3) / 1e9
            isolated = {"achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                        "launches_timed": ki["launches"], "avg_launch_us": 1e3 * ki["ms"] / ki["launches"],
                        "alg_bytes_per_launch": ki["bytes"] / ki["launches"], "iterations": itl,
                        "timing": f"HIP events on every {a.isolated_every}-th tridiagonal SYMV launch of "
                                  f"iterations {itl}, each fenced against the other stream groups (the kernel "
                                  f"with the device to itself)"}
            log(f"isolated SYMV launches (iterations {itl}): {isolated['avg_launch_us']:.2f} us, "
                f"{ach / 1e3:.2f} TB/s")
    if a.launch_log:
        write_launch_log()
    with_graphs = None
    if a.graph_leg:
        with_graphs = leg(0, 0, "batched FP64, HIP graphs, tuned launch configuration", graphs=True,
                          use_tuning=True)
        with_graphs["form"] = ("batched FP64 eigensolve and contractions, tridiagonalisation captured into HIP "
                               "graphs (replayed when a block set returns), tuned configuration")
        st_g = gpu.stage_times()
        with_graphs["graph_captures"] = int(st_g["graph_captures"]) - gpu_graphs[0]
        with_graphs["graph_replays"] = int(st_g["graph_replays"]) - gpu_graphs[1]
    two_stage = None
    if a.two_stage_leg and a.tridiag == "one-stage":
        gpu.load().nrg_gpu_set_tridiag_method(1)
        try:
            two_stage = leg(0, 0, "batched FP64, two-stage tridiagonalisation")
        finally:
            gpu.load().nrg_gpu_set_tridiag_method(0)
        two_stage["form"] = ("batched FP64 eigensolve with the two-stage tridiagonalisation (band of 32 by panel QRs "
                             "and matrix-core updates, bulge chasing one workgroup per block), FP64 contractions")
    if a.plain:
        # one block per eigensolver pass (the per-block schedule of the
        # reference's dsyev loop) with FP64 contraction products
        plain = leg(0, 1, "plain FP64")
        plain["form"] = "plain FP64: one block per eigensolver pass, FP64 contraction products"
    if a.split == 0 and a.split_leg > 0:
        batched_split = leg(a.split_leg, 0, "batched split")
        batched_split["form"] = (f"batched eigensolve (FP64) + split-precision contraction products "
                                 f"({a.split_leg} slices for the invariant matrices, {a.split_spectral} for "
                                 f"the spectral-side families, batches under {a.split_min_flops:.0e} flops in "
                                 f"FP64)")
    h.close()

    # executed work of the timed steps, per step
    mfma = sum(work[f]["flops"] for f in MFMA_FAMILIES) / a.steps
    vec = work["tridiag_symv"]["flops"] / a.steps
    eig_flops = sum(work[f]["flops"] for f in EIG_FAMILIES) / a.steps
    eig_s = (stages.get("gpu_tridiag", 0) + stages.get("gpu_dc", 0) + stages.get("gpu_back", 0)) / a.steps
    executed = {
        "mfma_fp64_flops_per_step": mfma,
        "vector_fp64_flops_per_step": vec,
        "mfma_fp64_tflops_over_step": mfma / (ms_step * 1e-3) / 1e12,
        "pct_fp64_mfma_peak_over_step": 100.0 * mfma / (ms_step * 1e-3) / (FP64_PEAK_TFS * 1e12),
        "eigensolver_fp64_tflops": eig_flops / max(eig_s, 1e-12) / 1e12,
        "eigensolver_pct_fp64_peak": 100.0 * eig_flops / max(eig_s, 1e-12) / (FP64_PEAK_TFS * 1e12),
        "note": "flops actually issued in the timed steps (every launch counted): GEMM products 2mnk "
                "(lower-triangle SYR2K halved) on FP64 matrix cores, the tridiagonal SYMV's 2m^2 per column "
                "on the vector units; eigensolver rate = (SYMV + SYR2K + D&C merge + back-transform flops) "
                "over the tridiagonalisation + D&C + back-transform time",
        "per_family": {f: {"launches": v["launches"] / a.steps, "flops": v["flops"] / a.steps,
                           "bytes": v["bytes"] / a.steps} for f, v in work.items() if v["launches"]},
        "per_family_unit": "launches, flops and bytes per timed step",
    }
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "iter/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
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
                   "sort": a.sort, "tridiag_schedule": a.tridiag_align, "tridiag": a.tridiag, "symv": a.symv,
                   "form": ("batched eigensolve (FP64) + split-precision contraction GEMMs "
                            f"({a.split} slices on int8 matrix cores)") if a.split > 0 else
                           "batched eigensolve (FP64) + FP64 contraction GEMMs"},
        "stages_s_per_step": {k: v / a.steps for k, v in stages.items()},
        "executed_flops": executed,
        "pct_mfma_peak": executed["pct_fp64_mfma_peak_over_step"],
        "plain_fp64": plain,
        "batched_split": batched_split,
        "batched_over_plain": (value / plain["value"]) if plain else None,
        "batched_split_over_plain": (batched_split["value"] / plain["value"]) if (plain and batched_split) else None,
        "launch_options": {"graphs": a.graphs, "tuning": ("cache " + os.path.relpath(a.tuning_cache, ROOT))
                           if tuned else "off", "graph_captures": gpu_graphs[0], "graph_replays": gpu_graphs[1]},
        "batched_no_graphs_default_config": untuned,
        "batched_graphs_tuned": with_graphs,
        "batched_two_stage": two_stage,
    }
    ks = {f: v for f, v in ks.items() if v.get("launches") and v.get("ms", 0) > 0}
    if ks:
        fam = max(ks, key=lambda f: ks[f]["ms"])
        k = ks[fam]
        in_place = None
        if fam == "tridiag_symv" and isolated:
            # The SYMV's own roofline comes from the fenced launches of the
            # isolated leg.  An in-place launch's event interval also holds
            # the time its dispatch waits for CU slots that the other stream
            # groups' kernels hold (rocprofv3 over the same launches shows
            # about half of it), so it is reported beside, not as, the
            # kernel's duration.
            ach_in = k["bytes"] / (k["ms"] * 1e-3) / 1e9
            in_place = {"achieved": ach_in, "frac": ach_in / HBM_PEAK_GBS, "unit": "GB/s",
                        "launches_timed": k["launches"], "avg_launch_us": 1e3 * k["ms"] / k["launches"],
                        "alg_bytes_per_launch": k["bytes"] / k["launches"],
                        "timing": f"HIP events on every {a.profile_every}-th launch of the family over the "
                                  f"timed steps, timed in place: the interval runs from the dispatch to the end "
                                  f"of the kernel and includes waiting for CU slots held by the other stream "
                                  f"groups' kernels"}
            k = {"bytes": isolated["alg_bytes_per_launch"] * isolated["launches_timed"],
                 "ms": isolated["avg_launch_us"] * isolated["launches_timed"] * 1e-3,
                 "launches": isolated["launches_timed"], "flops": 0.0}
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
        roof["timing"] = (isolated["timing"] + "; achieved = algorithmic bytes / event time of those launches"
                          if in_place else
                          f"HIP events on every {a.profile_every}-th launch of the family over the timed steps "
                          f"(achieved = algorithmic bytes or flops / event time of those launches)")
        roof["avg_launch_us"] = 1e3 * k["ms"] / max(k["launches"], 1)
        roof["alg_bytes_per_launch"] = k["bytes"] / max(k["launches"], 1)
        if in_place:
            roof["iterations"] = isolated["iterations"]
            roof["in_place"] = in_place
        else:
            roof["alg_flops_per_launch"] = k["flops"] / max(k["launches"], 1)
        tr = profiled_traffic(fam, a.workload)
        if tr:
            roof["traffic"] = tr["traffic_bytes_per_launch"]
            roof["traffic_sample"] = {"alg_bytes_per_launch": tr.get("alg_bytes_per_launch"),
                                      "traffic_over_algorithmic": tr.get("traffic_over_algorithmic"),
                                      "source": tr["source"]}
        if fam == "tridiag_symv":
            # the same launches with the device to themselves (the isolated
            # leg), and the whole family's algorithmic bytes over the time of
            # the tridiagonalisations in the timed steps (the stream groups'
            # launches overlap, so one launch's event interval under-states
            # the bandwidth the family draws)
            tt = stages.get("gpu_tridiag", 0.0)
            if tt > 0:
                agg = work["tridiag_symv"]["bytes"] / tt / 1e9
                roof["family_over_tridiag"] = {
                    "achieved": agg, "frac": agg / HBM_PEAK_GBS, "unit": "GB/s",
                    "bytes_per_step": work["tridiag_symv"]["bytes"] / a.steps,
                    "tridiag_ms_per_step": 1e3 * tt / a.steps,
                    "timing": "algorithmic bytes of every tridiagonal SYMV launch in the timed steps / the "
                              "event-timed span of the tridiagonalisations (which also runs the column kernels "
                              "and trailing updates)"}
        out["roofline"] = roof
        # sampled launches scaled to all launches of the family: kernel time
        # summed over launches (launches of different streams overlap in the
        # wall time of a step)
        out["kernel_ms_per_step_est"] = {f: v["ms"] * a.profile_every / a.steps for f, v in ks.items()
                                         if v["launches"]}
    else:
        out["roofline"] = None
    out["cpu_baseline"] = cb
    if a.synthetic:
        log("synthetic 256-block batch")
        out["synthetic_batch"] = synthetic_batch(log)
    if a.spectrum_leg:
        if time.perf_counter() - t_start < a.leg_budget:
            out["workload_nk8000"] = workload_nk8000(gpu, models, Host, log,
                                                     (a.split_leg or 8, a.split_spectral, a.split_min_flops))
        else:
            out["workload_nk8000"] = {"skipped": f"over the {a.leg_budget:.0f} s budget for starting this leg"}
    log("done")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
