"""Summaries of rocprofv3 CSV output for the kernel profiles of this library.

    python -m nrgamd.profsum stats  DIR [--title T] [--command C]
        kernel table from the ``*kernel_stats.csv`` of a
        ``rocprofv3 --kernel-trace --stats -f csv`` run;
        with ``--log LOG --family F``: also the mean duration of the family's
        dispatches in the timed steps, located through the launch log that
        bench.py --launch-log wrote in the same run;
    python -m nrgamd.profsum mfma DIR --kernel REGEX [--mops COUNTER --scale S --peak P]
        matrix-core utilisation (MfmaUtil) and executed operations over the
        dispatch durations of a ``rocprofv3 --kernel-trace --pmc MfmaUtil
        SQ_INSTS_VALU_MFMA_MOPS_F64`` (or ..._I8) run;
    python -m nrgamd.profsum first-timed --log LOG --kernel K --ints I,J
        per-kernel ordinal of the first timed dispatch of a kernel variant;
    python -m nrgamd.profsum traffic --log LOG --copy-fetch DIR --copy-write DIR
        --copy-bytes B --variant REGEX FIRST FETCH_DIR WRITE_DIR [--variant ...]
        HBM bytes per dispatch of each kernel variant over the dispatches that
        ``rocprofv3 --pmc FETCH_SIZE`` / ``--pmc WRITE_SIZE`` runs counted
        (KiB per dispatch), each scaled by what the same counter reads for a
        plain device copy of B bytes (B read and B written per dispatch), next
        to the algorithmic bytes of the same dispatches from the launch log.

The traffic summary ends with ``traffic_bytes_per_launch``,
``alg_bytes_per_launch`` and ``launches`` lines that bench.py reads back.
"""
from __future__ import annotations

import argparse
import csv
import glob
import os
import re
import sys


def _rows(d, suffix):
    files = sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True))
    if not files:
        raise SystemExit(f"no *{suffix} under {d}")
    out = []
    for f in files:
        with open(f, newline="") as fh:
            out.extend(csv.DictReader(fh))
    return out


def _col(row, *names):
    low = {k.lower(): k for k in row}
    for n in names:
        if n.lower() in low:
            return row[low[n.lower()]]
    raise KeyError(f"none of {names} in {list(row)}")


def read_log(path):
    """Launch log of bench.py --launch-log: list of dicts."""
    with open(path, newline="") as fh:
        return list(csv.DictReader(fh))


def kernel_key(name):
    """(base name, template integers) of a demangled or mangled kernel name,
    e.g. 'void nrg::k_tr_symv<2, 16>(...)' and '_ZN3nrg9k_tr_symvILi2ELi16E...'
    both give ('k_tr_symv', (2, 16))."""
    m = re.search(r"(k_[a-z0-9_]+)", name)
    base = m.group(1) if m else name
    if "<" in name:
        t = re.search(r"<([^>]*)>", name)
        ints = tuple(int(x) for x in re.findall(r"-?\d+", t.group(1))) if t else ()
    else:
        ints = tuple(int(x) for x in re.findall(r"Li(-?\d+)E", name))
    return base, ints


def _ordinals(rows, pat):
    """Kernel-trace or counter rows of kernels matching pat, numbered per
    kernel name in dispatch order from 1 (the profiler's per-kernel dispatch
    count), as {name: [(ordinal, row), ...]}."""
    out = {}
    rows = sorted(rows, key=lambda r: int(_col(r, "Dispatch_Id", "Dispatch_ID")))
    for r in rows:
        name = _col(r, "Kernel_Name")
        if not re.search(pat, name):
            continue
        lst = out.setdefault(name, [])
        lst.append((len(lst) + 1, r))
    return out


def family_window(d, log, family, events_only=False, phase="timed"):
    """Mean duration (us) of the family's dispatches in one phase of the
    bench's launch log ("timed": the timed steps, "isolated": the fenced
    leg after them; events_only: only those that carried the bench's
    timing events)."""
    rows = _rows(d, "kernel_trace.csv")
    want = {}
    for r in read_log(log):
        if r["family"] == family and r["phase"] == phase:
            if events_only and r.get("events", "1") != "1":
                continue
            want.setdefault(kernel_key(r["kernel"]), set()).add(int(r["ordinal"]))
    tot, cnt = 0.0, 0
    for name, lst in _ordinals(rows, "|".join(sorted({k[0] for k in want}))).items():
        sel = want.get(kernel_key(name), set())
        for o, r in lst:
            if o in sel:
                tot += int(_col(r, "End_Timestamp")) - int(_col(r, "Start_Timestamp"))
                cnt += 1
    return tot / max(cnt, 1) / 1e3, cnt


def stats(d, title, command, top, log=None, family=None):
    rows = _rows(d, "kernel_stats.csv")
    recs = []
    for r in rows:
        recs.append((_col(r, "Name", "KernelName"), int(_col(r, "Calls")), float(_col(r, "TotalDurationNs")),
                     float(_col(r, "AverageNs")), float(_col(r, "Percentage"))))
    recs.sort(key=lambda x: -x[2])
    lines = [f"# {title}", "", "Command (repo root as cwd):", "", "    " + command, "",
             "Kernel summary (`*kernel_stats.csv` of the run; durations in microseconds):", "",
             "| kernel | calls | total us | avg us | % |", "|---|---:|---:|---:|---:|"]
    for name, calls, tot, avg, pct in recs[:top]:
        short = re.sub(r"\(.*$", "", name.replace("(anonymous namespace)::", ""))
        lines.append(f"| `{short}` | {calls} | {tot / 1e3:.0f} | {avg / 1e3:.2f} | {pct:.1f} |")
    if log and family:
        us_all, n_all = family_window(d, log, family)
        us_in, n_in = family_window(d, log, family, events_only=True)
        us, n = family_window(d, log, family, events_only=True, phase="isolated")
        lines += ["", f"Family `{family}` (dispatches located through the launch log of the same run): in the "
                      f"timed steps {n_all} dispatches, mean duration {us_all:.2f} us (the {n_in} of them that "
                      f"carried the bench's in-place timing events: {us_in:.2f} us); in the isolated leg after "
                      f"the timed steps, the {n} dispatches that carried the bench's events (each fenced against "
                      f"the other stream groups, the kernel with the device to itself): mean duration "
                      f"{us:.2f} us.", "",
                  f"isolated_family_avg_us = {us:.3f}", f"isolated_family_launches = {n}",
                  f"timed_family_events_avg_us = {us_in:.3f}", f"timed_family_events_launches = {n_in}",
                  f"timed_family_all_avg_us = {us_all:.3f}", f"timed_family_all_launches = {n_all}"]
    return "\n".join(lines) + "\n"


def _counter(d, counter, kernel_re, first=1):
    """{ordinal: value} of `counter` for dispatches of the kernel matching
    kernel_re (one kernel name).  The counted dispatches are the per-kernel
    ordinals first, first+1, ... (--kernel-iteration-range [first-...])."""
    rows = [r for r in _rows(d, "counter_collection.csv") if _col(r, "Counter_Name") == counter]
    found = _ordinals(rows, kernel_re)
    if len(found) != 1:
        raise SystemExit(f"{counter}: /{kernel_re}/ matches {sorted(found)} under {d}")
    (name, lst), = found.items()
    return {first - 1 + o: float(_col(r, "Counter_Value")) for o, r in lst}, name


def traffic(a):
    cf, _ = _counter(a.copy_fetch, "FETCH_SIZE", a.copy_kernel)
    cw, _ = _counter(a.copy_write, "WRITE_SIZE", a.copy_kernel)
    f_fetch = a.copy_bytes / (1024.0 * max(cf.values()))
    f_write = a.copy_bytes / (1024.0 * max(cw.values()))
    log = read_log(a.log)
    lines = [f"# {a.title}", "", "Commands (repo root as cwd):", ""]
    lines += ["    " + c for c in a.command]
    lines += ["", f"Copy calibration: {a.copy_bytes:.0f} bytes read and written per copy dispatch; the counters "
                  f"read {max(cf.values()):.0f} KiB (FETCH_SIZE) and {max(cw.values()):.0f} KiB (WRITE_SIZE): "
                  f"factors {f_fetch:.4f} and {f_write:.4f} bytes per counted byte.", "",
              "| kernel | dispatches counted | fetch B/dispatch | write B/dispatch | HBM B/dispatch | "
              "algorithmic B/dispatch | traffic / algorithmic | timed launches |",
              "|---|---:|---:|---:|---:|---:|---:|---:|"]
    tot_t = tot_a = tot_n = 0.0
    for regex, first, fdir, wdir in a.variant:
        first = int(first)
        kf, name = _counter(fdir, "FETCH_SIZE", regex, first)
        kw, _ = _counter(wdir, "WRITE_SIZE", regex, first)
        key = kernel_key(name)
        common = sorted(set(kf) & set(kw))
        mine = [r for r in log if kernel_key(r["kernel"]) == key]
        alg = {int(r["ordinal"]): float(r["bytes"]) for r in mine}
        timed = sum(1 for r in mine if r["phase"] == "timed")
        common = [o for o in common if o in alg]
        if not common:
            raise SystemExit(f"no counted dispatches of /{regex}/ found in the launch log")
        fb = sum(kf[o] for o in common) / len(common) * 1024.0 * f_fetch
        wb = sum(kw[o] for o in common) / len(common) * 1024.0 * f_write
        ab = sum(alg[o] for o in common) / len(common)
        lines.append(f"| `{regex}` | {len(common)} (ordinals {common[0]}-{common[-1]}) | {fb:.0f} | {wb:.0f} | "
                     f"{fb + wb:.0f} | {ab:.0f} | {(fb + wb) / ab:.3f} | {timed} |")
        # family figure: each variant weighted by its launches in the timed steps
        tot_t += timed * (fb + wb)
        tot_a += timed * ab
        tot_n += timed
    lines += ["", "Family figures weight each variant by its launches in the timed steps of the logged run.",
              "", f"traffic_bytes_per_launch = {tot_t / tot_n:.0f}",
              f"alg_bytes_per_launch = {tot_a / tot_n:.0f}", f"traffic_over_algorithmic = {tot_t / tot_a:.4f}",
              f"launches = {tot_n:.0f}", ""]
    return "\n".join(lines)


def mfma(d, kernel_re, counters, scale, peak_tflops, title, command):
    """Matrix-core figures of the dispatches of kernels matching kernel_re
    in a ``rocprofv3 --kernel-trace --pmc MfmaUtil <MOPS counter>`` run:
    per kernel name the dispatch count, mean MfmaUtil (percent of SIMD
    cycles with the matrix core busy), executed operations (the MOPS counter
    x `scale`) over the summed dispatch durations, against `peak_tflops`."""
    rows = _rows(d, "counter_collection.csv")
    trace = {int(_col(r, "Dispatch_Id", "Dispatch_ID")): r for r in _rows(d, "kernel_trace.csv")}
    per = {}
    for r in rows:
        name = _col(r, "Kernel_Name")
        if not re.search(kernel_re, name):
            continue
        did = int(_col(r, "Dispatch_Id", "Dispatch_ID"))
        rec = per.setdefault(re.sub(r"\(.*$", "", name), {}).setdefault(did, {})
        rec[_col(r, "Counter_Name")] = rec.get(_col(r, "Counter_Name"), 0.0) + float(_col(r, "Counter_Value"))
    lines = [f"# {title}", "", "Commands (repo root as cwd):", ""] + ["    " + c for c in command] + [
        "", f"Operations = {counters[1]} x {scale:g} per dispatch; time = the dispatch's kernel-trace "
            f"duration; peak {peak_tflops:g} TFLOP/s (or TOP/s).", "",
        "| kernel | dispatches | mean MfmaUtil % | executed T(FL)OP | summed duration ms | achieved T(FL)OP/s | "
        "fraction of peak |", "|---|---:|---:|---:|---:|---:|---:|"]
    for name, recs in sorted(per.items()):
        ops = dur = util = 0.0
        n = 0
        for did, rec in recs.items():
            if did not in trace or counters[1] not in rec:
                continue
            t = trace[did]
            dur += (int(_col(t, "End_Timestamp")) - int(_col(t, "Start_Timestamp"))) * 1e-9
            ops += rec[counters[1]] * scale
            util += rec.get(counters[0], 0.0)
            n += 1
        if n == 0:
            continue
        ach = ops / dur / 1e12 if dur > 0 else 0.0
        lines.append(f"| `{name}` | {n} | {util / n:.1f} | {ops / 1e12:.3f} | {dur * 1e3:.2f} | {ach:.2f} | "
                     f"{ach / peak_tflops:.3f} |")
    return "\n".join(lines) + "\n"


def first_timed(log, base, ints, skip=0):
    """Per-kernel ordinal of the first dispatch in the timed steps of the
    kernel `base<ints...>` (template integers as a prefix), plus `skip`."""
    for r in read_log(log):
        b, k = kernel_key(r["kernel"])
        if b == base and k[:len(ints)] == tuple(ints) and r["phase"] == "timed":
            return int(r["ordinal"]) + skip
    raise SystemExit(f"no timed dispatch of {base}<{ints}> in {log}")


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("--title", default="rocprofv3 kernel trace")
    s.add_argument("--command", default="")
    s.add_argument("--top", type=int, default=30)
    s.add_argument("--log", default=None)
    s.add_argument("--family", default=None)
    f = sub.add_parser("first-timed")
    f.add_argument("--log", required=True)
    f.add_argument("--kernel", required=True, help="base name, e.g. k_tr_symv")
    f.add_argument("--ints", default="", help="leading template integers, e.g. 2,16")
    f.add_argument("--skip", type=int, default=0)
    mf = sub.add_parser("mfma")
    mf.add_argument("dir")
    mf.add_argument("--kernel", required=True, help="kernel name regex")
    mf.add_argument("--util", default="MfmaUtil")
    mf.add_argument("--mops", default="SQ_INSTS_VALU_MFMA_MOPS_F64")
    mf.add_argument("--scale", type=float, default=512.0)
    mf.add_argument("--peak", type=float, default=78.6)
    mf.add_argument("--title", default="matrix-core counters")
    mf.add_argument("--command", action="append", default=[])
    t = sub.add_parser("traffic")
    t.add_argument("--log", required=True)
    t.add_argument("--copy-fetch", required=True)
    t.add_argument("--copy-write", required=True)
    t.add_argument("--copy-bytes", type=float, required=True)
    t.add_argument("--copy-kernel", default="k_stream_copy")
    t.add_argument("--variant", nargs=4, action="append", required=True,
                   metavar=("REGEX", "FIRST", "FETCH_DIR", "WRITE_DIR"))
    t.add_argument("--title", default="HBM traffic per dispatch")
    t.add_argument("--command", action="append", default=[])
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        sys.stdout.write(stats(a.dir, a.title, a.command, a.top, a.log, a.family))
    elif a.cmd == "mfma":
        sys.stdout.write(mfma(a.dir, a.kernel, (a.util, a.mops), a.scale, a.peak, a.title, a.command))
    elif a.cmd == "first-timed":
        ints = [int(x) for x in a.ints.split(",") if x.strip()]
        print(first_timed(a.log, a.kernel, ints, a.skip))
    else:
        sys.stdout.write(traffic(a))


if __name__ == "__main__":
    main()
