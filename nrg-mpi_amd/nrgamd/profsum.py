"""Summaries of rocprofv3 CSV output for the kernel profiles of this library.

    python -m nrgamd.profsum stats  DIR [--title T] [--command C]
        kernel table from the ``*kernel_stats.csv`` of a
        ``rocprofv3 --kernel-trace --stats -f csv`` run;
    python -m nrgamd.profsum traffic --fetch DIR --write DIR --copy-fetch DIR
        --copy-write DIR --copy-bytes B --kernel NAME --last N [--alg-bytes A]
        HBM bytes per dispatch of NAME over its last N dispatches, from
        ``rocprofv3 --pmc FETCH_SIZE`` and ``--pmc WRITE_SIZE`` runs (KiB per
        dispatch), each scaled by what the same counter reads for a plain
        device copy of B bytes (B read and B written per dispatch).

The traffic summary ends with a ``traffic_bytes_per_launch = X`` line and a
``launches = N`` line that bench.py reads back.
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


def stats(d, title, command, top):
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
        short = re.sub(r"\(.*$", "", name)
        lines.append(f"| `{short}` | {calls} | {tot / 1e3:.0f} | {avg / 1e3:.2f} | {pct:.1f} |")
    return "\n".join(lines) + "\n"


def _counter(d, counter, kernel_re):
    vals = []
    pat = re.compile(kernel_re)
    rows = _rows(d, "counter_collection.csv")
    rows.sort(key=lambda r: int(_col(r, "Dispatch_Id", "Dispatch_ID")))
    for r in rows:
        if _col(r, "Counter_Name") != counter:
            continue
        if not pat.search(_col(r, "Kernel_Name")):
            continue
        vals.append(float(_col(r, "Counter_Value")))
    if not vals:
        raise SystemExit(f"no {counter} rows for /{kernel_re}/ under {d}")
    return vals


def traffic(a):
    cf = _counter(a.copy_fetch, "FETCH_SIZE", a.copy_kernel)
    cw = _counter(a.copy_write, "WRITE_SIZE", a.copy_kernel)
    f_fetch = a.copy_bytes / (1024.0 * max(cf))
    f_write = a.copy_bytes / (1024.0 * max(cw))
    kf = _counter(a.fetch, "FETCH_SIZE", a.kernel)[-a.last:]
    kw = _counter(a.write, "WRITE_SIZE", a.kernel)[-a.last:]
    if len(kf) != a.last or len(kw) != a.last:
        raise SystemExit(f"fewer than {a.last} dispatches of {a.kernel} ({len(kf)}, {len(kw)})")
    fetch_b = sum(kf) / a.last * 1024.0 * f_fetch
    write_b = sum(kw) / a.last * 1024.0 * f_write
    t = fetch_b + write_b
    lines = [f"# {a.title}", "", "Commands (repo root as cwd):", ""]
    lines += ["    " + c for c in a.command]
    lines += ["", f"Copy calibration: {a.copy_bytes} bytes read and written per copy dispatch; the counters read "
                  f"{max(cf):.0f} KiB (FETCH_SIZE) and {max(cw):.0f} KiB (WRITE_SIZE), factors "
                  f"{f_fetch:.4f} and {f_write:.4f} bytes per counted byte.", "",
              f"`{a.kernel}`, last {a.last} dispatches (the timed steps):", "",
              "| quantity | bytes per dispatch |", "|---|---:|",
              f"| FETCH_SIZE (scaled) | {fetch_b:.0f} |", f"| WRITE_SIZE (scaled) | {write_b:.0f} |",
              f"| HBM traffic | {t:.0f} |"]
    if a.alg_bytes:
        lines.append(f"| algorithmic bytes | {a.alg_bytes:.0f} |")
        lines += ["", f"traffic / algorithmic = {t / a.alg_bytes:.3f}"]
    lines += ["", f"traffic_bytes_per_launch = {t:.0f}", f"launches = {a.last}", ""]
    return "\n".join(lines)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("--title", default="rocprofv3 kernel trace")
    s.add_argument("--command", default="")
    s.add_argument("--top", type=int, default=30)
    t = sub.add_parser("traffic")
    for k in ("fetch", "write", "copy_fetch", "copy_write"):
        t.add_argument("--" + k.replace("_", "-"), dest=k, required=True)
    t.add_argument("--copy-bytes", type=float, required=True)
    t.add_argument("--copy-kernel", default="copyBuffer|copy")
    t.add_argument("--kernel", required=True)
    t.add_argument("--last", type=int, required=True)
    t.add_argument("--alg-bytes", type=float, default=0.0, help="algorithmic bytes per dispatch")
    t.add_argument("--title", default="HBM traffic per dispatch")
    t.add_argument("--command", action="append", default=[])
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        sys.stdout.write(stats(a.dir, a.title, a.command, a.top))
    else:
        sys.stdout.write(traffic(a))


if __name__ == "__main__":
    main()
