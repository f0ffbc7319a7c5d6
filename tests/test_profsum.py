"""The profile summariser's join of a rocprofv3 kernel trace with the
bench's launch log (nrg-mpi_amd/nrgamd/profsum.py): dispatches are matched
by kernel and per-kernel ordinal, and the mean duration is taken over the
requested phase (timed steps or the isolated leg) and, on request, only
over the launches that carried timing events.  Synthetic CSV files, CPU
only."""
import csv
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "nrg-mpi_amd"))

from nrgamd import profsum  # noqa: E402

SYMV = "void nrg::k_tr_symv<2, 16, 4, 1>(nrg::TrArgs, int, int, int, int)"
COL = "nrg::k_tr_col(nrg::TrArgs, int, int, int, int, int, int, int)"
SYMV_MANGLED = "_ZN3nrg9k_tr_symvILi2ELi16ELi4ELi1EEEvNS_6TrArgsEiiii"


def write_trace(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Dispatch_Id", "Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        for r in rows:
            w.writerow(r)


def write_log(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["family", "kernel", "threads", "ordinal", "bytes", "flops", "phase", "events"])
        for r in rows:
            w.writerow(r)


def test_family_window_by_phase_and_events(tmp_path):
    d = tmp_path / "trace"
    d.mkdir()
    # SYMV dispatches 1..6 (durations 10, 20, ..., 60 us) interleaved with
    # column kernels; dispatch ids out of order in the file
    rows, did, t = [], 0, 0
    for i in range(6):
        did += 1
        rows.append([did, COL, t, t + 5000])
        t += 10000
        did += 1
        rows.append([did, SYMV, t, t + 10000 * (i + 1)])
        t += 100000
    write_trace(d / "run_kernel_trace.csv", list(reversed(rows)))
    log = tmp_path / "log.csv"
    # the log names the kernel by its mangled name; ordinals 1..6 of the SYMV
    write_log(log, [
        ["tridiag_symv", SYMV_MANGLED, 1024, 1, 1e6, 0, "warmup", 0],
        ["tridiag_symv", SYMV_MANGLED, 1024, 2, 1e6, 0, "timed", 1],
        ["tridiag_symv", SYMV_MANGLED, 1024, 3, 1e6, 0, "timed", 0],
        ["tridiag_symv", SYMV_MANGLED, 1024, 4, 1e6, 0, "leg", 0],
        ["tridiag_symv", SYMV_MANGLED, 1024, 5, 1e6, 0, "isolated", 1],
        ["tridiag_symv", SYMV_MANGLED, 1024, 6, 1e6, 0, "isolated", 0],
    ])
    us, n = profsum.family_window(str(d), str(log), "tridiag_symv")
    assert n == 2 and abs(us - 25.0) < 1e-9            # ordinals 2, 3: 20 and 30 us
    us, n = profsum.family_window(str(d), str(log), "tridiag_symv", events_only=True)
    assert n == 1 and abs(us - 20.0) < 1e-9
    us, n = profsum.family_window(str(d), str(log), "tridiag_symv", events_only=True, phase="isolated")
    assert n == 1 and abs(us - 50.0) < 1e-9
    us, n = profsum.family_window(str(d), str(log), "tridiag_symv", phase="isolated")
    assert n == 2 and abs(us - 55.0) < 1e-9


def test_kernel_key_demangled_and_mangled():
    assert profsum.kernel_key(SYMV) == ("k_tr_symv", (2, 16, 4, 1))
    assert profsum.kernel_key(SYMV_MANGLED) == ("k_tr_symv", (2, 16, 4, 1))
    assert profsum.kernel_key(COL) == ("k_tr_col", ())
