"""The launch-configuration tuner of the batched tridiagonalisation
(nrg-mpi_amd/csrc/tuner.cpp, host C++): compiled with g++ into a small
driver that feeds it synthetic timings, so the selection rule is checked
without a GPU.  Checked: the histogram key, exploration of every candidate
(default first) three times each, the choice by lowest mean time (a
candidate with one lucky sample but a higher mean is not taken), the cache
file round trip into mode 2, and that keys absent from the cache get the
default.  CPU only."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "nrg-mpi_amd", "csrc")

DRIVER = r"""
#include "tuner.h"
#include <cstdio>
#include <cstdlib>
#include <string>
using namespace nrg;
static void say(const char* what, const TuneConfig& c) { std::printf("%s %d %d\n", what, c.groups, c.symv_wide_wg); }
int main(int argc, char** argv) {
  const std::string path = argv[1];
  const int n[3] = {4142, 3438, 700};
  const std::string k = Tuner::key(n, 3);
  std::printf("key %s\n", k.c_str());
  Tuner t;
  t.set(1, path);
  // synthetic times: groups 3 is fastest on average; (4, 256) gets one
  // lucky sample below everyone's mean but is slower on average
  auto ms = [](const TuneConfig& c, int i) {
    if (c.groups == 3) return 100.0 + i;          // mean 101
    if (c.groups == 1) return 150.0;
    if (c.symv_wide_wg == 256) return i == 0 ? 90.0 : 130.0;   // mean 116.7, minimum 90
    return 110.0;
  };
  int seen[8] = {0};
  for (int it = 0; it < 15; ++it) {
    bool ex = false;
    const TuneConfig c = t.choose(k, &ex);
    if (!ex) { std::printf("explored-too-early %d\n", it); return 1; }
    const auto cand = Tuner::candidates();
    int idx = -1;
    for (size_t j = 0; j < cand.size(); ++j) if (cand[j] == c) idx = (int)j;
    std::printf("explore %d %d %d\n", it, idx, seen[idx]);
    t.record(k, c, ms(c, seen[idx]++));
  }
  bool ex = true;
  say("chosen", t.choose(k, &ex));
  std::printf("exploring %d\n", (int)ex);
  Tuner u;
  u.set(2, path);
  say("cached", u.choose(k, &ex));
  const int other[1] = {64};
  say("absent", u.choose(Tuner::key(other, 1), &ex));
  return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not found")
def test_tuner_explores_then_keeps_the_lowest_mean(tmp_path):
    src = tmp_path / "drv.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "drv"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", CSRC, str(src), os.path.join(CSRC, "tuner.cpp"), "-o",
                    str(exe)], check=True, capture_output=True, text=True)
    cache = tmp_path / "cache.txt"
    out = subprocess.run([str(exe), str(cache)], check=True, capture_output=True, text=True).stdout.split("\n")
    # key: block count, largest block in units of 64, sum n^3 on a quarter-octave scale
    assert out[0].startswith("key b3_m65_c")
    # candidates in order (the default first), each three times
    explore = [tuple(int(x) for x in l.split()[1:]) for l in out if l.startswith("explore")]
    assert [e[1:] for e in explore] == [(c, s) for c in range(5) for s in range(3)]
    # lowest mean (groups 3), not the lowest single sample (4, 256)
    assert "chosen 3 512" in out
    assert "exploring 0" in out
    # the cache file carries means and sample counts; mode 2 reads it back
    lines = [l.split() for l in cache.read_text().splitlines() if l and not l.startswith("#")]
    by = {(int(l[1]), int(l[2])): (float(l[3]), int(l[4])) for l in lines}
    assert by[(3, 512)] == (101.0, 3)
    assert abs(by[(4, 256)][0] - 350.0 / 3.0) < 1e-6
    assert "cached 3 512" in out
    # a block set the cache has not seen gets the default configuration
    assert "absent 4 512" in out
