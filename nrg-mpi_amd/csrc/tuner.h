// Autotuner of the batched tridiagonalisation, keyed on the block-size
// histogram of an iteration.
//
// The launch configuration that is fastest depends on the block set: how
// many stream groups share the device (the column kernel of one group runs
// while the SYMV of another streams its trailing matrices) and from which
// launch size on the SYMV uses 32-column panels.  An NRG run meets a few
// block sets again and again (the steady state alternates between two block
// tables), so the tuner times each candidate configuration on successive
// iterations with the same histogram key (kMinSamples times each), then
// keeps the one with the lowest mean time: the mean, not the minimum, so
// that a configuration timed more often is not favoured by its luckiest
// sample.  The measurements persist in a text cache file, one line per
// (key, candidate):
//     <key> <groups> <symv_wide_wg> <mean_ms> <samples>
// Modes: 0 off (the defaults), 1 explore unmeasured candidates and update the
// cache, 2 use the cache only (defaults for keys it does not hold).
#pragma once
#include <map>
#include <string>
#include <vector>

namespace nrg {

struct TuneConfig {
  int groups = 2;
  int symv_wide_wg = 512;
  bool operator==(const TuneConfig& o) const { return groups == o.groups && symv_wide_wg == o.symv_wide_wg; }
};

class Tuner {
 public:
  // candidates, the default first
  static std::vector<TuneConfig> candidates();
  void set(int mode, const std::string& path);
  int mode() const { return mode_; }
  // histogram key: block count, largest block in units of 64, sum n^3 on a
  // quarter-octave scale
  static std::string key(const int* n, int nblk);
  // configuration for this block set; *exploring is true while an untimed
  // candidate is being measured
  TuneConfig choose(const std::string& k, bool* exploring);
  void record(const std::string& k, const TuneConfig& c, double ms);
  size_t entries() const { return t_.size(); }

 private:
  static constexpr int kMinSamples = 3;
  struct Stat {
    double sum = 0;
    int samples = 0;
    double mean() const { return samples ? sum / samples : 0.0; }
  };
  void load();
  void save() const;
  int mode_ = 0;
  std::string path_;
  std::map<std::string, std::map<std::pair<int, int>, Stat>> t_;
};

}  // namespace nrg
