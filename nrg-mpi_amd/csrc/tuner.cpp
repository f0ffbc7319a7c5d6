// Autotuner of the batched tridiagonalisation (see tuner.h).
#include "tuner.h"
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>

namespace nrg {

std::vector<TuneConfig> Tuner::candidates() {
  return {TuneConfig{4, 512}, TuneConfig{1, 512}, TuneConfig{2, 512}, TuneConfig{3, 512}, TuneConfig{4, 256}};
}

void Tuner::set(int mode, const std::string& path) {
  mode_ = std::max(0, std::min(mode, 2));
  path_ = path;
  t_.clear();
  if (mode_ && !path_.empty()) load();
}

std::string Tuner::key(const int* n, int nblk) {
  int nmax = 0;
  double s3 = 0;
  for (int b = 0; b < nblk; ++b) {
    nmax = std::max(nmax, n[b]);
    s3 += (double)n[b] * n[b] * n[b];
  }
  std::ostringstream o;
  o << "b" << nblk << "_m" << (nmax + 63) / 64 << "_c" << (s3 > 0 ? (int)std::lround(4.0 * std::log2(s3)) : 0);
  return o.str();
}

TuneConfig Tuner::choose(const std::string& k, bool* exploring) {
  *exploring = false;
  const std::vector<TuneConfig> cand = candidates();
  if (mode_ == 0) return cand[0];
  auto it = t_.find(k);
  if (mode_ == 1) {
    for (const TuneConfig& c : cand) {
      const bool seen = it != t_.end() && it->second.count({c.groups, c.symv_wide_wg});
      if (!seen || it->second.at({c.groups, c.symv_wide_wg}).samples < kMinSamples) {
        *exploring = true;
        return c;
      }
    }
  }
  if (it == t_.end() || it->second.empty()) return cand[0];
  TuneConfig best = cand[0];
  double bt = 1e300;
  for (const auto& kv : it->second)
    if (kv.second.samples > 0 && kv.second.mean() < bt) {
      bt = kv.second.mean();
      best = TuneConfig{kv.first.first, kv.first.second};
    }
  return best;
}

void Tuner::record(const std::string& k, const TuneConfig& c, double ms) {
  if (mode_ != 1 || !(ms > 0)) return;
  Stat& s = t_[k][{c.groups, c.symv_wide_wg}];
  s.sum += ms;
  s.samples += 1;
  if (!path_.empty()) save();
}

void Tuner::load() {
  std::ifstream in(path_);
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream is(line);
    std::string k;
    int g, wg, ns;
    double ms;
    if (!(is >> k >> g >> wg >> ms >> ns) || k.empty() || k[0] == '#') continue;
    Stat& s = t_[k][{g, wg}];
    s.sum = ms * ns;
    s.samples = ns;
  }
}

void Tuner::save() const {
  const std::string tmp = path_ + ".part";
  {
    std::ofstream o(tmp);
    o.precision(10);   // means keep accumulating samples across runs
    o << "# key groups symv_wide_wg mean_ms samples (batched tridiagonalisation, MI355X)\n";
    for (const auto& kv : t_)
      for (const auto& c : kv.second)
        o << kv.first << " " << c.first.first << " " << c.first.second << " " << c.second.mean() << " "
          << c.second.samples << "\n";
  }
  std::rename(tmp.c_str(), path_.c_str());
}

}  // namespace nrg
