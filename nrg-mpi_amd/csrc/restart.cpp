// Restart file of the CFS spectrum phase (the role of saveRestart /
// loadRestart, module.F90:1723-1890, and their use in main.F90:165-221).
//
// The reference writes its restore arrays (every iteration's full
// eigenvectors, operators, reduced density and bins) as one unformatted
// stream, "restart.dat", before every CFS step, and a restarted run skips the
// NRG loop and the reduced-density sweep and resumes the spikes at the saved
// iteration.  The CFS state of this library is its device-resident history
// (kept columns, spectral products, reduced density per block), so the file
// holds that state, in the same kind of stream: each record preceded by its
// size, led by the reference's mode word MODE_RESTART_CFS = 1002 and the
// iteration.  Only history entries at or after the saved iteration are
// written: the spikes of the earlier ones are already in the saved bins.
//
// Layout (little endian):
//   int32 mode (1002), int32 iteration, int32 nstate, int32 state[nstate]
//   int32 nv, nmat, nbin; vector<double> bins
//   int32 nentries, then per history entry:
//     int32 it, nsub, nfull, rho_ready; int64 wsize
//     vector<int32> sub, basis_in, bn, kept, rho_zero
//     vector<int64> bvoff, uoff, roff
//     vector<double> E, U, Ed, W, rho     (U, Ed, W, rho from the device)
//     int32 ngroups, per group int32 K, L, cK, l0, dL, im, side, zero_t,
//                              double factor, int64 woff
// where vector<T> is int64 count followed by count values.
#include "cfs.h"
#include <cstdio>

namespace nrg {

namespace {

constexpr int32_t kModeRestartCfs = 1002;   // MODE_RESTART_CFS, module.F90:1720

struct File {
  FILE* f = nullptr;
  std::string path;
  File(const char* p, const char* mode) : f(std::fopen(p, mode)), path(p) {}
  ~File() { if (f) std::fclose(f); }
  void put(const void* p, size_t n) {
    if (n && std::fwrite(p, 1, n, f) != n) throw Error(NRG_ERR_STATE, "restart: write to " + path + " failed");
  }
  void get(void* p, size_t n) {
    if (n && std::fread(p, 1, n, f) != n) throw Error(NRG_ERR_STATE, "restart: " + path + " is truncated");
  }
  template <typename T> void put1(T v) { put(&v, sizeof(T)); }
  template <typename T> T get1() { T v; get(&v, sizeof(T)); return v; }
  template <typename T> void putv(const std::vector<T>& v) {
    put1<int64_t>((int64_t)v.size());
    put(v.data(), v.size() * sizeof(T));
  }
  template <typename T> void getv(std::vector<T>& v) {
    const int64_t n = get1<int64_t>();
    NRG_REQUIRE(n >= 0 && n < (int64_t(1) << 40), "restart: bad record size in " + path);
    v.resize((size_t)n);
    get(v.data(), v.size() * sizeof(T));
  }
  // a device buffer's first `n` values, staged through host memory
  void putd(const DevBuf<double>& d, hipStream_t s) {
    const int64_t n = d.p ? (int64_t)d.n : 0;
    put1<int64_t>(n);
    std::vector<double> h((size_t)std::max<int64_t>(n, 0));
    if (n) {
      NRG_HIP(hipMemcpyAsync(h.data(), d.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
      NRG_HIP(hipStreamSynchronize(s));
    }
    put(h.data(), h.size() * sizeof(double));
  }
  void getd(DevBuf<double>& d, hipStream_t s) {
    std::vector<double> h;
    getv(h);
    d.reserve(std::max<size_t>(h.size(), 1));
    d.n = h.size();
    if (!h.empty()) {
      NRG_HIP(hipMemcpyAsync(d.p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, s));
      NRG_HIP(hipStreamSynchronize(s));
    }
  }
};

}  // namespace

void save_restart(Context& C, const char* path, int iteration, int nstate, const int32_t* state) {
  Cfs& F = cfs();
  NRG_REQUIRE(path && *path, "save_restart: no file name");
  NRG_REQUIRE(F.nbin > 0, "save_restart: the spectrum phase has not been set up");
  NRG_REQUIRE(nstate >= 0 && (nstate == 0 || state), "save_restart: bad host state");
  const std::string tmp = std::string(path) + ".part";
  {
    File o(tmp.c_str(), "wb");
    NRG_REQUIRE(o.f != nullptr, "save_restart: cannot open " + tmp);
    o.put1<int32_t>(kModeRestartCfs);
    o.put1<int32_t>(iteration);
    o.put1<int32_t>(nstate);
    o.put(state, sizeof(int32_t) * (size_t)nstate);
    o.put1<int32_t>(C.nv);
    o.put1<int32_t>(C.nmat);
    o.put1<int32_t>(F.nbin);
    DevBuf<double> bins_view;
    bins_view.p = F.bins.p;
    bins_view.n = (size_t)2 * C.nmat * F.nbin;
    o.putd(bins_view, C.stream);
    bins_view.p = nullptr;   // not owned
    int32_t ne = 0;
    for (const HistEntry* h : F.hist) ne += (h->it >= iteration) ? 1 : 0;
    o.put1<int32_t>(ne);
    for (const HistEntry* h : F.hist) {
      if (h->it < iteration) continue;
      o.put1<int32_t>(h->it);
      o.put1<int32_t>(h->nsub);
      o.put1<int32_t>(h->nfull);
      o.put1<int32_t>(h->rho_ready ? 1 : 0);
      o.put1<int64_t>(h->wsize);
      o.putv(h->sub);
      o.putv(h->basis_in);
      o.putv(std::vector<int32_t>(h->bn.begin(), h->bn.end()));
      o.putv(std::vector<int32_t>(h->kept.begin(), h->kept.end()));
      o.putv(std::vector<int32_t>(h->rho_zero.begin(), h->rho_zero.end()));
      o.putv(h->bvoff);
      o.putv(h->uoff);
      o.putv(h->roff);
      o.putv(h->E);
      o.putd(h->U, C.stream);
      o.putd(h->Ed, C.stream);
      o.putd(h->W, C.stream);
      o.putd(h->rho, C.stream);
      o.put1<int32_t>((int32_t)h->groups.size());
      for (const SpikeGroup& g : h->groups) {
        const int32_t iv[8] = {g.K, g.L, g.cK, g.l0, g.dL, g.im, g.side, g.zero_t};
        o.put(iv, sizeof(iv));
        o.put1<double>(g.factor);
        o.put1<int64_t>(g.woff);
      }
    }
  }
  // replace the previous file only once the new one is complete
  NRG_REQUIRE(std::rename(tmp.c_str(), path) == 0, std::string("save_restart: cannot rename to ") + path);
}

void load_restart(Context& C, const char* path, int* iteration, int nstate, int32_t* state) {
  Cfs& F = cfs();
  NRG_REQUIRE(path && *path, "load_restart: no file name");
  File in(path, "rb");
  NRG_REQUIRE(in.f != nullptr, std::string("restart file not found: ") + path);
  NRG_REQUIRE(in.get1<int32_t>() == kModeRestartCfs, std::string("broken restart file: ") + path);
  *iteration = in.get1<int32_t>();
  const int32_t ns = in.get1<int32_t>();
  NRG_REQUIRE(ns == nstate, "load_restart: host state size differs from the file's");
  in.get(state, sizeof(int32_t) * (size_t)ns);
  const int32_t nv = in.get1<int32_t>(), nmat = in.get1<int32_t>(), nbin = in.get1<int32_t>();
  NRG_REQUIRE(nv == C.nv && nmat == C.nmat, "load_restart: the file belongs to another model");
  F.clear();
  F.nbin = nbin;
  in.getd(F.bins, C.stream);
  NRG_REQUIRE(F.bins.n == (size_t)2 * nmat * nbin, "load_restart: bin record size");
  const int32_t ne = in.get1<int32_t>();
  for (int e = 0; e < ne; ++e) {
    HistEntry* h = new HistEntry;
    try {
      h->it = in.get1<int32_t>();
      h->nsub = in.get1<int32_t>();
      h->nfull = in.get1<int32_t>();
      h->rho_ready = in.get1<int32_t>() != 0;
      h->wsize = in.get1<int64_t>();
      std::vector<int32_t> v;
      in.getv(h->sub);
      in.getv(h->basis_in);
      in.getv(v); h->bn.assign(v.begin(), v.end());
      in.getv(v); h->kept.assign(v.begin(), v.end());
      in.getv(v); h->rho_zero.assign(v.begin(), v.end());
      in.getv(h->bvoff);
      in.getv(h->uoff);
      in.getv(h->roff);
      in.getv(h->E);
      in.getd(h->U, C.stream);
      in.getd(h->Ed, C.stream);
      in.getd(h->W, C.stream);
      in.getd(h->rho, C.stream);
      const int32_t ng = in.get1<int32_t>();
      NRG_REQUIRE(ng >= 0, "load_restart: bad group count");
      h->groups.resize(ng);
      for (SpikeGroup& g : h->groups) {
        int32_t iv[8];
        in.get(iv, sizeof(iv));
        g.K = iv[0]; g.L = iv[1]; g.cK = iv[2]; g.l0 = iv[3]; g.dL = iv[4]; g.im = iv[5]; g.side = iv[6];
        g.zero_t = iv[7];
        g.factor = in.get1<double>();
        g.woff = in.get1<int64_t>();
      }
      NRG_REQUIRE((int)h->sub.size() == h->nsub * (nv + 3) && (int)h->bn.size() == h->nsub,
                  "load_restart: inconsistent block table");
    } catch (...) {
      h->release();
      delete h;
      throw;
    }
    F.hist.push_back(h);
  }
}

}  // namespace nrg
