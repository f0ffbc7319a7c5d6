// Optional per-kernel-family timing.  A profiled launch carries its own
// start/stop events in the dispatch (hipExtLaunchKernel), so timing adds no
// separate event records to the stream.  Off by default; `mask` selects the
// families.  Each API call collects the event pairs that have completed by
// its end (without waiting, so that calls keep their asynchrony), and
// reading the statistics waits for the rest; a collected pair adds its
// duration, a launch and the algorithmic bytes/flops the launch site
// declared.
#pragma once
#include <string>
#include "common.h"
#include <hip/hip_ext.h>
#include <array>

namespace nrg {

enum KFamily : int {
  KF_ASSEMBLE = 0,
  KF_TR_COL,
  KF_TR_SYMV,
  KF_GEMM_TRIDIAG,
  KF_DC_LEAF,
  KF_DC_PREP,
  KF_DC_SECULAR,
  KF_DC_VECTORS,
  KF_GEMM_DC,
  KF_DC_FINISH,
  KF_BT_SMALL,
  KF_GEMM_BACK,
  KF_GEMM_INV,
  KF_GEMM_SPEC,
  KF_GEMM_CFSW,
  KF_GEMM_CFSX,
  KF_CFS_BIN,
  KF_GEMM_RHO,
  KF_RHO_GATHER,
  KF_TR_BAND,
  KF_BT_BAND,
  KF_OTHER,
  KF_COUNT
};

struct KStat {
  double ms = 0, bytes = 0, flops = 0;
  long long launches = 0;
};

struct Profiler {
  bool on = false;
  uint32_t mask = 0xffffffffu;   // families timed while on
  uint32_t every = 1;            // time every `every`-th launch of a family
  // a timed SYMV launch of the multi-stream tridiagonalisation runs with
  // the device to itself (fenced against the other stream groups) when set;
  // otherwise it is timed in place, sharing the device as it does untimed
  bool exclusive = true;
  std::array<uint64_t, KF_COUNT> seen{};
  // launch log: one record per launch of the logged families, with the
  // launch's ordinal among launches of the same kernel function (the
  // numbering a profiler's per-kernel dispatch counter uses)
  struct LogRec { int fam, threads, kid, timed; uint64_t nth; double bytes, flops; };
  struct KernelSeen { const void* fn; uint64_t count; std::string name; };
  uint32_t log_mask = 0;
  std::vector<LogRec> log;
  std::vector<KernelSeen> nth;   // kernels in order of first launch (kid = index)
  uint64_t next_nth(const void* k, hipStream_t s, int& kid) {
    for (size_t i = 0; i < nth.size(); ++i)
      if (nth[i].fn == k) { kid = (int)i; return ++nth[i].count; }
    const char* nm = hipKernelNameRefByPtr(k, s);
    nth.push_back(KernelSeen{k, 1, nm ? std::string(nm) : std::string()});
    kid = (int)nth.size() - 1;
    return 1;
  }
  std::array<KStat, KF_COUNT> st{};
  // work issued per family since the last reset, counted for every launch
  // (timed or not): launches, algorithmic bytes and flops as declared at the
  // launch sites
  std::array<double, KF_COUNT> w_launches{}, w_bytes{}, w_flops{};
  struct Pending { hipEvent_t a, b; int fam; double bytes, flops; };
  std::vector<Pending> pend;
  std::vector<hipEvent_t> pool;

  hipEvent_t get() {
    if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
    hipEvent_t e;
    NRG_HIP(hipEventCreate(&e));
    return e;
  }
  bool timed(int fam) {
    if (!on || !((mask >> fam) & 1u)) return false;
    return (seen[fam]++ % every) == 0;
  }
  // whether the next launch of `fam` will be timed (does not count it)
  bool peek(int fam) const {
    if (!on || !((mask >> fam) & 1u)) return false;
    return (seen[fam] % every) == 0;
  }
  // collect the pairs whose stop event has completed, keep the others
  void drain_ready() {
    size_t k = 0;
    for (size_t i = 0; i < pend.size(); ++i) {
      const Pending p = pend[i];
      if (hipEventQuery(p.b) != hipSuccess) { pend[k++] = p; continue; }
      float ms = 0;
      NRG_HIP(hipEventElapsedTime(&ms, p.a, p.b));
      KStat& s = st[p.fam];
      s.ms += ms; s.bytes += p.bytes; s.flops += p.flops; s.launches += 1;
      pool.push_back(p.a);
      pool.push_back(p.b);
    }
    pend.resize(k);
  }
  // call after the device has been synchronised
  void drain() {
    for (auto& p : pend) {
      float ms = 0;
      NRG_HIP(hipEventElapsedTime(&ms, p.a, p.b));
      KStat& k = st[p.fam];
      k.ms += ms; k.bytes += p.bytes; k.flops += p.flops; k.launches += 1;
      pool.push_back(p.a);
      pool.push_back(p.b);
    }
    pend.clear();
  }
  void reset() { st = {}; seen = {}; }
};

Profiler& profiler();

// Launch `kernel`, timed when the profiler is on for family `fam`.
template <typename F, typename... Args>
inline void plaunch(int fam, double bytes, double flops, F kernel, const dim3& grid, const dim3& block,
                    uint32_t shm, hipStream_t s, Args... args) {
  Profiler& P = profiler();
  const bool tm = P.timed(fam);
  P.w_launches[fam] += 1.0;
  P.w_bytes[fam] += bytes;
  P.w_flops[fam] += flops;
  if (P.log_mask) {
    int kid = 0;
    const uint64_t k = P.next_nth((const void*)kernel, s, kid);   // counted for every kernel
    if ((P.log_mask >> fam) & 1u)
      P.log.push_back(Profiler::LogRec{fam, (int)(block.x * block.y * block.z), kid, tm ? 1 : 0, k, bytes, flops});
  }
  if (tm) {
    const hipEvent_t a = P.get(), b = P.get();
    hipExtLaunchKernelGGL(kernel, grid, block, shm, s, a, b, 0, args...);
    P.pend.push_back(Profiler::Pending{a, b, fam, bytes, flops});
  } else {
    hipLaunchKernelGGL(kernel, grid, block, shm, s, args...);
  }
  // a launch the runtime rejects (resources, configuration) is reported at
  // its own call site
  NRG_HIP(hipGetLastError());
}

}  // namespace nrg
