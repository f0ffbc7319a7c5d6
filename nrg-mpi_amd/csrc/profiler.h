// Optional per-kernel-family timing with HIP events recorded on the
// library's stream around each launch.  Off by default; when on, each API
// call drains the recorded events at its end (after its own synchronisation)
// and accumulates duration, launch count and the algorithmic bytes/flops the
// launch sites declare.
#pragma once
#include "common.h"
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
  KF_OTHER,
  KF_COUNT
};

struct KStat {
  double ms = 0, bytes = 0, flops = 0;
  long long launches = 0;
};

struct Profiler {
  bool on = false;
  std::array<KStat, KF_COUNT> st{};
  struct Pending { hipEvent_t a, b; int fam; double bytes, flops; };
  std::vector<Pending> pend;
  std::vector<hipEvent_t> pool;

  hipEvent_t get() {
    if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
    hipEvent_t e;
    NRG_HIP(hipEventCreate(&e));
    return e;
  }
  // returns an event pair index, or -1 when off
  int begin(hipStream_t s) {
    if (!on) return -1;
    Pending p{get(), get(), 0, 0, 0};
    NRG_HIP(hipEventRecord(p.a, s));
    pend.push_back(p);
    return (int)pend.size() - 1;
  }
  void end(int id, hipStream_t s, int fam, double bytes, double flops) {
    if (id < 0) return;
    Pending& p = pend[id];
    p.fam = fam; p.bytes = bytes; p.flops = flops;
    NRG_HIP(hipEventRecord(p.b, s));
  }
  // call after the stream has been synchronised
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
  void reset() { st = {}; }
};

Profiler& profiler();

}  // namespace nrg
