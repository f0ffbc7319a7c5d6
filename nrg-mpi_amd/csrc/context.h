// Library context: one device, one stream, persistent device buffers.
#pragma once
#include "common.h"
#include "eig.h"
#include "gemm.h"
#include "tuner.h"

namespace nrg {

struct Timers {
  // per-stage milliseconds of the last call (host wall clock around synchronised stages)
  double assemble = 0, eig = 0, tridiag = 0, dc = 0, back = 0, invariant = 0, spectral = 0;
  double rho = 0, cfs = 0;
};

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool initialised = false;
  std::string last_error;
  BatchedEigensolver eig;
  EigStats eig_stats;
  GemmRunner gemm;
  Timers timers;
  int fault_iteration = -1000000;   // test hook: fail diagonalisation at this iteration
  int batch_limit = 0;              // max blocks per eigensolver launch sequence (0: all)
  bool u_lazy = false;              // d_U columns are back-transformed on demand (ensure_columns)
  // stages that return without waiting for the device are timed with event
  // pairs, read when the stage times are asked for
  hipEvent_t ev_back[2] = {nullptr, nullptr}, ev_inv[2] = {nullptr, nullptr};
  bool back_timed = false, inv_timed = false;
  double last_back_ms = 0, last_inv_ms = 0;   // the last completed spans
  // diagonalisation stage boundaries: before and after the assembly, after
  // the eigensolve (read after the eigenvalue download has been waited for)
  hipEvent_t ev_dz[3] = {nullptr, nullptr, nullptr};
  // this device's copy of the no-allocation entry point's scratch array
  // (hipGetSymbolAddress at init)
  char* scratch = nullptr;
  // tridiagonalisation launch options: graphs (nrg_gpu_set_graphs) and the
  // histogram-keyed tuner (nrg_gpu_set_tuning)
  bool graphs_on = false;
  int groups_override = 0;   // stream groups of the tridiagonalisation (0: tuner / default)
  Tuner tuner;

  // ---- model constants (set once)
  int nv = 0, nops = 0, nchain = 0, ncm = 0, ncoef = 0, nmat = 0;
  std::vector<int32_t> operation;       // [nops][nv+nchain] row-major (op, column)
  std::vector<int32_t> hop_valid, hop_type, hop_chan, hop_swap;  // [nops*nops]
  std::vector<double> hop_coef;
  std::vector<int32_t> inv_a, inv_b, inv_t;  // invariant table (deduplicated, last wins)
  std::vector<double> inv_c;
  std::vector<double> coef_spec;        // [nmat][nops]
  std::vector<int32_t> cons;            // chargeStart, chargeEnd, spinStart, spinEnd
  std::vector<int32_t> cons_diff_spec;  // [nmat][nv]
  std::vector<int32_t> pos_op, neg_op;  // [nmat][nv] label shifts of the spectral operators
  DevBuf<int32_t> d_op, d_hop;          // d_hop: [valid,type,chan,swap,bcs] x nops^2
  DevBuf<double> d_hop_coef;
  // BCS pairing: channel + 1 of the last forward match of (op_i, op_j) in
  // coefficient_BCS (0: none), and chain_BCS(row, ch) column-major, rows 0..
  std::vector<int32_t> bcs_chan;        // [nops*nops]
  std::vector<double> chain_bcs;
  int bcs_rows = 0;

  // ---- current iteration (device resident)
  int nsub = 0, nfull = 0;
  std::vector<int32_t> sub;             // nsub*(nv+3)
  std::vector<int32_t> basis_in;        // nfull*(nv+3)
  std::vector<int> bn;                  // block sizes
  std::vector<int64_t> boff, bvoff;     // sum n^2 / sum n offsets
  std::vector<double> E_block;          // shifted eigenvalues in block order (as returned)
  DevBuf<int32_t> d_basis_in;
  DevBuf<double> d_H, d_U, d_w, d_Eold;
  // invariant matrices of the kept basis (dense k x k x ncm / nmat, column-major)
  int kold = 0;
  DevBuf<double> d_inv, d_inv_next;
  int kold_spec = 0;
  DevBuf<double> d_spec, d_spec_next;
  DevBuf<double> d_tmp;
  DevBuf<int32_t> d_gidx;
  DevBuf<double> d_gscale;
  // host staging of the gather lists and scales, kept between calls so
  // their pages stay mapped and they do not regrow every iteration
  std::vector<int32_t> h_gidx;
  std::vector<double> h_gscale;
};

Context& ctx();

// Make columns [0, hi[b]) of every block of d_U eigenvectors of the full
// matrices (hi == nullptr: all columns).  Only the batched solve of
// nrg_gpu_diagonalize without an eigenvector output defers its
// back-transform; every reader of d_U calls this first.
inline void ensure_columns(Context& C, const int32_t* hi) {
  if (!C.u_lazy) return;
  std::vector<int> h(C.nsub);
  bool all = true;
  for (int b = 0; b < C.nsub; ++b) {
    h[b] = hi ? std::min(std::max((int)hi[b], 0), C.bn[b]) : C.bn[b];
    all = all && h[b] == C.bn[b];
  }
  if (!C.back_timed) NRG_HIP(hipEventRecord(C.ev_back[0], C.stream));   // first transform of the iteration
  C.eig.back_columns(h.data(), C.stream, nullptr);
  NRG_HIP(hipEventRecord(C.ev_back[1], C.stream));
  C.back_timed = true;
  if (all) C.u_lazy = false;
}

}  // namespace nrg
