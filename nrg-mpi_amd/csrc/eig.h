// Batched symmetric eigensolver for the (Q,Sz) blocks of one NRG iteration.
//
// Replaces one dsyev('V','U') call per block (diagonalization.F90:266-269)
// by a single batched pipeline over all blocks of the iteration:
//   1. blocked Householder tridiagonalisation (panel columns with a
//      multi-workgroup SYMV, trailing SYR2K on the FP64 matrix cores);
//   2. divide-and-conquer tridiagonal eigensolver (QL leaves, rank-one
//      merges with deflation, bisection-solved secular equations,
//      Gu-Eisenstat eigenvectors, merge products on the FP64 matrix cores);
//   3. blocked WY back-transform of the eigenvectors (FP64 matrix cores).
// Blocks are packed back to back: block b is an n_b x n_b column-major
// matrix at element offset aoff[b] (= sum of n^2 of earlier blocks) and owns
// n_b eigenvalues at voff[b] (= sum of n of earlier blocks), the layout of the
// reference's eigenvector/eigenvalue arrays (prepareBasis.F90:97-110).
#pragma once
#include "common.h"
#include "gemm.h"
#include "profiler.h"
#include <array>

namespace nrg {

struct EigStats {
  double ms_tridiag = 0, ms_dc = 0, ms_back = 0;
  double ms_tridiag_issue = 0;   // host time spent issuing the tridiagonalisation's launches
  int max_n = 0, nblk = 0, dc_levels = 0;
  double flops_alg = 0;   // sum (10/3) n^3 over blocks
};

class BatchedEigensolver {
 public:
  // Solve all blocks.  On entry A holds the matrices (full storage); on exit
  // Z holds eigenvectors (ascending eigenvalue order, one column each) and w
  // the eigenvalues.  A is overwritten with Householder data.  A and Z are
  // device buffers of sum n^2 doubles; w of sum n.  With defer_back the
  // columns of Z stay the eigenvectors of the tridiagonal matrices until
  // back_columns() transforms them (A and the reflector scalars are kept
  // until the next solve).
  void solve(int nblk, const int* n, const int64_t* aoff, const int64_t* voff,
             double* A, double* Z, double* w, bool defer_back, hipStream_t s,
             EigStats* st = nullptr);
  // Back-transform (and chop) columns [done_b, hi[b]) of every block b
  // (hi == nullptr: all columns); columns already transformed are skipped.
  // *ms (if not null) accumulates the time.
  void back_columns(const int* hi, hipStream_t s, double* ms = nullptr);
  // Stage times (tridiagonalisation, divide and conquer, back-transform) of
  // the last solve that was given an EigStats, from its events; waits for
  // the last of them (callers read them after waiting for the device).
  void stats(EigStats* st);
  // Columns of block b that hold eigenvectors of the full matrix.
  int columns_done(int b) const { return done_[b]; }
  void release();
  // Column schedule of the tridiagonalisation: all blocks start together
  // (false) or end together (true, the default).
  void set_align_end(bool on) { align_end_ = on; }
  // Blocks split over this many streams in the tridiagonalisation (1..4).
  void set_groups(int g) { groups_ = std::max(1, std::min(g, MAX_TR_GROUPS)); }
  int groups() const { return groups_; }
  // 32-column SYMV panels only for launches with at least `wg` of them and
  // `bytes` of traffic (both per group count); 16-column panels otherwise.
  void set_symv_wide(int wg, double bytes) { symv_wide_wg_ = std::max(1, wg); symv_wide_bytes_ = bytes; }
  int symv_wide_wg() const { return symv_wide_wg_; }
  // Capture the tridiagonalisation's launch sequence as a HIP graph and
  // replay it for a repeated block set (off by default).
  void set_graphs(bool on) { use_graphs_ = on; }
  // Two-stage tridiagonalisation (dense -> band of 32 -> tridiagonal, sbr.hip)
  // instead of the one-stage reduction.
  void set_two_stage(bool on) { two_stage_ = on; }
  // SYMV of the one-stage reduction: row parts added with LDS atomics (0)
  // or in register/LDS row chunks in a fixed order (1)
  void set_symv_kind(int kind) { symv_kind_ = kind; }
  // trailing size from which the last columns of every block are reduced in
  // LDS by one launch (0: column steps to the end; at most 192)
  void set_finish(int m) { finish_m_ = std::max(0, std::min(m, 192)); }
  bool two_stage() const { return two_stage_; }
  long long graph_captures() const { return graph_captures_; }
  long long graph_replays() const { return graph_replays_; }

 private:
  void tridiag(hipStream_t s);
  void tridiag_two_stage(hipStream_t s);
  void back_q2(const int* lo, const int* hi, hipStream_t s);
  void divide_conquer(hipStream_t s);
  void back_transform(const int* lo, const int* hi, hipStream_t s);
  void panel_factors(hipStream_t s);
  void materialize(const int* hi, hipStream_t s);

  int nblk_ = 0;
  std::vector<int> n_;
  std::vector<int64_t> aoff_, voff_;
  std::vector<int> order_;          // blocks by n descending
  int64_t tot_n_ = 0, tot_n2_ = 0;
  double *A_ = nullptr, *Z_ = nullptr, *w_ = nullptr;

  DevBuf<int32_t> d_n_, d_order_;          // sizes, size order
  bool align_end_ = true;                   // end-aligned column schedule (see tridiag)
  DevBuf<int64_t> d_aoff_, d_voff_;
  DevBuf<double> d_d_, d_e_, d_tau_, d_y_, d_P_, d_part_;
  // divide and conquer
  DevBuf<double> d_ws_;
  DevBuf<int32_t> d_iws_;
  DevBuf<int32_t> d_tasks_;
  DevBuf<int32_t> d_k_;
  std::vector<int32_t> h_k_;
  // deferred top merge of each block (lev 0: the block is a leaf)
  // a block's deferred top merge: size m (first half m1), k non-deflated
  // columns grouped as k1 first-half, k3 both-halves, k - k1 - k3 second-half
  struct TopMerge { int lev = 0, m = 0, m1 = 0, k = 0, k1 = 0, k3 = 0; int64_t wo = 0, io = 0; };
  std::vector<TopMerge> top_;
  std::vector<int> mat_;                    // columns of Z holding tridiagonal eigenvectors
  DevBuf<double> d_usel_;
  DevBuf<int32_t> d_mtask_;
  // back-transform
  DevBuf<double> d_V_, d_T_, d_W_, d_X2_;   // panel V, [G | T], W = V T^T, T V^T Z
  DevBuf<int64_t> d_rng_;
  // reflector panels: V at vpo[b] + p n NBB, [G | T] at index tpo[b] + p
  std::vector<int64_t> h_vpo_;
  std::vector<int32_t> h_tpo_, h_plist_;
  DevBuf<int64_t> d_vpo_;
  DevBuf<int32_t> d_tpo_, d_plist_;
  hipEvent_t ev_bt_ = nullptr;   // panel factors built (second stream)
  GemmRunner gemm_bt_;           // their V^T V products
  GemmBatch batch_bt_;
  bool bt_wait_ = false;
  std::vector<int64_t> h_rng_;
  std::vector<int> done_;           // transformed columns per block
  GemmRunner gemm_;
  GemmBatch batch_;
  // block groups of the tridiagonalisation: per group a stream (group 0:
  // the caller's), a slot table and a trailing-update runner
  static constexpr int MAX_TR_GROUPS = 4;
  int groups_ = 2;
  hipStream_t sg_[MAX_TR_GROUPS] = {};
  hipEvent_t evg_[MAX_TR_GROUPS] = {};
  hipEvent_t ev_fork_ = nullptr;
  hipEvent_t evf_[MAX_TR_GROUPS] = {};   // fences around timed SYMV launches
  hipEvent_t ev_st_[4] = {};             // stage boundaries of the last timed solve (stats())
  DevBuf<char> d_slotg_[MAX_TR_GROUPS];
  GemmRunner gemmg_[MAX_TR_GROUPS];
  GemmBatch batchg_[MAX_TR_GROUPS];
  int symv_wide_wg_ = 512;
  double symv_wide_bytes_ = 64.0e6;
  // captured tridiagonalisations: launch sequence of one block set, with the
  // work its launches declare (added to the issued-work counters on replay)
  struct TrGraph {
    std::vector<int64_t> key;
    hipGraphExec_t exec = nullptr;
    std::array<double, KF_COUNT> launches{}, bytes{}, flops{};
    long long used = 0;
  };
  static constexpr size_t kMaxGraphs = 4;
  std::vector<TrGraph> graphs_;
  bool use_graphs_ = false;
  // two-stage reduction: band, stage-2 reflectors and their offsets
  bool two_stage_ = false;
  int symv_kind_ = 0;
  int finish_m_ = 192;
  DevBuf<char> d_fin_;   // tail finisher tasks
  bool solved_two_stage_ = false;   // how the current solve reduced (its back-transform)
  int bt_off_ = 1;                  // row offset of a stage-1 reflector's unit element
  DevBuf<double> d_tq_, d_band_, d_v2_, d_t2_;
  DevBuf<int64_t> d_sboff_;
  std::vector<int64_t> h_sboff_;
  DevBuf<int32_t> d_sbch_, d_err_;
  long long graph_clock_ = 0, graph_captures_ = 0, graph_replays_ = 0;
};

}  // namespace nrg
