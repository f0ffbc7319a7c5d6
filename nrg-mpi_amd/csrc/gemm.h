// Batched, variable-size FP64 GEMM on the CDNA4 FP64 matrix cores.
//
//   C_t = alpha_t * opA_t * opB_t + beta_t * C_t      for every task t
//
// Matrices are column-major.  Each stored operand may carry a row-gather
// index list (stored row r of the operand is fetched from row idx[r]) and the
// inner dimension may carry a per-k scale applied to opB.  This covers the
// plain products of the eigensolver (trailing updates, divide-and-conquer
// merges, back-transform) and the gathered products of the invariant-matrix,
// spectral and reduced-density contractions, whose coefficient matrices are
// signed partial permutations or per-operator dense sub-blocks.
#pragma once
#include "common.h"
#include <cstring>
#include "profiler.h"

namespace nrg {

struct GemmTask {
  int m, n, k;
  int lda, ldb, ldc;
  int transA, transB;       // 0: N, 1: T (of the stored matrix)
  int64_t offA, offB, offC; // element offsets into the base pointers
  int64_t offGA, offGB;     // offsets into the gather arrays, -1 = identity
  int64_t offS;             // offset into the k-scale array, -1 = none
  double alpha, beta;
  int lower;                // 1: square C, only tiles on or below the diagonal
  int pad;
};

struct GemmBatch {
  std::vector<GemmTask> tasks;
  void add(const GemmTask& t) {
    if (t.m > 0 && t.n > 0) tasks.push_back(t);
  }
  void clear() { tasks.clear(); }
};

// One split-K reduction: C(offC, m x n, ldc) = alpha sum of `parts` partial
// products at wsoff (m x n each, ld m) + beta C.
struct SplitRed {
  int64_t offC, wsoff;
  int m, n, ldc, parts;
  double alpha, beta;
};

// Runs all tasks of the batch in one launch.  A/B/C/gather/scale base
// pointers are shared by all tasks; tasks address them through offsets.
// Buffers for the task table live in the caller-owned workspace.
struct GemmRunner {
  // task table and tile prefix sums, uploaded together in one copy (two
  // slots: the unsplit and the split-K launch of one call)
  DevBuf<char> d_pack[2];
  std::vector<char> h_pack;
  std::vector<int32_t> h_tile_start;
  GemmBatch plain_, part_;
  std::vector<SplitRed> h_red_;
  DevBuf<SplitRed> d_red_;
  DevBuf<double> d_splitw_;
  void upload_tables(const GemmBatch& b, hipStream_t s, const GemmTask*& tk, const int32_t*& ts, int slot = 0) {
    const size_t nt = b.tasks.size();
    const size_t tb = (nt * sizeof(GemmTask) + 255) / 256 * 256;
    h_pack.resize(tb + h_tile_start.size() * sizeof(int32_t));
    std::memcpy(h_pack.data(), b.tasks.data(), nt * sizeof(GemmTask));
    std::memcpy(h_pack.data() + tb, h_tile_start.data(), h_tile_start.size() * sizeof(int32_t));
    d_pack[slot].upload(h_pack.data(), h_pack.size(), s);
    tk = reinterpret_cast<const GemmTask*>(d_pack[slot].p);
    ts = reinterpret_cast<const int32_t*>(d_pack[slot].p + tb);
  }
  void launch(const GemmBatch& b, const double* A, const double* B, double* C,
              const int32_t* gatherA, const int32_t* gatherB, const double* kscale,
              hipStream_t s, int family, int slot);
  // split > 0: the invariant-matrix products run as split-precision
  // products with `split` slices; split_spec > 0: the spectral-side families
  // (spectral invariant, CFS weights, rho) with `split_spec` slices.  A batch
  // of fewer than split_min_flops FP64-equivalent flops stays on the FP64
  // kernel (the slicing passes would cost more than they save).
  int split = 0;
  int split_spec = 0;
  double split_min_flops = 0.0;
  void run(const GemmBatch& b, const double* A, const double* B, double* C,
           const int32_t* gatherA, const int32_t* gatherB, const double* kscale,
           hipStream_t s, int family = KF_OTHER);
  // FP64 product from int8 matrix-core products of `slices` (3..8) signed
  // 7-bit digits per element (gemm_split.hip); digit planes and exponents
  // in the workspaces below
  DevBuf<int8_t> d_planes_;
  DevBuf<int32_t> d_exp_;
  double* oz_C_ = nullptr;
  void run_split(const GemmBatch& b, const double* A, const double* B, double* C,
                 const int32_t* gatherA, const int32_t* gatherB, const double* kscale,
                 int slices, hipStream_t s, int family = KF_OTHER);
  // Recording: between begin() and flush(), run() appends its task tables
  // (and split-K reduction records) to one host buffer and its launches to a
  // list; flush() uploads the buffer in one copy and issues the launches in
  // the recorded order on its stream.  Saves a copy per product where a
  // caller issues many small batched products in a row.
  struct RLaunch {
    int kind;                  // 0: GEMM, 1: split-K reduction
    int fast, tv, g, family;   // g: gather/scale bits of the fast kernel
    unsigned grid, gx;
    int nt;
    size_t tab, tst;           // byte offsets of the task table / tile starts (or reduction records)
    double bytes, flops;
    const double *A, *B;
    double* C;                 // nullptr: the split-K workspace
    const int32_t *gA, *gB;
    const double* ks;
  };
  bool rec_ = false;
  std::vector<char> r_host_;
  DevBuf<char> r_dev_;
  std::vector<RLaunch> r_l_;
  int64_t r_ws_ = 0;
  void begin() { rec_ = true; r_host_.clear(); r_l_.clear(); r_ws_ = 0; }
  void flush(hipStream_t s);
  // flush() in two parts, for callers that interleave the recorded launches
  // with other work: upload the tables once, then issue recorded launches
  // [i0, i1) where they belong (recorded() before and after a run() gives
  // its range); end_replay() drops the record
  size_t recorded() const { return r_l_.size(); }
  void upload_recorded(hipStream_t s);
  void issue(size_t i0, size_t i1, hipStream_t s);
  void end_replay() { r_l_.clear(); }
  size_t r_append(const void* p, size_t n) {
    const size_t o = (r_host_.size() + 255) / 256 * 256;
    r_host_.resize(o + n);
    std::memcpy(r_host_.data() + o, p, n);
    return o;
  }
  void release() {
    d_pack[0].release(); d_pack[1].release(); d_red_.release(); d_splitw_.release(); r_dev_.release();
    d_planes_.release(); d_exp_.release();
  }
  // Split of a batch into plain tasks and split-K parts (plain_, part_,
  // h_red_); returns the partial-product doubles the parts need.
  int64_t decompose(const GemmBatch& b);
  // run() (FP64 form) with every device-side table and the split-K partials
  // in a caller-provided device workspace: the same decomposition and
  // launches, no device allocation.  workspace_bytes(b) is what it needs.
  size_t workspace_bytes(const GemmBatch& b);
  void run_into(const GemmBatch& b, const double* A, const double* B, double* C, char* ws, size_t ws_bytes,
                hipStream_t s, int family = KF_OTHER);
  char* ws_ = nullptr;           // set during run_into: tables go here instead of the DevBufs
  size_t ws_cap_ = 0, ws_used_ = 0;
  hipStream_t ws_stream_ = nullptr;
  char* ws_stage(const void* p, size_t n);
};

inline GemmTask gemm_task(int m, int n, int k, int transA, int transB,
                          int64_t offA, int lda, int64_t offB, int ldb,
                          int64_t offC, int ldc, double alpha, double beta) {
  GemmTask t;
  t.m = m; t.n = n; t.k = k;
  t.lda = lda; t.ldb = ldb; t.ldc = ldc;
  t.transA = transA; t.transB = transB;
  t.offA = offA; t.offB = offB; t.offC = offC;
  t.offGA = -1; t.offGB = -1; t.offS = -1;
  t.alpha = alpha; t.beta = beta;
  t.lower = 0; t.pad = 0;
  return t;
}

}  // namespace nrg
