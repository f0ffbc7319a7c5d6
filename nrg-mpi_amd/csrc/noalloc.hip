// The batched FP64 product of nrg_gpu_dgemm_batched for callers that cannot
// allocate device memory (code under stream capture, tight serving loops,
// harnesses that forbid hipMalloc): operands and results are device arrays
// the caller owns, and every other device buffer the kernels need -- the
// task table, the tile starts, the split-K reduction records and the
// split-K partial products -- lives in one statically allocated array.
//
// The array is file-scope device memory, so each GPU the library runs on
// holds its own copy (kScratchBytes) until the process ends.  Host code
// never uses the array's name as an address: nrg_gpu_init looks up each
// device's copy with hipGetSymbolAddress once, with that device current,
// and keeps the pointer in the context.  Every call on a GPU shares that
// GPU's array and leaves the last call's tables in it; calls that overlap on
// one GPU race (the library is not reentrant in any case).
#include "noalloc.h"

namespace nrg {

static __device__ __align__(256) char scratch[kScratchBytes];

void scratch_setup(Context& C) {
  void* p = nullptr;
  NRG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  NRG_REQUIRE(p != nullptr, "no-allocation scratch: symbol address not found");
  C.scratch = static_cast<char*>(p);
}

void dgemm_batched_noalloc(Context& C, int ntask, const int32_t* mnk, int transa, int transb, const double* A,
                           const int64_t* offA, const double* B, const int64_t* offB, double* Cm,
                           const int64_t* offC) {
  NRG_REQUIRE(C.scratch != nullptr, "dgemm_batched_noalloc: the scratch array is not set up (nrg_gpu_init)");
  NRG_REQUIRE(ntask >= 0 && mnk && offA && offB && offC, "dgemm_batched_noalloc: bad arguments");
  NRG_REQUIRE(ntask <= kNoallocMaxTasks, "dgemm_batched_noalloc: at most " + std::to_string(kNoallocMaxTasks) +
                                            " products per call");
  if (ntask == 0) return;
  NRG_REQUIRE(A && B && Cm, "dgemm_batched_noalloc: null device array");
  GemmBatch b;
  for (int i = 0; i < ntask; ++i) {
    const int m = mnk[3 * i], n = mnk[3 * i + 1], k = mnk[3 * i + 2];
    NRG_REQUIRE(m > 0 && n > 0 && k > 0, "dgemm_batched_noalloc: sizes must be positive");
    NRG_REQUIRE(offA[i] >= 0 && offB[i] >= 0 && offC[i] >= 0, "dgemm_batched_noalloc: negative offset");
    const int lda = transa ? k : m, ldb = transb ? n : k;
    b.add(gemm_task(m, n, k, transa, transb, offA[i], lda, offB[i], ldb, offC[i], m, 1.0, 0.0));
  }
  const size_t need = C.gemm.workspace_bytes(b);
  NRG_REQUIRE(need <= kScratchBytes, "dgemm_batched_noalloc: the products need " + std::to_string(need) +
                                         " bytes of scratch (task tables and split-K partials), more than the " +
                                         std::to_string(kScratchBytes) + " the no-allocation entry point holds");
  C.gemm.run_into(b, A, B, Cm, C.scratch, kScratchBytes, C.stream);
  NRG_HIP(hipStreamSynchronize(C.stream));
}

}  // namespace nrg
