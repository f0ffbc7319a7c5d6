#pragma once
#include "context.h"

namespace nrg {
// scratch array of the no-allocation entry point, per device
constexpr size_t kScratchBytes = size_t(128) << 20;
constexpr int kNoallocMaxTasks = 4096;
// look up this device's copy of the array (device current)
void scratch_setup(Context& C);
// nrg_gpu_dgemm_batched (FP64 form) on device arrays, no device allocation
void dgemm_batched_noalloc(Context& C, int ntask, const int32_t* mnk, int transa, int transb, const double* A,
                           const int64_t* offA, const double* B, const int64_t* offB, double* Cm,
                           const int64_t* offC);
}  // namespace nrg
