#pragma once
#include "common.h"

namespace nrg {
// nbytes copied device to device reps times by k_stream_copy; true when the
// copy arrived
bool copy_calibration(int64_t nbytes, int reps, hipStream_t s);
}  // namespace nrg
