// Plain streaming copy of known size.  Under `rocprofv3 --pmc FETCH_SIZE`
// (or WRITE_SIZE) its counter reading per dispatch, against the bytes it
// reads and writes, gives the factor that turns the same counters of the
// NRG kernels into HBM bytes.
#include "calib.h"

namespace nrg {

// each thread copies 16-byte vectors in a grid-stride loop: every byte of
// src is read once and every byte of dst written once
__global__ __launch_bounds__(256) void k_stream_copy(double2* __restrict__ dst, const double2* __restrict__ src,
                                                     int64_t n2) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += stride) dst[i] = src[i];
}

__global__ void k_iota_f64(double* p, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = (double)i;
}

bool copy_calibration(int64_t nbytes, int reps, hipStream_t s) {
  NRG_REQUIRE(nbytes >= 16 && nbytes % 16 == 0, "copy calibration: bytes must be a positive multiple of 16");
  NRG_REQUIRE(reps >= 1 && reps <= 100, "copy calibration: 1..100 repetitions");
  const int64_t n = nbytes / 8;
  DevBuf<double> src, dst;
  src.reserve((size_t)n);
  dst.reserve((size_t)n);
  k_iota_f64<<<2048, 256, 0, s>>>(src.p, n);
  NRG_HIP(hipGetLastError());
  for (int r = 0; r < reps; ++r) {
    k_stream_copy<<<8192, 256, 0, s>>>(reinterpret_cast<double2*>(dst.p), reinterpret_cast<const double2*>(src.p),
                                       n / 2);
    NRG_HIP(hipGetLastError());
  }
  double tail[2] = {0.0, 0.0};
  NRG_HIP(hipMemcpyAsync(tail, dst.p + (n - 2), 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  NRG_HIP(hipStreamSynchronize(s));
  src.release();
  dst.release();
  return tail[0] == (double)(n - 2) && tail[1] == (double)(n - 1);
}

}  // namespace nrg
