// Shared device/host helpers for the NRG HIP library (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>
#include <stdexcept>

namespace nrg {

// Library-wide error: thrown inside the library, converted to a status code
// and a message at the C-ABI boundary.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

enum Status : int {
  NRG_OK = 0,
  NRG_ERR_HIP = 1,
  NRG_ERR_ARG = 2,
  NRG_ERR_STATE = 3,
  NRG_ERR_NUMERIC = 4,
  NRG_ERR_INJECTED = 5,
};

#define NRG_HIP(call)                                                          \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      throw ::nrg::Error(::nrg::NRG_ERR_HIP,                                   \
                         std::string(#call) + ": " + hipGetErrorString(e_) +   \
                             " (" __FILE__ ":" + std::to_string(__LINE__) + ")"); \
  } while (0)

#define NRG_REQUIRE(cond, msg)                                                 \
  do {                                                                         \
    if (!(cond)) throw ::nrg::Error(::nrg::NRG_ERR_ARG, std::string(msg));     \
  } while (0)

// A growable device buffer (never shrinks; reallocation drops contents).
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t cap = 0;
  size_t n = 0;
  void reserve(size_t count) {
    if (count <= cap) { n = count; return; }
    if (p) NRG_HIP(hipFree(p));
    p = nullptr;
    cap = 0;
    size_t c = count + count / 8 + 64;
    NRG_HIP(hipMalloc(&p, c * sizeof(T)));
    cap = c;
    n = count;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr; cap = 0; n = 0;
  }
  void upload(const T* h, size_t count, hipStream_t s) {
    reserve(count);
    if (count) NRG_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T>& v, hipStream_t s) { upload(v.data(), v.size(), s); }
  void download(T* h, size_t count, hipStream_t s) const {
    if (count) NRG_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
  }
  void zero(hipStream_t s);
};

// Fill kernel used instead of hipMemsetAsync so that the same code is safe
// inside captured graphs.
__global__ void k_fill_f64(double* p, double v, size_t n);
__global__ void k_fill_i32(int32_t* p, int32_t v, size_t n);
void fill(double* p, double v, size_t n, hipStream_t s);
void fill(int32_t* p, int32_t v, size_t n, hipStream_t s);

template <typename T>
inline void DevBuf<T>::zero(hipStream_t s) {
  if (n) fill(reinterpret_cast<T*>(p), T(0), n, s);
}

inline int ceil_div(long long a, long long b) { return (int)((a + b - 1) / b); }

}  // namespace nrg
