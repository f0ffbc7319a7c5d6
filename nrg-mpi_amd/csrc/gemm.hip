// Batched variable-size FP64 GEMM on v_mfma_f64_16x16x4f64 (gfx950).
//
// One workgroup (256 lanes = 4 wave64) computes a 64x64 tile of C; each wave
// owns a 32x32 quadrant held as 2x2 MFMA 16x16 accumulators.  The K loop
// stages 16-deep slices of opA (64x16) and opB (16x64) through LDS.  Tiles of
// all tasks are flattened into one grid; a workgroup finds its task by binary
// search over the tile prefix sums, so blocks of very different sizes share
// one launch.
#include "gemm.h"

namespace nrg {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int TM = 64, TN = 64, TK = 16;
constexpr int LDS_STRIDE = TM + 1;   // odd stride in doubles: spreads k-rows over banks

__global__ __launch_bounds__(256) void k_gemm_f64(
    const GemmTask* __restrict__ tasks, const int32_t* __restrict__ tile_start, int ntasks,
    const double* __restrict__ A, const double* __restrict__ B, double* __restrict__ C,
    const int32_t* __restrict__ gA, const int32_t* __restrict__ gB,
    const double* __restrict__ ks) {
  __shared__ double As[TK * LDS_STRIDE];
  __shared__ double Bs[TK * LDS_STRIDE];
  const int bid = blockIdx.x;
  // binary search: last task with tile_start <= bid
  int lo = 0, hi = ntasks - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (tile_start[mid] <= bid) lo = mid; else hi = mid - 1;
  }
  const GemmTask t = tasks[lo];
  const int tiles_m = (t.m + TM - 1) / TM;
  const int local = bid - tile_start[lo];
  int m0, n0;
  if (t.lower) {
    // tiles (ti, tj), tj <= ti, enumerated row by row: local = ti(ti+1)/2 + tj
    int ti = (int)((sqrt(8.0 * (double)local + 1.0) - 1.0) * 0.5);
    while ((int64_t)ti * (ti + 1) / 2 > local) --ti;
    while ((int64_t)(ti + 1) * (ti + 2) / 2 <= local) ++ti;
    m0 = ti * TM;
    n0 = (local - ti * (ti + 1) / 2) * TN;
  } else {
    m0 = (local % tiles_m) * TM;
    n0 = (local / tiles_m) * TN;
  }
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;

  const double* Ab = A + t.offA;
  const double* Bb = B + t.offB;
  const int32_t* ga = (t.offGA >= 0) ? gA + t.offGA : nullptr;
  const int32_t* gb = (t.offGB >= 0) ? gB + t.offGB : nullptr;
  const double* sc = (t.offS >= 0) ? ks + t.offS : nullptr;

  d4 acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};

  for (int k0 = 0; k0 < t.k; k0 += TK) {
    // ---- stage opA(m0:m0+64, k0:k0+16) into As[kk][i]
    for (int e = tid; e < TM * TK; e += 256) {
      int i, kk;
      if (!t.transA) { i = e % TM; kk = e / TM; }   // stored rows = i: i fastest
      else { kk = e % TK; i = e / TK; }             // stored rows = k: k fastest
      const int gi = m0 + i, gk = k0 + kk;
      double v = 0.0;
      if (gi < t.m && gk < t.k) {
        if (!t.transA) {
          const int r = ga ? ga[gi] : gi;
          v = Ab[(int64_t)r + (int64_t)gk * t.lda];
        } else {
          const int r = ga ? ga[gk] : gk;
          v = Ab[(int64_t)r + (int64_t)gi * t.lda];
        }
      }
      As[kk * LDS_STRIDE + i] = v;
    }
    // ---- stage opB(k0:k0+16, n0:n0+64) into Bs[kk][j]
    for (int e = tid; e < TN * TK; e += 256) {
      int j, kk;
      if (!t.transB) { kk = e % TK; j = e / TK; }   // stored rows = k
      else { j = e % TN; kk = e / TN; }             // stored rows = j
      const int gj = n0 + j, gk = k0 + kk;
      double v = 0.0;
      if (gj < t.n && gk < t.k) {
        if (!t.transB) {
          const int r = gb ? gb[gk] : gk;
          v = Bb[(int64_t)r + (int64_t)gj * t.ldb];
        } else {
          const int r = gb ? gb[gj] : gj;
          v = Bb[(int64_t)r + (int64_t)gk * t.ldb];
        }
        if (sc) v *= sc[gk];
      }
      Bs[kk * LDS_STRIDE + j] = v;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < TK / 4; ++s) {
      const int kk = s * 4 + (lane >> 4);
      const int li = lane & 15;
      double a0 = As[kk * LDS_STRIDE + wm * 32 + li];
      double a1 = As[kk * LDS_STRIDE + wm * 32 + 16 + li];
      double b0 = Bs[kk * LDS_STRIDE + wn * 32 + li];
      double b1 = Bs[kk * LDS_STRIDE + wn * 32 + 16 + li];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
    __syncthreads();
  }
  // ---- epilogue: D[i][j] lives in lane j + 16*(i%4), register i/4
  // (lane layout of v_mfma_f64_16x16x4f64 checked on gfx950 by test_dgemm)
  double* Cb = C + t.offC;
  for (int mi = 0; mi < 2; ++mi)
    for (int ni = 0; ni < 2; ++ni)
      for (int r = 0; r < 4; ++r) {
        const int gi = m0 + wm * 32 + mi * 16 + (lane >> 4) + 4 * r;
        const int gj = n0 + wn * 32 + ni * 16 + (lane & 15);
        if (gi < t.m && gj < t.n) {
          double* c = Cb + (int64_t)gi + (int64_t)gj * t.ldc;
          double v = t.alpha * acc[mi][ni][r];
          if (t.beta != 0.0) v += t.beta * (*c);
          *c = v;
        }
      }
}

void GemmRunner::run(const GemmBatch& b, const double* A, const double* B, double* C,
                     const int32_t* gatherA, const int32_t* gatherB, const double* kscale,
                     hipStream_t s, int family) {
  const int nt = (int)b.tasks.size();
  if (nt == 0) return;
  h_tile_start.resize(nt);
  long long total = 0;
  double flops = 0, bytes = 0;
  for (int i = 0; i < nt; ++i) {
    const GemmTask& t = b.tasks[i];
    NRG_REQUIRE(t.m > 0 && t.n > 0 && t.k >= 0, "gemm: bad task shape");
    NRG_REQUIRE(!t.lower || t.m == t.n, "gemm: a lower-triangle task must be square");
    const double half = t.lower ? 0.5 : 1.0;
    flops += half * 2.0 * t.m * t.n * t.k;
    bytes += 8.0 * ((double)t.m * t.k + (double)t.k * t.n + half * (t.beta != 0.0 ? 2.0 : 1.0) * t.m * t.n);
    h_tile_start[i] = (int32_t)total;
    const long long tm = ceil_div(t.m, TM);
    total += t.lower ? tm * (tm + 1) / 2 : tm * ceil_div(t.n, TN);
  }
  NRG_REQUIRE(total < (1LL << 31), "gemm: too many tiles");
  d_tasks.upload(b.tasks, s);
  d_tile_start.upload(h_tile_start, s);
  const int pid = profiler().begin(s);
  hipLaunchKernelGGL(k_gemm_f64, dim3((unsigned)total), dim3(256), 0, s, d_tasks.p,
                     d_tile_start.p, nt, A, B, C, gatherA, gatherB, kscale);
  profiler().end(pid, s, family, bytes, flops);
  NRG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
__global__ void k_fill_f64(double* p, double v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t st = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) p[i] = v;
}
__global__ void k_fill_i32(int32_t* p, int32_t v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t st = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) p[i] = v;
}
void fill(double* p, double v, size_t n, hipStream_t s) {
  if (!n) return;
  unsigned g = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_fill_f64, dim3(g), dim3(256), 0, s, p, v, n);
}
void fill(int32_t* p, int32_t v, size_t n, hipStream_t s) {
  if (!n) return;
  unsigned g = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_fill_i32, dim3(g), dim3(256), 0, s, p, v, n);
}

}  // namespace nrg
