// Split-precision batched GEMM: FP64 products from FP16 matrix-core products
// (Ozaki-style error-free slicing; Ozaki, Ogita, Oishi & Rump, Numer.
// Algorithms 59 (2012) 95-118).
//
// Every row i of opA is scaled by 2^-ea_i (ea_i the binary exponent of its
// largest magnitude, so that the scaled row lies in (-1, 1)) and written in
// S signed base-2^7 digits,
//     opA(i,k) = 2^ea_i * sum_{s<S} a_s(i,k) 2^{-7(s+1)}  (+ a remainder below
//                2^{-7S}),
// with |a_0| <= 128 and |a_s| <= 64 for s > 0 (round-to-nearest digits);
// every column j of opB likewise with eb_j.  The digits are small integers,
// exact in FP16.  Products of digit matrices run on v_mfma_f32_16x16x32_f16:
// per 32-deep K slice and per diagonal d = s + t, the sum of the d+1 digit
// products is an integer below (d+3) 2^12 * 32 per element, so an FP32
// accumulator that is flushed to FP64 every 256 K (< 2^24 for d <= 13) holds
// it exactly.  Pairs with s + t >= S are dropped (the triangular truncation
// of the scheme).  The FP64 result is
//     C(i,j) = alpha 2^{ea_i + eb_j} sum_d 2^{-7(d+2)} P_d(i,j) + beta C(i,j).
//
// Tiles, task table, gathers, k-scales and the lower-triangle option are
// those of the FP64 kernel (gemm.hip); one workgroup (4 waves) computes a
// 64 x 64 tile, each wave a 32 x 32 quadrant.
#include "gemm.h"

namespace nrg {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d4s __attribute__((ext_vector_type(4)));

constexpr int SP_T = 64;          // tile edge
constexpr int SP_K = 32;          // K slice (one 16x16x32 MFMA step)
constexpr int SP_KS = SP_K + 8;   // LDS row stride in halves (k fastest)
constexpr int SP_FLUSH = 8;       // slices per FP32 accumulation window (256 K)
constexpr int SP_BETA = 7;        // bits per digit
constexpr int SP_MAXS = 8;

template <int S>
__global__ __launch_bounds__(256) void k_gemm_split(
    const GemmTask* __restrict__ tasks, const int32_t* __restrict__ tile_start, int ntasks,
    const double* __restrict__ A, const double* __restrict__ B, double* __restrict__ C,
    const int32_t* __restrict__ gA, const int32_t* __restrict__ gB, const double* __restrict__ ks) {
  __shared__ __attribute__((aligned(16))) _Float16 As[S][SP_T][SP_KS];
  __shared__ __attribute__((aligned(16))) _Float16 Bs[S][SP_T][SP_KS];
  __shared__ double s_sa[SP_T], s_sb[SP_T];    // 2^-ea_i, 2^-eb_j
  __shared__ int s_ea[SP_T], s_eb[SP_T];
  __shared__ double s_red[4][SP_T];
  const int bid = blockIdx.x;
  int lo = 0, hi = ntasks - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tile_start[mid] <= bid) lo = mid; else hi = mid - 1;
  }
  const GemmTask t = tasks[lo];
  const int tiles_m = (t.m + SP_T - 1) / SP_T;
  const int local = bid - tile_start[lo];
  int m0, n0;
  if (t.lower) {
    int ti = (int)((sqrt(8.0 * (double)local + 1.0) - 1.0) * 0.5);
    while ((int64_t)ti * (ti + 1) / 2 > local) --ti;
    while ((int64_t)(ti + 1) * (ti + 2) / 2 <= local) ++ti;
    m0 = ti * SP_T;
    n0 = (local - ti * (ti + 1) / 2) * SP_T;
  } else {
    m0 = (local % tiles_m) * SP_T;
    n0 = (local / tiles_m) * SP_T;
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

  auto opA = [&](int gi, int gk) -> double {   // caller checks bounds
    if (!t.transA) {
      const int r = ga ? ga[gi] : gi;
      return Ab[(int64_t)r + (int64_t)gk * t.lda];
    }
    const int r = ga ? ga[gk] : gk;
    return Ab[(int64_t)r + (int64_t)gi * t.lda];
  };
  auto opB = [&](int gk, int gj) -> double {
    double v;
    if (!t.transB) {
      const int r = gb ? gb[gk] : gk;
      v = Bb[(int64_t)r + (int64_t)gj * t.ldb];
    } else {
      const int r = gb ? gb[gj] : gj;
      v = Bb[(int64_t)r + (int64_t)gk * t.ldb];
    }
    return sc ? v * sc[gk] : v;
  };

  // ---- pass 1: row maxima of opA(m0:m0+64, :) and column maxima of
  // opB(:, n0:n0+64); thread (r = tid & 63, phase = tid >> 6) takes k = phase mod 4
  {
    const int r = tid & 63, ph = tid >> 6;
    double ma = 0.0, mb = 0.0;
    const int gi = m0 + r, gj = n0 + r;
    for (int k = ph; k < t.k; k += 4) {
      if (gi < t.m) ma = fmax(ma, fabs(opA(gi, k)));
      if (gj < t.n) mb = fmax(mb, fabs(opB(k, gj)));
    }
    s_red[ph][r] = ma;
    __syncthreads();
    if (ph == 0) ma = fmax(fmax(s_red[0][r], s_red[1][r]), fmax(s_red[2][r], s_red[3][r]));
    __syncthreads();
    s_red[ph][r] = mb;
    __syncthreads();
    if (ph == 0) {
      mb = fmax(fmax(s_red[0][r], s_red[1][r]), fmax(s_red[2][r], s_red[3][r]));
      int e;
      (void)frexp(ma, &e);   // ma = f 2^e, f in [0.5, 1): |row| / 2^e < 1
      if (ma == 0.0) e = 0;
      s_ea[r] = e;
      s_sa[r] = ldexp(1.0, -e);
      (void)frexp(mb, &e);
      if (mb == 0.0) e = 0;
      s_eb[r] = e;
      s_sb[r] = ldexp(1.0, -e);
    }
    __syncthreads();
  }

  // ---- pass 2: digits of 32-deep slices in LDS, FP16 products per diagonal
  f4 acc[S][2][2];
  d4s acc64[2][2];
#pragma unroll
  for (int d = 0; d < S; ++d)
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[d][a][b] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc64[a][b] = d4s{0.0, 0.0, 0.0, 0.0};

  auto flush = [&]() {
#pragma unroll
    for (int d = 0; d < S; ++d) {
      const double w = ldexp(1.0, -SP_BETA * (d + 2));
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
#pragma unroll
          for (int r = 0; r < 4; ++r) acc64[a][b][r] += w * (double)acc[d][a][b][r];
          acc[d][a][b] = f4{0.f, 0.f, 0.f, 0.f};
        }
    }
  };

  // element e = tid + 256 u of the 64 x 32 slices, index order chosen so a
  // wave reads contiguous stored memory; the next slice's values are
  // fetched into registers while the matrix cores work on the current one
  double ra[8], rb[8];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = tid + 256 * u;
      int i, kk;
      if (!t.transA) { i = e & 63; kk = e >> 6; } else { kk = e & 31; i = e >> 5; }
      const int gi = m0 + i, gk = k0 + kk;
      ra[u] = (gi < t.m && gk < t.k) ? opA(gi, gk) : 0.0;
      int j, kb;
      if (!t.transB) { kb = e & 31; j = e >> 5; } else { j = e & 63; kb = e >> 6; }
      const int gj = n0 + j, gk2 = k0 + kb;
      rb[u] = (gj < t.n && gk2 < t.k) ? opB(gk2, gj) : 0.0;
    }
  };
  auto digits = [&]() {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = tid + 256 * u;
      int i, kk;
      if (!t.transA) { i = e & 63; kk = e >> 6; } else { kk = e & 31; i = e >> 5; }
      double r = ra[u] * s_sa[i];
#pragma unroll
      for (int s = 0; s < S; ++s) {
        const double dg = rint(ldexp(r, SP_BETA));
        r = ldexp(r, SP_BETA) - dg;
        As[s][i][kk] = (_Float16)dg;
      }
      int j, kb;
      if (!t.transB) { kb = e & 31; j = e >> 5; } else { j = e & 63; kb = e >> 6; }
      double q = rb[u] * s_sb[j];
#pragma unroll
      for (int s = 0; s < S; ++s) {
        const double dg = rint(ldexp(q, SP_BETA));
        q = ldexp(q, SP_BETA) - dg;
        Bs[s][j][kb] = (_Float16)dg;
      }
    }
  };

  // lane layout of v_mfma_f32_16x16x32_f16: A row (lane & 15), k 8*(lane>>4)..+7;
  // B column (lane & 15), same k; D row 4*(lane>>4)+r, column lane & 15
  const int li = lane & 15, kg = (lane >> 4) * 8;
  if (t.k > 0) {
    fetch(0);
    digits();
    __syncthreads();
  }
  int slice = 0;
  for (int k0 = 0; k0 < t.k; k0 += SP_K, ++slice) {
    const bool more = (k0 + SP_K < t.k);
    if (more) fetch(k0 + SP_K);
#pragma unroll
    for (int sa = 0; sa < S; ++sa) {
      h8 av[2];
#pragma unroll
      for (int mi = 0; mi < 2; ++mi) av[mi] = *(const h8*)&As[sa][wm * 32 + mi * 16 + li][kg];
#pragma unroll
      for (int tt = 0; sa + tt < S; ++tt) {
        h8 bv[2];
#pragma unroll
        for (int ni = 0; ni < 2; ++ni) bv[ni] = *(const h8*)&Bs[tt][wn * 32 + ni * 16 + li][kg];
#pragma unroll
        for (int mi = 0; mi < 2; ++mi)
#pragma unroll
          for (int ni = 0; ni < 2; ++ni)
            acc[sa + tt][mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_f16(av[mi], bv[ni], acc[sa + tt][mi][ni], 0, 0, 0);
      }
    }
    if ((slice + 1) % SP_FLUSH == 0) flush();
    if (more) {
      __syncthreads();
      digits();
      __syncthreads();
    }
  }
  flush();

  double* Cb = C + t.offC;
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < 2; ++ni)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int il = wm * 32 + mi * 16 + (lane >> 4) * 4 + r;
        const int jl = wn * 32 + ni * 16 + (lane & 15);
        const int gi = m0 + il, gj = n0 + jl;
        if (gi < t.m && gj < t.n) {
          double* c = Cb + (int64_t)gi + (int64_t)gj * t.ldc;
          double v = t.alpha * ldexp(acc64[mi][ni][r], s_ea[il] + s_eb[jl]);
          if (t.beta != 0.0) v += t.beta * (*c);
          *c = v;
        }
      }
}

void GemmRunner::run_split(const GemmBatch& b, const double* A, const double* B, double* C,
                           const int32_t* gatherA, const int32_t* gatherB, const double* kscale,
                           int slices, hipStream_t s, int family) {
  const int nt = (int)b.tasks.size();
  if (nt == 0) return;
  NRG_REQUIRE(slices >= 3 && slices <= SP_MAXS, "gemm: split slices must be 3..8");
  NRG_REQUIRE(!rec_, "gemm: split-precision products are not recorded");
  h_tile_start.resize(nt);
  long long total = 0;
  double flops = 0, bytes = 0;
  for (int i = 0; i < nt; ++i) {
    const GemmTask& t = b.tasks[i];
    NRG_REQUIRE(t.m > 0 && t.n > 0 && t.k >= 0, "gemm: bad task shape");
    NRG_REQUIRE(!t.lower || t.m == t.n, "gemm: a lower-triangle task must be square");
    const double half = t.lower ? 0.5 : 1.0;
    flops += half * 2.0 * t.m * t.n * t.k;   // FP64-equivalent algorithmic flops
    bytes += 8.0 * ((double)t.m * t.k + (double)t.k * t.n + half * (t.beta != 0.0 ? 2.0 : 1.0) * t.m * t.n);
    h_tile_start[i] = (int32_t)total;
    const long long tm = ceil_div(t.m, SP_T);
    total += t.lower ? tm * (tm + 1) / 2 : tm * ceil_div(t.n, SP_T);
  }
  NRG_REQUIRE(total < (1LL << 31), "gemm: too many tiles");
  const GemmTask* tk;
  const int32_t* ts;
  upload_tables(b, s, tk, ts);
  const dim3 g((unsigned)total), blk(256);
  switch (slices) {
    case 3: plaunch(family, bytes, flops, k_gemm_split<3>, g, blk, 0, s, tk, ts, nt, A, B, C, gatherA, gatherB, kscale); break;
    case 4: plaunch(family, bytes, flops, k_gemm_split<4>, g, blk, 0, s, tk, ts, nt, A, B, C, gatherA, gatherB, kscale); break;
    case 5: plaunch(family, bytes, flops, k_gemm_split<5>, g, blk, 0, s, tk, ts, nt, A, B, C, gatherA, gatherB, kscale); break;
    case 6: plaunch(family, bytes, flops, k_gemm_split<6>, g, blk, 0, s, tk, ts, nt, A, B, C, gatherA, gatherB, kscale); break;
    case 7: plaunch(family, bytes, flops, k_gemm_split<7>, g, blk, 0, s, tk, ts, nt, A, B, C, gatherA, gatherB, kscale); break;
    default: plaunch(family, bytes, flops, k_gemm_split<8>, g, blk, 0, s, tk, ts, nt, A, B, C, gatherA, gatherB, kscale); break;
  }
  NRG_HIP(hipGetLastError());
}

}  // namespace nrg
