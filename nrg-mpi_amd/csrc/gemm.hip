// Batched variable-size FP64 GEMM on v_mfma_f64_16x16x4f64 (gfx950).
//
// One workgroup (256 lanes = 4 wave64) computes a 64x64 tile of C; each wave
// owns a 32x32 quadrant held as 2x2 MFMA 16x16 accumulators.  The K loop
// stages 16-deep slices of opA (64x16) and opB (16x64) through LDS.  Tiles of
// all tasks are flattened into one grid, so blocks of very different sizes
// share one launch; the tile table holds the first tile of every task
// followed by the task of every tile, so a workgroup finds its task with one
// load whatever the number of tasks.
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
  const int lo = tile_start[ntasks + bid];   // task of this tile
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

// ---------------------------------------------------------------------------
// Fast path: the task table is searched in LDS, the K loop runs in 32-deep
// slices and the next slice's operands are fetched into registers while the
// matrix cores work on the current one.  G selects gathers and the k-scale
// along K (bit 0: opA = A^T with its stored rows gathered, bit 1: opB = B
// with its stored rows gathered, bit 2: opB scaled per k), the forms of the
// invariant-matrix, spectral and reduced-density contractions: K is the
// lane-contiguous index of those loads, so each lane needs one gather index
// (and one scale) per slice, fetched one slice ahead of the operands.
constexpr int TK2 = 32;

template <int TA, int TB, int G>
__global__ __launch_bounds__(256, 3) void k_gemm_f64_fast(
    const GemmTask* __restrict__ tasks, const int32_t* __restrict__ tile_start, int ntasks,
    const double* __restrict__ A, const double* __restrict__ B, double* __restrict__ C,
    const int32_t* __restrict__ gA, const int32_t* __restrict__ gB, const double* __restrict__ ks) {
  static_assert(!(G & 1) || TA == 1, "A is gathered along K only when stored K-major");
  static_assert(!(G & 6) || TB == 0, "B is gathered and scaled along K only when stored K-major");
  // A and B slices; after the K loop the same space holds the C tile
  // (64 columns of LDS_STRIDE rows) for the column-wise epilogue
  static_assert(2 * TK2 == TN, "the C tile fits in the A/B slices");
  __shared__ double ABs[2 * TK2 * LDS_STRIDE];
  double* As = ABs;
  double* Bs = ABs + TK2 * LDS_STRIDE;
  const int bid = blockIdx.x;
  const int ti = tile_start[ntasks + bid];   // task of this tile
  const GemmTask t = tasks[ti];
  const int tiles_m = (t.m + TM - 1) / TM;
  const int local = bid - tile_start[ti];
  int m0, n0;
  if (t.lower) {
    int tr = (int)((sqrt(8.0 * (double)local + 1.0) - 1.0) * 0.5);
    while ((int64_t)tr * (tr + 1) / 2 > local) --tr;
    while ((int64_t)(tr + 1) * (tr + 2) / 2 <= local) ++tr;
    m0 = tr * TM;
    n0 = (local - tr * (tr + 1) / 2) * TN;
  } else {
    m0 = (local % tiles_m) * TM;
    n0 = (local / tiles_m) * TN;
  }
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // uniform over the wave
  const int wm = wave & 1, wn = wave >> 1;
  const double* Ab = A + t.offA;
  const double* Bb = B + t.offB;

  // element e = tid + 256 u of the 64 x 32 slice of opA / 32 x 64 of opB
  // (index order chosen so a wave reads contiguous stored memory).  Each
  // load is a uniform row base (scalar registers) plus one per-lane 32-bit
  // byte offset shared by all 16 loads of the slice.
  double ra[8], rb[8];
  // lane part of the element indices: (contiguous index, strided index)
  const int a_c = !TA ? (tid & 63) : (tid & 31), a_s = !TA ? (tid >> 6) : (tid >> 5);
  const int b_c = TB ? (tid & 63) : (tid & 31), b_s = TB ? (tid >> 6) : (tid >> 5);
  const uint32_t a_off = (uint32_t)(a_s * t.lda + a_c) * 8u;
  const uint32_t b_off = (uint32_t)(b_s * t.ldb + b_c) * 8u;
  // strides of u along the strided index: 4 (64-wide contiguous) or 8
  constexpr int a_su = !TA ? 4 : 8, b_su = TB ? 4 : 8;
  // gathers and k-scale of this task (identity where the task has none):
  // stored row of slice element k0 + (tid & 31), and its scale, for the
  // slice being fetched (_c) and the one after it (_n)
  const int32_t* gat = ((G & 1) && t.offGA >= 0) ? gA + t.offGA : nullptr;
  const int32_t* gbt = ((G & 2) && t.offGB >= 0) ? gB + t.offGB : nullptr;
  const double* kst = ((G & 4) && t.offS >= 0) ? ks + t.offS : nullptr;
  int ga_c = 0, gb_c = 0, ga_n = 0, gb_n = 0;
  double ks_c = 1.0, ks_n = 1.0;
  auto load_idx = [&](int k0) {
    const int kk = min(k0 + (tid & 31), max(t.k - 1, 0));
    if (G & 1) ga_n = gat ? gat[kk] : kk;
    if (G & 2) gb_n = gbt ? gbt[kk] : kk;
    if (G & 4) ks_n = kst ? kst[kk] : 1.0;
  };
  auto fetch = [&](int k0) {
    if (G) { ga_c = ga_n; gb_c = gb_n; ks_c = ks_n; }
    // gathered operands: row bases without k0, the stored row in the lane offset
    const uint32_t a_lo = (G & 1) ? (uint32_t)(a_s * t.lda + ga_c) * 8u : a_off;
    const uint32_t b_lo = (G & 2) ? (uint32_t)(b_s * t.ldb + gb_c) * 8u : b_off;
    const int ka = (G & 1) ? 0 : k0, kb = (G & 2) ? 0 : k0;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      // opA: TA = 0 -> row i = m0 + a_c, k = k0 + a_s + 4u; TA = 1 -> k = k0 + a_c, row = m0 + a_s + 8u
      const int sa = a_s + a_su * u;
      const int gi = !TA ? m0 + a_c : m0 + sa, gk = !TA ? k0 + sa : k0 + a_c;
      const char* pa = (const char*)(Ab + (!TA ? (int64_t)(k0 + a_su * u) * t.lda + m0
                                               : (int64_t)(m0 + a_su * u) * t.lda + ka));
      double v = 0.0;
      if (gi < t.m && gk < t.k) v = *(const double*)(pa + a_lo);
      ra[u] = v;
      // opB: TB = 1 -> col j = n0 + b_c, k = k0 + b_s + 4u; TB = 0 -> k = k0 + b_c, col = n0 + b_s + 8u
      const int sb = b_s + b_su * u;
      const int gj = TB ? n0 + b_c : n0 + sb, gk2 = TB ? k0 + sb : k0 + b_c;
      const char* pb = (const char*)(Bb + (TB ? (int64_t)(k0 + b_su * u) * t.ldb + n0
                                              : (int64_t)(n0 + b_su * u) * t.ldb + kb));
      double w = 0.0;
      if (gj < t.n && gk2 < t.k) w = *(const double*)(pb + b_lo);
      rb[u] = w;
    }
    if (G) load_idx(k0 + TK2);   // the next slice's rows, in flight with these operands
  };
  auto store = [&]() {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = tid + 256 * u;
      int i, kk;
      if (!TA) { i = e & 63; kk = e >> 6; } else { kk = e & 31; i = e >> 5; }
      As[kk * LDS_STRIDE + i] = ra[u];
      int j, kb;
      if (!TB) { kb = e & 31; j = e >> 5; } else { j = e & 63; kb = e >> 6; }
      Bs[kb * LDS_STRIDE + j] = (G & 4) ? rb[u] * ks_c : rb[u];
    }
  };

  // C columns of this wave (16 wave .. 16 wave + 15), row = lane: fetched
  // into the slice registers while the last slice is multiplied
  const bool readc = (t.beta != 0.0);
  double* Cb = C + t.offC;
  const uint32_t c_off = (uint32_t)lane * 8u;   // row m0 + lane of a column
  auto c_col = [&](int u) { return (char*)(Cb + (int64_t)(n0 + 16 * wave + u) * t.ldc + m0); };
  auto fetch_c = [&]() {
    const int gi = m0 + lane;
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int gj = n0 + 16 * wave + u;
      double v = 0.0;
      if (gi < t.m && gj < t.n) v = *(const double*)(c_col(u) + c_off);
      (u < 8 ? ra[u] : rb[u - 8]) = v;
    }
  };

  d4 acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};

  if (t.k > 0) {
    if (G) load_idx(0);
    fetch(0);
    store();
    __syncthreads();
  } else if (readc) {
    fetch_c();
  }
  for (int k0 = 0; k0 < t.k; k0 += TK2) {
    const bool more = (k0 + TK2 < t.k);
    if (more) fetch(k0 + TK2);   // in flight while this slice is multiplied
    else if (readc) fetch_c();
#pragma unroll
    for (int s = 0; s < TK2 / 4; ++s) {
      const int kk = s * 4 + (lane >> 4);
      const int li = lane & 15;
      const double a0 = As[kk * LDS_STRIDE + wm * 32 + li];
      const double a1 = As[kk * LDS_STRIDE + wm * 32 + 16 + li];
      const double b0 = Bs[kk * LDS_STRIDE + wn * 32 + li];
      const double b1 = Bs[kk * LDS_STRIDE + wn * 32 + 16 + li];
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
    if (more) {
      __syncthreads();
      store();
      __syncthreads();
    }
  }
  // epilogue: the accumulators go to LDS in the matrix-core layout and
  // come back as whole 64-row columns, so that every C store (and the C
  // prefetch above) moves one contiguous column per wave instruction
  __syncthreads();   // every wave is done with the A/B slices
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < 2; ++ni)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int il = wm * 32 + mi * 16 + (lane >> 4) + 4 * r;
        const int jl = wn * 32 + ni * 16 + (lane & 15);
        ABs[jl * LDS_STRIDE + il] = acc[mi][ni][r];
      }
  __syncthreads();
  const int gi = m0 + lane;
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int jl = 16 * wave + u;
    const int gj = n0 + jl;
    if (gi < t.m && gj < t.n) {
      double v = t.alpha * ABs[jl * LDS_STRIDE + lane];
      if (readc) v += t.beta * (u < 8 ? ra[u] : rb[u - 8]);
      *(double*)(c_col(u) + c_off) = v;
    }
  }
}



// Fast-kernel instance for transposition code tv = 2 transA + transB and
// gather/scale bits g; nullptr where the fast kernel has no such form (a
// gather along M or N, a scale along a strided index): the general kernel
// runs those.
using FastGemm = void (*)(const GemmTask*, const int32_t*, int, const double*, const double*, double*,
                          const int32_t*, const int32_t*, const double*);
static FastGemm fast_kernel(int tv, int g) {
  switch (tv * 8 + g) {
    case 0: return k_gemm_f64_fast<0, 0, 0>;
    case 2: return k_gemm_f64_fast<0, 0, 2>;
    case 4: return k_gemm_f64_fast<0, 0, 4>;
    case 6: return k_gemm_f64_fast<0, 0, 6>;
    case 8: return k_gemm_f64_fast<0, 1, 0>;
    case 16: return k_gemm_f64_fast<1, 0, 0>;
    case 17: return k_gemm_f64_fast<1, 0, 1>;
    case 18: return k_gemm_f64_fast<1, 0, 2>;
    case 19: return k_gemm_f64_fast<1, 0, 3>;
    case 20: return k_gemm_f64_fast<1, 0, 4>;
    case 21: return k_gemm_f64_fast<1, 0, 5>;
    case 22: return k_gemm_f64_fast<1, 0, 6>;
    case 23: return k_gemm_f64_fast<1, 0, 7>;
    case 24: return k_gemm_f64_fast<1, 1, 0>;
    case 25: return k_gemm_f64_fast<1, 1, 1>;
    default: return nullptr;
  }
}

// Deterministic split-K reduction: C = alpha sum_c W_c + beta C, partials
// summed in chunk order.  Grid: (element chunks, records).
__global__ void k_splitk_reduce(const SplitRed* __restrict__ recs, const double* __restrict__ W, double* C) {
  const SplitRed r = recs[blockIdx.y];
  const int64_t mn = (int64_t)r.m * r.n;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < mn; e += (int64_t)gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int c = 0; c < r.parts; ++c) acc += W[r.wsoff + c * mn + e];
    const int i = (int)(e % r.m), j = (int)(e / r.m);
    double* cp = C + r.offC + i + (int64_t)j * r.ldc;
    double v = r.alpha * acc;
    if (r.beta != 0.0) v += r.beta * (*cp);
    *cp = v;
  }
}

int64_t GemmRunner::decompose(const GemmBatch& b) {
  plain_.clear();
  part_.clear();
  h_red_.clear();
  // A launch of fewer than 512 tiles leaves most of the 256 CUs idle while
  // each tile runs its whole K loop: then every task with K >= 256 is split
  // into parts of 128-255, and the unsplit rest joins the split launch.
  int64_t tot = 0;
  for (const GemmTask& t : b.tasks) {
    const int64_t tm = ceil_div(t.m, TM);
    tot += t.lower ? tm * (tm + 1) / 2 : tm * ceil_div(t.n, TN);
  }
  const bool thin = tot < 512;
  int64_t wsz = 0;
  auto split = [&](const GemmTask& t, int parts) {
    const int chunk = (int)(ceil_div(t.k, parts * TK2) * TK2);
    parts = ceil_div(t.k, chunk);
    SplitRed r;
    r.offC = t.offC; r.wsoff = wsz; r.m = t.m; r.n = t.n; r.ldc = t.ldc; r.parts = parts;
    r.alpha = t.alpha; r.beta = t.beta;
    h_red_.push_back(r);
    for (int c = 0; c < parts; ++c) {
      const int k0 = c * chunk, kc = std::min(chunk, t.k - k0);
      GemmTask u = t;
      u.k = kc;
      if (!t.transA) u.offA += (int64_t)k0 * t.lda;
      else if (t.offGA >= 0) u.offGA += k0;
      else u.offA += k0;
      if (t.transB) u.offB += (int64_t)k0 * t.ldb;
      else if (t.offGB >= 0) u.offGB += k0;
      else u.offB += k0;
      if (t.offS >= 0) u.offS += k0;
      u.offC = wsz + (int64_t)c * t.m * t.n;
      u.ldc = t.m;
      u.alpha = 1.0;
      u.beta = 0.0;
      part_.add(u);
    }
    wsz += (int64_t)parts * t.m * t.n;
  };
  for (const GemmTask& t : b.tasks) {
    const int64_t tiles = (int64_t)ceil_div(t.m, TM) * ceil_div(t.n, TN);
    int parts = 1;
    if (!t.lower && t.k >= 1024 && tiles <= 16) parts = std::min(16, t.k / 512);
    if (thin && !t.lower && t.k >= 256) parts = std::max(parts, std::min(32, t.k / 128));
    if (parts < 2) plain_.add(t);
    else split(t, parts);
  }
  if (thin && !part_.tasks.empty() && !plain_.tasks.empty()) {
    const GemmBatch rest = plain_;
    plain_.clear();
    for (const GemmTask& t : rest.tasks) {
      if (t.lower || t.k == 0) plain_.add(t);
      else split(t, 1);
    }
  }
  return wsz;
}

void GemmRunner::run(const GemmBatch& b, const double* A, const double* B, double* C,
                     const int32_t* gatherA, const int32_t* gatherB, const double* kscale,
                     hipStream_t s, int family) {
  const int nt = (int)b.tasks.size();
  if (nt == 0) return;
  const int sl = (family == KF_GEMM_INV) ? split
                 : (family == KF_GEMM_SPEC || family == KF_GEMM_CFSW || family == KF_GEMM_CFSX ||
                    family == KF_GEMM_RHO) ? split_spec : 0;
  if (sl > 0) {
    double fl = 0.0;
    if (split_min_flops > 0.0)
      for (const GemmTask& t : b.tasks) fl += (t.lower ? 1.0 : 2.0) * t.m * t.n * (double)t.k;
    if (fl >= split_min_flops) {
      run_split(b, A, B, C, gatherA, gatherB, kscale, sl, s, family);
      return;
    }
  }
  // Products with a long K and few output tiles (the back-transform's
  // V^T V and V^T Z) are split along K into partial products written to a
  // workspace and summed in a fixed order, so that they fill the device.
  const int64_t wsz = decompose(b);
  if (!plain_.tasks.empty()) launch(plain_, A, B, C, gatherA, gatherB, kscale, s, family, 0);
  if (!part_.tasks.empty()) {
    int64_t mx = 0;
    for (const SplitRed& r : h_red_) mx = std::max<int64_t>(mx, (int64_t)r.m * r.n);
    const unsigned gx = (unsigned)std::min<int64_t>(ceil_div(mx, 256), 256);
    if (rec_) {
      // the partial products go to the shared workspace (C = nullptr)
      r_ws_ = std::max(r_ws_, wsz);
      launch(part_, A, B, nullptr, gatherA, gatherB, kscale, s, family, 1);
      RLaunch r{};
      r.kind = 1; r.family = family; r.gx = gx; r.nt = (int)h_red_.size();
      r.tab = r_append(h_red_.data(), h_red_.size() * sizeof(SplitRed));
      r.C = C;
      r_l_.push_back(r);
      return;
    }
    d_splitw_.reserve(wsz);
    launch(part_, A, B, d_splitw_.p, gatherA, gatherB, kscale, s, family, 1);
    d_red_.upload(h_red_, s);
    plaunch(family, 0, 0, k_splitk_reduce, dim3(gx, (unsigned)h_red_.size()), dim3(256), 0, s,
            (const SplitRed*)d_red_.p, (const double*)d_splitw_.p, C);
  }
  NRG_HIP(hipGetLastError());
}

void GemmRunner::launch(const GemmBatch& b, const double* A, const double* B, double* C,
                        const int32_t* gatherA, const int32_t* gatherB, const double* kscale,
                        hipStream_t s, int family, int slot) {
  const int nt = (int)b.tasks.size();
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
  // after the tile starts, the task of every tile (a workgroup's lookup)
  h_tile_start.resize((size_t)nt + (size_t)total);
  for (int i = 0; i < nt; ++i) {
    const int64_t e = (i + 1 < nt) ? h_tile_start[i + 1] : total;
    std::fill(h_tile_start.begin() + nt + h_tile_start[i], h_tile_start.begin() + nt + e, (int32_t)i);
  }
  bool uniform = true;
  int gbits = 0;
  for (int i = 0; i < nt; ++i) {
    const GemmTask& t = b.tasks[i];
    uniform = uniform && t.transA == b.tasks[0].transA && t.transB == b.tasks[0].transB;
    if (gatherA && t.offGA >= 0) gbits |= 1;
    if (gatherB && t.offGB >= 0) gbits |= 2;
    if (kscale && t.offS >= 0) gbits |= 4;
  }
  const int tv = b.tasks[0].transA * 2 + b.tasks[0].transB;
  const FastGemm fk = uniform ? fast_kernel(tv, gbits) : nullptr;
  if (rec_) {
    RLaunch r{};
    r.kind = 0; r.fast = fk ? 1 : 0; r.tv = tv; r.g = gbits; r.family = family;
    r.grid = (unsigned)total; r.nt = nt;
    r.tab = r_append(b.tasks.data(), (size_t)nt * sizeof(GemmTask));
    r.tst = r_append(h_tile_start.data(), h_tile_start.size() * sizeof(int32_t));
    r.bytes = bytes; r.flops = flops;
    r.A = A; r.B = B; r.C = (slot == 1) ? nullptr : C;
    r.gA = gatherA; r.gB = gatherB; r.ks = kscale;
    r_l_.push_back(r);
    return;
  }
  const GemmTask* tk;
  const int32_t* ts;
  if (ws_) {
    // tables in the caller's workspace (run_into)
    const size_t tb = (size_t)nt * sizeof(GemmTask), sb = h_tile_start.size() * sizeof(int32_t);
    tk = reinterpret_cast<const GemmTask*>(ws_stage(b.tasks.data(), tb));
    ts = reinterpret_cast<const int32_t*>(ws_stage(h_tile_start.data(), sb));
  } else {
    upload_tables(b, s, tk, ts, slot);
  }
  const dim3 g((unsigned)total), blk(256);
  if (fk) {
    plaunch(family, bytes, flops, fk, g, blk, 0, s, tk, ts, nt, A, B, C, gatherA, gatherB, kscale);
  } else {
    plaunch(family, bytes, flops, k_gemm_f64, g, blk, 0, s, tk, ts, nt, A, B, C, gatherA, gatherB, kscale);
  }
}

char* GemmRunner::ws_stage(const void* p, size_t n) {
  const size_t o = (ws_used_ + 255) / 256 * 256;
  NRG_REQUIRE(o + n <= ws_cap_, "gemm: workspace too small for the task tables");
  // the copy has consumed the host table when this returns (the caller's
  // staging vectors are reused by the next launch); each table of one call
  // has its own offset, so no copy overwrites a table a queued kernel reads
  if (n) {
    NRG_HIP(hipMemcpyAsync(ws_ + o, p, n, hipMemcpyHostToDevice, ws_stream_));
    NRG_HIP(hipStreamSynchronize(ws_stream_));
  }
  ws_used_ = o + n;
  return ws_ + o;
}

size_t GemmRunner::workspace_bytes(const GemmBatch& b) {
  const int64_t wsz = decompose(b);
  auto al = [](size_t n) { return (n + 255) / 256 * 256; };
  size_t n = 0;
  for (const GemmBatch* x : {&plain_, &part_}) {
    if (x->tasks.empty()) continue;
    // task table, tile starts and the task of every tile
    size_t tiles = 0;
    for (const GemmTask& t : x->tasks) {
      const size_t tm = ceil_div(t.m, TM);
      tiles += t.lower ? tm * (tm + 1) / 2 : tm * ceil_div(t.n, TN);
    }
    n += al(x->tasks.size() * sizeof(GemmTask)) + al((x->tasks.size() + tiles) * sizeof(int32_t));
  }
  if (!part_.tasks.empty()) n += al(h_red_.size() * sizeof(SplitRed)) + al((size_t)wsz * sizeof(double));
  return n;
}

void GemmRunner::run_into(const GemmBatch& b, const double* A, const double* B, double* C, char* ws,
                          size_t ws_bytes, hipStream_t s, int family) {
  NRG_REQUIRE(!rec_, "gemm: run_into while recording");
  if (b.tasks.empty()) return;
  const size_t need = workspace_bytes(b);
  NRG_REQUIRE(need <= ws_bytes, "gemm: the products need " + std::to_string(need) +
                                    " bytes of workspace, more than the " + std::to_string(ws_bytes) + " given");
  const int64_t wsz = decompose(b);
  ws_ = ws; ws_cap_ = ws_bytes; ws_used_ = 0; ws_stream_ = s;
  try {
    if (!plain_.tasks.empty()) launch(plain_, A, B, C, nullptr, nullptr, nullptr, s, family, 0);
    if (!part_.tasks.empty()) {
      // partial products after the tables, then the fixed-order reduction
      double* W = nullptr;
      {
        const GemmBatch parts = part_;
        const std::vector<SplitRed> red = h_red_;
        const SplitRed* dred = reinterpret_cast<const SplitRed*>(ws_stage(red.data(), red.size() * sizeof(SplitRed)));
        const size_t o = (ws_used_ + 255) / 256 * 256;
        NRG_REQUIRE(o + (size_t)wsz * sizeof(double) <= ws_cap_, "gemm: workspace too small for split-K partials");
        W = reinterpret_cast<double*>(ws_ + o);
        ws_used_ = o + (size_t)wsz * sizeof(double);
        launch(parts, A, B, W, nullptr, nullptr, nullptr, s, family, 1);
        int64_t mx = 0;
        for (const SplitRed& r : red) mx = std::max<int64_t>(mx, (int64_t)r.m * r.n);
        const unsigned gx = (unsigned)std::min<int64_t>(ceil_div(mx, 256), 256);
        plaunch(family, 0, 0, k_splitk_reduce, dim3(gx, (unsigned)red.size()), dim3(256), 0, s, dred,
                (const double*)W, C);
      }
    }
  } catch (...) {
    ws_ = nullptr;
    throw;
  }
  ws_ = nullptr;
  NRG_HIP(hipGetLastError());
}

void GemmRunner::flush(hipStream_t s) {
  upload_recorded(s);
  issue(0, r_l_.size(), s);
  r_l_.clear();
}

void GemmRunner::upload_recorded(hipStream_t s) {
  rec_ = false;
  if (r_l_.empty()) return;
  if (r_ws_ > 0) d_splitw_.reserve(r_ws_);
  r_dev_.upload(r_host_.data(), r_host_.size(), s);
}

void GemmRunner::issue(size_t i0, size_t i1, hipStream_t s) {
  const char* base = r_dev_.p;
  const dim3 blk(256);
  for (size_t i = i0; i < i1 && i < r_l_.size(); ++i) {
    const RLaunch& r = r_l_[i];
    if (r.kind == 1) {
      plaunch(r.family, 0, 0, k_splitk_reduce, dim3(r.gx, (unsigned)r.nt), blk, 0, s,
              reinterpret_cast<const SplitRed*>(base + r.tab), (const double*)d_splitw_.p, r.C);
      continue;
    }
    const GemmTask* tk = reinterpret_cast<const GemmTask*>(base + r.tab);
    const int32_t* ts = reinterpret_cast<const int32_t*>(base + r.tst);
    double* C = r.C ? r.C : d_splitw_.p;
    const dim3 g(r.grid);
    if (r.fast) {
      plaunch(r.family, r.bytes, r.flops, fast_kernel(r.tv, r.g), g, blk, 0, s, tk, ts, r.nt, r.A, r.B, C, r.gA,
              r.gB, r.ks);
    } else {
      plaunch(r.family, r.bytes, r.flops, k_gemm_f64, g, blk, 0, s, tk, ts, r.nt, r.A, r.B, C, r.gA, r.gB, r.ks);
    }
  }
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
