// Split-precision batched GEMM: FP64 products from int8 matrix-core products
// (Ozaki-style error-free slicing: Ozaki, Ogita, Oishi & Rump, Numer.
// Algorithms 59 (2012) 95-118; the integer form of Ootomo, Ozaki & Yokota,
// Int. J. High Perform. Comput. Appl. 38 (2024)).
//
// Every row i of opA is scaled by 2^-Ea_i (Ea_i one above the binary
// exponent of the row's largest magnitude, so that the scaled row lies in
// (-1/2, 1/2)) and written in S signed base-2^7 digits,
//     opA(i,k) = 2^Ea_i * sum_{s<S} a_s(i,k) 2^{-7(s+1)}  (+ a remainder of
//                at most 2^{-7S-1}),
// |a_s| <= 64 (round-to-nearest digits); every column j of opB likewise with
// Eb_j.  The digits are stored once per product as int8 planes, padded to
// multiples of 64 in rows and K: plane s of opA is [rows][K] with K
// contiguous, plane t of opB is [columns][K].  The diagonal sums
//     P_d(i,j) = sum_{s+t=d} sum_k a_s(i,k) b_t(k,j),   d < S,
// run on v_mfma_i32_16x16x64_i8 with int32 accumulation, exact while
// (d+1) 64^2 K < 2^31: the accumulators are flushed to FP64 every 16384 K.
// Pairs with s + t >= S are dropped (the triangular truncation of the
// scheme).  The FP64 result is
//     C(i,j) = alpha 2^{Ea_i + Eb_j} sum_d 2^{-7(d+2)} P_d(i,j) + beta C(i,j).
// S digits carry 7S - 1 bits of each operand (S = 8: 55 bits).
//
// Per task: an exponent pass over the rows of opA and columns of opB
// (gathers and the k-scale applied as in the FP64 kernel), a slicing pass
// writing the planes, then the product over 64 x 64 output tiles (4 waves,
// a 32 x 32 quadrant each) with the tile table, lower-triangle option and
// task layout of the FP64 kernel (gemm.hip).  Batches are processed in
// chunks of tasks whose planes fit a bounded workspace.
#include "gemm.h"

namespace nrg {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int OZ_T = 64;           // tile edge and padding unit (rows and K)
constexpr int OZ_BETA = 7;         // bits per digit
constexpr int OZ_MAXS = 8;
constexpr int OZ_FLUSH = 256;      // 64-deep K chunks per int32 window (16384 K)
constexpr int64_t OZ_PLANE_BUDGET = int64_t(3) << 30;   // plane bytes per chunk of tasks

// one task of a split batch: padded sizes and workspace offsets
struct OzTask {
  int mp, np, kp, pad;
  int64_t pa, pb;   // byte offsets of plane 0 of opA / opB (plane s at + s * rows_p * kp)
  int64_t ea, eb;   // int32 offsets of the row / column exponents
};

// a unit of the exponent and slicing passes: 64 rows of opA (side 0) or 64
// columns of opB (side 1) of one task, and for slicing one 64-deep K block
struct OzUnit {
  int task, side, r0, k0;
};

struct OzOps {
  const GemmTask* tasks;
  const OzTask* oz;
  const double* A;
  const double* B;
  const int32_t* gA;
  const int32_t* gB;
  const double* ks;
  int8_t* planes;
  int32_t* ex;
};

// element (r, k) of op(A) (side 0, r a row) or of op(B)^T (side 1, r a
// column), zero outside the task
__device__ inline double oz_elem(const OzOps& o, const GemmTask& t, int side, int r, int k) {
  if (side == 0) {
    if (r >= t.m || k >= t.k) return 0.0;
    const int32_t* ga = (t.offGA >= 0) ? o.gA + t.offGA : nullptr;
    const double* Ab = o.A + t.offA;
    if (!t.transA) return Ab[(int64_t)(ga ? ga[r] : r) + (int64_t)k * t.lda];
    return Ab[(int64_t)(ga ? ga[k] : k) + (int64_t)r * t.lda];
  }
  if (r >= t.n || k >= t.k) return 0.0;
  const int32_t* gb = (t.offGB >= 0) ? o.gB + t.offGB : nullptr;
  const double* Bb = o.B + t.offB;
  double v;
  if (!t.transB) v = Bb[(int64_t)(gb ? gb[k] : k) + (int64_t)r * t.ldb];
  else v = Bb[(int64_t)(gb ? gb[r] : r) + (int64_t)k * t.ldb];
  return (t.offS >= 0) ? v * o.ks[t.offS + k] : v;
}

// whether k is the stored row index of the operand (consecutive k are
// adjacent, or gathered rows of one column): then a wave walks k of one
// row; otherwise the 64 lanes take 64 rows at one k
__device__ inline bool oz_k_fastest(const GemmTask& t, int side) {
  return side == 0 ? (t.transA != 0) : (t.transB == 0);
}

// exponents: Ea_r = 1 + binary exponent of max_k |op(r, k)| (0 for a zero row)
__global__ __launch_bounds__(256) void k_oz_exponents(OzOps o, const OzUnit* __restrict__ units) {
  const OzUnit u = units[blockIdx.x];
  const GemmTask t = o.tasks[u.task];
  const OzTask z = o.oz[u.task];
  __shared__ double s_red[4][OZ_T];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  if (oz_k_fastest(t, u.side)) {
    // wave w takes rows w, w+4, ...: lanes walk k
    for (int rr = w; rr < OZ_T; rr += 4) {
      double m = 0.0;
      for (int k = lane; k < t.k; k += 64) m = fmax(m, fabs(oz_elem(o, t, u.side, u.r0 + rr, k)));
      for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
      if (lane == 0) s_red[0][rr] = m;
    }
    __syncthreads();
  } else {
    double m = 0.0;
    for (int k = w; k < t.k; k += 4) m = fmax(m, fabs(oz_elem(o, t, u.side, u.r0 + lane, k)));
    s_red[w][lane] = m;
    __syncthreads();
    if (w == 0) s_red[0][lane] = fmax(fmax(s_red[0][lane], s_red[1][lane]), fmax(s_red[2][lane], s_red[3][lane]));
    __syncthreads();
  }
  if (tid < OZ_T) {
    const double m = s_red[0][tid];
    int e = 0;
    if (m > 0.0) (void)frexp(m, &e);   // m = f 2^e, f in [0.5, 1)
    o.ex[(u.side ? z.eb : z.ea) + u.r0 + tid] = (m > 0.0) ? e + 1 : 0;
  }
}

// digits of one 64 x 64 block (64 rows r0.. at K block k0..) into the planes
template <int S>
__global__ __launch_bounds__(256) void k_oz_slice(OzOps o, const OzUnit* __restrict__ units) {
  const OzUnit u = units[blockIdx.x];
  const GemmTask t = o.tasks[u.task];
  const OzTask z = o.oz[u.task];
  __shared__ __attribute__((aligned(16))) int8_t s_d[S][OZ_T][OZ_T];
  const int tid = threadIdx.x;
  const int32_t* ex = o.ex + (u.side ? z.eb : z.ea);
  const bool kf = oz_k_fastest(t, u.side);
#pragma unroll 4
  for (int q = 0; q < 16; ++q) {
    const int e = tid + 256 * q;
    const int rr = kf ? (e >> 6) : (e & 63);
    const int kk = kf ? (e & 63) : (e >> 6);
    double x = ldexp(oz_elem(o, t, u.side, u.r0 + rr, u.k0 + kk), -ex[u.r0 + rr]);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const double y = ldexp(x, OZ_BETA);
      const double dg = rint(y);
      x = y - dg;
      s_d[s][rr][kk] = (int8_t)dg;
    }
  }
  __syncthreads();
  // plane rows of 64 bytes, 16 bytes per thread
  const int rows_p = u.side ? z.np : z.mp;
  int8_t* base = o.planes + (u.side ? z.pb : z.pa);
  const int rr = tid >> 2, c = (tid & 3) * 16;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const v4i v = *reinterpret_cast<const v4i*>(&s_d[s][rr][c]);
    *reinterpret_cast<v4i*>(base + (int64_t)s * rows_p * z.kp + (int64_t)(u.r0 + rr) * z.kp + u.k0 + c) = v;
  }
}

// Product kernel: one workgroup of 4 waves per 64 x 64 output tile, each
// wave a 32 x 32 quadrant (2 x 2 tiles of v_mfma_i32_16x16x64_i8, one int32
// accumulator set per diagonal d < S).  The digit planes of a 64-deep K
// chunk (64 rows of every A plane, 64 columns of every B plane: 8S KB) are
// copied into LDS once per workgroup by direct global-to-LDS loads
// (global_load_lds_dwordx4: each wave instruction fills 1 KB, 16 plane rows
// of 64 bytes, in lane order), and the four waves read their fragments from
// LDS (16 rows x 64 bytes per ds_read_b128, contiguous, no bank conflicts):
// every fragment crosses L2 once per workgroup instead of once per wave.
// WIN: K spans more than one int32 window, so the diagonal sums are flushed
// to FP64 every OZ_FLUSH chunks; otherwise they are combined in FP64 once,
// in the epilogue (the same order of additions as a single flush).
typedef __attribute__((address_space(3))) void oz_lds_t;
typedef __attribute__((address_space(1))) void oz_glb_t;
constexpr int OZ_PL = OZ_T * OZ_T;   // bytes of one plane's 64 x 64 chunk

// Buffering by slice count (measured on the synthetic batch): with S <= 6
// one buffer per workgroup and two workgroups per CU (one computes while the
// other's chunk streams in) is faster; with S >= 7 the larger footprint
// favours one workgroup per CU with a double buffer.
template <int S>
constexpr int oz_nbuf() { return S <= 6 ? 1 : 2; }
template <int S>
constexpr size_t oz_lds_bytes() { return (size_t)oz_nbuf<S>() * 2 * S * OZ_PL; }

template <int S, bool WIN, int NBUF = oz_nbuf<S>()>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3 - NBUF, 3 - NBUF))) void k_oz_gemm_lds(
    const GemmTask* __restrict__ tasks, const OzTask* __restrict__ oz, const int32_t* __restrict__ tile_start,
    int ntasks, const int8_t* __restrict__ planes, const int32_t* __restrict__ ex, double* __restrict__ C) {
  extern __shared__ __attribute__((aligned(16))) int8_t s_pl[];   // [2 buffers][A planes, B planes][64][64]
  constexpr int BUF = 2 * S * OZ_PL;
  const int bid = blockIdx.x;
  int lo = 0, hi = ntasks - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tile_start[mid] <= bid) lo = mid; else hi = mid - 1;
  }
  const GemmTask t = tasks[lo];
  const OzTask z = oz[lo];
  const int tiles_m = z.mp / OZ_T;
  const int local = bid - tile_start[lo];
  int m0, n0;
  if (t.lower) {
    int ti = (int)((sqrt(8.0 * (double)local + 1.0) - 1.0) * 0.5);
    while ((int64_t)ti * (ti + 1) / 2 > local) --ti;
    while ((int64_t)(ti + 1) * (ti + 2) / 2 <= local) ++ti;
    m0 = ti * OZ_T;
    n0 = (local - ti * (ti + 1) / 2) * OZ_T;
  } else {
    m0 = (local % tiles_m) * OZ_T;
    n0 = (local / tiles_m) * OZ_T;
  }
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int li = lane & 15, kq = (lane >> 4) * 16;
  const int64_t pa = (int64_t)z.mp * z.kp, pb = (int64_t)z.np * z.kp;
  // staging of chunk c into buffer `buf`: wave w copies rows 16w..16w+15 of
  // every A and B plane (lane l: bytes 16 (l & 3).. of row 16w + (l >> 2)),
  // one 1 KB instruction per plane; byte offsets into the planes fit 32
  // bits (the plane workspace is bounded by OZ_PLANE_BUDGET)
  const int row = 16 * wave + (lane >> 2);
  const uint32_t oa = (uint32_t)(z.pa + (int64_t)(m0 + row) * z.kp) + 16 * (lane & 3);
  const uint32_t ob = (uint32_t)(z.pb + (int64_t)(n0 + row) * z.kp) + 16 * (lane & 3);
  auto stage = [&](int c, int buf) {
    const uint32_t ko = (uint32_t)c * OZ_T;
    int8_t* l0 = s_pl + buf * BUF + wave * 1024;
#pragma unroll
    for (int u = 0; u < 2 * S; ++u) {
      const int opnd = u / S, pl = u % S;
      const uint32_t o = (opnd ? ob + (uint32_t)(pl * pb) : oa + (uint32_t)(pl * pa)) + ko;
      __builtin_amdgcn_global_load_lds((oz_glb_t*)(planes + o), (oz_lds_t*)(l0 + u * OZ_PL), 16, 0, 0);
    }
  };

  v4i acc[S][2][2];
#pragma unroll
  for (int d = 0; d < S; ++d)
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[d][a][b] = v4i{0, 0, 0, 0};
  double acc64[WIN ? 2 : 1][2][4];
  if constexpr (WIN) {
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc64[a][b][r] = 0.0;
  }
  auto flush = [&]() {
#pragma unroll
    for (int d = 0; d < S; ++d) {
      const double w = ldexp(1.0, -OZ_BETA * (d + 2));
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
#pragma unroll
          for (int r = 0; r < 4; ++r) acc64[a][b][r] += w * (double)acc[d][a][b][r];
          acc[d][a][b] = v4i{0, 0, 0, 0};
        }
    }
  };

  const int nchunk = z.kp / OZ_T;
  // fragment offsets inside a buffer: A plane s, sub-tile mi; B plane t, sub-tile ni
  const int aoff = (wm * 32 + li) * OZ_T + kq;
  const int boff = S * OZ_PL + (wn * 32 + li) * OZ_T + kq;
  if constexpr (NBUF == 2) {
    stage(0, 0);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
  }
  for (int c = 0; c < nchunk; ++c) {
    int buf = 0;
    if constexpr (NBUF == 2) {
      buf = c & 1;
      if (c + 1 < nchunk) stage(c + 1, buf ^ 1);
    } else {
      stage(c, 0);
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
    }
    const int8_t* sb = s_pl + buf * BUF;
    v4i bv[S][2];
#pragma unroll
    for (int tt = 0; tt < S; ++tt)
#pragma unroll
      for (int ni = 0; ni < 2; ++ni)
        bv[tt][ni] = *reinterpret_cast<const v4i*>(sb + boff + tt * OZ_PL + ni * 16 * OZ_T);
#pragma unroll
    for (int sa = 0; sa < S; ++sa) {
      v4i av[2];
#pragma unroll
      for (int mi = 0; mi < 2; ++mi) av[mi] = *reinterpret_cast<const v4i*>(sb + aoff + sa * OZ_PL + mi * 16 * OZ_T);
#pragma unroll
      for (int tt = 0; sa + tt < S; ++tt)
#pragma unroll
        for (int mi = 0; mi < 2; ++mi)
#pragma unroll
          for (int ni = 0; ni < 2; ++ni)
            acc[sa + tt][mi][ni] =
                __builtin_amdgcn_mfma_i32_16x16x64_i8(av[mi], bv[tt][ni], acc[sa + tt][mi][ni], 0, 0, 0);
    }
    if constexpr (WIN) {
      if ((c + 1) % OZ_FLUSH == 0) flush();
    }
    // this wave's copies of chunk c+1 have landed; after the barrier every
    // wave's have, and no wave reads buffer `buf` any more
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
  }
  if constexpr (WIN) flush();

  // epilogue: the exponents of this lane's 8 rows and 2 columns (padded
  // rows and columns have them too) and, with beta != 0, the 16 old C
  // values (clamped addresses) are read before any is used
  double* Cb = C + t.offC;
  const int32_t* ea = ex + z.ea;
  const int32_t* eb = ex + z.eb;
  const int gi0 = m0 + wm * 32 + (lane >> 4) * 4, gj0 = n0 + wn * 32 + (lane & 15);
  int er[2][4], ec[2];
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int r = 0; r < 4; ++r) er[mi][r] = ea[gi0 + mi * 16 + r];
#pragma unroll
  for (int ni = 0; ni < 2; ++ni) ec[ni] = eb[gj0 + ni * 16];
  double cold[2][2][4] = {};
  if (t.beta != 0.0) {
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
      for (int ni = 0; ni < 2; ++ni)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int gi = min(gi0 + mi * 16 + r, t.m - 1), gj = min(gj0 + ni * 16, t.n - 1);
          cold[mi][ni][r] = Cb[(int64_t)gi + (int64_t)gj * t.ldc];
        }
  }
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < 2; ++ni)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        double sum;
        if constexpr (WIN) {
          sum = acc64[mi][ni][r];
        } else {
          sum = 0.0;
#pragma unroll
          for (int d = 0; d < S; ++d) sum += ldexp(1.0, -OZ_BETA * (d + 2)) * (double)acc[d][mi][ni][r];
        }
        const int gi = gi0 + mi * 16 + r, gj = gj0 + ni * 16;
        if (gi < t.m && gj < t.n) {
          double v = t.alpha * ldexp(sum, er[mi][r] + ec[ni]);
          if (t.beta != 0.0) v += t.beta * cold[mi][ni][r];
          Cb[(int64_t)gi + (int64_t)gj * t.ldc] = v;
        }
      }
}

template <int S>
static void oz_launch(int family, const std::vector<GemmTask>& tk, const std::vector<OzTask>& oz,
                      const std::vector<int32_t>& tstart, int64_t tiles, const std::vector<OzUnit>& eunits,
                      const std::vector<OzUnit>& sunits, OzOps o, GemmRunner& R, hipStream_t s, double flops,
                      double bytes, int64_t kmax) {
  // tables in one upload: tasks, oz records, tile starts, exponent units, slice units
  auto al = [](size_t n) { return (n + 255) / 256 * 256; };
  const size_t b0 = al(tk.size() * sizeof(GemmTask)), b1 = al(oz.size() * sizeof(OzTask)),
               b2 = al(tstart.size() * sizeof(int32_t)), b3 = al(eunits.size() * sizeof(OzUnit)),
               b4 = al(sunits.size() * sizeof(OzUnit));
  R.h_pack.assign(b0 + b1 + b2 + b3 + b4, 0);
  char* h = R.h_pack.data();
  std::memcpy(h, tk.data(), tk.size() * sizeof(GemmTask));
  std::memcpy(h + b0, oz.data(), oz.size() * sizeof(OzTask));
  std::memcpy(h + b0 + b1, tstart.data(), tstart.size() * sizeof(int32_t));
  std::memcpy(h + b0 + b1 + b2, eunits.data(), eunits.size() * sizeof(OzUnit));
  std::memcpy(h + b0 + b1 + b2 + b3, sunits.data(), sunits.size() * sizeof(OzUnit));
  R.d_pack[0].upload(R.h_pack.data(), R.h_pack.size(), s);
  const char* d = R.d_pack[0].p;
  o.tasks = reinterpret_cast<const GemmTask*>(d);
  o.oz = reinterpret_cast<const OzTask*>(d + b0);
  const int32_t* dts = reinterpret_cast<const int32_t*>(d + b0 + b1);
  const OzUnit* deu = reinterpret_cast<const OzUnit*>(d + b0 + b1 + b2);
  const OzUnit* dsu = reinterpret_cast<const OzUnit*>(d + b0 + b1 + b2 + b3);
  double sl_bytes = 0;
  for (const OzUnit& u : sunits) { (void)u; sl_bytes += 64.0 * 64.0 * (8.0 + S); }
  plaunch(KF_OTHER, 0, 0, k_oz_exponents, dim3((unsigned)eunits.size()), dim3(256), 0, s, o, deu);
  plaunch(KF_OTHER, sl_bytes, 0, k_oz_slice<S>, dim3((unsigned)sunits.size()), dim3(256), 0, s, o, dsu);
  const uint32_t shm = (uint32_t)oz_lds_bytes<S>();
  if (kmax > OZ_FLUSH * OZ_T)
    plaunch(family, bytes, flops, k_oz_gemm_lds<S, true>, dim3((unsigned)tiles), dim3(256), shm, s, o.tasks, o.oz,
            dts, (int)tk.size(), (const int8_t*)o.planes, (const int32_t*)o.ex, R.oz_C_);
  else
    plaunch(family, bytes, flops, k_oz_gemm_lds<S, false>, dim3((unsigned)tiles), dim3(256), shm, s, o.tasks, o.oz,
            dts, (int)tk.size(), (const int8_t*)o.planes, (const int32_t*)o.ex, R.oz_C_);
}

void GemmRunner::run_split(const GemmBatch& b, const double* A, const double* B, double* C,
                           const int32_t* gatherA, const int32_t* gatherB, const double* kscale,
                           int slices, hipStream_t s, int family) {
  const int nt = (int)b.tasks.size();
  if (nt == 0) return;
  NRG_REQUIRE(slices >= 3 && slices <= OZ_MAXS, "gemm: split slices must be 3..8");
  NRG_REQUIRE(!rec_, "gemm: split-precision products are not recorded");
  auto pad = [](int x) { return (int)(((int64_t)std::max(x, 1) + OZ_T - 1) / OZ_T * OZ_T); };
  int i0 = 0;
  while (i0 < nt) {
    // a chunk of tasks whose planes fit the budget (at least one task)
    std::vector<GemmTask> tk;
    std::vector<OzTask> oz;
    std::vector<int32_t> tstart;
    std::vector<OzUnit> eu, su;
    int64_t pbytes = 0, nex = 0, tiles = 0, kmax = 0;
    double flops = 0, bytes = 0;
    int i = i0;
    for (; i < nt; ++i) {
      const GemmTask& t = b.tasks[i];
      NRG_REQUIRE(t.m > 0 && t.n > 0 && t.k >= 0, "gemm: bad task shape");
      NRG_REQUIRE(!t.lower || t.m == t.n, "gemm: a lower-triangle task must be square");
      OzTask z{};
      z.mp = pad(t.m); z.np = pad(t.n); z.kp = pad(t.k);
      kmax = std::max<int64_t>(kmax, z.kp);
      const int64_t need = (int64_t)slices * ((int64_t)z.mp + z.np) * z.kp;
      if (i > i0 && pbytes + need > OZ_PLANE_BUDGET) break;
      z.pa = pbytes;
      z.pb = pbytes + (int64_t)slices * z.mp * z.kp;
      pbytes += need;
      z.ea = nex;
      z.eb = nex + z.mp;
      nex += z.mp + z.np;
      const int q = (int)tk.size();
      for (int r = 0; r < z.mp; r += OZ_T) eu.push_back(OzUnit{q, 0, r, 0});
      for (int r = 0; r < z.np; r += OZ_T) eu.push_back(OzUnit{q, 1, r, 0});
      for (int k0 = 0; k0 < z.kp; k0 += OZ_T) {
        for (int r = 0; r < z.mp; r += OZ_T) su.push_back(OzUnit{q, 0, r, k0});
        for (int r = 0; r < z.np; r += OZ_T) su.push_back(OzUnit{q, 1, r, k0});
      }
      const double half = t.lower ? 0.5 : 1.0;
      flops += half * 2.0 * t.m * t.n * t.k;   // FP64-equivalent algorithmic flops
      bytes += (double)slices * ((double)z.mp + z.np) * z.kp + half * 8.0 * (t.beta != 0.0 ? 2.0 : 1.0) * t.m * t.n;
      tstart.push_back((int32_t)tiles);
      const int64_t tm = z.mp / OZ_T;
      tiles += t.lower ? tm * (tm + 1) / 2 : tm * (z.np / OZ_T);
      tk.push_back(t);
      oz.push_back(z);
    }
    NRG_REQUIRE(tiles < (1LL << 31) && (int64_t)su.size() < (1LL << 31), "gemm: split batch too large");
    d_planes_.reserve((size_t)pbytes);
    d_exp_.reserve((size_t)std::max<int64_t>(nex, 1));
    OzOps o{nullptr, nullptr, A, B, gatherA, gatherB, kscale, d_planes_.p, d_exp_.p};
    oz_C_ = C;
    switch (slices) {
      case 3: oz_launch<3>(family, tk, oz, tstart, tiles, eu, su, o, *this, s, flops, bytes, kmax); break;
      case 4: oz_launch<4>(family, tk, oz, tstart, tiles, eu, su, o, *this, s, flops, bytes, kmax); break;
      case 5: oz_launch<5>(family, tk, oz, tstart, tiles, eu, su, o, *this, s, flops, bytes, kmax); break;
      case 6: oz_launch<6>(family, tk, oz, tstart, tiles, eu, su, o, *this, s, flops, bytes, kmax); break;
      case 7: oz_launch<7>(family, tk, oz, tstart, tiles, eu, su, o, *this, s, flops, bytes, kmax); break;
      default: oz_launch<8>(family, tk, oz, tstart, tiles, eu, su, o, *this, s, flops, bytes, kmax); break;
    }
    // the chunk's tables and planes are reused by the next chunk
    if (i < nt) NRG_HIP(hipStreamSynchronize(s));
    i0 = i;
  }
  NRG_HIP(hipGetLastError());
}

}  // namespace nrg
