// Two-stage tridiagonalisation of the batched eigensolver: dense -> band ->
// tridiagonal, with the back-transform of the requested eigenvector columns.
//
// Published methods followed (the code is written from the mathematics):
//  * successive band reduction, stage 1: blocked Householder QR of each
//    panel below the band and the two-sided compact-WY update
//    A22 <- Q^T A22 Q = A22 - V W^T - W V^T, W = A22 V T - 1/2 V T^T V^T A22 V T
//    (Bischof, Lang & Sun 2000; Schreiber & Van Loan 1989 for the WY form);
//  * stage 2: bulge chasing of the band to tridiagonal form with
//    Householder reflectors of the band width, one sweep per column
//    (Lang 1993; the task decomposition of Haidar, Ltaief & Dongarra 2011).
//
// Stage 1 moves the trailing matrix once per panel of NBS = 32 columns
// through matrix-core products (the one-stage reduction streams it once per
// column through a bandwidth-bound SYMV).  Stage 2 touches only the band.
//
// Stage 1, panel p of a block (columns c0 = 32 p .. c0 + 31, rows r0 = c0 +
// 32 .. n - 1, m = n - r0 >= 2 rows below the band):
//   * kb + 1 column launches (kb = min(32, m) reflectors), each over the
//     64-row tiles of every active block: launch jc completes reflector
//     jc - 1 from the partial sums the previous launch recorded (norm of the
//     column below its diagonal, its products with the other panel columns)
//     and the copy of row r0 + jc - 1 it made (ping-pong buffers: tile 0
//     rewrites that row during the launch that reads it),
//     applies it to the panel columns to its right, stores v in place below
//     the band, builds column jc - 1 of the compact-WY factor T, and records
//     the partial sums of the updated column jc (one launch per column: the
//     products are formed from the unscaled column, v^T x = x_0 + scal
//     sum_{r>0} x_r y_r, so they need not wait for the norm);
//   * Z = A22 V (lower storage: the 64 x 64 diagonal tiles are made symmetric
//     first, then per 64-row tile I the products A22[I, :I+1] V[:I+1] and
//     A22[I+1:, I]^T V[I+1:]), S = V^T Z, Q = -1/2 T^T S T (one workgroup per
//     block), W = [Z V] [T; Q], and the lower-triangle update
//     A22 -= [V W] [W V]^T, all on the FP64 matrix cores.
// Per block panel buffer P (n x 5 NBS, ld n): [V | W | V | Z | V], so that
// [Z V] and [V W], [W V] are contiguous operands.
//
// Stage 2 (one workgroup of 16 waves per block): the band (lower, 2 NBS
// rows per column to hold the bulges) is reduced by sweeps j = 0 .. n - 3.
// Task (j, k) works on rows Rk = [j + 1 + 32 k, +32): for k = 0 it
// annihilates column j below its subdiagonal; for k >= 1 it applies the
// previous reflector from the right to the bulge block B[Rk, R(k-1)],
// annihilates that block's first column, applies the new reflector from the
// left to the rest of it; then the two-sided update of the diagonal block
// B[Rk, Rk].  Task (j, k) may start once task (j - 1, k + 1) is done (the
// tasks of one wavefront 2j + k touch disjoint data).  Wave w runs the
// sweeps j = w mod 16 in order and waits on an LDS progress word of wave
// w - 1; the waves of one workgroup are co-resident, so every wait ends (a
// spin bound turns a stall into an error instead of a hang).
#include "eig.h"
#include <algorithm>
#include <cmath>

namespace nrg {

constexpr int NBS = 32;            // band width
constexpr int SBR = 64;            // rows per workgroup of the panel kernel
constexpr int SBREC = NBS + 1;     // record: norm, 32 column sums
constexpr int SBG = 6;             // thread groups that sum the records
constexpr int PCOL = 5 * NBS;      // panel buffer columns
constexpr int LB = 2 * NBS;        // band storage rows per column
constexpr int SB2_WAVES = 8;
constexpr long long SB2_SPIN = 1LL << 26;   // bound of one dependency wait

__device__ inline double sb_wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// panels of a block: p with m = n - 32 (p + 1) >= 2
__host__ __device__ inline int sb_panels(int n) { return n >= NBS + 2 ? (n - 2) / NBS : 0; }
// sum_{m=1}^{M} ceil(m / 32): stage-2 task offsets
__host__ __device__ inline int64_t sb_F(int64_t M) {
  if (M <= 0) return 0;
  const int64_t q = M / NBS, r = M % NBS;
  return NBS * q * (q + 1) / 2 + r * (q + 1);
}
// first task of sweep j of a block of n
__host__ __device__ inline int64_t sb_task0(int n, int j) { return sb_F(n - 1) - sb_F(n - 1 - j); }
__host__ __device__ inline int64_t sb_tasks(int n) { return n >= 3 ? sb_task0(n, n - 2) : 0; }

struct SbSlot {
  int32_t n, coff, b, ntile;   // ntile: tiles of the first panel (record stride)
  int64_t aoff, voff, pP, pRec, pTQ, pad;
};
static_assert(sizeof(SbSlot) == 64, "slot record is 64 bytes");

struct SbArgs {
  const SbSlot* slot;
  double* A;
  double* tau;
  double* P;
  double* rec;
  double* TQ;    // per block [T; Q] (64 x 32, ld 64) and S (32 x 32) behind it
};

template <class F>
__device__ inline bool sb_flat_map(int L, int act, F tiles, int& q_out, int& t_out) {
  const int lane = threadIdx.x & 63;
  int base = 0;
  for (int q0 = 0; q0 < act; q0 += 64) {
    const int q = q0 + lane;
    const int t = (q < act) ? tiles(q) : 0;
    int s = t;
    for (int o = 1; o < 64; o <<= 1) {
      const int u = __shfl_up(s, o, 64);
      if (lane >= o) s += u;
    }
    const unsigned long long hit = __ballot(base + s > L);
    if (hit) {
      const int f = __ffsll((unsigned long long)hit) - 1;
      const int sf = __shfl(s, f, 64), tf = __shfl(t, f, 64);
      q_out = q0 + f;
      t_out = L - (base + sf - tf);
      return true;
    }
    base += __shfl(s, 63, 64);
  }
  return false;
}

// ---------------------------------------------------------------------------
// stage 1: panel column kernel.  256 threads: lane = row of the 64-row tile,
// wave w holds panel columns 8w .. 8w+7 of that row in registers.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_sb_col(SbArgs a, int t, int jc, int act) {
  int qs, tile;
  auto tiles = [&](int q) {
    const SbSlot& s = a.slot[q];
    const int p = t - s.coff;
    if (p < 0 || p >= sb_panels(s.n)) return 0;
    const int m = s.n - (p + 1) * NBS;
    if (jc > min(NBS, m)) return 0;
    return (m + SBR - 1) / SBR;
  };
  if (!sb_flat_map((int)blockIdx.x, act, tiles, qs, tile)) return;
  const SbSlot sl = a.slot[qs];
  const int n = sl.n;
  const int p = t - sl.coff;
  const int c0 = p * NBS, r0 = c0 + NBS, m = n - r0, kb = min(NBS, m);
  const int64_t ld = n;
  double* A = a.A + sl.aoff;
  double* P = a.P + sl.pP;
  double* TQ = a.TQ + sl.pTQ;
  const int ntile = (m + SBR - 1) / SBR;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int r = r0 + tile * SBR + lane;
  const bool rv = r < n;
  __shared__ double s_sum[SBREC];
  __shared__ double s_grp[SBG][SBREC];
  __shared__ double s_row[NBS];
  __shared__ double s_x[SBR];
  __shared__ double s_red[4][8];
  __shared__ double s_v[SBR];

  double x[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) x[u] = rv ? A[r + (int64_t)(c0 + 8 * wv + u) * ld] : 0.0;

  if (jc >= 1) {
    const int j = jc - 1;
    // complete reflector j from the records of launch j (fixed tile order)
    // (SBG groups of 33 threads each sum every SBG-th tile, then the groups
    // are added in a fixed order: the loads of one thread are few and issued
    // together)
    const double* rc = a.rec + sl.pRec + (int64_t)(j & 1) * sl.ntile * SBREC;
    if (tid < SBG * SBREC) {
      const int g = tid / SBREC, comp = tid - g * SBREC;
      double s0 = 0.0, s1 = 0.0;
      int q = g;
      for (; q + SBG < ntile; q += 2 * SBG) {
        s0 += rc[(int64_t)q * SBREC + comp];
        s1 += rc[(int64_t)(q + SBG) * SBREC + comp];
      }
      if (q < ntile) s0 += rc[(int64_t)q * SBREC + comp];
      s_grp[g][comp] = s0 + s1;
    } else if (tid >= 224 && tid < 224 + NBS) {
      // row r0 + j of the panel as launch j left it, from the copy launch j
      // made (tile 0 rewrites the row itself during this launch)
      s_row[tid - 224] = a.rec[sl.pRec + 2 * (int64_t)sl.ntile * SBREC + (j & 1) * NBS + (tid - 224)];
    }
    __syncthreads();
    if (tid < SBREC) {
      double s = 0.0;
#pragma unroll
      for (int g = 0; g < SBG; ++g) s += s_grp[g][tid];
      s_sum[tid] = s;
    }
    __syncthreads();
    const double xn2 = s_sum[0];
    const double alpha = s_row[j];
    double beta, tau, scal;
    if (xn2 == 0.0) {
      beta = alpha; tau = 0.0; scal = 0.0;
    } else {
      beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
      tau = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
    // v_j of this row (1 at r0 + j, x_r scal below, 0 above); the column j
    // value of this row sits in wave j / 8, register j % 8
    if (wv == (j >> 3)) {
      double xr = 0.0;
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (u == (j & 7)) xr = x[u];
      const double v = (!rv || r < r0 + j) ? 0.0 : (r == r0 + j) ? 1.0 : xr * scal;
      s_v[lane] = v;
    }
    __syncthreads();
    const double vr = s_v[lane];
    // apply H_j = I - tau v v^T to the columns right of j; store v_j in place
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int col = 8 * wv + u;
      if (col > j) {
        const double wd = s_row[col] + scal * s_sum[1 + col];   // v_j^T x_col
        x[u] -= tau * vr * wd;
      } else if (col == j) {
        if (r == r0 + j) x[u] = beta;
        else if (r > r0 + j) x[u] = vr;
      }
    }
    if (tile == 0 && tid == 0) a.tau[sl.voff + c0 + j] = tau;
    if (tile == 0 && tid < NBS) {
      // column j of T: T(0:j, j) = -tau T(0:j, 0:j) G(0:j, j), T(j, j) = tau,
      // G(l, j) = v_l^T v_j = x_l(r0 + j) + scal sum_{r > r0+j} v_l(r) x_r
      const int i = tid;
      double acc = 0.0;
      if (i < j)
        for (int l = i; l < j; ++l) acc += TQ[i + l * 64] * (s_row[l] + scal * s_sum[1 + l]);
      if (i < j) TQ[i + j * 64] = -tau * acc;
      else if (i == j) TQ[i + j * 64] = tau;
      else TQ[i + j * 64] = 0.0;
    }
    if (rv) {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int col = 8 * wv + u;
        if (col >= j) A[r + (int64_t)(c0 + col) * ld] = x[u];
      }
    }
  }

  if (jc < kb) {
    // row r0 + jc as this launch leaves it, for the next launch
    if (tile == 0 && lane == jc) {
      double* rowc = a.rec + sl.pRec + 2 * (int64_t)sl.ntile * SBREC + (jc & 1) * NBS;
#pragma unroll
      for (int u = 0; u < 8; ++u) rowc[8 * wv + u] = x[u];
    }
    // partial sums of the updated column jc over this tile's rows r > r0 + jc
    if (wv == (jc >> 3)) {
      double xr = 0.0;
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (u == (jc & 7)) xr = x[u];
      s_x[lane] = (rv && r > r0 + jc) ? xr : 0.0;
    }
    __syncthreads();
    const double xr = s_x[lane];
    double* rc = a.rec + sl.pRec + (int64_t)(jc & 1) * sl.ntile * SBREC + (int64_t)tile * SBREC;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int col = 8 * wv + u;
      const double s = sb_wave_sum(col == jc ? xr * xr : xr * x[u]);
      if (lane == 0) s_red[wv][u] = s;
    }
    __syncthreads();
    if (tid < NBS) {
      const double s = s_red[tid >> 3][tid & 7];
      if (tid == jc) rc[0] = s;
      rc[1 + tid] = (tid == jc) ? 0.0 : s;
    }
  } else {
    // jc == kb: the panel's explicit V into P columns [0,32), [64,96), [128,160)
    // (unit at r0 + col, v below, zero above and for columns >= kb)
    if (rv) {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int col = 8 * wv + u;
        const double v = (col >= kb || r < r0 + col) ? 0.0 : (r == r0 + col) ? 1.0 : x[u];
        P[r + (int64_t)col * ld] = v;
        P[r + (int64_t)(2 * NBS + col) * ld] = v;
        P[r + (int64_t)(4 * NBS + col) * ld] = v;
      }
    }
    if (tile == 0 && tid >= kb && tid < NBS) {
      a.tau[sl.voff + c0 + tid] = 0.0;
      for (int i = 0; i < NBS; ++i) TQ[i + tid * 64] = 0.0;
    }
  }
}

// symmetric 64 x 64 diagonal tiles of the trailing matrix (lower -> upper),
// tiles aligned at r0: grid flattened over the active blocks' tiles
__global__ __launch_bounds__(256) void k_sb_symdiag(SbArgs a, int t, int act) {
  int qs, tile;
  auto tiles = [&](int q) {
    const SbSlot& s = a.slot[q];
    const int p = t - s.coff;
    if (p < 0 || p >= sb_panels(s.n)) return 0;
    return (s.n - (p + 1) * NBS + 63) / 64;
  };
  if (!sb_flat_map((int)blockIdx.x, act, tiles, qs, tile)) return;
  const SbSlot sl = a.slot[qs];
  const int n = sl.n;
  const int r0 = (t - sl.coff + 1) * NBS;
  double* A = a.A + sl.aoff;
  const int d0 = r0 + tile * 64;
  const int w = min(64, n - d0);
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int i = e & 63, j = e >> 6;   // upper element (i, j), i < j, from (j, i)
    if (i < j && j < w) A[(d0 + i) + (int64_t)(d0 + j) * n] = A[(d0 + j) + (int64_t)(d0 + i) * n];
  }
}

// Q = -1/2 T^T S T per active block (32 x 32), into rows 32..63 of [T; Q]
__global__ __launch_bounds__(256) void k_sb_q(SbArgs a, int t, int act) {
  const int q = blockIdx.x;
  if (q >= act) return;
  const SbSlot sl = a.slot[q];
  const int p = t - sl.coff;
  if (p < 0 || p >= sb_panels(sl.n)) return;
  double* TQ = a.TQ + sl.pTQ;
  const double* S = TQ + 64 * NBS;   // 32 x 32, ld 32
  __shared__ double sT[NBS][NBS + 1], sS[NBS][NBS + 1], sU[NBS][NBS + 1];
  for (int e = threadIdx.x; e < NBS * NBS; e += 256) {
    const int i = e % NBS, j = e / NBS;
    sT[i][j] = TQ[i + j * 64];
    sS[i][j] = S[i + j * NBS];
  }
  __syncthreads();
  // U = S T
  for (int e = threadIdx.x; e < NBS * NBS; e += 256) {
    const int i = e % NBS, j = e / NBS;
    double acc = 0.0;
    for (int l = 0; l <= j; ++l) acc += sS[i][l] * sT[l][j];
    sU[i][j] = acc;
  }
  __syncthreads();
  // Q = -1/2 T^T U
  for (int e = threadIdx.x; e < NBS * NBS; e += 256) {
    const int i = e % NBS, j = e / NBS;
    double acc = 0.0;
    for (int l = 0; l <= i; ++l) acc += sT[l][i] * sU[l][j];
    TQ[NBS + i + j * 64] = -0.5 * acc;
  }
}

// band of every block (lower, LB rows per column: the band's NBS + 1
// diagonals, zeros for the bulges) from A after stage 1
__global__ void k_sb_band(const int32_t* nn, const int64_t* aoff, const int64_t* boff, const double* A,
                          double* band) {
  const int b = blockIdx.y;
  const int n = nn[b];
  const double* Ab = A + aoff[b];
  double* Bb = band + boff[b];
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < (int64_t)n * LB;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(e / LB), d = (int)(e % LB);
    Bb[e] = (d <= NBS && c + d < n) ? Ab[(c + d) + (int64_t)c * n] : 0.0;
  }
}

// ---------------------------------------------------------------------------
// stage 2: bulge chasing, one workgroup (16 waves) per block
// ---------------------------------------------------------------------------
struct Sb2Args {
  const int32_t* n;
  const int64_t* boff;    // band offsets (n LB per block)
  const int64_t* v2off;   // reflector offsets (NBS per task)
  const int64_t* t2off;   // tau offsets (one per task)
  const int64_t* voff;
  double* band;
  double* v2;
  double* tau2;
  double* d;
  double* e;
  int32_t* err;
};

__global__ __launch_bounds__(SB2_WAVES * 64) void k_sb2st(Sb2Args a) {
  const int b = blockIdx.x;
  const int n = a.n[b];
  double* B = a.band + a.boff[b];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int c = lane & 31, h = lane >> 5;   // column-owner layout: rows 16h .. 16h+15 of column c
  __shared__ int s_prog[SB2_WAVES];
  __shared__ int s_abort;
  __shared__ double s_T[SB2_WAVES][NBS][NBS + 1];
  __shared__ double s_vec[SB2_WAVES][3][NBS];   // v, previous v, u / wv
  if (tid < SB2_WAVES) s_prog[tid] = -1;
  if (tid == 0) s_abort = 0;
  __syncthreads();
  double (*T)[NBS + 1] = s_T[wv];
  double* sv = s_vec[wv][0];
  double* svp = s_vec[wv][1];
  double* su = s_vec[wv][2];
  const int nsweep = n - 2;   // sweeps j = 0 .. n - 3
  for (int j = wv; j < nsweep; j += SB2_WAVES) {
    const int nst = (n - 1 - j + NBS - 1) / NBS;
    const int prevw = (wv + SB2_WAVES - 1) % SB2_WAVES;
    const int nprev = (j > 0) ? (n - j + NBS - 1) / NBS : 0;   // tasks of sweep j - 1
    double tau_p = 0.0;
    for (int k = 0; k < nst; ++k) {
      // wait: sweep j - 1 has done task k + 1 (or all of its tasks)
      if (j > 0) {
        const int need = min(k + 2, nprev);
        long long spin = 0;
        while (true) {
          const int pv = __hip_atomic_load(&s_prog[prevw], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          if ((pv >> 10) > j - 1 || ((pv >> 10) == j - 1 && (pv & 1023) >= need)) break;
          if (__hip_atomic_load(&s_abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
          if (++spin > SB2_SPIN) {
            __hip_atomic_store(&s_abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (lane == 0) atomicExch(a.err, 1);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
        // agent-scope acquire: invalidates this CU's L1, so that the band
        // lines another wave of the workgroup stored are not served from a
        // copy this wave cached before
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      }
      if (__hip_atomic_load(&s_abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
      const int r0 = j + 1 + k * NBS;
      const int L = min(NBS, n - r0);
      const int64_t to = a.t2off[b] + sb_task0(n, j) + k;
      double* vout = a.v2 + a.v2off[b] + (sb_task0(n, j) + k) * NBS;
      // loads: X = B[Rk, R(k-1)] (k >= 1) or column j (k = 0), and the
      // diagonal block D = B[Rk, Rk] (lower; the upper half mirrored)
      double x[16], dd[16];
      const int p0 = r0 - NBS;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int i = 16 * h + q;
        x[q] = 0.0;
        if (i < L) {
          if (k >= 1) x[q] = B[(int64_t)(p0 + c) * LB + (NBS + i - c)];
          else if (c == 0) x[q] = B[(int64_t)j * LB + 1 + i];
        }
        dd[q] = 0.0;
        if (i < L && c < L) dd[q] = (i >= c) ? B[(int64_t)(r0 + c) * LB + (i - c)] : B[(int64_t)(r0 + i) * LB + (c - i)];
      }
      if (k >= 1) {
        // right-apply H_{k-1} = I - tau_p vp vp^T: X[i, :] -= tau_p (X[i,:] . vp) vp
        const double vpc = svp[c];
#pragma unroll
        for (int q = 0; q < 16; ++q) T[16 * h + q][c] = x[q] * vpc;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        {
          // lane (row c, half h) sums 16 columns of row c
          double s = 0.0;
#pragma unroll
          for (int q = 0; q < 16; ++q) s += T[c][16 * h + q];
          s += __shfl_xor(s, 32, 64);
          if (h == 0) su[c] = s;
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
#pragma unroll
        for (int q = 0; q < 16; ++q) x[q] -= tau_p * su[16 * h + q] * vpc;
      }
      // Householder of column 0 of X (rows 0 .. L-1), held by lanes 0 and 32
      double n2 = 0.0;
      if (c == 0) {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          const int i = 16 * h + q;
          if (i >= 1 && i < L) n2 += x[q] * x[q];
        }
      }
      const double xn2 = __shfl(n2, 0, 64) + __shfl(n2, 32, 64);
      const double alpha = __shfl(x[0], 0, 64);
      double beta, tau, scal;
      if (xn2 == 0.0) {
        beta = alpha; tau = 0.0; scal = 0.0;
      } else {
        beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
        tau = (beta - alpha) / beta;
        scal = 1.0 / (alpha - beta);
      }
      if (c == 0) {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          const int i = 16 * h + q;
          sv[i] = (i >= L) ? 0.0 : (i == 0) ? 1.0 : x[q] * scal;
        }
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      double vr[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) vr[q] = sv[16 * h + q];
      const double vc = sv[c];
      if (k >= 1) {
        // left-apply to columns 1..31: X[:, c] -= tau v (v . X[:, c])
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < 16; ++q) s += vr[q] * x[q];
        s += __shfl_xor(s, 32, 64);
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          const int i = 16 * h + q;
          if (c == 0) x[q] = (i == 0) ? beta : 0.0;
          else x[q] -= tau * vr[q] * s;
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          const int i = 16 * h + q;
          if (i < L) B[(int64_t)(p0 + c) * LB + (NBS + i - c)] = x[q];
        }
      } else if (c == 0) {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          const int i = 16 * h + q;
          if (i < L) B[(int64_t)j * LB + 1 + i] = (i == 0) ? beta : 0.0;
        }
      }
      // two-sided update of D: y = D v, w = tau y - 1/2 tau^2 (v.y) v,
      // D -= v w^T + w v^T
      double y = 0.0;
#pragma unroll
      for (int q = 0; q < 16; ++q) y += dd[q] * vr[q];
      y += __shfl_xor(y, 32, 64);                       // y_c
      const double vy = sb_wave_sum((h == 0) ? vc * y : 0.0);
      const double wc = tau * y - 0.5 * tau * tau * vy * vc;
      if (h == 0) su[c] = wc;
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int i = 16 * h + q;
        if (i < L && c < L && i >= c) {
          const double nv = dd[q] - vr[q] * wc - su[i] * vc;
          B[(int64_t)(r0 + c) * LB + (i - c)] = nv;
        }
      }
      // reflector out (back-transform) and kept for the next task's right-apply
      if (h == 0) vout[c] = vc;
      if (lane == 0) a.tau2[to] = tau;
      if (h == 0) svp[c] = vc;
      tau_p = tau;
      // the task's band stores must have completed before its progress word
      // is published: the workgroup-scope release orders the LDS store after
      // earlier LDS operations only, so the store counter is drained first
      __builtin_amdgcn_s_waitcnt(0);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      if (lane == 0)
        __hip_atomic_store(&s_prog[wv], (j << 10) | (k + 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      __builtin_amdgcn_wave_barrier();
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    if (lane == 0)
      __hip_atomic_store(&s_prog[wv], (j << 10) | 1023, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __syncthreads();
  // tridiagonal out
  for (int cc = tid; cc < n; cc += SB2_WAVES * 64) {
    a.d[a.voff[b] + cc] = B[(int64_t)cc * LB];
    a.e[a.voff[b] + cc] = (cc + 1 < n) ? B[(int64_t)cc * LB + 1] : 0.0;
  }
}

// ---------------------------------------------------------------------------
// stage-2 back-transform: Z[:, cols] <- Q2 Z[:, cols] (sweeps in reverse
// order; the tasks of one sweep touch disjoint rows).  One workgroup per
// (block, chunk of nc columns); the chunk sits in LDS.
// ---------------------------------------------------------------------------
struct Sb2Back {
  const int32_t* n;
  const int64_t* aoff;
  const int64_t* v2off;
  const int64_t* t2off;
  const double* v2;
  const double* tau2;
  double* Z;
  const int32_t* chunks;   // per chunk: block, first column, columns
};

constexpr int SB2B_LDS = 152 * 1024;

constexpr int SB2B_K = 4;   // reflectors of one sweep per half-wave and round
__global__ __launch_bounds__(1024) void k_sb2st_back(Sb2Back a) {
  extern __shared__ double zs[];
  const int b = a.chunks[3 * blockIdx.x], col0 = a.chunks[3 * blockIdx.x + 1], nc = a.chunks[3 * blockIdx.x + 2];
  const int n = a.n[b];
  double* Z = a.Z + a.aoff[b] + (int64_t)col0 * n;
  const int tid = threadIdx.x;
  for (int e = tid; e < n * nc; e += 1024) zs[e] = Z[(e % nc) * (int64_t)n + e / nc];   // zs[row * nc + col]
  __syncthreads();
  // half-wave h applies the reflectors k = h + 32 i of each sweep to all nc
  // columns, SB2B_K of them per round; the first round of the next sweep is
  // loaded while the current sweep is applied (for n <= 4097 a sweep has
  // one round)
  const int half = tid >> 5, hl = tid & 31;
  const double* v2 = a.v2 + a.v2off[b];
  const double* t2 = a.tau2 + a.t2off[b];
  auto load = [&](int j, int round, double (&vv)[SB2B_K], double (&tt)[SB2B_K]) {
    const int nst = (n - 1 - j + NBS - 1) / NBS;
    const int64_t t0 = sb_task0(n, j);
#pragma unroll
    for (int i = 0; i < SB2B_K; ++i) {
      const int k = half + 32 * (SB2B_K * round + i);
      const int r0 = j + 1 + k * NBS;
      vv[i] = (k < nst && hl < n - r0) ? v2[(t0 + k) * NBS + hl] : 0.0;
      tt[i] = (k < nst) ? t2[t0 + k] : 0.0;
    }
  };
  double vn[SB2B_K], tn[SB2B_K];
  if (n >= 3) load(n - 3, 0, vn, tn);
  for (int j = n - 3; j >= 0; --j) {
    const int nst = (n - 1 - j + NBS - 1) / NBS;
    const int rounds = (nst + 32 * SB2B_K - 1) / (32 * SB2B_K);
    double vc[SB2B_K], tc[SB2B_K];
#pragma unroll
    for (int i = 0; i < SB2B_K; ++i) { vc[i] = vn[i]; tc[i] = tn[i]; }
    if (j > 0) load(j - 1, 0, vn, tn);
    for (int round = 0; round < rounds; ++round) {
      if (round > 0) load(j, round, vc, tc);
#pragma unroll
      for (int i = 0; i < SB2B_K; ++i) {
        const int k = half + 32 * (SB2B_K * round + i);
        if (k >= nst) break;
        const int r0 = j + 1 + k * NBS;
        const bool in = hl < n - r0;
        for (int cc = 0; cc < nc; ++cc) {
          const double z = in ? zs[(r0 + hl) * nc + cc] : 0.0;
          double sdot = vc[i] * z;
          for (int o = 16; o > 0; o >>= 1) sdot += __shfl_xor(sdot, o, 32);
          if (in) zs[(r0 + hl) * nc + cc] = z - tc[i] * sdot * vc[i];
        }
      }
    }
    __syncthreads();
  }
  for (int e = tid; e < n * nc; e += 1024) Z[(e % nc) * (int64_t)n + e / nc] = zs[e];
}

// ---------------------------------------------------------------------------
// host: stage 1 and stage 2 for all blocks
// ---------------------------------------------------------------------------
void BatchedEigensolver::tridiag_two_stage(hipStream_t s) {
  const int nblk = nblk_;
  // per block offsets: panel buffer, records, [T;Q|S], band, stage-2 reflectors
  std::vector<int64_t> poff(nblk), roff(nblk), tqoff(nblk), boff(nblk), v2off(nblk), t2off(nblk);
  int64_t racc = 0, bacc = 0, vacc = 0, tacc = 0;
  int Pmax = 0;
  for (int b = 0; b < nblk; ++b) {
    const int n = n_[b];
    poff[b] = voff_[b] * PCOL;
    roff[b] = racc;
    racc += 2 * (int64_t)((std::max(n - NBS, 1) + SBR - 1) / SBR) * SBREC + 2 * NBS;   // + two row copies
    tqoff[b] = (int64_t)b * (64 * NBS + NBS * NBS);
    boff[b] = bacc;
    bacc += (int64_t)n * LB;
    v2off[b] = vacc;
    t2off[b] = tacc;
    vacc += sb_tasks(n) * NBS;
    tacc += sb_tasks(n);
    Pmax = std::max(Pmax, sb_panels(n));
  }
  d_P_.reserve(tot_n_ * PCOL);
  d_part_.reserve(std::max<int64_t>(racc, 1));
  d_tq_.reserve((int64_t)nblk * (64 * NBS + NBS * NBS));
  d_band_.reserve(std::max<int64_t>(bacc, 1));
  d_v2_.reserve(std::max<int64_t>(vacc, 1));
  d_t2_.reserve(std::max<int64_t>(tacc, 1));
  h_sboff_.resize(3 * (size_t)nblk);
  for (int b = 0; b < nblk; ++b) {
    h_sboff_[b] = boff[b];
    h_sboff_[nblk + b] = v2off[b];
    h_sboff_[2 * nblk + b] = t2off[b];
  }
  d_sboff_.upload(h_sboff_.data(), h_sboff_.size(), s);

  // end-aligned panel schedule: block q starts at step Pmax - P_q
  const int ng = std::max(1, std::min(groups_, std::min(nblk, MAX_TR_GROUPS)));
  for (int g = 1; g < ng; ++g)
    if (!sg_[g]) {
      NRG_HIP(hipStreamCreateWithFlags(&sg_[g], hipStreamNonBlocking));
      NRG_HIP(hipEventCreateWithFlags(&evg_[g], hipEventDisableTiming));
    }
  if (ng > 1 && !ev_fork_) NRG_HIP(hipEventCreateWithFlags(&ev_fork_, hipEventDisableTiming));
  struct Group {
    std::vector<int> q;          // slots (size order)
    std::vector<int> coff;
    SbArgs sa;
    hipStream_t s;
    GemmRunner* gemm;
    GemmBatch* batch;
    std::vector<size_t> rec;     // recorded launch ranges: 5 per step
  };
  Group grp[MAX_TR_GROUPS];
  for (int g = 0; g < ng; ++g) {
    Group& G = grp[g];
    std::vector<SbSlot> slots;
    for (int q = g; q < nblk; q += ng) {
      const int b = order_[q];
      G.q.push_back(q);
      SbSlot r{};
      r.n = n_[b];
      r.coff = align_end_ ? Pmax - sb_panels(n_[b]) : 0;
      r.b = b;
      r.ntile = (std::max(n_[b] - NBS, 1) + SBR - 1) / SBR;
      r.aoff = aoff_[b]; r.voff = voff_[b]; r.pP = poff[b]; r.pRec = roff[b]; r.pTQ = tqoff[b]; r.pad = 0;
      G.coff.push_back(r.coff);
      slots.push_back(r);
    }
    G.s = g ? sg_[g] : s;
    G.gemm = &gemmg_[g];
    G.batch = &batchg_[g];
    d_slotg_[g].upload(reinterpret_cast<const char*>(slots.data()), slots.size() * sizeof(SbSlot), s);
    G.sa = SbArgs{reinterpret_cast<const SbSlot*>(d_slotg_[g].p), A_, d_tau_.p, d_P_.p, d_part_.p, d_tq_.p};
  }
  // the GEMMs of every step are recorded up front (their shapes depend only
  // on the schedule) and uploaded once per group
  auto step_of = [&](Group& G, int i, int t) { return t - G.coff[i]; };
  for (int g = 0; g < ng; ++g) {
    Group& G = grp[g];
    G.gemm->begin();
    G.rec.assign(1, 0);
    for (int t = 0; t < Pmax; ++t) {
      // 1. Z[I] = A22[I, :I+1] V[:I+1]
      G.batch->clear();
      for (size_t i = 0; i < G.q.size(); ++i) {
        const int b = order_[G.q[i]], n = n_[b], p = step_of(G, (int)i, t);
        if (p < 0 || p >= sb_panels(n)) continue;
        const int r0 = (p + 1) * NBS, m = n - r0;
        for (int I = 0; I * 64 < m; ++I) {
          const int mi = std::min(64, m - 64 * I), ki = std::min(64 * (I + 1), m);
          G.batch->add(gemm_task(mi, NBS, ki, 0, 0, aoff_[b] + (r0 + 64 * I) + (int64_t)r0 * n, n, poff[b] + r0, n,
                                 poff[b] + (r0 + 64 * I) + (int64_t)(3 * NBS) * n, n, 1.0, 0.0));
        }
      }
      G.gemm->run(*G.batch, A_, d_P_.p, d_P_.p, nullptr, nullptr, nullptr, G.s, KF_GEMM_TRIDIAG);
      G.rec.push_back(G.gemm->recorded());
      // 2. Z[I] += A22[I+1:, I]^T V[I+1:]
      G.batch->clear();
      for (size_t i = 0; i < G.q.size(); ++i) {
        const int b = order_[G.q[i]], n = n_[b], p = step_of(G, (int)i, t);
        if (p < 0 || p >= sb_panels(n)) continue;
        const int r0 = (p + 1) * NBS, m = n - r0;
        for (int I = 0; 64 * (I + 1) < m; ++I) {
          const int mi = 64, kk = m - 64 * (I + 1);
          G.batch->add(gemm_task(mi, NBS, kk, 1, 0, aoff_[b] + (r0 + 64 * (I + 1)) + (int64_t)(r0 + 64 * I) * n, n,
                                 poff[b] + r0 + 64 * (I + 1), n, poff[b] + (r0 + 64 * I) + (int64_t)(3 * NBS) * n, n,
                                 1.0, 1.0));
        }
      }
      G.gemm->run(*G.batch, A_, d_P_.p, d_P_.p, nullptr, nullptr, nullptr, G.s, KF_GEMM_TRIDIAG);
      G.rec.push_back(G.gemm->recorded());
      // 3. S = V^T Z
      G.batch->clear();
      for (size_t i = 0; i < G.q.size(); ++i) {
        const int b = order_[G.q[i]], n = n_[b], p = step_of(G, (int)i, t);
        if (p < 0 || p >= sb_panels(n)) continue;
        const int r0 = (p + 1) * NBS, m = n - r0;
        G.batch->add(gemm_task(NBS, NBS, m, 1, 0, poff[b] + r0, n, poff[b] + r0 + (int64_t)(3 * NBS) * n, n,
                               tqoff[b] + 64 * NBS, NBS, 1.0, 0.0));
      }
      G.gemm->run(*G.batch, d_P_.p, d_P_.p, d_tq_.p, nullptr, nullptr, nullptr, G.s, KF_GEMM_TRIDIAG);
      G.rec.push_back(G.gemm->recorded());
      // 4. W = [Z V] [T; Q]
      G.batch->clear();
      for (size_t i = 0; i < G.q.size(); ++i) {
        const int b = order_[G.q[i]], n = n_[b], p = step_of(G, (int)i, t);
        if (p < 0 || p >= sb_panels(n)) continue;
        const int r0 = (p + 1) * NBS, m = n - r0;
        G.batch->add(gemm_task(m, NBS, 2 * NBS, 0, 0, poff[b] + r0 + (int64_t)(3 * NBS) * n, n, tqoff[b], 64,
                               poff[b] + r0 + (int64_t)NBS * n, n, 1.0, 0.0));
      }
      G.gemm->run(*G.batch, d_P_.p, d_tq_.p, d_P_.p, nullptr, nullptr, nullptr, G.s, KF_GEMM_TRIDIAG);
      G.rec.push_back(G.gemm->recorded());
      // 5. A22 -= [V W] [W V]^T, lower tiles
      G.batch->clear();
      for (size_t i = 0; i < G.q.size(); ++i) {
        const int b = order_[G.q[i]], n = n_[b], p = step_of(G, (int)i, t);
        if (p < 0 || p >= sb_panels(n)) continue;
        const int r0 = (p + 1) * NBS, m = n - r0;
        GemmTask tk = gemm_task(m, m, 2 * NBS, 0, 1, poff[b] + r0, n, poff[b] + r0 + (int64_t)NBS * n, n,
                                aoff_[b] + r0 + (int64_t)r0 * n, n, -1.0, 1.0);
        tk.lower = 1;
        G.batch->add(tk);
      }
      G.gemm->run(*G.batch, d_P_.p, d_P_.p, A_, nullptr, nullptr, nullptr, G.s, KF_GEMM_TRIDIAG);
      G.rec.push_back(G.gemm->recorded());
    }
    G.gemm->upload_recorded(s);
  }
  if (ng > 1) {
    NRG_HIP(hipEventRecord(ev_fork_, s));
    for (int g = 1; g < ng; ++g) NRG_HIP(hipStreamWaitEvent(sg_[g], ev_fork_, 0));
  }
  for (int t = 0; t < Pmax; ++t)
    for (int g = 0; g < ng; ++g) {
      Group& G = grp[g];
      // active slots: coff <= t (slots are in size order: coff non-decreasing)
      int act = 0;
      while (act < (int)G.q.size() && G.coff[act] <= t) ++act;
      int tiles = 0, kbmax = 0, dtiles = 0;
      double cbytes = 0;
      for (int i = 0; i < act; ++i) {
        const int n = n_[order_[G.q[i]]], p = t - G.coff[i];
        if (p >= sb_panels(n)) continue;
        const int m = n - (p + 1) * NBS;
        tiles += ceil_div(m, SBR);
        dtiles += ceil_div(m, 64);
        kbmax = std::max(kbmax, std::min(NBS, m));
        cbytes += 16.0 * NBS * m;
      }
      if (tiles == 0) continue;
      for (int jc = 0; jc <= kbmax; ++jc)
        plaunch(KF_TR_COL, cbytes, 0, k_sb_col, dim3(tiles), dim3(256), 0, G.s, G.sa, t, jc, act);
      plaunch(KF_TR_COL, 0, 0, k_sb_symdiag, dim3(dtiles), dim3(256), 0, G.s, G.sa, t, act);
      const size_t* R = G.rec.data() + 5 * (size_t)t;
      G.gemm->issue(R[0], R[1], G.s);
      G.gemm->issue(R[1], R[2], G.s);
      G.gemm->issue(R[2], R[3], G.s);
      plaunch(KF_TR_COL, 0, 0, k_sb_q, dim3(act), dim3(256), 0, G.s, G.sa, t, act);
      G.gemm->issue(R[3], R[4], G.s);
      G.gemm->issue(R[4], R[5], G.s);
    }
  for (int g = 0; g < ng; ++g) grp[g].gemm->end_replay();
  // stage 2 per group, right after its stage 1
  int32_t* err = d_err_.p;
  if (!err) {
    d_err_.reserve(1);
    err = d_err_.p;
  }
  fill(err, 0, 1, s);
  for (int g = 1; g < ng; ++g) {
    NRG_HIP(hipEventRecord(evg_[g], sg_[g]));
    NRG_HIP(hipStreamWaitEvent(s, evg_[g], 0));
  }
  {
    int nmax = n_[order_[0]];
    hipLaunchKernelGGL(k_sb_band, dim3((unsigned)std::max(1, std::min(4096, ceil_div((int64_t)nmax * LB, 256))), nblk),
                       dim3(256), 0, s, (const int32_t*)d_n_.p, (const int64_t*)d_aoff_.p, (const int64_t*)d_sboff_.p,
                       (const double*)A_, d_band_.p);
    NRG_HIP(hipGetLastError());
  }
  Sb2Args a2{d_n_.p, d_sboff_.p, d_sboff_.p + nblk, d_sboff_.p + 2 * nblk, d_voff_.p, d_band_.p, d_v2_.p, d_t2_.p,
             d_d_.p, d_e_.p, err};
  double f2 = 0;
  for (int b = 0; b < nblk; ++b) f2 += 6.0 * (double)NBS * n_[b] * (double)n_[b];
  plaunch(KF_TR_BAND, 0, f2, k_sb2st, dim3(nblk), dim3(SB2_WAVES * 64), 0, s, a2);
  int32_t herr = 0;
  NRG_HIP(hipMemcpyAsync(&herr, err, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  NRG_HIP(hipStreamSynchronize(s));
  NRG_REQUIRE(herr == 0, "eigensolver: band reduction stalled (a dependency wait exceeded its bound)");
}

void BatchedEigensolver::back_q2(const int* lo, const int* hi, hipStream_t s) {
  std::vector<int32_t> ch;
  size_t shm = 0;
  for (int b = 0; b < nblk_; ++b) {
    const int n = n_[b];
    if (n < 3 || hi[b] <= lo[b]) continue;
    const int ncmax = std::max(1, std::min(8, (int)(SB2B_LDS / (8 * (size_t)n))));
    for (int c = lo[b]; c < hi[b]; c += ncmax) {
      const int nc = std::min(ncmax, hi[b] - c);
      ch.push_back(b);
      ch.push_back(c);
      ch.push_back(nc);
      shm = std::max(shm, (size_t)n * nc * sizeof(double));
    }
  }
  if (ch.empty()) return;
  d_sbch_.upload(ch.data(), ch.size(), s);
  Sb2Back a{d_n_.p, d_aoff_.p, d_sboff_.p + nblk_, d_sboff_.p + 2 * nblk_, d_v2_.p, d_t2_.p, Z_, d_sbch_.p};
  plaunch(KF_BT_BAND, 0, 0, k_sb2st_back, dim3((unsigned)(ch.size() / 3)), dim3(1024), (uint32_t)shm, s, a);
}

}  // namespace nrg
