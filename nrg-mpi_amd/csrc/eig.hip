// Batched symmetric eigensolver: see eig.h for the pipeline.
//
// Published methods followed (all code here is written from the mathematics):
//  * blocked Householder tridiagonalisation with the panel recurrence of
//    Dongarra, Sorensen & Hammarling (1989) (the LAPACK dsytrd/dlatrd
//    scheme), lower storage;
//  * Cuppen's divide and conquer (1981) with the deflation rules of Dongarra
//    & Sorensen (1987) and the eigenvector formula of Gu & Eisenstat (1995);
//  * compact WY accumulation of reflectors (Schreiber & Van Loan 1989).
#include "eig.h"
#include <climits>
#include "profiler.h"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>

namespace nrg {

constexpr int NB = 32;     // tridiagonalisation panel width
constexpr int RA = 64;     // rows per workgroup, column-update kernel
constexpr int CG = 16;     // narrowest SYMV panel (16 waves x 1 column); sizes the partial buffers
constexpr int USTRIDE = 2 * NB + 1;  // column-kernel record: V^T x, W^T x, norm
constexpr int LEAF = 32;   // divide-and-conquer leaf size
constexpr int DC_LDS_M = 8192;  // merges up to this size run the deflation scan in LDS (128 KB)
constexpr int SEC_LDS_M = 4000; // secular solves stage poles and weights in LDS up to this size
constexpr int NBB = 128;   // back-transform panel width (Z is streamed once per panel)
constexpr int MAXN = 16384;
#ifndef SYMV_U
#define SYMV_U 4   // column loads in flight per lane in the SYMV
#endif

__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// DPP lane move of a double (two 32-bit moves), all lanes of the row enabled
template <int CTRL>
__device__ inline double dpp_move(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Sum over each 16-lane row of the wave, in every lane of the row: quad
// exchanges (lane ^ 1, lane ^ 2), then rotations of the row by 4 and 8.
// Register-to-register moves only (no LDS-pipe permutes); the caller keeps
// all 64 lanes active.
__device__ inline double row_sum16(double v) {
  v += dpp_move<0xB1>(v);    // quad_perm [1,0,3,2]
  v += dpp_move<0x4E>(v);    // quad_perm [2,3,0,1]
  v += dpp_move<0x124>(v);   // row_ror:4
  v += dpp_move<0x128>(v);   // row_ror:8
  return v;
}

__device__ inline double lane_value(double v, int l) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __builtin_bit_cast(double, (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Sum over the 64 lanes of a fully active wave: row sums, then the first
// lane of each row read into scalars.  The same value in every lane, and
// the same order of additions in every wave.
__device__ inline double wave_sum_full(double v) {
  v = row_sum16(v);
  return (lane_value(v, 0) + lane_value(v, 16)) + (lane_value(v, 32) + lane_value(v, 48));
}

// ===========================================================================
// 1. tridiagonalisation
// ===========================================================================
// Per-block panel buffer P (n x 4NB, column-major, ld n):
//   [0,NB) V, [NB,2NB) W, [2NB,3NB) W copy, [3NB,4NB) V copy
// so that the trailing update is one product [V W] * [W V]^T with K = 2NB.
//
// One column c takes two launches.  The column kernel (64 rows per
// workgroup) finishes w of column c-1, updates column c with the panel so
// far and records, per workgroup, the partial norm of the column below the
// subdiagonal and the partial products V^T x, W^T x with that column x.  The
// SYMV kernel forms the reflector v (v_0 = 1, v_i = x_i * scal), its
// workgroup 0 completes u1 = V^T v and u2 = W^T v, and all workgroups
// compute y = A v of the trailing block from its lower triangle only:
// y_i = sum_{j<=i} A(i,j) v_j + sum_{j>i} A(j,i) v_j.  Workgroup p takes the
// PW columns [PW p, PW p + PW) (PW = 32 while the launch is bandwidth-bound,
// 16 once it is latency-bound: 16 waves, 2 or 1 columns each), stages v of
// its rows in LDS, streams each column's lower part contiguously, keeps the
// column dot in registers and adds the row parts
// A(i,j) v_j into an LDS vector with LDS atomics; that vector is written as
// py[p][i] (rows i >= PW p) and y_i = sum_{p <= i/PW} py[p][i] is completed
// in the next column kernel (16 waves share the sum, fixed order), as is
// y.v from one partial per SYMV workgroup.  The reductions across
// workgroups have a fixed order; the LDS atomics of one workgroup's waves do
// not, so results agree from run to run only to rounding.
// Slot q of the size order (blocks by decreasing n): everything a
// workgroup needs to address its block, in one 64-byte record so that the
// workgroup-to-block map and the block's pointers cost one load.
struct TrSlot {
  int32_t n;       // block size
  int32_t coff;    // step at which the slot starts: column c = step - coff
  int32_t b;       // block index
  int32_t pad;
  int64_t aoff;    // matrix offset
  int64_t voff;    // vector offset (d, e, tau)
  int64_t pP;      // panel buffer offset
  int64_t pRec;    // column / SYMV record offset
  int64_t pPy;     // SYMV partial-vector offset
  int64_t pad2;
};
static_assert(sizeof(TrSlot) == 64, "slot record is 64 bytes");

struct TrArgs {
  const TrSlot* slot;
  double* A;
  double* d;
  double* e;
  double* tau;
  double* py;              // per block: ceil(n/CG) partial vectors of length n
  double* P;
  double* part;            // per block: column records, SYMV records, u1/u2
};

// Pins values that are uniform over the workgroup in scalar registers at
// this point (an empty asm statement that reads them): the slot record and
// the kernel's base pointers are then loaded in one round before the
// workgroup's early exit instead of partly after it.
#define TR_PIN(...) asm volatile("" ::__VA_ARGS__)

// Flattened launch over the active blocks (slots q < act of the size order):
// slot q owns tiles(q) consecutive workgroups.  Maps workgroup L to its slot
// and tile with a wave-level prefix scan; every wave does it (no barrier).
template <class F>
__device__ inline bool flat_map(int L, int act, F tiles, int& q_out, int& t_out) {
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

// record layout inside a block's `part` region
__host__ __device__ inline int64_t colrec_size(int n) { return (int64_t)((n + RA - 1) / RA + 1) * USTRIDE; }
__host__ __device__ inline int64_t symrec_size(int n) { return (int64_t)(n + CG - 1) / CG + 1; }
__host__ __device__ inline int64_t part_size(int n) { return colrec_size(n) + symrec_size(n) + 2 * NB; }

// Column kernel.  Workgroup = RA = 64 rows (r = c + 64 blockIdx.x + lane),
// 16 waves:
//   jp >= 0: finalise w of column c-1 (panel slot jp) for rows >= c;
//   always:  copy the Householder vector of column c-1 into A (rows >= c+1);
//   jj >= 0: update column c with the jj previous panel columns, write it
//            back, record d[c], the partial norm of rows >= c+2 and the
//            partial products sum V[r,q] x_r, sum W[r,q] x_r (q < jj).
__global__ __launch_bounds__(1024) void k_tr_col(TrArgs a, int t, int jp, int jj, int vslot_prev, int pw, int act,
                                                 int tmax) {
  int qs, tile;
  auto tiles = [&](int q) {
    const int cq = t - a.slot[q].coff, nq = a.slot[q].n;
    return (cq >= 0 && cq < nq) ? (nq - cq + RA - 1) / RA : 0;
  };
  TrSlot sl;
  if (tmax > 0) {
    // tile-major order, deepest tiles first: a tile's share of the SYMV
    // partial sums grows with its row index
    qs = (int)blockIdx.x % act;
    sl = a.slot[qs];
    TR_PIN("s"(sl.n), "s"(sl.coff), "s"(sl.aoff), "s"(sl.voff), "s"(sl.pP), "s"(sl.pRec), "s"(sl.pPy),
           "s"(a.A), "s"(a.P), "s"(a.part), "s"(a.py), "s"(a.tau), "s"(a.d));
    const int cq = t - sl.coff;
    tile = ((cq >= 0 && cq < sl.n) ? (sl.n - cq + RA - 1) / RA : 0) - 1 - (int)blockIdx.x / act;
    if (tile < 0) return;
  } else {
    if (!flat_map((int)blockIdx.x, act, tiles, qs, tile)) return;
    sl = a.slot[qs];
  }
  const int c = t - sl.coff;
  if (c == 0) jp = -1;   // a block starting at a panel boundary has no previous column
  const int n = sl.n;
  const int r0 = c + tile * RA;
  const int64_t ld = n;
  double* A = a.A + sl.aoff;
  double* P = a.P + sl.pP;
  double* colrec = a.part + sl.pRec;
  const double* symrec = colrec + colrec_size(n);
  const double* ubuf = symrec + symrec_size(n);
  const double* py = a.py + sl.pPy;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wv = tid >> 6;
  const int r = r0 + lane;
  __shared__ double sV[NB][RA], sW[NB][RA];   // panel columns of this workgroup's rows
  __shared__ double su1[NB], su2[NB], sVc[NB], sWc[NB];
  __shared__ double s_y[16][64];
  __shared__ double s_red[16];
  __shared__ double s_w[64], s_x[64];
  __shared__ double s_pp[16][64];  // panel sums: waves 0-7 for w, 8-15 for the column

  // ---- every global read in one round: panel rows (into registers first,
  // stored to LDS after the other reads are issued), row c, the SYMV
  // partials of column c-1 (y, y.v, u1, u2), the column and the reflector
  const int nvc = max(jp + 1, jj);          // V columns used (w reads V[:, jp])
  const int nwc = (jp >= 0) ? jp : max(jj, 0);  // W columns read (W[:, jp] is formed here)
  static_assert(NB * RA == 2 * 1024, "two panel elements per thread");
  double pv[2], pwc[2];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int e = tid + 1024 * u;
    const int col = e / RA, rr = r0 + (e - col * RA);
    pv[u] = (col < nvc && rr < n) ? P[rr + (int64_t)col * ld] : 0.0;
    pwc[u] = (col < nwc && rr < n) ? P[rr + (int64_t)(NB + col) * ld] : 0.0;
  }
  // row c of V and W, and u1 = V^T v, u2 = W^T v of column c-1 (completed by
  // the SYMV kernel): one value per thread, kept in a register until every
  // read is issued
  int rslot = -1;
  double rv = 0.0;
  if (tid < nvc) {
    rslot = 0; rv = P[c + (int64_t)tid * ld];
  } else if (tid >= 64 && tid - 64 < nwc) {
    rslot = 1; rv = P[c + (int64_t)(NB + tid - 64) * ld];
  } else if (jp >= 0 && tid >= 128 && tid - 128 < jp) {
    rslot = 2; rv = ubuf[tid - 128];
  } else if (jp >= 0 && tid >= 192 && tid - 192 < jp) {
    rslot = 3; rv = ubuf[NB + tid - 192];
  }
  const bool copy = (c >= 1 && c - 1 <= n - 2 && r >= c + 1 && r < n);
  double a_rc = 0.0, vsl = 0.0, y_c = 0.0, tp = 0.0;
  if (wv == 0) {
    if (jj >= 0 && r < n) a_rc = A[r + c * ld];
    if (copy) vsl = P[r + vslot_prev * ld];
    if (jp >= 0) {
      tp = a.tau[sl.voff + c - 1];
      y_c = py[0];   // row c has one partial
    }
  }
  // The first round of the SYMV partials below is straight-line loads from
  // clamped (always valid) addresses, selected afterwards: no loop and no
  // branch, so that it is issued together with every read above; loops
  // handle the rare remainders after the first values are consumed.
  double ty[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  double yvp = 0.0;
  int npan = 0, nq = 0;
  const int rr = min(r, n - 1);          // a valid row for every lane
  const double* ycol = py + (rr - c);
  if (jp >= 0) {
    // y.v: one partial per SYMV workgroup of column c-1 (m = n - c >= 1)
    npan = (n - c + pw - 1) / pw;
    yvp = symrec[min(tid, npan - 1)];
    if (tid >= npan) yvp = 0.0;
    // y of row r: the 16 waves split the partial vectors py[q][r-c],
    // q <= (r-c)/pw; 8 loads per lane cover the first 128 partials of a row
    const int nqr = (rr - c) / pw + 1;
    nq = (r < n) ? nqr : 0;
#pragma unroll
    for (int u = 0; u < 8; ++u) ty[u] = ycol[(int64_t)min(wv + 16 * u, nqr - 1) * n];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (wv + 16 * u >= nq) ty[u] = 0.0;
  }
  if (jp >= 0) {
    for (int q = tid + 1024; q < npan; q += 1024) yvp += symrec[q];
    for (int q0 = wv + 128; q0 < nq; q0 += 128) {
      double x[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int q = q0 + 16 * u;
        x[u] = (q < nq) ? ycol[(int64_t)q * n] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) ty[u] += x[u];
    }
  }
  if (rslot == 0) sVc[tid] = rv;
  else if (rslot == 1) sWc[tid - 64] = rv;
  else if (rslot == 2) su1[tid - 128] = rv;
  else if (rslot == 3) su2[tid - 192] = rv;
  if (jp >= 0) {
    s_y[wv][lane] = ((ty[0] + ty[1]) + (ty[2] + ty[3])) + ((ty[4] + ty[5]) + (ty[6] + ty[7]));
    const double s = wave_sum_full(yvp);
    if (lane == 0) s_red[wv] = s;
  }
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int e = tid + 1024 * u;
    const int col = e / RA, row = e - col * RA;
    sV[col][row] = pv[u];
    sW[col][row] = pwc[u];
  }
  __syncthreads();

  // ---- panel sums of w (waves 0-7, i < jp) and of the column update
  // (waves 8-15, i < jj, i != jp): wave w takes i = w mod 8, at most 4 terms,
  // all LDS reads issued together; wave 0 adds the 8 partials below
  static_assert(NB == 32, "4 panel terms per wave");
  {
    const int par = wv & 7;
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int i = par + 8 * u;
      if (wv < 8) {
        if (i < jp) acc += sW[i][lane] * su1[i] + sV[i][lane] * su2[i];
      } else if (i < jj && i != jp) {
        acc += sV[i][lane] * sWc[i] + sW[i][lane] * sVc[i];
      }
    }
    s_pp[wv][lane] = acc;
  }
  __syncthreads();
  // ---- wave 0: w of column c-1, the Householder vector, column c
  if (wv == 0) {
    double w_r = 0.0, wc = 0.0;
    if (jp >= 0) {
      // w_pre.v = tau (y.v - 2 u1.u2);  alpha = -tau/2 * w_pre.v; w at row c
      double uu = 0.0, cc = 0.0;
      if (lane < jp) {
        uu = su1[lane] * su2[lane];
        cc = sWc[lane] * su1[lane] + sVc[lane] * su2[lane];
      }
      uu = wave_sum_full(uu);
      cc = wave_sum_full(cc);
      double yv = 0.0;
#pragma unroll
      for (int w = 0; w < 16; ++w) yv += s_red[w];
      const double al = -0.5 * tp * tp * (yv - 2.0 * uu);
      wc = tp * (y_c - cc) + al * sVc[jp];
      double acc = 0.0, pp = 0.0;
#pragma unroll
      for (int w = 0; w < 16; ++w) acc += s_y[w][lane];
#pragma unroll
      for (int w = 0; w < 8; ++w) pp += s_pp[w][lane];
      acc -= pp;
      w_r = tp * acc + al * sV[jp][lane];
      if (r < n) {
        P[r + (NB + jp) * ld] = w_r;
        P[r + (2 * NB + jp) * ld] = w_r;
      }
    }
    if (copy) A[r + (c - 1) * ld] = vsl;
    if (jj >= 0) {
      double cp = 0.0;
#pragma unroll
      for (int w = 8; w < 16; ++w) cp += s_pp[w][lane];
      double col = a_rc - cp;
      if (jp >= 0 && jp < jj) col -= sV[jp][lane] * wc + w_r * sVc[jp];
      if (r < n) {
        A[r + c * ld] = col;
        if (r == c) a.d[sl.voff + c] = col;
      }
      const double x = (r < n && r >= c + 2) ? col : 0.0;
      s_x[lane] = x;
      s_w[lane] = w_r;
      const double sq = wave_sum_full(x * x);
      if (lane == 0) colrec[(int64_t)tile * USTRIDE + 2 * NB] = sq;
    }
  }
  if (jj < 0) return;
  __syncthreads();
  // ---- partial V^T x (slots q < jj) and W^T x (slots NB + q): product
  // k = tid / 16 over the rows g, g + 16, g + 32, g + 48 (g = tid mod 16),
  // summed over the 16 lanes of its DPP row; a wave holds products 4 wv..+3
  // and skips as a whole when none of them exists
  if (4 * wv < 2 * jj) {
    const int k = tid >> 4, g = tid & 15;
    const double* src = (k < jj) ? sV[k] : (k - jj == jp) ? s_w : sW[min(k - jj, NB - 1)];
    double t = 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u) t += src[g + 16 * u] * s_x[g + 16 * u];
    t = row_sum16(t);
    if (g == 0 && k < 2 * jj) colrec[(int64_t)tile * USTRIDE + (k < jj ? k : NB + (k - jj))] = t;
  }
}

// Reflector + lower-triangle SYMV kernel, one workgroup (16 waves) per
// PW = 16 CPW columns of the trailing block of one block.  VL = 1 stages v
// of rows [j0, m) in LDS next to the row parts (2 doubles of LDS per row);
// VL = 0 reads v from the updated column in global memory (1 double per
// row), for trailing blocks whose two LDS vectors would not fit.
template <int CPW, int NW, int U, int VL>
__global__ __launch_bounds__(NW * 64, 8) void k_tr_symv(TrArgs a, int t, int jj, int act, int panel_major) {
  constexpr int PW = CPW * NW;
  extern __shared__ double ax[];   // row parts, rows [j0, m), then (VL) v of those rows
  int qs, pan;
  auto tiles = [&](int q) {
    const int cq = t - a.slot[q].coff, mq = a.slot[q].n - cq - 1;
    return (cq >= 0 && mq > 0) ? (mq + PW - 1) / PW : 0;
  };
  TrSlot sl;
  if (panel_major) {
    // workgroup L takes panel L / act of slot L % act: the longest panels of
    // all blocks are dispatched first (the trailing sizes of the blocks are
    // nearly equal under the end-aligned schedule)
    qs = (int)blockIdx.x % act;
    pan = (int)blockIdx.x / act;
    sl = a.slot[qs];
    TR_PIN("s"(sl.n), "s"(sl.coff), "s"(sl.aoff), "s"(sl.voff), "s"(sl.pP), "s"(sl.pRec), "s"(sl.pPy),
           "s"(a.A), "s"(a.P), "s"(a.part), "s"(a.py), "s"(a.tau), "s"(a.e));
    const int cq = t - sl.coff, mq = sl.n - cq - 1;
    if (pan >= ((cq >= 0 && mq > 0) ? (mq + PW - 1) / PW : 0)) return;
  } else {
    if (!flat_map((int)blockIdx.x, act, tiles, qs, pan)) return;
    sl = a.slot[qs];
  }
  const int c = t - sl.coff;
  const int n = sl.n;
  const int m = n - c - 1;
  const int j0 = pan * PW;
  const int64_t ld = n;
  const double* A = a.A + sl.aoff;
  double* P = a.P + sl.pP;
  const double* colrec = a.part + sl.pRec;
  double* symrec = a.part + sl.pRec + colrec_size(n);
  double* ubuf = symrec + symrec_size(n);
  double* py = a.py + sl.pPy + (int64_t)pan * n;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wv = tid >> 6;
  __shared__ double s_red[NW];
  __shared__ double s_u[NW][64];

  // ---- one round of reads before anything waits: the column norms and
  // alpha (the reflector), the raw column c of the rows this workgroup
  // stages (v = scal x, once scal is known), the u records (panel 0) and
  // the first chunk of the wave's first column.  Clamped, always valid
  // addresses; out-of-range values are discarded where they are used.
  const int ncol = (n - c + RA - 1) / RA;   // column-kernel records of column c
  const double* vcol = A + (c + 1) + (int64_t)c * ld;
  const double* At = A + (c + 1) + (int64_t)(c + 1) * ld;   // At[i + j*ld] = A(c+1+i, c+1+j)
  double nrm = colrec[(int64_t)min(lane, ncol - 1) * USTRIDE + 2 * NB];
  const double alpha = A[(c + 1) + c * ld];
  constexpr int VS = 2;   // staged rows per thread in the first round
  double vraw[VS];
#pragma unroll
  for (int u = 0; u < VS; ++u) vraw[u] = VL ? vcol[min(j0 + tid + NW * 64 * u, m - 1)] : 0.0;
  // record slot of u1 (lanes < jj) and u2 (lanes jj..2jj-1); a valid
  // dummy slot for the other lanes, so that the loads need no lane branch
  const int uslot = (lane >= 2 * jj) ? 0 : (lane < jj) ? lane : NB + (lane - jj);
  double urec[2] = {0.0, 0.0};
  if (pan == 0) {
#pragma unroll
    for (int u = 0; u < 2; ++u) urec[u] = colrec[(int64_t)min(wv + NW * u, ncol - 1) * USTRIDE + uslot];
  }
  const int jf = min(j0 + wv * CPW, m - 1);   // the wave's first column (clamped)
  double xf[U];
#pragma unroll
  for (int u = 0; u < U; ++u) xf[u] = At[min(jf + lane + 64 * u, m - 1) + (int64_t)jf * ld];

  // reflector from the updated column c (identical in every workgroup)
  double xn2;
  {
    if (lane >= ncol) nrm = 0.0;
    for (int q = lane + 64; q < ncol; q += 64) nrm += colrec[(int64_t)q * USTRIDE + 2 * NB];
    xn2 = wave_sum_full(nrm);   // every wave forms the same sum
  }
  double beta, tau, scal;
  if (xn2 == 0.0) {
    beta = alpha; tau = 0.0; scal = 0.0;
  } else {
    beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
    tau = (beta - alpha) / beta;
    scal = 1.0 / (alpha - beta);
  }
  // v_i = 1 (i = 0), A(c+1+i, c) * scal (i > 0): read from the column
  auto vv = [&](int i) { return (i == 0) ? 1.0 : vcol[i] * scal; };
  if (pan == 0) {
    if (tid == 0) {
      a.e[sl.voff + c] = beta;
      a.tau[sl.voff + c] = tau;
    }
    // u1_q = V[c+1,q] + scal sum_{r>=c+2} V[r,q] x_r (same for W): the
    // waves split the records, lane = slot, fixed order
    double s = 0.0;
    if (lane < 2 * jj) {
#pragma unroll
      for (int u = 0; u < 2; ++u)
        if (wv + NW * u < ncol) s += urec[u];
      for (int q = wv + 2 * NW; q < ncol; q += NW) s += colrec[(int64_t)q * USTRIDE + uslot];
    }
    s_u[wv][lane] = s;
  }
  double* sv = ax + (m - j0);   // VL: v_i of rows [j0, m), staged once per workgroup
#pragma unroll
  for (int u = 0; u < VS; ++u) {
    const int i = j0 + tid + NW * 64 * u;
    if (i < m) {
      ax[i - j0] = 0.0;
      if (VL) sv[i - j0] = (i == 0) ? 1.0 : vraw[u] * scal;
    }
  }
  for (int i = j0 + tid + NW * 64 * VS; i < m; i += NW * 64) {
    ax[i - j0] = 0.0;
    if (VL) sv[i - j0] = vv(i);
  }
  // v_i for a row i in [j0, m)
  auto vr = [&](int i) { return VL ? sv[i - j0] : vv(i); };
  __syncthreads();
  if (pan == 0 && tid < 2 * jj) {
    const int col = (tid < jj) ? tid : NB + (tid - jj);
    double s = 0.0;
    for (int w = 0; w < NW; ++w) s += s_u[w][tid];
    ubuf[col] = P[(c + 1) + (int64_t)col * ld] + scal * s;
  }

  for (int q = 0; q < CPW; ++q) {
    const int j = j0 + wv * CPW + q;
    if (j >= m) break;
    const double* col = At + (int64_t)j * ld;
    const double vj = vr(j);
    // U loads in flight per lane, all issued before the first use; two
    // accumulators keep the kernel within 64 VGPRs (two workgroups per CU)
    double s0 = 0.0, s1 = 0.0;
    // rows i + 64 u of column j, those below m
    auto use_guarded = [&](const double (&x)[U], int i) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int r = i + 64 * u;
        if (r < m) {
          if (u & 1) s1 += x[u] * vr(r);
          else s0 += x[u] * vr(r);
          if (r > j) atomicAdd(&ax[r - j0], x[u] * vj);
        }
      }
    };
    int i = j + lane;
    if (q == 0) {
      // the chunk read before the reflector
      use_guarded(xf, i);
      i += 64 * U;
    }
    for (; i + 64 * (U - 1) < m; i += 64 * U) {
      double x[U];
#pragma unroll
      for (int u = 0; u < U; ++u) x[u] = col[i + 64 * u];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (u & 1) s1 += x[u] * vr(i + 64 * u);
        else s0 += x[u] * vr(i + 64 * u);
        if (u > 0 || i > j) atomicAdd(&ax[i + 64 * u - j0], x[u] * vj);
      }
    }
    if (i < m) {
      // the last rows of the column as one guarded batch (one round trip)
      double x[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        x[u] = 0.0;
        if (i + 64 * u < m) x[u] = col[i + 64 * u];
      }
      use_guarded(x, i);
    }
    const double t = wave_sum_full(s0 + s1);
    if (lane == 0) atomicAdd(&ax[j - j0], t);
  }
  __syncthreads();
  // row parts out, y.v partial, v into the panel buffer (panel rows)
  double yv = 0.0;
  for (int i = j0 + tid; i < m; i += NW * 64) {
    const double t = ax[i - j0];
    py[i] = t;
    yv += t * vr(i);
  }
  if (tid < PW && j0 + tid < m) {
    const int i = j0 + tid;
    const int r = c + 1 + i;
    const double vi = vr(i);
    P[r + jj * ld] = vi;
    P[r + (3 * NB + jj) * ld] = vi;
  }
  yv = wave_sum_full(yv);
  if (lane == 0) s_red[wv] = yv;
  __syncthreads();
  if (tid == 0) {
    double t = 0.0;
    for (int w = 0; w < NW; ++w) t += s_red[w];
    symrec[pan] = t;
  }
}

// Tail finisher: the last columns of every block (a trailing block of at
// most FIN_MX rows, whose lower triangle fits one workgroup's LDS packed by
// columns: 148 KB) reduced in one launch, one workgroup per block, by the
// unblocked two-sided Householder reduction (dsytd2, lower): per column the
// reflector from the column below the diagonal, p = tau A22 v (4 lanes per
// row), w = p - (tau/2)(p.v) v, A22 -= v w^T + w v^T.  The reflectors,
// d, e and tau are stored as the blocked steps store them.  Starts at a
// panel boundary, after that panel's trailing update: the trailing block is
// then the fully updated matrix.  Replaces the last ~FIN_MX column steps,
// which are latency-bound launches (two kernels per step and group).
constexpr int FIN_MX = 192;
struct FinTask {
  int32_t n, c0;      // block size, first column left to reduce
  int64_t aoff, voff;
};

__global__ __launch_bounds__(1024) void k_tr_finish(const FinTask* __restrict__ tasks, double* A, double* d,
                                                    double* e, double* tau) {
  __shared__ double L[FIN_MX * (FIN_MX + 1) / 2];
  __shared__ double sv[FIN_MX], sw[FIN_MX];
  __shared__ double s_par[4];
  const FinTask tk = tasks[blockIdx.x];
  const int n = tk.n, c0 = tk.c0, m = n - c0;
  if (m < 1) return;
  double* Ab = A + tk.aoff;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  // column j of the packed lower triangle starts at j m - j (j - 1) / 2
  auto pk = [&](int i, int j) { return j * m - ((j * (j - 1)) >> 1) + (i - j); };
  // the trailing lower triangle into LDS: wave wv takes columns wv + 16 q,
  // lanes the rows; four columns' rows are in flight together
  for (int j0 = wv; j0 < m; j0 += 64) {
    double x[4][3];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const int j = j0 + 16 * q, i = j + lane + 64 * r;
        x[q][r] = (j < m && i < m) ? Ab[(c0 + i) + (int64_t)(c0 + j) * n] : 0.0;
      }
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const int j = j0 + 16 * q, i = j + lane + 64 * r;
        if (j < m && i < m) L[pk(i, j)] = x[q][r];
      }
  }
  __syncthreads();
  for (int k = 0; k < m; ++k) {
    const int c = c0 + k;
    const double* colk = &L[pk(k, k)];   // colk[r] = A(k + r, k)
    if (tid == 0) d[tk.voff + c] = colk[0];
    if (k == m - 1) break;
    const int len = m - k - 1;           // rows k+1 .. m-1
    if (wv == 0) {
      double s2 = 0.0;
      for (int r = 2 + lane; r <= len; r += 64) {
        const double x = colk[r];
        s2 += x * x;
      }
      s2 = wave_sum_full(s2);
      if (lane == 0) {
        const double alpha = colk[1];
        double beta, tq, scal;
        if (s2 == 0.0) {
          beta = alpha; tq = 0.0; scal = 0.0;
        } else {
          beta = -copysign(sqrt(alpha * alpha + s2), alpha);
          tq = (beta - alpha) / beta;
          scal = 1.0 / (alpha - beta);
        }
        s_par[0] = beta; s_par[1] = tq; s_par[2] = scal;
        e[tk.voff + c] = beta;
        tau[tk.voff + c] = tq;
      }
    }
    __syncthreads();
    const double tq = s_par[1], scal = s_par[2];
    // v (v_{k+1} = 1) into LDS and below the subdiagonal of column c
    for (int r = tid; r < len; r += 1024) {
      const int i = k + 1 + r;
      const double vi = (r == 0) ? 1.0 : colk[1 + r] * scal;
      sv[i] = vi;
      if (r > 0) Ab[(c0 + i) + (int64_t)c * n] = vi;
    }
    __syncthreads();
    if (tq == 0.0) continue;   // H = I (uniform over the workgroup)
    // p = tau A22 v: row k+1+tid/4, the 4 lanes of a row take every 4th column
    {
      const int row = k + 1 + (tid >> 2), q = tid & 3;
      double acc = 0.0;
      if (row < m)
        for (int j = k + 1 + q; j < m; j += 4) acc += ((row >= j) ? L[pk(row, j)] : L[pk(j, row)]) * sv[j];
      acc += __shfl_xor(acc, 1, 64);
      acc += __shfl_xor(acc, 2, 64);
      if (row < m && q == 0) sw[row] = tq * acc;
    }
    __syncthreads();
    if (wv == 0) {
      double sdot = 0.0;
      for (int i = k + 1 + lane; i < m; i += 64) sdot += sw[i] * sv[i];
      sdot = wave_sum_full(sdot);
      if (lane == 0) s_par[3] = -0.5 * tq * sdot;
    }
    __syncthreads();
    const double a2 = s_par[3];
    for (int i = k + 1 + tid; i < m; i += 1024) sw[i] += a2 * sv[i];
    __syncthreads();
    // A22 -= v w^T + w v^T on the lower triangle: columns per wave, rows per lane
    for (int j = k + 1 + wv; j < m; j += 16) {
      const double vj = sv[j], wj = sw[j];
      double* col = &L[pk(j, j)];
      for (int i = j + lane; i < m; i += 64) col[i - j] -= sv[i] * wj + sw[i] * vj;
    }
    __syncthreads();
  }
}

// Row-chunk SYMV: the same product as k_tr_symv without LDS atomics.  The
// workgroup walks the rows of its panel in chunks of SR = 256; in a chunk
// each wave streams its CPW columns (4 rows per lane and column, the next
// chunk's loads issued before this chunk's arithmetic), keeps the column
// dots in registers and sums its row parts A(i,j) v_j over its columns in
// registers; the 16 waves' row parts go to a double-buffered LDS array and
// 256 threads add them in a fixed wave order (one barrier per chunk).  The
// first chunk's rows wait in LDS for the column dots (they hold the panel's
// own rows).  Results are the same from run to run bit for bit.
constexpr int SR = 256;
template <int CPW, int VL>
__global__ __launch_bounds__(1024) void k_tr_symv_rows(TrArgs a, int t, int jj, int act, int panel_major) {
  constexpr int NW = 16, PW = CPW * NW;
  extern __shared__ double ax[];   // VL: v of rows [j0, m)
  __shared__ double s_part[2][NW][SR];
  __shared__ double s_first[SR];
  __shared__ double s_cd[PW];
  __shared__ double s_red[NW];
  __shared__ double s_u[NW][64];
  int qs, pan;
  auto tiles = [&](int q) {
    const int cq = t - a.slot[q].coff, mq = a.slot[q].n - cq - 1;
    return (cq >= 0 && mq > 0) ? (mq + PW - 1) / PW : 0;
  };
  TrSlot sl;
  if (panel_major) {
    qs = (int)blockIdx.x % act;
    pan = (int)blockIdx.x / act;
    sl = a.slot[qs];
    TR_PIN("s"(sl.n), "s"(sl.coff), "s"(sl.aoff), "s"(sl.voff), "s"(sl.pP), "s"(sl.pRec), "s"(sl.pPy),
           "s"(a.A), "s"(a.P), "s"(a.part), "s"(a.py), "s"(a.tau), "s"(a.e));
    const int cq = t - sl.coff, mq = sl.n - cq - 1;
    if (pan >= ((cq >= 0 && mq > 0) ? (mq + PW - 1) / PW : 0)) return;
  } else {
    if (!flat_map((int)blockIdx.x, act, tiles, qs, pan)) return;
    sl = a.slot[qs];
  }
  const int c = t - sl.coff;
  const int n = sl.n;
  const int m = n - c - 1;
  const int j0 = pan * PW;
  const int64_t ld = n;
  const double* A = a.A + sl.aoff;
  double* P = a.P + sl.pP;
  const double* colrec = a.part + sl.pRec;
  double* symrec = a.part + sl.pRec + colrec_size(n);
  double* ubuf = symrec + symrec_size(n);
  double* py = a.py + sl.pPy + (int64_t)pan * n;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wv = tid >> 6;

  // reflector from the updated column c (identical in every workgroup)
  const int ncol = (n - c + RA - 1) / RA;
  double xn2;
  {
    double sl2 = 0.0;
    for (int q = lane; q < ncol; q += 64) sl2 += colrec[(int64_t)q * USTRIDE + 2 * NB];
    xn2 = wave_sum_full(sl2);
  }
  const double alpha = A[(c + 1) + c * ld];
  double beta, tau, scal;
  if (xn2 == 0.0) {
    beta = alpha; tau = 0.0; scal = 0.0;
  } else {
    beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
    tau = (beta - alpha) / beta;
    scal = 1.0 / (alpha - beta);
  }
  const double* vcol = A + (c + 1) + (int64_t)c * ld;
  auto vv = [&](int i) { return (i == 0) ? 1.0 : vcol[i] * scal; };
  if (pan == 0) {
    if (tid == 0) {
      a.e[sl.voff + c] = beta;
      a.tau[sl.voff + c] = tau;
    }
    double s = 0.0;
    if (lane < 2 * jj) {
      const int slot = (lane < jj) ? lane : NB + (lane - jj);
      for (int q = wv; q < ncol; q += NW) s += colrec[(int64_t)q * USTRIDE + slot];
    }
    s_u[wv][lane] = s;
  }
  if (VL)
    for (int i = j0 + tid; i < m; i += NW * 64) ax[i - j0] = vv(i);
  auto vr = [&](int i) { return VL ? ax[i - j0] : vv(i); };
  __syncthreads();
  if (pan == 0 && tid < 2 * jj) {
    const int col = (tid < jj) ? tid : NB + (tid - jj);
    double s = 0.0;
    for (int w = 0; w < NW; ++w) s += s_u[w][tid];
    ubuf[col] = P[(c + 1) + (int64_t)col * ld] + scal * s;
  }

  const double* At = A + (c + 1) + (int64_t)(c + 1) * ld;   // At[i + j*ld] = A(c+1+i, c+1+j)
  int jq[CPW];
  double vjq[CPW], cd[CPW];
#pragma unroll
  for (int q = 0; q < CPW; ++q) {
    jq[q] = j0 + wv * CPW + q;
    vjq[q] = (jq[q] < m) ? vr(jq[q]) : 0.0;
    cd[q] = 0.0;
  }
  const int nch = (m - j0 + SR - 1) / SR;
  double xn[CPW][4];
  auto load = [&](int ch) {
    const int base = j0 + ch * SR;
#pragma unroll
    for (int q = 0; q < CPW; ++q)
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int i = base + lane + 64 * u;
        xn[q][u] = (jq[q] < m && i >= jq[q] && i < m) ? At[i + (int64_t)jq[q] * ld] : 0.0;
      }
  };
  load(0);
  double yvp = 0.0;
  for (int ch = 0; ch < nch; ++ch) {
    double x[CPW][4];
#pragma unroll
    for (int q = 0; q < CPW; ++q)
#pragma unroll
      for (int u = 0; u < 4; ++u) x[q][u] = xn[q][u];
    if (ch + 1 < nch) load(ch + 1);
    const int base = j0 + ch * SR;
    double (*Pp)[SR] = s_part[ch & 1];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int i = base + lane + 64 * u;
      const double vi = (i < m) ? vr(i) : 0.0;
      double ra = 0.0;
#pragma unroll
      for (int q = 0; q < CPW; ++q) {
        cd[q] += x[q][u] * vi;
        if (i > jq[q]) ra += x[q][u] * vjq[q];
      }
      Pp[wv][lane + 64 * u] = ra;
    }
    __syncthreads();
    if (tid < SR) {
      double s = 0.0;
#pragma unroll
      for (int w = 0; w < NW; ++w) s += Pp[w][tid];
      const int i = base + tid;
      if (ch == 0) s_first[tid] = s;
      else if (i < m) {
        py[i] = s;
        yvp += s * vr(i);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < CPW; ++q) {
    const double tq = wave_sum_full(cd[q]);
    if (lane == 0) s_cd[wv * CPW + q] = tq;
  }
  __syncthreads();
  if (tid < SR) {
    const int i = j0 + tid;
    if (i < m) {
      const double s = s_first[tid] + ((tid < PW) ? s_cd[tid] : 0.0);
      py[i] = s;
      yvp += s * vr(i);
    }
  }
  if (tid < PW && j0 + tid < m) {
    const int i = j0 + tid;
    const int r = c + 1 + i;
    const double vi = vr(i);
    P[r + jj * ld] = vi;
    P[r + (3 * NB + jj) * ld] = vi;
  }
  yvp = wave_sum_full(yvp);
  if (lane == 0) s_red[wv] = yvp;
  __syncthreads();
  if (tid == 0) {
    double t2 = 0.0;
    for (int w = 0; w < NW; ++w) t2 += s_red[w];
    symrec[pan] = t2;
  }
}

// ===========================================================================
// 2. divide and conquer
// ===========================================================================
// integer task records
//   leaf:  [b, s, m]
//   merge: [b, s, mid, e, wsoff_lo, wsoff_hi, iwsoff_lo, iwsoff_hi]
struct DcArgs {
  const int32_t* n;
  const int64_t* aoff;
  const int64_t* voff;
  double* d;   // diagonal (torn)
  const double* e;
  double* w;   // eigenvalues of the current subproblems
  double* Z;
  double* ws;
  int32_t* iws;
  int32_t* kout;
};

__device__ inline int64_t join64(int32_t lo, int32_t hi) {
  return (int64_t)(uint32_t)lo | ((int64_t)hi << 32);
}

__global__ void k_dc_tear(const int32_t* bnd, int nb_, const int32_t* n, const int64_t* voff,
                          double* d, const double* e) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb_) return;
  const int b = bnd[2 * i], mid = bnd[2 * i + 1];
  const int64_t o = voff[b];
  const double r = fabs(e[o + mid - 1]);
  d[o + mid - 1] -= r;
  d[o + mid] -= r;
}

// QL with implicit shifts on a leaf (m <= LEAF), one wave per leaf.  The
// diagonal and off-diagonal live in registers (lane j holds d_j and e_j);
// the scalar recurrence reads them with readlane, so every lane runs it with
// the same values, and the lane that owns an entry takes its update.  Each
// lane applies the plane rotations to its own row of the leaf eigenvector
// matrix in LDS, carrying in a register the entry that the next rotation
// of the sweep reuses.  No LDS traffic or barrier on the recurrence.
__global__ __launch_bounds__(64) void k_dc_leaf(DcArgs a, const int32_t* tasks, int ntask) {
  const int t = blockIdx.x;
  if (t >= ntask) return;
  const int b = tasks[3 * t], s = tasks[3 * t + 1], m = tasks[3 * t + 2];
  const int lane = threadIdx.x;
  const int n = a.n[b];
  const int64_t vo = a.voff[b];
  __shared__ double zl[LEAF][LEAF + 1];   // row `lane` is touched by that lane only
  double dr = 0.0, er = 0.0;
  if (lane < m) {
    dr = a.d[vo + s + lane];
    er = (lane < m - 1) ? a.e[vo + s + lane] : 0.0;
    for (int j = 0; j < m; ++j) zl[lane][j] = (j == lane) ? 1.0 : 0.0;
  }
  const double eps = 2.220446049250313e-16;
  for (int l = 0; l < m; ++l) {
    int iter = 0;
    while (true) {
      // first mm >= l with a negligible e_mm relative to its two diagonal
      // neighbours (m - 1 when there is none)
      const double dn = __shfl_down(dr, 1, 64);
      const bool neg = lane >= l && lane < m - 1 && fabs(er) <= eps * (fabs(dr) + fabs(dn));
      const unsigned long long hit = __ballot(neg);
      const int mm = hit ? __ffsll(hit) - 1 : m - 1;
      if (mm == l) break;
      if (++iter > 90) break;
      const double dl0 = lane_value(dr, l), el = lane_value(er, l);
      double g = (lane_value(dr, l + 1) - dl0) / (2.0 * el);
      double r = hypot(g, 1.0);
      g = lane_value(dr, mm) - dl0 + el / (g + copysign(r, g));
      double sn = 1.0, cs = 1.0, p = 0.0;
      double zc = (lane < m) ? zl[lane][mm] : 0.0;   // entry i + 1 of the rotation at i
      int last = -1;
      bool early = false;
      for (int i = mm - 1; i >= l; --i) {
        const double ei = lane_value(er, i);
        const double f = sn * ei;
        const double bb = cs * ei;
        r = hypot(f, g);
        if (lane == i + 1) er = r;
        if (r == 0.0) {
          if (lane == i + 1) dr -= p;
          if (lane == mm) er = 0.0;
          early = true;
          break;
        }
        sn = f / r;
        cs = g / r;
        g = lane_value(dr, i + 1) - p;
        r = (lane_value(dr, i) - g) * sn + 2.0 * cs * bb;
        p = sn * r;
        if (lane == i + 1) dr = g + p;
        g = cs * r - bb;
        if (lane < m) {
          const double zi = zl[lane][i];
          zl[lane][i + 1] = sn * zi + cs * zc;
          zc = cs * zi - sn * zc;
        }
        last = i;
      }
      if (last >= 0 && lane < m) zl[lane][last] = zc;
      if (early) continue;
      if (lane == l) { dr -= p; er = g; }
      if (lane == mm) er = 0.0;
    }
  }
  // selection sort ascending, columns follow: the minimum of d over
  // [i, m), first index on ties, by a butterfly over the wave
  for (int i = 0; i < m - 1; ++i) {
    double v = (lane >= i && lane < m) ? dr : __builtin_inf();
    int k = lane;
    for (int o = 32; o > 0; o >>= 1) {
      const double v2 = __shfl_xor(v, o, 64);
      const int k2 = __shfl_xor(k, o, 64);
      if (v2 < v || (v2 == v && k2 < k)) { v = v2; k = k2; }
    }
    const int kmin = __builtin_amdgcn_readfirstlane(k);
    if (kmin != i) {
      const double di = lane_value(dr, i);
      if (lane == i) dr = v;
      else if (lane == kmin) dr = di;
      if (lane < m) {
        const double tz = zl[lane][i];
        zl[lane][i] = zl[lane][kmin];
        zl[lane][kmin] = tz;
      }
    }
  }
  if (lane < m) {
    a.w[vo + s + lane] = dr;
    double* Z = a.Z + a.aoff[b];
    for (int j = 0; j < m; ++j) Z[(int64_t)(s + lane) + (int64_t)(s + j) * n] = zl[lane][j];
  }
}

// merge workspace layout (doubles, per merge, m = e - s):
//   ds[m] zs[m] dl[m] zl[m] dv[m] zh[m] tr[m] rc[m] rs[m] | Qcd[m*m] | U[m*m] | Qt[m*m]
// ints: col[m] nd[m] df[m] org[m] rp[2m] | sorted-deflated order[m] | nrot, pad |
//   ndt[m] perm[m] inv[m] half[m] (at 8m + 8: support of each non-deflated
//   column, its position in the grouped order and back, the half of each
//   sorted pole)
__device__ inline int dws_vec(int m) { return 9 * m; }
__host__ __device__ inline int64_t dws_int(int m) { return 12 * (int64_t)m + 16; }

// The non-deflated columns of Qcd and the rows of U are stored grouped by
// support (LAPACK dlaed2's column types): first the columns with nonzeros
// only in the rows of the first half (bit 1), then those with both halves
// (bits 1|2: a deflation rotation mixed two columns from different halves),
// then those of the second half only (bit 2), each group in ascending pole
// order.  Qcd U then splits into two products on the halves' rows, over the
// columns [0, k1 + k3) and [k1, k): about half the flops of the dense one.
// kout holds k per merge at [t] and (k1, k3) at [ntask + 2t].

__global__ __launch_bounds__(256) void k_dc_merge_prep(DcArgs a, const int32_t* tasks, int ntask) {
  const int t = blockIdx.x;
  const int32_t* tk = tasks + 8 * t;
  const int b = tk[0], s = tk[1], mid = tk[2], e = tk[3];
  const int64_t wo = join64(tk[4], tk[5]);
  const int64_t io = join64(tk[6], tk[7]);
  const int m = e - s, m1 = mid - s, m2 = e - mid;
  const int n = a.n[b];
  const int64_t vo = a.voff[b];
  double* Z = a.Z + a.aoff[b];
  double* W = a.ws + wo;
  int32_t* I = a.iws + io;
  double *ds = W, *zs = W + m, *dl = W + 2 * m, *zl = W + 3 * m, *dv = W + 4 * m;
  double *rc = W + 7 * m, *rsn = W + 8 * m;
  double* Qcd = W + dws_vec(m);
  int32_t *col = I, *nd = I + m, *df = I + 2 * m, *rp = I + 4 * m;
  int32_t *ndt = I + 8 * m + 8, *perm = I + 9 * m + 8, *inv = I + 10 * m + 8, *hf = I + 11 * m + 8;
  const int tid = threadIdx.x;
  extern __shared__ double dc_shm[];
  __shared__ double s_red[4];
  __shared__ int s_k, s_nrot;
  __shared__ double s_rho;
  __shared__ int s_cnt[3][256];
  __shared__ int s_tot[3];

  const double beta = a.e[vo + mid - 1];
  const double rho = fabs(beta);
  const double sg = (beta >= 0.0) ? 1.0 : -1.0;
  // merge ranks of the two ascending lists (ties: left first), searched in
  // LDS when they fit (instantiated per memory kind, as the scan below)
  auto merge_ranks = [&](const auto* D1, const auto* D2) {
    for (int i = tid; i < m; i += 256) {
      double v, z;
      int rank, cidx;
      if (i < m1) {
        v = D1[i];
        z = Z[(int64_t)(mid - 1) + (int64_t)(s + i) * n];
        int lo = 0, hi = m2;
        while (lo < hi) { int md = (lo + hi) >> 1; if (D2[md] < v) lo = md + 1; else hi = md; }
        rank = i + lo;
        cidx = s + i;
      } else {
        const int j = i - m1;
        v = D2[j];
        z = sg * Z[(int64_t)mid + (int64_t)(mid + j) * n];
        int lo = 0, hi = m1;
        while (lo < hi) { int md = (lo + hi) >> 1; if (D1[md] <= v) lo = md + 1; else hi = md; }
        rank = j + lo;
        cidx = mid + j;
      }
      ds[rank] = v;
      zs[rank] = z * 0.7071067811865476;  // z / sqrt(2): |z| = 1
      col[rank] = cidx;
      hf[rank] = (i < m1) ? 1 : 2;
    }
  };
  if (m <= DC_LDS_M) {
    // the two ascending lists are adjacent in w
    for (int i = tid; i < m; i += 256) dc_shm[i] = a.w[vo + s + i];
    __syncthreads();
    merge_ranks(dc_shm, dc_shm + m1);
  } else {
    merge_ranks(a.w + vo + s, a.w + vo + mid);
  }
  __syncthreads();
  double mx = 0.0;
  for (int i = tid; i < m; i += 256) mx = fmax(mx, fmax(fabs(ds[i]), fabs(zs[i])));
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
  if ((tid & 63) == 0) s_red[tid >> 6] = mx;
  // stage the sorted poles, weights and halves in LDS for the serial scan
  const bool use_lds = (m <= DC_LDS_M);
  unsigned char* shf = reinterpret_cast<unsigned char*>(dc_shm + 2 * m);
  if (use_lds)
    for (int i = tid; i < m; i += 256) { dc_shm[i] = ds[i]; dc_shm[m + i] = zs[i]; shf[i] = (unsigned char)hf[i]; }
  __syncthreads();
  // The scan and the copies after it are written once and instantiated for
  // LDS and for global poles/weights: with one pointer that may be either,
  // the compiler emits flat accesses, and a flat load waits for every global
  // store before it (the scan's index stores), one memory latency per step.
  auto scan = [&](auto* sds, auto* szs, const auto* shh) {
    const double rp2 = 2.0 * rho;
    const double tol = 8.0 * 2.220446049250313e-16 * fmax(fmax(s_red[0], s_red[1]), fmax(s_red[2], s_red[3]));
    s_rho = rp2;
    // serial deflation scan in ascending order.  The pole and weight of
    // the running column pj are carried in registers (written back when pj
    // is settled) and those of the next position are fetched one step
    // ahead, so the loop-carried chain is arithmetic only.
    // tp: halves holding the running column's nonzeros
    int k = 0, nd_def = 0, nrot = 0, pj = -1, tp = 0;
    double zp = 0.0, dp = 0.0;
    double zn = szs[0], dn = sds[0];
    int hn = shh[0];
    for (int j = 0; j < m; ++j) {
      const double zj = zn, dj = dn;
      const int hj = hn;
      if (j + 1 < m) { zn = szs[j + 1]; dn = sds[j + 1]; hn = shh[j + 1]; }
      if (rp2 * fabs(zj) <= tol) { df[nd_def++] = j; continue; }
      if (pj < 0) { pj = j; zp = zj; dp = dj; tp = hj; continue; }
      const double dt = dj - dp;
      const double r2 = zj * zj + zp * zp;   // |z| <= 1: no overflow
      // |dt c s| <= tol with c = zj / r, s = -zp / r, r^2 = zj^2 + zp^2,
      // tested as |dt zj zp| <= tol r^2: the square root and the division
      // stay off the loop-carried chain unless the pair deflates
      if (fabs(dt * zj * zp) <= tol * r2) {
        const double tt = sqrt(r2);
        const double rt = 1.0 / tt;
        const double cv = zj * rt, sv = -zp * rt;
        szs[pj] = 0.0;
        rp[2 * nrot] = pj;
        rp[2 * nrot + 1] = j;
        rc[nrot] = cv;
        rsn[nrot] = sv;
        ++nrot;
        sds[pj] = dp * cv * cv + dj * sv * sv;
        df[nd_def++] = pj;
        pj = j;
        zp = tt;
        dp = dp * sv * sv + dj * cv * cv;
        tp |= hj;   // column j now carries the rotated-out column's rows too
      } else {
        szs[pj] = zp;
        sds[pj] = dp;
        ndt[k] = tp;
        nd[k++] = pj;
        pj = j;
        zp = zj;
        dp = dj;
        tp = hj;
      }
    }
    if (pj >= 0) {
      szs[pj] = zp;
      sds[pj] = dp;
      ndt[k] = tp;
      nd[k++] = pj;
    }
    s_k = k;
    s_nrot = nrot;
    a.kout[t] = k;
    I[7 * m] = nrot;
  };
  auto gather = [&](const auto* sds, const auto* szs) {
    const int k = s_k;
    for (int i = tid; i < k; i += 256) {
      dl[i] = sds[nd[i]];
      zl[i] = szs[nd[i]];
    }
    for (int i = tid; i < m - k; i += 256) dv[i] = sds[df[i]];
  };
  if (use_lds) {
    if (tid == 0) scan(dc_shm, dc_shm + m, shf);
    __syncthreads();
    gather(dc_shm, dc_shm + m);
  } else {
    if (tid == 0) scan(ds, zs, hf);
    __syncthreads();
    gather(ds, zs);
  }
  if (tid == 0) W[5 * m] = s_rho;  // zh[0] slot keeps rho until the z-hat pass
  // grouped order of the non-deflated columns: a stable partition of
  // [0, k) by support (first half, both, second half); thread tid takes a
  // contiguous range, the per-thread counts are scanned by one thread per group
  {
    const int k = s_k;
    const int per = (k + 255) / 256;
    const int i0 = min(k, tid * per), i1 = min(k, i0 + per);
    int c1 = 0, c3 = 0, c2 = 0;
    for (int i = i0; i < i1; ++i) {
      const int ty = ndt[i];
      c1 += (ty == 1);
      c3 += (ty == 3);
      c2 += (ty == 2);
    }
    s_cnt[0][tid] = c1;
    s_cnt[1][tid] = c3;
    s_cnt[2][tid] = c2;
    __syncthreads();
    if (tid < 3) {
      int acc = 0;
      for (int q = 0; q < 256; ++q) {
        const int v = s_cnt[tid][q];
        s_cnt[tid][q] = acc;
        acc += v;
      }
      s_tot[tid] = acc;
    }
    __syncthreads();
    const int k1 = s_tot[0], k3 = s_tot[1];
    int p1 = s_cnt[0][tid], p3 = k1 + s_cnt[1][tid], p2 = k1 + k3 + s_cnt[2][tid];
    for (int i = i0; i < i1; ++i) {
      const int ty = ndt[i];
      const int pos = (ty == 1) ? p1++ : (ty == 3) ? p3++ : p2++;
      perm[i] = pos;
      inv[pos] = i;
    }
    if (tid == 0) {
      a.kout[ntask + 2 * t] = k1;
      a.kout[ntask + 2 * t + 1] = k3;
    }
  }
  (void)Qcd;
}

// Apply the deflation rotations to the merged columns, one lane per row.
// The scan makes chains: a rotation (pj, j) finishes column pj and hands j on
// as the next pj, so the running column stays in a register and is stored
// when its chain ends.  Both columns of the next 16 rotations are fetched
// together (degenerate clusters make many short chains, whose first
// columns would otherwise each cost a load round trip).  Grid: (row chunks
// of 256, merges).
__global__ __launch_bounds__(256) void k_dc_rotate(DcArgs a, const int32_t* tasks, int ntask) {
  constexpr int CH = 1024;   // rotations staged in LDS per round
  constexpr int RQ = 16;     // rotations whose columns are loaded together
  __shared__ int32_t s_c[2 * CH];
  __shared__ double s_r[2 * CH];
  const int t = blockIdx.y;
  const int32_t* tk = tasks + 8 * t;
  const int b = tk[0], s = tk[1], e = tk[3];
  const int m = e - s;
  if (blockIdx.x * 256 >= m) return;
  const int32_t* I = a.iws + join64(tk[6], tk[7]);
  const int nrot = I[7 * m];
  if (nrot == 0) return;
  const int row = blockIdx.x * 256 + threadIdx.x;
  const int n = a.n[b];
  const double* W = a.ws + join64(tk[4], tk[5]);
  const int32_t *col = I, *rp = I + 4 * m;
  const double *rc = W + 7 * m, *rsn = W + 8 * m;
  double* Zr = a.Z + a.aoff[b] + (s + row);
  int cc = -1;        // carried column and its current value
  double cvl = 0.0;
  for (int c0q = 0; c0q < nrot; c0q += CH) {
    const int nch = min(CH, nrot - c0q);
    __syncthreads();   // the previous round's tables are consumed
    for (int q = threadIdx.x; q < nch; q += 256) {
      s_c[2 * q] = col[rp[2 * (c0q + q)]];
      s_c[2 * q + 1] = col[rp[2 * (c0q + q) + 1]];
      s_r[2 * q] = rc[c0q + q];
      s_r[2 * q + 1] = rsn[c0q + q];
    }
    __syncthreads();
    if (row >= m) continue;
    for (int q0 = 0; q0 < nch; q0 += RQ) {
      const int nq = min(RQ, nch - q0);
      // both columns of the next RQ rotations in one round of loads: a
      // rotation's first column is the running column of its chain (then
      // the register value replaces the loaded one) or a column that no
      // rotation has written yet, and its second column is a fresh scan
      // position, so the loads read current values
      double y[RQ], xp[RQ];
#pragma unroll
      for (int u = 0; u < RQ; ++u)
        if (u < nq) {
          y[u] = Zr[(int64_t)s_c[2 * (q0 + u) + 1] * n];
          xp[u] = Zr[(int64_t)s_c[2 * (q0 + u)] * n];
        }
#pragma unroll
      for (int u = 0; u < RQ; ++u)
        if (u < nq) {
          const int q = q0 + u;
          const int c0 = s_c[2 * q];
          double x;
          if (c0 == cc) {
            x = cvl;
          } else {
            if (cc >= 0) Zr[(int64_t)cc * n] = cvl;
            x = xp[u];
          }
          const double r = s_r[2 * q], sn = s_r[2 * q + 1];
          Zr[(int64_t)c0 * n] = r * x + sn * y[u];
          cvl = r * y[u] - sn * x;
          cc = s_c[2 * q + 1];
        }
    }
  }
  if (row < m && cc >= 0) Zr[(int64_t)cc * n] = cvl;
}

// Gather the rotated merge columns as [non-deflated (grouped order) |
// deflated] into Qcd (m x m, ld m).  Grid: (row chunks of 64, column chunks of DC_CCH,
// merges); each workgroup owns 64 rows and DC_CCH columns of one merge, its
// 4 waves take every 4th column (a top-level merge is thousands of
// workgroups, enough to stream the m^2 copy at bandwidth).
constexpr int DC_CCH = 256;
__global__ __launch_bounds__(256) void k_dc_rotgather(DcArgs a, const int32_t* tasks, int ntask) {
  const int t = blockIdx.z;
  const int32_t* tk = tasks + 8 * t;
  const int b = tk[0], s = tk[1], e = tk[3];
  const int m = e - s;
  const int row0 = blockIdx.x * 64;
  const int c0 = blockIdx.y * DC_CCH;
  if (row0 >= m || c0 >= m) return;
  const int c1 = min(m, c0 + DC_CCH);
  const int n = a.n[b];
  const int k = a.kout[t];
  double* Z = a.Z + a.aoff[b];
  double* W = a.ws + join64(tk[4], tk[5]);
  const int32_t* I = a.iws + join64(tk[6], tk[7]);
  const int32_t *col = I, *nd = I + m, *df = I + 2 * m, *inv = I + 10 * m + 8;
  double* Qcd = W + dws_vec(m);
  const int tid = threadIdx.x;
  const int rl = tid & 63, g = tid >> 6;
  const int row = row0 + rl;
  if (row >= m) return;
#pragma unroll 8
  for (int c = c0 + g; c < c1; c += 4) {
    const int srcc = (c < k) ? col[nd[inv[c]]] : col[df[c - k]];
    Qcd[row + (int64_t)c * m] = Z[(int64_t)(s + row) + (int64_t)srcc * n];
  }
}

// Secular equation for tau = lambda - d_origin: one wave per root, four
// roots per workgroup.  The 64 lanes split every pole sum (lane l takes poles
// l, l+64, ...) and combine it with an xor butterfly, which leaves the same
// bits in every lane, so the iteration below runs uniformly across the wave.
// Poles and weights are staged in LDS when they fit.
constexpr int SEC_ROOTS = 4;

__device__ inline void sec_sums(const double* dl, const double* zl, int k, int i, double dorg, double tau,
                                double rho, double& psi, double& dpsi, double& phi, double& dphi) {
  const int lane = threadIdx.x & 63;
  double a0 = 0.0, a1 = 0.0, b0 = 0.0, b1 = 0.0;
  for (int j = lane; j < k; j += 64) {
    const double r = 1.0 / ((dl[j] - dorg) - tau);
    const double w2 = rho * zl[j] * zl[j];
    const double wr = w2 * r;
    if (j <= i) { a0 += wr; a1 += wr * r; }
    else { b0 += wr; b1 += wr * r; }
  }
  psi = wave_sum(a0);
  dpsi = wave_sum(a1);
  phi = wave_sum(b0);
  dphi = wave_sum(b1);
}

__device__ inline double sec_value(const double* dl, const double* zl, int k, double dorg, double x,
                                   double rho) {
  const int lane = threadIdx.x & 63;
  double f = 0.0;
  for (int j = lane; j < k; j += 64) f += rho * zl[j] * zl[j] / ((dl[j] - dorg) - x);
  return 1.0 + wave_sum(f);
}

__global__ __launch_bounds__(64 * SEC_ROOTS) void k_dc_secular(DcArgs a, const int32_t* tasks, int ntask) {
  extern __shared__ double sec_shm[];
  const int t = blockIdx.y;
  const int k = a.kout[t];
  if (blockIdx.x * SEC_ROOTS >= k) return;
  const int i = blockIdx.x * SEC_ROOTS + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int32_t* tk = tasks + 8 * t;
  const int m = tk[3] - tk[1];
  const int64_t wo = join64(tk[4], tk[5]);
  const int64_t io = join64(tk[6], tk[7]);
  double* W = a.ws + wo;
  const double *dl = W + 2 * m, *zl = W + 3 * m;
  if (k <= SEC_LDS_M) {
    for (int j = threadIdx.x; j < k; j += 64 * SEC_ROOTS) { sec_shm[j] = dl[j]; sec_shm[k + j] = zl[j]; }
    __syncthreads();
    dl = sec_shm;
    zl = sec_shm + k;
  }
  if (i >= k) return;
  double* tr = W + 6 * m;
  int32_t* org = a.iws + io + 3 * m;
  const double rho = W[5 * m];
  int o;
  double lo, hi;
  if (i < k - 1) {
    const double gap = dl[i + 1] - dl[i];
    const double half = 0.5 * gap;
    const double f = sec_value(dl, zl, k, dl[i], half, rho);
    if (f > 0.0) { o = i; lo = 0.0; hi = half; }
    else { o = i + 1; lo = -(gap - half); hi = 0.0; }
  } else {
    o = k - 1; lo = 0.0; hi = rho;
    // make sure g(hi) > 0 (sum z^2 = 1 up to rounding)
    for (int it = 0; it < 8; ++it) {
      if (sec_value(dl, zl, k, dl[o], hi, rho) > 0.0) break;
      hi *= 2.0;
    }
  }
  // Safeguarded rational iteration (Bunch, Nielsen & Sorensen 1978): the
  // sums over poles left and right of the root are each replaced by one pole
  // term matching value and slope at tau; the two-pole model is solved in
  // closed form.  Steps leaving the bracket [lo, hi] fall back to bisection.
  const double dorg = dl[o];
  const double p = dl[i] - dorg;                       // left pole (shifted)
  const bool last = (i == k - 1);
  const double q = last ? 0.0 : dl[i + 1] - dorg;      // right pole (shifted)
  const double eps = 2.220446049250313e-16;
  double tau = lo + 0.5 * (hi - lo);
  for (int it = 0; it < 200; ++it) {
    double psi, dpsi, phi, dphi;
    sec_sums(dl, zl, k, i, dorg, tau, rho, psi, dpsi, phi, dphi);
    const double g = 1.0 + psi + phi;
    // converged once g is at the level of its own rounding error (the test
    // of LAPACK dlaed4, in this scaling: psi <= 0 <= phi, rho folded in)
    const double errb = 8.0 * (phi - psi) + 2.0 + 3.0 * fabs(tau) * (dpsi + dphi);
    if (fabs(g) <= eps * errb) break;
    if (g > 0.0) hi = tau; else lo = tau;
    // model: C + B1/(p-x) + B2/(q-x) = 0, u = p - x
    const double B1 = dpsi * (p - tau) * (p - tau);
    double C = 1.0 + psi - B1 / (p - tau);
    double xnew;
    if (last) {
      C += phi;
      xnew = (C != 0.0) ? p + B1 / C : 0.5 * (lo + hi);
    } else {
      const double B2 = dphi * (q - tau) * (q - tau);
      C += phi - B2 / (q - tau);
      const double G = q - p;
      // solved for the distance to the origin pole (p = 0 when o = i, q = 0
      // when o = i + 1), so that the new point carries no cancellation
      if (o == i) {
        // u = p - x: C u^2 + (C G + B1 + B2) u + B1 G = 0, u in (-G, 0)
        const double qa = C, qb = C * G + B1 + B2, qc = B1 * G;
        double u;
        if (qa == 0.0) {
          u = -qc / qb;
        } else {
          const double disc = fmax(qb * qb - 4.0 * qa * qc, 0.0);
          const double qq = -0.5 * (qb + copysign(sqrt(disc), qb));
          const double u1 = qq / qa, u2 = (qq != 0.0) ? qc / qq : u1;
          u = (u1 < 0.0 && u1 > -G) ? u1 : u2;
        }
        xnew = p - u;
      } else {
        // w = q - x: C w^2 + (B1 + B2 - C G) w - B2 G = 0, w in (0, G)
        const double qa = C, qb = B1 + B2 - C * G, qc = -B2 * G;
        double w;
        if (qa == 0.0) {
          w = -qc / qb;
        } else {
          const double disc = fmax(qb * qb - 4.0 * qa * qc, 0.0);
          const double qq = -0.5 * (qb + copysign(sqrt(disc), qb));
          const double w1 = qq / qa, w2 = (qq != 0.0) ? qc / qq : w1;
          w = (w1 > 0.0 && w1 < G) ? w1 : w2;
        }
        xnew = q - w;
      }
    }
    // a step below the rounding level of tau: converged (tested before the
    // bracket, which a converged step may touch)
    if (fabs(xnew - tau) <= 2.0 * eps * fabs(tau)) break;
    if (!(xnew > lo && xnew < hi)) xnew = lo + 0.5 * (hi - lo);
    if (xnew <= lo || xnew >= hi) break;          // bracket exhausted
    const double step = fabs(xnew - tau);
    tau = xnew;
    if (step <= 2.0 * eps * fabs(tau)) break;
  }
  if (tau == 0.0) tau = (o == i) ? hi : lo;     // never sit on the pole
  if (lane == 0) {
    tr[i] = tau;
    org[i] = o;
  }
}

// z-hat of the non-deflated roots (Gu-Eisenstat): one wave per root; the
// lanes split the k-1 ratio factors of the product and combine their
// partial products by a butterfly, so the serial chain is k/64 factors.
__global__ __launch_bounds__(256) void k_dc_zhat(DcArgs a, const int32_t* tasks, int ntask) {
  const int t = blockIdx.y;
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int k = a.kout[t];
  if (j >= k) return;   // uniform per wave
  const int32_t* tk = tasks + 8 * t;
  const int m = tk[3] - tk[1];
  double* W = a.ws + join64(tk[4], tk[5]);
  const int32_t* org = a.iws + join64(tk[6], tk[7]) + 3 * m;
  const double *dl = W + 2 * m, *zl = W + 3 * m, *tr = W + 6 * m;
  double* zh = W + 7 * m;   // rc slot is free after the rotations
  const double rho = W[5 * m];
  const double dj = dl[j];
  // factor i < k-1: (lambda_i - d_j) / (d_i - d_j) for i < j, over d_{i+1} - d_j for i >= j
  double p = 1.0;
  for (int i = lane; i < k - 1; i += 64) {
    const double den = (i < j) ? (dl[i] - dj) : (dl[i + 1] - dj);
    p *= ((dl[org[i]] - dj) + tr[i]) / den;
  }
  for (int o = 32; o > 0; o >>= 1) p *= __shfl_xor(p, o, 64);
  p *= (dl[org[k - 1]] - dj) + tr[k - 1];
  p /= rho;
  if (lane == 0) zh[j] = copysign(sqrt(fabs(p)), zl[j]);
}

// eigenvectors of the rank-one problem: one wave per column; the rows
// follow the grouped column order of Qcd.
__global__ __launch_bounds__(64) void k_dc_vectors(DcArgs a, const int32_t* tasks, int ntask) {
  const int t = blockIdx.y;
  const int i = blockIdx.x;
  const int k = a.kout[t];
  if (i >= k) return;
  const int32_t* tk = tasks + 8 * t;
  const int m = tk[3] - tk[1];
  double* W = a.ws + join64(tk[4], tk[5]);
  const int32_t* org = a.iws + join64(tk[6], tk[7]) + 3 * m;
  const int32_t* perm = a.iws + join64(tk[6], tk[7]) + 9 * m + 8;
  const double *dl = W + 2 * m, *tr = W + 6 * m, *zh = W + 7 * m;
  double* U = W + dws_vec(m) + (int64_t)m * m;
  const double dorg = dl[org[i]], ti = tr[i];
  double ss = 0.0;
  for (int j = threadIdx.x; j < k; j += 64) {
    const double v = zh[j] / ((dl[j] - dorg) - ti);
    U[perm[j] + (int64_t)i * k] = v;   // row of pole j in the grouped order
    ss += v * v;
  }
  ss = wave_sum(ss);
  const double inv = 1.0 / sqrt(ss);
  for (int j = threadIdx.x; j < k; j += 64) U[perm[j] + (int64_t)i * k] *= inv;
}

constexpr int FIN_T = 1024;   // threads of the final-ordering kernel (its bitonic sort)

// final ordering of one merge: lambda (ascending) merged with the sorted
// deflated values; columns from Qt (non-deflated) and Qcd (deflated).
__global__ __launch_bounds__(FIN_T) void k_dc_finish(DcArgs a, const int32_t* tasks, int ntask) {
  const int t = blockIdx.x;
  const int32_t* tk = tasks + 8 * t;
  const int b = tk[0], s = tk[1], e = tk[3];
  const int m = e - s;
  const int k = a.kout[t];
  const int n = a.n[b];
  const int64_t vo = a.voff[b];
  double* W = a.ws + join64(tk[4], tk[5]);
  int32_t* I = a.iws + join64(tk[6], tk[7]);
  const int32_t* org = I + 3 * m;
  int32_t* dord = I + 6 * m;
  const double *dl = W + 2 * m, *dv = W + 4 * m, *tr = W + 6 * m;
  double* sdv = W + 8 * m;  // sorted deflated values (rs slot is free now)
  const double* Qcd = W + dws_vec(m);
  const double* Qt = W + dws_vec(m) + 2 * (int64_t)m * m;
  double* Z = a.Z + a.aoff[b];
  const int md = m - k;
  const int tid = threadIdx.x;
  // sort of the deflated values, ties by index: (value, index) pairs in
  // LDS, bitonic network over the next power of two (padding +inf / INT_MAX
  // sorts last).  Merges too large for LDS rank each value against all.
  extern __shared__ double fin_shm[];
  if (md <= DC_LDS_M) {
    int P = 1;
    while (P < md) P <<= 1;
    double* kv = fin_shm;
    int32_t* ki = reinterpret_cast<int32_t*>(fin_shm + P);
    for (int i = tid; i < P; i += FIN_T) {
      kv[i] = (i < md) ? dv[i] : __builtin_inf();
      ki[i] = (i < md) ? i : INT_MAX;
    }
    __syncthreads();
    for (int len = 2; len <= P; len <<= 1) {
      for (int st = len >> 1; st > 0; st >>= 1) {
        for (int x = tid; x < P / 2; x += FIN_T) {
          // pair (i, i + st) with i having bit st clear; direction by bit len
          const int i = ((x & ~(st - 1)) << 1) | (x & (st - 1));
          const int j = i + st;
          const bool up = ((i & len) == 0);
          const double vi = kv[i], vj = kv[j];
          const int ii = ki[i], ij = ki[j];
          const bool gt = (vi > vj) || (vi == vj && ii > ij);   // (i) after (j)
          if (gt == up) {
            kv[i] = vj; kv[j] = vi;
            ki[i] = ij; ki[j] = ii;
          }
        }
        __syncthreads();
      }
    }
    for (int r = tid; r < md; r += FIN_T) {
      dord[r] = ki[r];
      sdv[r] = kv[r];
    }
  } else {
    for (int i = tid; i < md; i += FIN_T) {
      const double v = dv[i];
      int r = 0;
      for (int j = 0; j < md; ++j) {
        const double u = dv[j];
        r += (u < v) || (u == v && j < i);
      }
      dord[r] = i;
      sdv[r] = v;
    }
  }
  __syncthreads();
  // eigenvalues: lambda_i then deflated, merged by rank
  for (int i = tid; i < m; i += FIN_T) {
    double v;
    int fin, src;
    if (i < k) {
      v = dl[org[i]] + tr[i];
      int lo = 0, hi = md;
      while (lo < hi) { int mm = (lo + hi) >> 1; if (sdv[mm] < v) lo = mm + 1; else hi = mm; }
      fin = i + lo;
      src = -1 - i;
    } else {
      const int j = i - k;
      v = sdv[j];
      int lo = 0, hi = k;
      while (lo < hi) {
        int mm = (lo + hi) >> 1;
        const double lv = dl[org[mm]] + tr[mm];
        if (lv <= v) lo = mm + 1; else hi = mm;
      }
      fin = j + lo;
      src = dord[j];
    }
    a.w[vo + s + fin] = v;
    I[5 * m + fin] = src;   // rp slot reused: source column per final position
  }
  (void)Qcd; (void)Qt; (void)Z; (void)n;
}

// Write the merged eigenvectors in final (ascending) order into the merge's
// square region of Z.  Grid: (row chunks of 64, column chunks of DC_CCH,
// merges).
__global__ __launch_bounds__(256) void k_dc_scatter(DcArgs a, const int32_t* tasks, int ntask) {
  const int t = blockIdx.z;
  const int32_t* tk = tasks + 8 * t;
  const int b = tk[0], s = tk[1], e = tk[3];
  const int m = e - s;
  const int row = blockIdx.x * 64 + (threadIdx.x & 63);
  const int c0 = blockIdx.y * DC_CCH;
  if (row >= m || c0 >= m) return;
  const int c1 = min(m, c0 + DC_CCH);
  const int g = threadIdx.x >> 6;
  const int k = a.kout[t];
  const int n = a.n[b];
  double* Z = a.Z + a.aoff[b];
  const double* W = a.ws + join64(tk[4], tk[5]);
  const int32_t* I = a.iws + join64(tk[6], tk[7]);
  const double* Qcd = W + dws_vec(m);
  const double* Qt = W + dws_vec(m) + 2 * (int64_t)m * m;
#pragma unroll 8
  for (int c = c0 + g; c < c1; c += 4) {
    const int src = I[5 * m + c];
    const double v = (src < 0) ? Qt[row + (int64_t)(-1 - src) * m] : Qcd[row + (int64_t)(k + src) * m];
    Z[(int64_t)(s + row) + (int64_t)(s + c) * n] = v;
  }
}

// Deferred top merges.  A block's last merge (the whole block) keeps its
// workspace until the columns are asked for; final column c of the block is
// Qcd[:, :k] U[:, i] when it is root i of the secular equation
// (I[5m + c] = -1 - i) or the deflated column Qcd[:, k + src] (src >= 0).
// Task record (12 ints): b, lo, nc, k, m, pad, wo (2), io (2), uoff (2).
struct DcTop {
  const int32_t* n;
  const int64_t* aoff;
  const double* ws;
  const int32_t* iws;
  double* Z;
  double* usel;
};

// Usel (k x nc, ld k): column j = U[:, i] for a root, zero for a deflated
// column.  Grid: (chunks of 256 rows of U, nc, tasks).
__global__ __launch_bounds__(256) void k_dc_top_select(DcTop a, const int32_t* tasks) {
  const int32_t* tk = tasks + 12 * blockIdx.z;
  const int lo = tk[1], nc = tk[2], k = tk[3], m = tk[4];
  const int j = blockIdx.y;
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (j >= nc || r >= k) return;
  const double* W = a.ws + join64(tk[6], tk[7]);
  const int32_t* I = a.iws + join64(tk[8], tk[9]);
  const int src = I[5 * m + lo + j];
  double* us = a.usel + join64(tk[10], tk[11]);
  us[r + (int64_t)j * k] = (src < 0) ? W[dws_vec(m) + (int64_t)m * m + r + (int64_t)(-1 - src) * k] : 0.0;
}

// Deflated columns into Z (after the product wrote the root columns).
// Grid: (chunks of 256 rows, nc, tasks).
__global__ __launch_bounds__(256) void k_dc_top_deflated(DcTop a, const int32_t* tasks) {
  const int32_t* tk = tasks + 12 * blockIdx.z;
  const int b = tk[0], lo = tk[1], nc = tk[2], k = tk[3], m = tk[4];
  const int j = blockIdx.y;
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (j >= nc || r >= m) return;
  const int32_t* I = a.iws + join64(tk[8], tk[9]);
  const int src = I[5 * m + lo + j];
  if (src < 0) return;
  const double* W = a.ws + join64(tk[6], tk[7]);
  const int n = a.n[b];
  a.Z[a.aoff[b] + r + (int64_t)(lo + j) * n] = W[dws_vec(m) + r + (int64_t)(k + src) * m];
}

// ===========================================================================
// 3. back-transform
// ===========================================================================
// Reflector panels of all blocks: panel p (reflectors c0 = p NBB ..
// c0 + NBB - 1) of block b has its explicit V (rows c0+1..n-1, ld n) at
// vpo[b] + p n NBB and its [G | T] pair (2 NBB^2) at index tpo[b] + p.
struct BtArgs {
  const int32_t* n;
  const int64_t* aoff;
  const int64_t* voff;
  const double* A;
  const double* tau;
  double* V;
  double* T;
  const int64_t* vpo;
  const int32_t* tpo;
};

// Explicit reflector panel V (rows c0+1..n-1).  Grid: (chunks, panels);
// panel i of the list is block pl[2i], panel pl[2i+1].
__global__ void k_bt_buildv(BtArgs a, const int32_t* pl, int off) {
  const int b = pl[2 * blockIdx.y], p = pl[2 * blockIdx.y + 1];
  const int n = a.n[b];
  const int c0 = p * NBB;
  if (c0 >= n - off) return;
  const int nbp = min(NBB, n - off - c0);
  const int mr = n - c0 - off;
  const double* A = a.A + a.aoff[b];
  double* V = a.V + a.vpo[b] + (int64_t)p * n * NBB;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < mr * NBB; q += gridDim.x * blockDim.x) {
    const int r = q % mr, i = q / mr;
    const int row = c0 + off + r, c = c0 + i;
    double v = 0.0;
    if (i < nbp) {
      if (row == c + off) v = 1.0;
      else if (row > c + off) v = A[row + (int64_t)c * n];
    }
    V[r + (int64_t)i * n] = v;
  }
}

// T of the compact WY form from the Gram matrix G = V^T V, built column by
// column in LDS (NBB^2 doubles; gfx950 gives a workgroup up to 160 KB):
// T(0:i, i) = T(0:i, 0:i) * (-tau_i G(0:i, i)),  T(i, i) = tau_i.
// Row j of the product is shared by the TQ = 4 adjacent lanes 4j..4j+3
// (terms l = j + sub, j + sub + 4, ...), combined with two xor shuffles.
constexpr int TQ = 4;
__global__ __launch_bounds__(NBB * TQ) void k_bt_tmat(BtArgs a, const int32_t* pl, int off) {
  const int b = pl[2 * blockIdx.x], p = pl[2 * blockIdx.x + 1];
  const int n = a.n[b];
  const int c0 = p * NBB;
  if (c0 >= n - off) return;
  const int nbp = min(NBB, n - off - c0);
  const double* G = a.T + (int64_t)(a.tpo[b] + p) * 2 * NBB * NBB;
  double* T = a.T + (int64_t)(a.tpo[b] + p) * 2 * NBB * NBB + NBB * NBB;
  const double* tau = a.tau + a.voff[b] + c0;
  __shared__ double sT[NBB * NBB];
  __shared__ double tmp[NBB];
  const int j = threadIdx.x / TQ, sub = threadIdx.x % TQ;   // row j, term phase sub
  // G(j, i) is fetched one column ahead so that its latency overlaps the
  // previous column's work
  double gji = 0.0;
  for (int i = 0; i < nbp; ++i) {
    const double gcur = gji;
    if (sub == 0 && i + 1 < nbp && j < i + 1) gji = G[j + (i + 1) * NBB];
    if (sub == 0 && j < i) tmp[j] = -tau[i] * gcur;
    __syncthreads();
    double acc = 0.0;
    if (j < i)
      for (int l = j + sub; l < i; l += TQ) acc += sT[j + l * NBB] * tmp[l];
    acc += __shfl_xor(acc, 1, 64);
    acc += __shfl_xor(acc, 2, 64);
    if (sub == 0) {
      if (j < i) sT[j + i * NBB] = acc;
      else if (j == i) sT[i + i * NBB] = tau[i];
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < NBB * NBB; e += NBB * TQ) {
    const int r = e % NBB, i = e / NBB;
    T[e] = (i < nbp && r <= i) ? sT[r + i * NBB] : 0.0;
  }
}

// ===========================================================================
// chop |U| < 1e-14 (diagonalization.F90:273-280)
// ===========================================================================
// Grid: (chunks, ranges); range r is rng[2r] (start) and rng[2r+1] (length).
__global__ void k_chop(double* Z, const int64_t* rng, double thr) {
  double* z = Z + rng[2 * blockIdx.y];
  const int64_t n = rng[2 * blockIdx.y + 1];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    if (fabs(z[i]) < thr) z[i] = 0.0;
}

// ===========================================================================
// host driver
// ===========================================================================
static double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void BatchedEigensolver::stats(EigStats* st) {
  if (!st || !ev_st_[3]) return;
  NRG_HIP(hipEventSynchronize(ev_st_[3]));
  float a = 0.f, b = 0.f, c = 0.f;
  NRG_HIP(hipEventElapsedTime(&a, ev_st_[0], ev_st_[1]));
  NRG_HIP(hipEventElapsedTime(&b, ev_st_[1], ev_st_[2]));
  NRG_HIP(hipEventElapsedTime(&c, ev_st_[2], ev_st_[3]));
  st->ms_tridiag = a;
  st->ms_dc = b;
  st->ms_back = c;
}

void BatchedEigensolver::release() {
  for (TrGraph& e : graphs_) (void)hipGraphExecDestroy(e.exec);
  graphs_.clear();
  d_n_.release(); d_order_.release();
  for (int g = 0; g < MAX_TR_GROUPS; ++g) {
    d_slotg_[g].release();
    d_fin_.release();
    gemmg_[g].release();
    if (sg_[g]) {
      (void)hipStreamSynchronize(sg_[g]);
      (void)hipStreamDestroy(sg_[g]);
      (void)hipEventDestroy(evg_[g]);
      sg_[g] = nullptr;
      evg_[g] = nullptr;
    }
  }
  if (ev_fork_) { (void)hipEventDestroy(ev_fork_); ev_fork_ = nullptr; }
  if (ev_bt_) { (void)hipEventDestroy(ev_bt_); ev_bt_ = nullptr; }
  bt_wait_ = false;
  d_vpo_.release(); d_tpo_.release(); d_plist_.release(); gemm_bt_.release();
  for (int g = 0; g < MAX_TR_GROUPS; ++g)
    if (evf_[g]) { (void)hipEventDestroy(evf_[g]); evf_[g] = nullptr; }
  for (hipEvent_t& e : ev_st_)
    if (e) { (void)hipEventDestroy(e); e = nullptr; }
  d_aoff_.release(); d_voff_.release();
  d_d_.release(); d_e_.release(); d_tau_.release(); d_y_.release(); d_P_.release(); d_part_.release();
  d_ws_.release(); d_iws_.release(); d_tasks_.release(); d_k_.release(); d_usel_.release(); d_mtask_.release();
  d_V_.release(); d_T_.release(); d_W_.release(); d_X2_.release(); d_rng_.release(); gemm_.release();
  nblk_ = 0;
}

void BatchedEigensolver::solve(int nblk, const int* n, const int64_t* aoff, const int64_t* voff,
                               double* A, double* Z, double* w, bool defer_back, hipStream_t s,
                               EigStats* st) {
  nblk_ = nblk;
  n_.assign(n, n + nblk);
  aoff_.assign(aoff, aoff + nblk);
  voff_.assign(voff, voff + nblk);
  A_ = A; Z_ = Z; w_ = w;
  tot_n_ = 0; tot_n2_ = 0;
  int nmax = 0;
  for (int b = 0; b < nblk; ++b) {
    NRG_REQUIRE(n_[b] >= 1 && n_[b] <= MAXN, "eigensolver: block size out of range");
    tot_n_ += n_[b];
    tot_n2_ += (int64_t)n_[b] * n_[b];
    nmax = std::max(nmax, n_[b]);
  }
  if (nblk == 0) return;
  order_.resize(nblk);
  for (int b = 0; b < nblk; ++b) order_[b] = b;
  std::stable_sort(order_.begin(), order_.end(), [&](int x, int y) { return n_[x] > n_[y]; });
  d_n_.upload(n_.data(), nblk, s);
  d_order_.upload(order_.data(), nblk, s);
  d_aoff_.upload(aoff_.data(), nblk, s);
  d_voff_.upload(voff_.data(), nblk, s);
  d_d_.reserve(tot_n_);
  d_e_.reserve(tot_n_);
  d_tau_.reserve(tot_n_);
  fill(d_e_.p, 0.0, tot_n_, s);
  fill(d_tau_.p, 0.0, tot_n_, s);
  // Stage times come from events on the caller's stream, read by
  // stats() once the caller has waited for the device: timing adds no
  // synchronisation, so the host issues the next stage's launches while the
  // device still runs the previous one.
  if (st) {
    for (hipEvent_t& e : ev_st_)
      if (!e) NRG_HIP(hipEventCreate(&e));
    NRG_HIP(hipEventRecord(ev_st_[0], s));
  }
  const double t0 = now_ms();
  solved_two_stage_ = two_stage_;
  bt_off_ = two_stage_ ? 32 : 1;
  if (two_stage_) tridiag_two_stage(s);
  else tridiag(s);
  if (st) {
    st->ms_tridiag_issue = now_ms() - t0;
    NRG_HIP(hipEventRecord(ev_st_[1], s));
  }
  // the back-transform's panel factors overlap the divide and conquer
  if (sg_[1]) {
    NRG_HIP(hipEventRecord(ev_fork_, s));
    NRG_HIP(hipStreamWaitEvent(sg_[1], ev_fork_, 0));
    panel_factors(sg_[1]);
    if (!ev_bt_) NRG_HIP(hipEventCreateWithFlags(&ev_bt_, hipEventDisableTiming));
    NRG_HIP(hipEventRecord(ev_bt_, sg_[1]));
    bt_wait_ = true;
  } else {
    panel_factors(s);
    bt_wait_ = false;
  }
  divide_conquer(s);
  if (st) NRG_HIP(hipEventRecord(ev_st_[2], s));
  done_.assign(nblk, 0);
  if (!defer_back) back_columns(nullptr, s, nullptr);
  if (st) {
    NRG_HIP(hipEventRecord(ev_st_[3], s));
    st->max_n = nmax;
    st->nblk = nblk;
    double f = 0;
    for (int b = 0; b < nblk; ++b) f += (10.0 / 3.0) * (double)n_[b] * n_[b] * n_[b];
    st->flops_alg = f;
  }
}

// Whether a SYMV launch over trailing blocks of up to m rows fits the
// 160 KB LDS of a gfx950 workgroup with `per_row` doubles of dynamic LDS per
// row (2: row parts and v, 1: row parts only), counting the kernel's static
// LDS as the runtime reports it.
static size_t symv_static_lds() {
  static size_t st = 0;
  static bool done = false;
  if (!done) {
    hipFuncAttributes fa{}, fb{};
    NRG_HIP(hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(k_tr_symv<2, 16, SYMV_U, 1>)));
    NRG_HIP(hipFuncGetAttributes(&fb, reinterpret_cast<const void*>(k_tr_symv<1, 16, SYMV_U, 0>)));
    st = std::max(fa.sharedSizeBytes, fb.sharedSizeBytes);
    done = true;
  }
  return st;
}
static bool symv_lds_fits(int m, int per_row) {
  return (size_t)per_row * (size_t)std::max(m, 1) * sizeof(double) + symv_static_lds() <= 160 * 1024;
}
// the same for the row-chunk SYMV, whose only dynamic LDS is v (VL = 1)
static bool symv_rows_v_fits(int m) {
  static size_t st = 0;
  if (!st) {
    hipFuncAttributes fa{}, fb{};
    NRG_HIP(hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(k_tr_symv_rows<2, 1>)));
    NRG_HIP(hipFuncGetAttributes(&fb, reinterpret_cast<const void*>(k_tr_symv_rows<1, 1>)));
    st = std::max(fa.sharedSizeBytes, fb.sharedSizeBytes);
  }
  return (size_t)std::max(m, 1) * sizeof(double) + st <= 160 * 1024;
}

void BatchedEigensolver::tridiag(hipStream_t s) {
  const int nblk = nblk_;
  const int nmax = n_[order_[0]];
  std::vector<int64_t> poff(3 * nblk);
  int64_t pacc = 0, yacc = 0;
  for (int b = 0; b < nblk; ++b) {
    poff[b] = voff_[b] * 4 * NB;
    poff[nblk + b] = pacc;
    pacc += part_size(n_[b]);
    poff[2 * nblk + b] = yacc;
    yacc += (int64_t)((n_[b] + CG - 1) / CG) * n_[b];
  }
  d_y_.reserve(std::max<int64_t>(yacc, 1));
  d_P_.reserve(tot_n_ * 4 * NB);
  d_part_.reserve(pacc);
  // Column schedule.  Step t processes column c_q = t - coff[q] of every
  // started slot q.  With end alignment the slots start at multiples of NB
  // so that all blocks reach their last columns together: the light steps
  // at the end of a block (small trailing matrices, latency-bound launches)
  // are shared by all blocks instead of being left to the largest one, and
  // every slot sees the same panel phase t mod NB.
  std::vector<int32_t> coff(nblk);
  for (int q = 0; q < nblk; ++q) coff[q] = align_end_ ? ((nmax - n_[order_[q]]) / NB) * NB : 0;
  NRG_REQUIRE(symv_kind_ != 0 || symv_lds_fits(nmax - 1, 1),
              "eigensolver: block too large for the SYMV stage (its row-part vector exceeds the LDS)");
  int T = 0;
  for (int q = 0; q < nblk; ++q) T = std::max(T, coff[q] + n_[order_[q]]);
  // The column steps stop at the panel boundary t0 from which every block's
  // remaining trailing block has at most finish_m_ rows; the tail finisher
  // reduces those in LDS (finish_m_ = 0: column steps to the end).
  const int t0 = (finish_m_ > 0) ? std::max(0, ceil_div(T - finish_m_, NB) * NB) : T;

  // Groups of blocks (balanced by sum n^3, below) follow the same schedule
  // on their own streams.  Each step of a block is a
  // latency-bound column kernel followed by a bandwidth-bound SYMV; with the
  // groups on separate streams the column kernel of one group runs while the
  // SYMV of another streams its trailing matrices.
  const int ng = std::max(1, std::min(groups_, std::min(nblk, MAX_TR_GROUPS)));
  struct Group {
    std::vector<int> q;        // global slots of this group, size order
    std::vector<int32_t> coff;
    TrArgs ta;
    GemmRunner* gemm;
    GemmBatch* batch;
    DevBuf<char>* d_slot;
    hipStream_t s;
    int pw_prev = CG;          // panel width of the previous column's SYMV launch
    std::vector<size_t> syr2k; // recorded trailing-update launches of panel t / NB: [syr2k[i], syr2k[i+1])
  };
  Group grp[MAX_TR_GROUPS];
  for (int g = 1; g < ng; ++g)
    if (!sg_[g]) {
      NRG_HIP(hipStreamCreateWithFlags(&sg_[g], hipStreamNonBlocking));
      NRG_HIP(hipEventCreateWithFlags(&evg_[g], hipEventDisableTiming));
    }
  if (ng > 1 && !ev_fork_) NRG_HIP(hipEventCreateWithFlags(&ev_fork_, hipEventDisableTiming));
  for (int g = 0; g < ng && ng > 1; ++g)
    if (!evf_[g]) NRG_HIP(hipEventCreateWithFlags(&evf_[g], hipEventDisableTiming));
  {
    // Blocks in size order go to the group with the least n^3 so far (the
    // SYMV bytes of a block are ~(4/3) n^3): the groups run independently
    // and share the bandwidth, so the heaviest sets the end.  Each group's
    // slots stay in size order (start steps non-decreasing).
    std::vector<double> load(ng, 0.0);
    for (int q = 0; q < nblk; ++q) {
      int g = 0;
      for (int h = 1; h < ng; ++h)
        if (load[h] < load[g]) g = h;
      const double nq = n_[order_[q]];
      load[g] += nq * nq * nq;
      grp[g].q.push_back(q);
    }
  }
  for (int g = 0; g < ng; ++g) {
    Group& G = grp[g];
    std::vector<TrSlot> slots(G.q.size());
    for (size_t i = 0; i < G.q.size(); ++i) {
      const int q = G.q[i], b = order_[q];
      G.coff.push_back(coff[q]);
      TrSlot& r = slots[i];
      r.n = n_[b]; r.coff = coff[q]; r.b = b; r.pad = 0;
      r.aoff = aoff_[b]; r.voff = voff_[b];
      r.pP = poff[b]; r.pRec = poff[nblk + b]; r.pPy = poff[2 * nblk + b]; r.pad2 = 0;
    }
    G.gemm = &gemmg_[g];
    G.batch = &batchg_[g];
    G.d_slot = &d_slotg_[g];
    G.s = g ? sg_[g] : s;
    G.d_slot->upload(reinterpret_cast<const char*>(slots.data()), slots.size() * sizeof(TrSlot), s);
    G.ta = TrArgs{reinterpret_cast<const TrSlot*>(G.d_slot->p), A_, d_d_.p, d_e_.p, d_tau_.p, d_y_.p, d_P_.p,
                  d_part_.p};
  }
  // trailing updates [V W][W V]^T of panel boundary t (t = NB, 2 NB, ...):
  // every group's products are recorded up front and their task tables
  // uploaded once; the step loop issues each panel's launches in place
  auto syr2k_batch = [&](Group& G, int t) {
    G.batch->clear();
    for (size_t i = 0; i < G.q.size(); ++i) {
      const int b = order_[G.q[i]];
      const int nb_ = n_[b];
      const int c = t - G.coff[i];
      const int mt = nb_ - c;
      if (c <= 0 || mt <= 0) continue;
      GemmTask tk = gemm_task(mt, mt, 2 * NB, 0, 1,
                              poff[b] + c, nb_,                       // [V W] rows c..
                              poff[b] + c + (int64_t)2 * NB * nb_, nb_, // [W V] rows c..
                              aoff_[b] + c + (int64_t)c * nb_, nb_, -1.0, 1.0);
      tk.lower = 1;   // only the lower triangle is read from here on
      G.batch->add(tk);
    }
  };
  for (int g = 0; g < ng; ++g) {
    Group& G = grp[g];
    G.gemm->begin();
    G.syr2k.assign(1, 0);
    for (int t = NB; t < T && t <= t0; t += NB) {
      syr2k_batch(G, t);
      G.gemm->run(*G.batch, d_P_.p, d_P_.p, A_, nullptr, nullptr, nullptr, G.s, KF_GEMM_TRIDIAG);
      G.syr2k.push_back(G.gemm->recorded());
    }
    G.gemm->upload_recorded(s);   // on the caller's stream, before the fork
  }

  // tail finisher tasks: every block with columns left at step t0 (blocks
  // not started by then are reduced whole)
  std::vector<FinTask> fin;
  if (t0 < T) {
    for (int q = 0; q < nblk; ++q) {
      const int b = order_[q];
      const int c0 = std::max(0, t0 - coff[q]);
      if (c0 >= n_[b]) continue;
      NRG_REQUIRE(n_[b] - c0 <= FIN_MX, "eigensolver: tail finisher block above its LDS size");
      fin.push_back(FinTask{n_[b], c0, aoff_[b], voff_[b]});
    }
    d_fin_.upload(reinterpret_cast<const char*>(fin.data()), fin.size() * sizeof(FinTask), s);
  }

  // From here on the work is launches only (every table is on the device):
  // with graphs on, the fork, the column steps and the join of this block
  // set are captured once and replayed when the same block set comes again
  // (the same sizes, offsets, schedule and buffer addresses, so the same
  // tables and launch arguments).
  const bool graph = use_graphs_ && !profiler().on && !profiler().log_mask;
  std::vector<int64_t> key;
  if (graph) {
    key.reserve(3 * nblk + 32);
    key.push_back(nblk);
    key.push_back(align_end_);
    key.push_back(ng);
    key.push_back(symv_wide_wg_);
    key.push_back(symv_kind_);
    key.push_back(finish_m_);
    key.push_back((int64_t)symv_wide_bytes_);
    for (int b = 0; b < nblk; ++b) { key.push_back(n_[b]); key.push_back(aoff_[b]); key.push_back(voff_[b]); }
    for (const void* p : {(const void*)A_, (const void*)d_d_.p, (const void*)d_e_.p, (const void*)d_tau_.p,
                          (const void*)d_y_.p, (const void*)d_P_.p, (const void*)d_part_.p})
      key.push_back((int64_t)(intptr_t)p);
    key.push_back((int64_t)(intptr_t)d_fin_.p);
    for (int g = 0; g < ng; ++g) {
      key.push_back((int64_t)(intptr_t)d_slotg_[g].p);
      key.push_back((int64_t)(intptr_t)gemmg_[g].r_dev_.p);
      key.push_back((int64_t)(intptr_t)(g ? sg_[g] : s));
    }
    for (TrGraph& e : graphs_)
      if (e.key == key) {
        NRG_HIP(hipGraphLaunch(e.exec, s));
        Profiler& P = profiler();
        for (int f = 0; f < KF_COUNT; ++f) {
          P.w_launches[f] += e.launches[f]; P.w_bytes[f] += e.bytes[f]; P.w_flops[f] += e.flops[f];
        }
        e.used = ++graph_clock_;
        ++graph_replays_;
        for (int g = 0; g < ng; ++g) grp[g].gemm->end_replay();
        NRG_HIP(hipGetLastError());
        return;
      }
    NRG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  }
  Profiler& Pw = profiler();
  const auto w0 = Pw.w_launches, b0 = Pw.w_bytes, f0 = Pw.w_flops;
  if (ng > 1) {
    // the other streams start after everything queued so far (inputs, tables)
    NRG_HIP(hipEventRecord(ev_fork_, s));
    for (int g = 1; g < ng; ++g) NRG_HIP(hipStreamWaitEvent(sg_[g], ev_fork_, 0));
  }

  // step t of group G; close_only: only close the previous panel (its last
  // w, its trailing update), as at the finisher's start step
  auto step = [&](Group& G, int t, bool close_only) {
    const int ns = (int)G.q.size();
    // slots of the group started by step t (coff is non-decreasing)
    int act;
    {
      int lo = 0, hi = ns;
      while (lo < hi) { int md = (lo + hi) / 2; if (G.coff[md] <= t) lo = md + 1; else hi = md; }
      act = lo;
    }
    auto rows_at = [&](int i, int tt) {   // trailing rows n - c of group slot i at step tt (<= 0: idle)
      const int nq = n_[order_[G.q[i]]];
      const int c = tt - G.coff[i];
      return (c >= 0 && c < nq) ? nq - c : 0;
    };
    auto colbytes = [&](int tt) {
      double b = 0;
      for (int i = 0; i < act; ++i) {
        const double m = rows_at(i, tt);
        b += 8.0 * (4.0 * m * (1 + NB / 8) + m * m / (2.0 * G.pw_prev));   // column, panel rows, y partials
      }
      return b;
    };
    const int jj = t % NB;   // panel position, the same for every slot
    int gA = 0, tA = 0;      // flattened grid: the row tiles of the active blocks
    for (int i = 0; i < act; ++i) {
      gA += ceil_div(rows_at(i, t), RA);
      tA = std::max(tA, ceil_div(rows_at(i, t), RA));
    }
    if (gA == 0) return;
    // tile-major order when the tile counts are nearly uniform
    const int tmA = (align_end_ && (int64_t)act * tA <= gA + gA / 8) ? tA : 0;
    const int gAl = tmA ? act * tA : gA;
    const TrArgs& ta = G.ta;
    if (jj == 0 && t > 0) {
      // close the previous panel: finalise its last w, trailing SYR2K
      plaunch(KF_TR_COL, colbytes(t), 0, k_tr_col, dim3(gAl), dim3(1024), 0, G.s, ta, t, NB - 1, -1, NB - 1,
              G.pw_prev, act, tmA);
      const int pi = t / NB - 1;
      G.gemm->issue(G.syr2k[pi], G.syr2k[pi + 1], G.s);
      if (close_only) return;
      plaunch(KF_TR_COL, colbytes(t), 0, k_tr_col, dim3(gAl), dim3(1024), 0, G.s, ta, t, -1, 0, NB - 1,
              G.pw_prev, act, tmA);
    } else {
      const int jp = (jj >= 1) ? jj - 1 : -1;
      const int vslot = (t >= 1) ? (t - 1) % NB : 0;
      plaunch(KF_TR_COL, colbytes(t), 0, k_tr_col, dim3(gAl), dim3(1024), 0, G.s, ta, t, jp, jj, vslot,
              G.pw_prev, act, tmA);
    }
    // algorithmic traffic: the lower triangle of the symmetric trailing
    // (n-c-1)^2 block once, plus the vectors
    double bytes = 0, flops = 0;
    int mB = 0;
    for (int i = 0; i < act; ++i) {
      const int m = rows_at(i, t) - 1;
      if (m <= 0) continue;
      bytes += 8.0 * (0.5 * (double)m * (m + 1.0) + 6.0 * m);
      flops += 2.0 * (double)m * m;
      mB = std::max(mB, m);
    }
    // 32-column panels (16 waves x 2 columns) while the launch is
    // bandwidth-bound and has enough of them to occupy the device,
    // 16-column panels (16 waves x 1 column) otherwise: a workgroup's
    // bandwidth is bounded by the loads it keeps in flight, so a launch of
    // few 32-column panels (the first columns, where one or two blocks are
    // active) runs faster as twice as many 16-column panels
    int g32 = 0;
    for (int i = 0; i < act; ++i) {
      const int m = rows_at(i, t) - 1;
      if (m > 0) g32 += ceil_div(m, 2 * CG);
    }
    const int pw = (bytes >= symv_wide_bytes_ / ng && g32 >= symv_wide_wg_ / ng) ? 2 * CG : CG;
    int gB = 0, tmax = 0;
    for (int i = 0; i < act; ++i) {
      const int m = rows_at(i, t) - 1;
      if (m > 0) gB += ceil_div(m, pw);
      if (m > 0) tmax = std::max(tmax, ceil_div(m, pw));
    }
    if (gB > 0) {
      // LDS: row parts and (VL) v of rows [j0, m) of the longest trailing
      // block (the row-chunk SYMV keeps only v there)
      const bool vl = symv_kind_ == 1 ? symv_rows_v_fits(mB) : symv_lds_fits(mB, 2);
      const uint32_t shB = symv_kind_ == 1 ? (uint32_t)(vl ? mB * sizeof(double) : 0)
                                           : (uint32_t)((vl ? 2 : 1) * mB * sizeof(double));
      // panel-major order when the panel counts are nearly uniform
      const int pm = (align_end_ && (int64_t)act * tmax <= gB + gB / 8) ? 1 : 0;
      const int g = pm ? act * tmax : gB;
      // A SYMV launch that carries timing events runs with the device to
      // itself: the other groups' streams are fenced around it, so that its
      // event interval is its execution and not also time spent waiting for
      // workgroup slots held by another stream's kernel.
      const int gi = (int)(&G - grp);
      const bool excl = (ng > 1) && profiler().exclusive && profiler().peek(KF_TR_SYMV);
      if (excl)
        for (int h = 0; h < ng; ++h)
          if (h != gi) {
            NRG_HIP(hipEventRecord(evf_[h], grp[h].s));
            NRG_HIP(hipStreamWaitEvent(G.s, evf_[h], 0));
          }
      if (symv_kind_ == 1) {
        if (pw == 2 * CG && vl)
          plaunch(KF_TR_SYMV, bytes, flops, k_tr_symv_rows<2, 1>, dim3(g), dim3(1024), shB, G.s, ta, t, jj, act, pm);
        else if (vl)
          plaunch(KF_TR_SYMV, bytes, flops, k_tr_symv_rows<1, 1>, dim3(g), dim3(1024), shB, G.s, ta, t, jj, act, pm);
        else if (pw == 2 * CG)
          plaunch(KF_TR_SYMV, bytes, flops, k_tr_symv_rows<2, 0>, dim3(g), dim3(1024), shB, G.s, ta, t, jj, act, pm);
        else
          plaunch(KF_TR_SYMV, bytes, flops, k_tr_symv_rows<1, 0>, dim3(g), dim3(1024), shB, G.s, ta, t, jj, act, pm);
      } else if (pw == 2 * CG && vl)
        plaunch(KF_TR_SYMV, bytes, flops, k_tr_symv<2, 16, SYMV_U, 1>, dim3(g), dim3(1024), shB, G.s, ta, t, jj, act, pm);
      else if (vl)
        plaunch(KF_TR_SYMV, bytes, flops, k_tr_symv<1, 16, SYMV_U, 1>, dim3(g), dim3(1024), shB, G.s, ta, t, jj, act, pm);
      else if (pw == 2 * CG)
        plaunch(KF_TR_SYMV, bytes, flops, k_tr_symv<2, 16, SYMV_U, 0>, dim3(g), dim3(1024), shB, G.s, ta, t, jj, act, pm);
      else
        plaunch(KF_TR_SYMV, bytes, flops, k_tr_symv<1, 16, SYMV_U, 0>, dim3(g), dim3(1024), shB, G.s, ta, t, jj, act, pm);
      if (excl) {
        NRG_HIP(hipEventRecord(evf_[gi], G.s));
        for (int h = 0; h < ng; ++h)
          if (h != gi) NRG_HIP(hipStreamWaitEvent(grp[h].s, evf_[gi], 0));
      }
      G.pw_prev = pw;
    }
  };
  for (int t = 0; t < t0; ++t)
    for (int g = 0; g < ng; ++g) step(grp[g], t, false);
  if (t0 > 0 && t0 < T)
    for (int g = 0; g < ng; ++g) step(grp[g], t0, true);
  for (int g = 0; g < ng; ++g) grp[g].gemm->end_replay();
  for (int g = 1; g < ng; ++g) {
    NRG_HIP(hipEventRecord(evg_[g], sg_[g]));
    NRG_HIP(hipStreamWaitEvent(s, evg_[g], 0));
  }
  if (!fin.empty()) {
    double fb = 0;
    for (const FinTask& f : fin) fb += 8.0 * (double)(f.n - f.c0) * (f.n - f.c0 + 1);
    plaunch(KF_TR_COL, fb, 0, k_tr_finish, dim3((unsigned)fin.size()), dim3(1024), 0, s,
            reinterpret_cast<const FinTask*>(d_fin_.p), A_, d_d_.p, d_e_.p, d_tau_.p);
  }
  if (graph) {
    hipGraph_t gr = nullptr;
    NRG_HIP(hipStreamEndCapture(s, &gr));
    TrGraph e;
    e.key = std::move(key);
    const hipError_t ie = hipGraphInstantiate(&e.exec, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    NRG_HIP(ie);
    for (int f = 0; f < KF_COUNT; ++f) {
      e.launches[f] = Pw.w_launches[f] - w0[f];
      e.bytes[f] = Pw.w_bytes[f] - b0[f];
      e.flops[f] = Pw.w_flops[f] - f0[f];
    }
    e.used = ++graph_clock_;
    NRG_HIP(hipGraphLaunch(e.exec, s));
    ++graph_captures_;
    if (graphs_.size() >= kMaxGraphs) {
      size_t old = 0;
      for (size_t i = 1; i < graphs_.size(); ++i)
        if (graphs_[i].used < graphs_[old].used) old = i;
      NRG_HIP(hipStreamSynchronize(s));   // the evicted graph may still run
      (void)hipGraphExecDestroy(graphs_[old].exec);
      graphs_.erase(graphs_.begin() + old);
    }
    graphs_.push_back(std::move(e));
  }
  NRG_HIP(hipGetLastError());
}

void BatchedEigensolver::divide_conquer(hipStream_t s) {
  const int nblk = nblk_;
  // copy the tridiagonal diagonal (tearing modifies it)
  std::vector<int32_t> leaves, bnd;
  std::vector<std::vector<int32_t>> levels;  // merge records per level
  std::vector<int> L(nblk);
  int maxL = 0;
  for (int b = 0; b < nblk; ++b) {
    const int n = n_[b];
    int l = 0;
    while ((n + (1 << l) - 1) >> l > LEAF) ++l;
    L[b] = l;
    maxL = std::max(maxL, l);
    const int nl = 1 << l;
    auto bd = [&](int i) { return (int)(((int64_t)i * n) >> l); };
    for (int i = 0; i < nl; ++i) {
      const int s0 = bd(i), s1 = bd(i + 1);
      if (s1 > s0) { leaves.push_back(b); leaves.push_back(s0); leaves.push_back(s1 - s0); }
      if (i > 0 && s0 > 0 && s0 < n) { bnd.push_back(b); bnd.push_back(s0); }
    }
  }
  levels.resize(maxL);
  // A block's last merge (level L[b], the whole block) is deferred: its
  // eigenvector product and final scatter run when columns are asked for
  // (materialize), so its workspace lives past the per-level range and its
  // records follow the level's other merges.
  std::vector<std::vector<int32_t>> tops(maxL);
  int64_t wsmax = 0, iwsmax = 0;
  for (int lev = 1; lev <= maxL; ++lev) {
    auto& rec = levels[lev - 1];
    int64_t wacc = 0, iacc = 0;
    for (int b = 0; b < nblk; ++b) {
      if (L[b] < lev) continue;
      const int n = n_[b];
      const int l = L[b];
      auto bd = [&](int i) { return (int)(((int64_t)i * n) >> l); };
      const int nm = 1 << (l - lev);
      for (int j = 0; j < nm; ++j) {
        const int s0 = bd(j << lev), mid = bd((j << lev) + (1 << (lev - 1))), e0 = bd((j + 1) << lev);
        if (!(s0 < mid && mid < e0)) continue;
        const int m = e0 - s0;
        if (lev == l) {
          for (int32_t v : {b, s0, mid, e0, 0, 0, 0, 0}) tops[lev - 1].push_back(v);
          continue;
        }
        rec.push_back(b); rec.push_back(s0); rec.push_back(mid); rec.push_back(e0);
        rec.push_back((int32_t)(wacc & 0xffffffff)); rec.push_back((int32_t)(wacc >> 32));
        rec.push_back((int32_t)(iacc & 0xffffffff)); rec.push_back((int32_t)(iacc >> 32));
        wacc += 9 * (int64_t)m + 3 * (int64_t)m * m + 8;
        iacc += dws_int(m);
      }
    }
    wsmax = std::max(wsmax, wacc);
    iwsmax = std::max(iwsmax, iacc);
  }
  top_.assign(nblk, TopMerge{});
  std::vector<int> nreg(maxL);
  int64_t wtop = wsmax, itop = iwsmax;
  for (int lev = 1; lev <= maxL; ++lev) {
    auto& rec = levels[lev - 1];
    nreg[lev - 1] = (int)rec.size() / 8;
    const auto& tr = tops[lev - 1];
    for (size_t i = 0; i < tr.size(); i += 8) {
      const int b = tr[i], m = tr[i + 3] - tr[i + 1];
      for (int q = 0; q < 4; ++q) rec.push_back(tr[i + q]);
      rec.push_back((int32_t)(wtop & 0xffffffff)); rec.push_back((int32_t)(wtop >> 32));
      rec.push_back((int32_t)(itop & 0xffffffff)); rec.push_back((int32_t)(itop >> 32));
      top_[b].lev = lev; top_[b].m = m; top_[b].m1 = tr[i + 2] - tr[i + 1]; top_[b].wo = wtop; top_[b].io = itop;
      wtop += 9 * (int64_t)m + 3 * (int64_t)m * m + 8;
      itop += dws_int(m);
    }
  }
  d_ws_.reserve(std::max<int64_t>(wtop, 1));
  d_iws_.reserve(std::max<int64_t>(itop, 1));
  // zero Z (off-diagonal blocks of the merge tree must be zero)
  fill(Z_, 0.0, tot_n2_, s);
  DcArgs da{d_n_.p, d_aoff_.p, d_voff_.p, d_d_.p, d_e_.p, w_, Z_, d_ws_.p, d_iws_.p, nullptr};
  // tear
  if (!bnd.empty()) {
    d_tasks_.upload(bnd.data(), bnd.size(), s);
    const int nb_ = (int)bnd.size() / 2;
    hipLaunchKernelGGL(k_dc_tear, dim3(ceil_div(nb_, 128)), dim3(128), 0, s, d_tasks_.p, nb_, d_n_.p,
                       d_voff_.p, d_d_.p, d_e_.p);
  }
  // leaves
  {
    DevBuf<int32_t>& tb = d_tasks_;
    tb.upload(leaves.data(), leaves.size(), s);
    const int nt = (int)leaves.size() / 3;
    plaunch(KF_DC_LEAF, 0, 0, k_dc_leaf, dim3(nt), dim3(64), 0, s, da, (const int32_t*)tb.p, nt);
  }
  for (int lev = 1; lev <= maxL; ++lev) {
    auto& rec = levels[lev - 1];
    const int nt = (int)rec.size() / 8;
    if (nt == 0) continue;
    d_tasks_.upload(rec.data(), rec.size(), s);
    d_k_.reserve(3 * nt);
    da.kout = d_k_.p;
    int mmax = 0;
    for (int i = 0; i < nt; ++i) mmax = std::max(mmax, rec[8 * i + 3] - rec[8 * i + 1]);
    {
      // poles and weights (2 doubles) and the half (1 byte) of each sorted pole
      const size_t ml = (size_t)std::min(mmax, DC_LDS_M);
      const size_t shm = 2 * sizeof(double) * ml + (ml + 7) / 8 * 8;
      const int32_t* tk = d_tasks_.p;
      plaunch(KF_DC_PREP, 0, 0, k_dc_merge_prep, dim3(nt), dim3(256), (uint32_t)shm, s, da, tk, nt);
      plaunch(KF_DC_PREP, 0, 0, k_dc_rotate, dim3(ceil_div(mmax, 256), nt), dim3(256), 0, s, da, tk, nt);
      plaunch(KF_DC_PREP, 0, 0, k_dc_rotgather, dim3(ceil_div(mmax, 64), ceil_div(mmax, DC_CCH), nt), dim3(256), 0,
              s, da, tk, nt);
    }
    h_k_.resize(3 * nt);
    d_k_.download(h_k_.data(), 3 * nt, s);
    NRG_HIP(hipStreamSynchronize(s));
    int kmax = 0;
    for (int i = 0; i < nt; ++i) kmax = std::max(kmax, h_k_[i]);
    if (kmax > 0) {
      const int32_t* tk = d_tasks_.p;
      plaunch(KF_DC_SECULAR, 0, 0, k_dc_secular, dim3(ceil_div(kmax, SEC_ROOTS), nt), dim3(64 * SEC_ROOTS),
              (uint32_t)(2 * sizeof(double) * (size_t)std::min(kmax, SEC_LDS_M)), s, da, tk, nt);
      plaunch(KF_DC_VECTORS, 0, 0, k_dc_zhat, dim3(ceil_div(kmax, 4), nt), dim3(256), 0, s, da, tk, nt);
      plaunch(KF_DC_VECTORS, 0, 0, k_dc_vectors, dim3(kmax, nt), dim3(64), 0, s, da, tk, nt);
      // Qt = Qcd U on the rows of each half, over the grouped columns with
      // support there: [0, k1 + k3) for the first half, [k1, k) for the
      // second.  Merges of at most 64 rows (one output tile) stay one dense
      // product: their launches are task-bound, not flop-bound.
      batch_.clear();
      for (int i = 0; i < nreg[lev - 1]; ++i) {
        const int k = h_k_[i];
        if (k == 0) continue;
        const int k1 = h_k_[nt + 2 * i], k3 = h_k_[nt + 2 * i + 1];
        const int m = rec[8 * i + 3] - rec[8 * i + 1], m1 = rec[8 * i + 2] - rec[8 * i + 1];
        const int64_t wo = (int64_t)(uint32_t)rec[8 * i + 4] | ((int64_t)rec[8 * i + 5] << 32);
        const int64_t base = wo + 9 * (int64_t)m, ub = base + (int64_t)m * m, qt = base + 2 * (int64_t)m * m;
        if (m <= 64) {
          batch_.add(gemm_task(m, k, k, 0, 0, base, m, ub, k, qt, m, 1.0, 0.0));
          continue;
        }
        batch_.add(gemm_task(m1, k, k1 + k3, 0, 0, base, m, ub, k, qt, m, 1.0, 0.0));
        batch_.add(gemm_task(m - m1, k, k - k1, 0, 0, base + m1 + (int64_t)k1 * m, m, ub + k1, k, qt + m1, m,
                             1.0, 0.0));
      }
      gemm_.run(batch_, d_ws_.p, d_ws_.p, d_ws_.p, nullptr, nullptr, nullptr, s, KF_GEMM_DC);
    }
    for (int i = nreg[lev - 1]; i < nt; ++i) {
      TopMerge& T = top_[rec[8 * i]];
      T.k = h_k_[i];
      T.k1 = h_k_[nt + 2 * i];
      T.k3 = h_k_[nt + 2 * i + 1];
    }
    const int32_t* tk = d_tasks_.p;
    // LDS: (value, index) pairs of the deflated values, padded to a power of two
    int P = 1;
    while (P < std::min(mmax, DC_LDS_M)) P <<= 1;
    plaunch(KF_DC_FINISH, 0, 0, k_dc_finish, dim3(nt), dim3(FIN_T),
            (uint32_t)((sizeof(double) + sizeof(int32_t)) * (size_t)P), s, da, tk, nt);
    if (nreg[lev - 1] > 0)
      plaunch(KF_DC_FINISH, 0, 0, k_dc_scatter, dim3(ceil_div(mmax, 64), ceil_div(mmax, DC_CCH), nreg[lev - 1]),
              dim3(256), 0, s, da, tk, nreg[lev - 1]);
  }
  mat_.resize(nblk);
  for (int b = 0; b < nblk; ++b) mat_[b] = (top_[b].lev > 0) ? 0 : n_[b];   // leaves write Z directly
  NRG_HIP(hipGetLastError());
}

// V, G = V^T V, T and W = V T^T of every reflector panel of every block.
// They depend only on the tridiagonalisation (A below the subdiagonal, tau),
// not on the columns transformed, so they are built once per solve, in four
// launches over all panels, on stream s (issued while the divide and conquer
// runs on the caller's stream when s is the second stream).
void BatchedEigensolver::panel_factors(hipStream_t s) {
  const int nblk = nblk_;
  h_vpo_.assign(nblk, 0);
  h_tpo_.assign(nblk, 0);
  h_plist_.clear();
  int64_t vacc = 0;
  int32_t tacc = 0;
  int64_t rows_max = 0;
  for (int b = 0; b < nblk; ++b) {
    const int n = n_[b];
    const int np = std::max(0, (n - bt_off_ + NBB - 1) / NBB);   // panels of reflectors 0..n-1-off
    h_vpo_[b] = vacc;
    h_tpo_[b] = tacc;
    vacc += (int64_t)np * n * NBB;
    tacc += np;
    for (int p = 0; p < np; ++p) {
      h_plist_.push_back(b);
      h_plist_.push_back(p);
    }
    rows_max = std::max<int64_t>(rows_max, n);
  }
  const int npl = (int)(h_plist_.size() / 2);
  if (npl == 0) return;
  d_V_.reserve(std::max<int64_t>(vacc, 1));
  d_T_.reserve((int64_t)tacc * 2 * NBB * NBB);
  d_vpo_.upload(h_vpo_.data(), nblk, s);
  d_tpo_.upload(h_tpo_.data(), nblk, s);
  d_plist_.upload(h_plist_.data(), h_plist_.size(), s);
  BtArgs ba{d_n_.p, d_aoff_.p, d_voff_.p, A_, d_tau_.p, d_V_.p, d_T_.p, d_vpo_.p, d_tpo_.p};
  hipLaunchKernelGGL(k_bt_buildv, dim3((unsigned)std::max<int64_t>(1, ceil_div(rows_max * NBB, 256 * 4)), npl),
                     dim3(256), 0, s, ba, (const int32_t*)d_plist_.p, bt_off_);
  // G = V^T V for every panel
  GemmBatch& bt = batch_bt_;
  bt.clear();
  for (int i = 0; i < npl; ++i) {
    const int b = h_plist_[2 * i], p = h_plist_[2 * i + 1];
    const int n = n_[b], c0 = p * NBB;
    const int mr = n - c0 - bt_off_;
    const int nbp = std::min(NBB, n - bt_off_ - c0);
    const int64_t vo = h_vpo_[b] + (int64_t)p * n * NBB;
    bt.add(gemm_task(nbp, nbp, mr, 1, 0, vo, n, vo, n, (int64_t)(h_tpo_[b] + p) * 2 * NBB * NBB, NBB, 1.0, 0.0));
  }
  gemm_bt_.run(bt, d_V_.p, d_V_.p, d_T_.p, nullptr, nullptr, nullptr, s, KF_GEMM_BACK);
  plaunch(KF_BT_SMALL, 0, 0, k_bt_tmat, dim3(npl), dim3(NBB * TQ), 0, s, ba, (const int32_t*)d_plist_.p, bt_off_);
  // W = V T^T for every panel (layout of V): the back-transform then forms
  // T V^T Z as one product W^T Z per panel
  d_W_.reserve(std::max<int64_t>(vacc, 1));
  bt.clear();
  for (int i = 0; i < npl; ++i) {
    const int b = h_plist_[2 * i], p = h_plist_[2 * i + 1];
    const int n = n_[b], c0 = p * NBB;
    const int mr = n - c0 - bt_off_;
    const int nbp = std::min(NBB, n - bt_off_ - c0);
    const int64_t vo = h_vpo_[b] + (int64_t)p * n * NBB;
    bt.add(gemm_task(mr, nbp, nbp, 0, 1, vo, n, (int64_t)(h_tpo_[b] + p) * 2 * NBB * NBB + NBB * NBB, NBB, vo, n,
                     1.0, 0.0));
  }
  gemm_bt_.run(bt, d_V_.p, d_T_.p, d_W_.p, nullptr, nullptr, nullptr, s, KF_GEMM_BACK);
  NRG_HIP(hipGetLastError());
}

void BatchedEigensolver::back_transform(const int* lo, const int* hi, hipStream_t s) {
  const int nblk = nblk_;
  const int nmax = n_[order_[0]];
  if (bt_wait_) {   // the panel factors were built on the second stream
    NRG_HIP(hipStreamWaitEvent(s, ev_bt_, 0));
    bt_wait_ = false;
  }
  d_X2_.reserve(tot_n_ * NBB);
  const int off = bt_off_;
  const int npan = std::max(0, (nmax - off + NBB - 1) / NBB);
  // the panels' products are recorded and issued after one table upload
  gemm_.begin();
  for (int p = npan - 1; p >= 0; --p) {
    const int c0 = p * NBB;
    // active: n - off > c0
    int act = 0;
    while (act < nblk && n_[order_[act]] > c0 + off) ++act;
    if (act == 0) continue;
    auto vo = [&](int b) { return h_vpo_[b] + (int64_t)p * n_[b] * NBB; };
    // X2 = T V^T Z(c0+off:, :nc) = W^T Z(c0+off:, :nc)
    batch_.clear();
    for (int q = 0; q < act; ++q) {
      const int b = order_[q];
      const int n = n_[b];
      const int mr = n - c0 - off;
      const int nbp = std::min(NBB, n - off - c0);
      const int nc = hi[b] - lo[b];
      if (nc <= 0) continue;
      batch_.add(gemm_task(nbp, nc, mr, 1, 0, vo(b), n, aoff_[b] + (int64_t)lo[b] * n + c0 + off, n,
                           voff_[b] * NBB, NBB, 1.0, 0.0));
    }
    if (batch_.tasks.empty()) continue;
    gemm_.run(batch_, d_W_.p, Z_, d_X2_.p, nullptr, nullptr, nullptr, s, KF_GEMM_BACK);
    // Z(c0+off:, :nc) -= V X2
    batch_.clear();
    for (int q = 0; q < act; ++q) {
      const int b = order_[q];
      const int n = n_[b];
      const int mr = n - c0 - off;
      const int nbp = std::min(NBB, n - off - c0);
      const int nc = hi[b] - lo[b];
      if (nc <= 0) continue;
      batch_.add(gemm_task(mr, nc, nbp, 0, 0, vo(b), n, voff_[b] * NBB, NBB,
                           aoff_[b] + (int64_t)lo[b] * n + c0 + off, n, -1.0, 1.0));
    }
    gemm_.run(batch_, d_V_.p, d_X2_.p, Z_, nullptr, nullptr, nullptr, s, KF_GEMM_BACK);
  }
  gemm_.flush(s);
  // chop the transformed columns (element ranges [aoff + lo n, aoff + hi n))
  h_rng_.clear();
  int64_t longest = 0;
  for (int b = 0; b < nblk; ++b) {
    if (hi[b] <= lo[b]) continue;
    const int64_t a0 = aoff_[b] + (int64_t)lo[b] * n_[b], len = (int64_t)(hi[b] - lo[b]) * n_[b];
    h_rng_.push_back(a0);
    h_rng_.push_back(len);
    longest = std::max(longest, len);
  }
  const int nr = (int)(h_rng_.size() / 2);
  if (nr > 0) {
    d_rng_.upload(h_rng_.data(), h_rng_.size(), s);
    const unsigned g = (unsigned)std::min<int64_t>((longest + 1023) / 1024, 4096);
    hipLaunchKernelGGL(k_chop, dim3(g, nr), dim3(256), 0, s, Z_, (const int64_t*)d_rng_.p, 1e-14);
  }
  NRG_HIP(hipGetLastError());
}

// Columns [mat_b, hi_b) of the deferred top merges into Z: Usel gathers the
// roots' secular vectors (zero for deflated columns), Z = Qcd[:, :k] Usel on
// the matrix cores, then the deflated columns are copied.
void BatchedEigensolver::materialize(const int* hi, hipStream_t s) {
  std::vector<int32_t> tk;
  int64_t uacc = 0;
  int ncmax = 0, kmax = 0, mmax = 0;
  batch_.clear();
  for (int b = 0; b < nblk_; ++b) {
    const TopMerge& T = top_[b];
    if (T.lev <= 0 || hi[b] <= mat_[b]) continue;
    const int lo = mat_[b], nc = hi[b] - lo, k = T.k, m = T.m;
    for (int32_t v : {b, lo, nc, k, m, 0, (int32_t)(T.wo & 0xffffffff), (int32_t)(T.wo >> 32),
                      (int32_t)(T.io & 0xffffffff), (int32_t)(T.io >> 32), (int32_t)(uacc & 0xffffffff),
                      (int32_t)(uacc >> 32)})
      tk.push_back(v);
    if (k > 0) {
      // rows of each half over the grouped columns with support there
      const int m1 = T.m1, k1 = T.k1, k3 = T.k3;
      const int64_t q = T.wo + 9 * (int64_t)m, z0 = aoff_[b] + (int64_t)lo * n_[b];
      batch_.add(gemm_task(m1, nc, k1 + k3, 0, 0, q, m, uacc, k, z0, n_[b], 1.0, 0.0));
      batch_.add(gemm_task(m - m1, nc, k - k1, 0, 0, q + m1 + (int64_t)k1 * m, m, uacc + k1, k, z0 + m1, n_[b],
                           1.0, 0.0));
    }
    uacc += (int64_t)k * nc;
    ncmax = std::max(ncmax, nc);
    kmax = std::max(kmax, k);
    mmax = std::max(mmax, m);
    mat_[b] = hi[b];
  }
  if (tk.empty()) return;
  const int nt = (int)tk.size() / 12;
  d_usel_.reserve(std::max<int64_t>(uacc, 1));
  d_mtask_.upload(tk.data(), tk.size(), s);
  DcTop a{d_n_.p, d_aoff_.p, d_ws_.p, d_iws_.p, Z_, d_usel_.p};
  if (kmax > 0)
    plaunch(KF_DC_FINISH, 0, 0, k_dc_top_select, dim3(ceil_div(kmax, 256), ncmax, nt), dim3(256), 0, s, a,
            (const int32_t*)d_mtask_.p);
  if (!batch_.tasks.empty())
    gemm_.run(batch_, d_ws_.p, d_usel_.p, Z_, nullptr, nullptr, nullptr, s, KF_GEMM_DC);
  plaunch(KF_DC_FINISH, 0, 0, k_dc_top_deflated, dim3(ceil_div(mmax, 256), ncmax, nt), dim3(256), 0, s, a,
          (const int32_t*)d_mtask_.p);
  NRG_HIP(hipGetLastError());
}

void BatchedEigensolver::back_columns(const int* hi, hipStream_t s, double* ms) {
  NRG_REQUIRE(nblk_ > 0, "eigensolver: no solve to back-transform");
  std::vector<int> lo(done_), h(nblk_);
  bool any = false;
  for (int b = 0; b < nblk_; ++b) {
    h[b] = std::min(std::max(hi ? hi[b] : n_[b], done_[b]), n_[b]);
    any = any || h[b] > lo[b];
  }
  if (!any) return;
  const double t0 = now_ms();
  materialize(h.data(), s);
  if (solved_two_stage_) back_q2(lo.data(), h.data(), s);
  back_transform(lo.data(), h.data(), s);
  done_ = h;
  if (ms) { NRG_HIP(hipStreamSynchronize(s)); *ms += now_ms() - t0; }
}

}  // namespace nrg
