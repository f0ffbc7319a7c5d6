// Block-Hamiltonian assembly (diagonalization.F90:153-263) for all blocks of
// an iteration in one launch.
//
// Element (r, c) of block b (rows r = product states start_input-1+r):
//   diagonal:     E_old(ref_r) sqrt(Lambda) + sum_ch operation(op_r, nv+ch) eps_ch
//   off-diagonal: with p = min(r,c), q = max(r,c), the (op_p, op_q) entry of the
//                 hopping lookup (built on the host by replaying the
//                 coefficient search: the last match sets type, channel and
//                 coef, any reversed match swaps the references) gives
//                 (type, channel, coef, swap) and
//                 H = t_channel * coef * invariant_matrix(jref, iref, type);
//                 with a BCS chain, a forward BCS row of (op_p, op_q) and
//                 ref_p == ref_q override it with chain_BCS(it+1, channel)
//                 (diagonalization.F90:215-245).
// 64x64 tiles of all blocks are flattened into one grid.
#include "assemble.h"

namespace nrg {

struct AsmArgs {
  const int32_t* tile_start;  // per block
  const int32_t* bn;
  const int64_t* boff;
  const int64_t* bvoff;
  int nblk;
  const int32_t* basis;       // nfull*(nv+3)
  int nv;
  const double* Eold;
  const double* inv;          // (kold, kold, ncm) column-major
  int64_t kold;
  const int32_t* op;          // [nops][nv+nchain]
  const int32_t* hop;         // [5][nops*nops]: valid, type, chan, swap, BCS channel + 1
  const double* hop_coef;
  int nops, nchain;
  double cd[4];               // chain_diagonal(it+1, ch)
  double cnd[4];              // chain_nondiagonal(it, ch)
  double cbcs[4];             // chain_BCS(it+1, ch)
  int bcs;                    // BCS rows present
  double sqrtl;
  double* H;
};

__global__ __launch_bounds__(256) void k_assemble(AsmArgs a) {
  const int bid = blockIdx.x;
  int lo = 0, hi = a.nblk - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (a.tile_start[mid] <= bid) lo = mid; else hi = mid - 1;
  }
  const int b = lo;
  const int n = a.bn[b];
  const int tiles = (n + 63) / 64;
  const int local = bid - a.tile_start[b];
  const int r0 = (local % tiles) * 64, c0 = (local / tiles) * 64;
  const int rs = a.nv + 3;
  const int32_t* rec = a.basis + a.bvoff[b] * rs;
  double* H = a.H + a.boff[b];
  const int r = r0 + (threadIdx.x & 63);
  if (r >= n) return;
  const int op_r = rec[r * rs + a.nv + 2] - 1;
  const int ref_r = rec[r * rs + a.nv] - 1;
  const int nn = a.nops * a.nops;
  for (int i = 0; i < 16; ++i) {
    const int c = c0 + (threadIdx.x >> 6) + 4 * i;
    if (c >= n) break;
    double v = 0.0;
    if (c == r) {
      v = a.Eold[ref_r] * a.sqrtl;
      for (int ch = 0; ch < a.nchain; ++ch) v += (double)a.op[op_r * (a.nv + a.nchain) + a.nv + ch] * a.cd[ch];
    } else {
      const int op_c = rec[c * rs + a.nv + 2] - 1;
      const int ref_c = rec[c * rs + a.nv] - 1;
      int op_p, op_q, ref_p, ref_q;
      if (r < c) { op_p = op_r; op_q = op_c; ref_p = ref_r; ref_q = ref_c; }
      else { op_p = op_c; op_q = op_r; ref_p = ref_c; ref_q = ref_r; }
      const int idx = op_p * a.nops + op_q;
      if (a.hop[idx]) {
        const int t = a.hop[nn + idx], ch = a.hop[2 * nn + idx];
        int iref = ref_p, jref = ref_q;
        if (a.hop[3 * nn + idx]) { iref = ref_q; jref = ref_p; }
        v = a.cnd[ch] * a.hop_coef[idx] * a.inv[(int64_t)jref + (int64_t)iref * a.kold + (int64_t)t * a.kold * a.kold];
      }
      if (a.bcs) {
        const int cb = a.hop[4 * nn + idx];
        if (cb > 0 && ref_p == ref_q) v = a.cbcs[cb - 1];
      }
    }
    H[(int64_t)r + (int64_t)c * n] = v;
  }
}

void assemble(Context& C, const double* chain_diag, const double* chain_nondiag, double sqrtl,
              const double* chain_bcs) {
  const int nblk = C.nsub;
  std::vector<int32_t> ts(nblk);
  long long tot = 0;
  for (int b = 0; b < nblk; ++b) {
    ts[b] = (int32_t)tot;
    const long long t = (C.bn[b] + 63) / 64;
    tot += t * t;
  }
  NRG_REQUIRE(C.nchain <= 4, "assemble: at most 4 chains");
  static DevBuf<int32_t> d_ts, d_bn;
  static DevBuf<int64_t> d_boff, d_bvoff;
  d_ts.upload(ts, C.stream);
  d_bn.upload(C.bn.data(), nblk, C.stream);
  d_boff.upload(C.boff.data(), nblk, C.stream);
  d_bvoff.upload(C.bvoff.data(), nblk, C.stream);
  AsmArgs a;
  a.tile_start = d_ts.p; a.bn = d_bn.p; a.boff = d_boff.p; a.bvoff = d_bvoff.p; a.nblk = nblk;
  a.basis = C.d_basis_in.p; a.nv = C.nv; a.Eold = C.d_Eold.p; a.inv = C.d_inv.p; a.kold = C.kold;
  a.op = C.d_op.p; a.hop = C.d_hop.p; a.hop_coef = C.d_hop_coef.p; a.nops = C.nops; a.nchain = C.nchain;
  for (int ch = 0; ch < 4; ++ch) {
    a.cd[ch] = ch < C.nchain ? chain_diag[ch] : 0.0;
    a.cnd[ch] = ch < C.nchain ? chain_nondiag[ch] : 0.0;
    a.cbcs[ch] = (chain_bcs && ch < C.nchain) ? chain_bcs[ch] : 0.0;
  }
  a.bcs = chain_bcs ? 1 : 0;
  a.sqrtl = sqrtl;
  a.H = C.d_H.p;
  double bytes = 0;
  for (int b = 0; b < nblk; ++b) bytes += 8.0 * (double)C.bn[b] * C.bn[b];
  plaunch(KF_ASSEMBLE, bytes, 0, k_assemble, dim3((unsigned)tot), dim3(256), 0, C.stream, a);
  NRG_HIP(hipGetLastError());
}

}  // namespace nrg
