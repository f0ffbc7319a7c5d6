// Invariant-matrix contractions of the kept basis.
//
// invariantMatrixDGEMM.F90:154-332 builds, for every block pair (L,R) with
// dQ = 1 and |d2Sz| = 1, a dense coefficient matrix C_t (n_L x n_R) and
// forms U_L[:,1:kL]^T C_t U_R[:,1:kR] with two dgemm calls.  C_t(a,b) is
// nonzero only where rows a, b share the variation (in-block index of the
// parent state) and (op_a, op_b) is a table row of type t, so every row and
// column holds at most one entry: C_t is a signed partial permutation.  Here
// it is never materialised: each (L,R,t) is one gathered GEMM task whose
// inner dimension runs over the matched row pairs (a_i, b_i) with the table
// coefficient as a k-scale, and the result is written transposed straight
// into the dense invariant matrix inv(r, l, t) (:330).
//
// invariantMatrixForSpectrumDGEMM.F90:148-370 uses a block-dense C: rows with
// equal operation share the old spectral operator S(ref_b, ref_a).  Per op
// the rows of L and R map to contiguous ranges of the old kept basis, so
// X_o = c(o) S_sub^T U_R[b_o,:] is a strided GEMM and the result
// U_L[a,:]^T X is one gathered GEMM over the stacked rows.
#include "assemble.h"
#include "blockrows.h"

namespace nrg {

namespace {

BlockRows block_rows(const Context& C, int blk) {
  return nrg::block_rows(C.basis_in.data(), C.bvoff[blk], C.bn[blk], C.nv, C.nops);
}

std::vector<int64_t> kept_starts(const int32_t* kept, int nsub, int64_t& k) {
  std::vector<int64_t> st(nsub);
  k = 0;
  for (int b = 0; b < nsub; ++b) { st[b] = k; k += kept[b]; }
  return st;
}

}  // namespace

void invariant(Context& C, const int32_t* kept, double* out_host) {
  const int nsub = C.nsub;
  const int nv = C.nv;
  int64_t k = 0;
  const std::vector<int64_t> st = kept_starts(kept, nsub, k);
  std::vector<BlockRows> rows(nsub);
  for (int b = 0; b < nsub; ++b) rows[b] = block_rows(C, b);
  const int cs0 = C.cons[0] - 1, cs1 = C.cons[1], ss0 = C.cons[2] - 1, ss1 = C.cons[3];
  const bool sz_only = (C.cons[0] == C.cons[1] && C.cons[0] == 0);
  std::vector<int32_t>& gidx = C.h_gidx;
  std::vector<double>& gsc = C.h_gscale;
  gidx.clear();
  gsc.clear();
  GemmBatch batch;
  // table rows of each type, in table order
  std::vector<std::vector<int>> qt(C.ncm);
  for (size_t q = 0; q < C.inv_t.size(); ++q)
    if (C.inv_t[q] >= 0 && C.inv_t[q] < C.ncm) qt[C.inv_t[q]].push_back((int)q);
  for (int L = 0; L < nsub; ++L) {
    if (kept[L] == 0) continue;
    const int32_t* labL = &C.sub[L * (nv + 3)];
    for (int R = 0; R < nsub; ++R) {
      if (kept[R] == 0) continue;
      const int32_t* labR = &C.sub[R * (nv + 3)];
      int dsz = 0, dq = 0;
      for (int i = ss0; i < ss1; ++i) dsz += labL[i] - labR[i];
      if (sz_only) {
        if (std::abs(dsz) != 1) continue;
      } else {
        for (int i = cs0; i < cs1; ++i) dq += labL[i] - labR[i];
        if (dq != 1 || std::abs(dsz) != 1) continue;
      }
      for (int t = 0; t < C.ncm; ++t) {
        // matched row pairs (a_i, b_i) of all table rows of type t
        int K = 0;
        for (int q : qt[t])
          K += (int)std::min(rows[L].by_op[C.inv_a[q]].size(), rows[R].by_op[C.inv_b[q]].size());
        if (K == 0) continue;
        // A gathers rows b of U_R, B gathers rows a of U_L: two index lists
        // [b_1..b_K | a_1..a_K], and the coefficient of each pair
        const int64_t g0 = (int64_t)gidx.size();
        const int64_t s0 = (int64_t)gsc.size();
        gidx.resize(g0 + 2 * (int64_t)K);
        gsc.resize(s0 + K);
        int32_t* pb = gidx.data() + g0;
        int32_t* pa = pb + K;
        double* pc = gsc.data() + s0;
        for (int q : qt[t]) {
          const auto& ra = rows[L].by_op[C.inv_a[q]];
          const auto& rb = rows[R].by_op[C.inv_b[q]];
          const size_t np = std::min(ra.size(), rb.size());
          std::copy(ra.begin(), ra.begin() + np, pa);
          std::copy(rb.begin(), rb.begin() + np, pb);
          std::fill(pc, pc + np, C.inv_c[q]);
          pa += np;
          pb += np;
          pc += np;
        }
        GemmTask tk = gemm_task(kept[R], kept[L], K, 1, 0, C.boff[R], C.bn[R], C.boff[L], C.bn[L],
                                st[R] + st[L] * k + (int64_t)t * k * k, (int)k, 1.0, 0.0);
        tk.offGA = g0;
        tk.offGB = g0 + K;
        tk.offS = s0;
        batch.add(tk);
      }
    }
  }
  const int64_t ksz = k * k * C.ncm;
  C.d_inv_next.reserve(std::max<int64_t>(ksz, 1));
  fill(C.d_inv_next.p, 0.0, ksz, C.stream);
  if (!batch.tasks.empty()) {
    C.d_gidx.upload(gidx, C.stream);
    C.d_gscale.upload(gsc, C.stream);
    C.gemm.run(batch, C.d_U.p, C.d_U.p, C.d_inv_next.p, C.d_gidx.p, C.d_gidx.p, C.d_gscale.p, C.stream, KF_GEMM_INV);
  }
  std::swap(C.d_inv, C.d_inv_next);
  C.kold = (int)k;
  if (out_host) {
    C.d_inv.download(out_host, ksz, C.stream);
    NRG_HIP(hipStreamSynchronize(C.stream));
  }
}

void invariant_spectrum(Context& C, const int32_t* kept, double* out_host) {
  const int nsub = C.nsub;
  const int nv = C.nv;
  NRG_REQUIRE(C.nmat > 0 && !C.coef_spec.empty(), "invariant_spectrum: spectral tables not set");
  NRG_REQUIRE(C.d_spec.p != nullptr, "invariant_spectrum: no spectral operator on the device");
  int64_t k = 0;
  const std::vector<int64_t> st = kept_starts(kept, nsub, k);
  std::vector<BlockRows> rows(nsub);
  for (int b = 0; b < nsub; ++b) rows[b] = block_rows(C, b);
  const int64_t ko = C.kold_spec;
  std::vector<int32_t> gidx;
  GemmBatch bx, bm;
  int64_t xacc = 0;
  for (int L = 0; L < nsub; ++L) {
    if (kept[L] == 0) continue;
    const int32_t* labL = &C.sub[L * (nv + 3)];
    for (int R = 0; R < nsub; ++R) {
      if (kept[R] == 0) continue;
      const int32_t* labR = &C.sub[R * (nv + 3)];
      for (int im = 0; im < C.nmat; ++im) {
        bool ok = true;
        for (int i = 0; i < nv; ++i) ok = ok && (labL[i] - labR[i] == C.cons_diff_spec[im * nv + i]);
        if (!ok) continue;
        // stacked rows of L (per op with a counterpart in R)
        int Ma = 0;
        for (int o = 0; o < C.nops; ++o)
          if (!rows[L].by_op[o].empty() && !rows[R].by_op[o].empty()) Ma += (int)rows[L].by_op[o].size();
        if (Ma == 0) continue;
        const int64_t xoff = xacc;
        xacc += (int64_t)Ma * kept[R];
        int arow = 0;
        const int64_t ga0 = (int64_t)gidx.size();
        for (int o = 0; o < C.nops; ++o) {
          const auto& ra = rows[L].by_op[o];
          const auto& rb = rows[R].by_op[o];
          if (ra.empty() || rb.empty()) continue;
          const auto& refa = rows[L].ref_by_op[o];
          const auto& refb = rows[R].ref_by_op[o];
          for (size_t q = 1; q < refa.size(); ++q)
            NRG_REQUIRE(refa[q] == refa[0] + (int)q, "invariant_spectrum: parent range not contiguous");
          for (size_t q = 1; q < refb.size(); ++q)
            NRG_REQUIRE(refb[q] == refb[0] + (int)q, "invariant_spectrum: parent range not contiguous");
          // X_o(a, r) = c(im,o) sum_b S(ref_b, ref_a) U_R(b, r)
          const int64_t gb = (int64_t)gidx.size();
          gidx.insert(gidx.end(), rb.begin(), rb.end());
          GemmTask tx = gemm_task((int)ra.size(), kept[R], (int)rb.size(), 1, 0,
                                  (int64_t)(refb[0] - 1) + (int64_t)(refa[0] - 1) * ko + (int64_t)im * ko * ko,
                                  (int)ko, C.boff[R], C.bn[R], xoff + arow, Ma,
                                  C.coef_spec[im * C.nops + o], 0.0);
          tx.offGB = gb;
          bx.add(tx);
          arow += (int)ra.size();
        }
        // rows a (stacked, same order) for the second product
        const int64_t gaS = (int64_t)gidx.size();
        for (int o = 0; o < C.nops; ++o) {
          const auto& ra = rows[L].by_op[o];
          if (ra.empty() || rows[R].by_op[o].empty()) continue;
          gidx.insert(gidx.end(), ra.begin(), ra.end());
        }
        (void)ga0;
        GemmTask tm = gemm_task(kept[R], kept[L], Ma, 1, 0, xoff, Ma, C.boff[L], C.bn[L],
                                st[R] + st[L] * k + (int64_t)im * k * k, (int)k, 1.0, 0.0);
        tm.offGB = gaS;
        bm.add(tm);
      }
    }
  }
  const int64_t ksz = k * k * C.nmat;
  C.d_spec_next.reserve(std::max<int64_t>(ksz, 1));
  fill(C.d_spec_next.p, 0.0, ksz, C.stream);
  if (!bm.tasks.empty()) {
    C.d_tmp.reserve(xacc);
    C.d_gidx.upload(gidx, C.stream);
    C.gemm.run(bx, C.d_spec.p, C.d_U.p, C.d_tmp.p, nullptr, C.d_gidx.p, nullptr, C.stream, KF_GEMM_SPEC);
    C.gemm.run(bm, C.d_tmp.p, C.d_U.p, C.d_spec_next.p, nullptr, C.d_gidx.p, nullptr, C.stream, KF_GEMM_SPEC);
  }
  std::swap(C.d_spec, C.d_spec_next);
  C.kold_spec = (int)k;
  if (out_host) C.d_spec.download(out_host, ksz, C.stream);
  NRG_HIP(hipStreamSynchronize(C.stream));
}

}  // namespace nrg
