// extern "C" entry points of libnrg_gpu.so (declared in include/nrg_gpu.h).
#include "../../include/nrg_gpu.h"
#include "assemble.h"
#include "noalloc.h"
#include "calib.h"
#include "cfs.h"
#include "context.h"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

namespace nrg {

Context& ctx() {
  static Context c;
  return c;
}

Profiler& profiler() {
  static Profiler p;
  return p;
}

static double wall_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <typename F>
static int guarded(F&& f) {
  Context& C = ctx();
  try {
    f(C);
    if (profiler().on && C.stream) profiler().drain_ready();
    C.last_error.clear();
    return NRG_OK;
  } catch (const Error& e) {
    C.last_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    C.last_error = e.what();
    return NRG_ERR_STATE;
  }
}

static void need_init(Context& C) {
  if (!C.initialised) throw Error(NRG_ERR_STATE, "nrg_gpu_init has not been called");
}

// Parse subspace records into block sizes and packed offsets.
static void set_blocks(Context& C, int nsub, const int32_t* sub, int nfull) {
  const int rs = C.nv + 3;
  C.nsub = nsub;
  C.nfull = nfull;
  C.sub.assign(sub, sub + (size_t)nsub * rs);
  C.bn.resize(nsub);
  C.boff.resize(nsub);
  C.bvoff.resize(nsub);
  int64_t o2 = 0, o1 = 0;
  for (int b = 0; b < nsub; ++b) {
    const int n = sub[b * rs + C.nv];
    const int64_t cnt = sub[b * rs + C.nv + 1];
    const int64_t si = sub[b * rs + C.nv + 2];
    NRG_REQUIRE(n >= 1, "subspace dimension < 1");
    NRG_REQUIRE(si - 1 == o1, "subspace start_input is not the running sum of dimensions");
    // count_eigenvector is int32 in the reference and may have wrapped: the
    // packed offset is recomputed in 64 bits and checked modulo 2^32
    NRG_REQUIRE((uint32_t)(cnt - 1) == (uint32_t)o2, "subspace count_eigenvector is not the running sum of dimension^2");
    C.bn[b] = n;
    C.boff[b] = o2;
    C.bvoff[b] = o1;
    o2 += (int64_t)n * n;
    o1 += n;
  }
  NRG_REQUIRE(o1 == nfull, "sum of subspace dimensions differs from numberOfBasis");
}

}  // namespace nrg

using namespace nrg;

extern "C" {

int nrg_gpu_init(int device) {
  return guarded([&](Context& C) {
    if (C.initialised) return;
    int nd = 0;
    NRG_HIP(hipGetDeviceCount(&nd));
    NRG_REQUIRE(device >= 0 && device < nd, "nrg_gpu_init: no such device");
    NRG_HIP(hipSetDevice(device));
    C.device = device;
    scratch_setup(C);
    NRG_HIP(hipStreamCreateWithFlags(&C.stream, hipStreamNonBlocking));
    NRG_HIP(hipEventCreate(&C.ev0));
    NRG_HIP(hipEventCreate(&C.ev1));
    for (hipEvent_t& e : C.ev_dz) NRG_HIP(hipEventCreate(&e));
    for (int i = 0; i < 2; ++i) {
      NRG_HIP(hipEventCreate(&C.ev_back[i]));
      NRG_HIP(hipEventCreate(&C.ev_inv[i]));
    }
    C.initialised = true;
  });
}

int nrg_gpu_finalize(void) {
  return guarded([&](Context& C) {
    if (!C.initialised) return;
    (void)hipStreamSynchronize(C.stream);
    cfs().release();
    C.eig.release();
    C.u_lazy = false;   // no solve left whose columns could be transformed
    C.back_timed = false;
    C.inv_timed = false;
    C.gemm.release();
    C.d_op.release(); C.d_hop.release(); C.d_hop_coef.release();
    C.d_basis_in.release(); C.d_H.release(); C.d_U.release(); C.d_w.release(); C.d_Eold.release();
    C.d_inv.release(); C.d_inv_next.release(); C.d_spec.release(); C.d_spec_next.release();
    C.d_tmp.release(); C.d_gidx.release(); C.d_gscale.release();
    (void)hipEventDestroy(C.ev0);
    (void)hipEventDestroy(C.ev1);
    for (hipEvent_t& e : C.ev_dz) {
      (void)hipEventDestroy(e);
      e = nullptr;
    }
    for (int i = 0; i < 2; ++i) {
      (void)hipEventDestroy(C.ev_back[i]);
      (void)hipEventDestroy(C.ev_inv[i]);
    }
    (void)hipStreamDestroy(C.stream);
    C.stream = nullptr;
    C.initialised = false;
    C.kold = 0;
    C.kold_spec = 0;
  });
}

const char* nrg_gpu_last_error(void) { return ctx().last_error.c_str(); }

// device copy of the pair tables read by the H assembly:
// [valid, type, channel, swap, BCS channel + 1] x nops^2 and the coefficients
static void upload_hop_table(Context& C) {
  const int nn = C.nops * C.nops;
  std::vector<int32_t> hop(5 * nn);
  for (int i = 0; i < nn; ++i) {
    hop[i] = C.hop_valid[i]; hop[nn + i] = C.hop_type[i]; hop[2 * nn + i] = C.hop_chan[i];
    hop[3 * nn + i] = C.hop_swap[i]; hop[4 * nn + i] = C.bcs_chan[i];
  }
  C.d_hop.upload(hop, C.stream);
  C.d_hop_coef.upload(C.hop_coef, C.stream);
  NRG_HIP(hipStreamSynchronize(C.stream));
}

int nrg_gpu_set_model(int nv, int nops, int nchain, int ncm, int ncoef, const int32_t* operation,
                      const int32_t* coef_diag_type, const double* coef_diag,
                      const int32_t* coef_inv_type, const double* coef_inv,
                      const int32_t* conservation_block, int nmat, const double* coef_spec,
                      const int32_t* cons_diff_spec) {
  return guarded([&](Context& C) {
    need_init(C);
    NRG_REQUIRE(nv >= 1 && nops >= 1 && nchain >= 1 && ncm >= 1 && ncoef >= 0, "set_model: bad sizes");
    NRG_REQUIRE(nchain <= 4, "set_model: at most 4 chains");
    C.nv = nv; C.nops = nops; C.nchain = nchain; C.ncm = ncm; C.ncoef = ncoef; C.nmat = nmat;
    const int nc = nv + nchain;
    C.operation.resize((size_t)nops * nc);
    for (int o = 0; o < nops; ++o)
      for (int j = 0; j < nc; ++j) C.operation[o * nc + j] = operation[o + j * nops];
    // hopping lookup: replay every coefficient row in order for each
    // (op_i, op_j); the last match sets (type, channel, coef), and a
    // reversed match swaps the references for good: a later forward match
    // does not restore them (diagonalization.F90:195-213)
    const int nn = nops * nops;
    C.hop_valid.assign(nn, 0); C.hop_type.assign(nn, 0); C.hop_chan.assign(nn, 0);
    C.hop_swap.assign(nn, 0); C.hop_coef.assign(nn, 0.0);
    for (int oi = 0; oi < nops; ++oi)
      for (int oj = 0; oj < nops; ++oj) {
        const int idx = oi * nops + oj;
        for (int ic = 0; ic < ncoef; ++ic) {
          const int a = coef_diag_type[ic] - 1, b = coef_diag_type[ic + ncoef] - 1;
          const int t = coef_diag_type[ic + 2 * ncoef] - 1, ch = coef_diag_type[ic + 3 * ncoef] - 1;
          int sw;
          if (oi == a && oj == b) sw = 0;
          else if (oi == b && oj == a) sw = 1;
          else continue;
          NRG_REQUIRE(t >= 0 && t < ncm && ch >= 0 && ch < nchain, "set_model: bad hopping row");
          C.hop_valid[idx] = 1; C.hop_type[idx] = t; C.hop_chan[idx] = ch; C.hop_swap[idx] |= sw;
          C.hop_coef[idx] = coef_diag[ic];
        }
      }
    // invariant table, deduplicated on (op_a, op_b, t): the later row wins
    // (invariantMatrixDGEMM.F90:248-253 overwrites the same slot)
    C.inv_a.clear(); C.inv_b.clear(); C.inv_t.clear(); C.inv_c.clear();
    for (int ic = 0; ic < ncoef; ++ic) {
      const int a = coef_inv_type[ic] - 1, b = coef_inv_type[ic + ncoef] - 1, t = coef_inv_type[ic + 2 * ncoef] - 1;
      NRG_REQUIRE(a >= 0 && a < nops && b >= 0 && b < nops && t >= 0 && t < ncm, "set_model: bad invariant row");
      bool found = false;
      for (size_t q = 0; q < C.inv_a.size(); ++q)
        if (C.inv_a[q] == a && C.inv_b[q] == b && C.inv_t[q] == t) { C.inv_c[q] = coef_inv[ic]; found = true; }
      if (!found) { C.inv_a.push_back(a); C.inv_b.push_back(b); C.inv_t.push_back(t); C.inv_c.push_back(coef_inv[ic]); }
    }
    C.cons = {conservation_block[0], conservation_block[2], conservation_block[1], conservation_block[3]};
    C.coef_spec.clear();
    C.cons_diff_spec.clear();
    if (nmat > 0 && coef_spec && cons_diff_spec) {
      NRG_REQUIRE(nmat <= ncm, "set_model: numberOfMatrix must not exceed numberOfConductionMatrix");
      C.coef_spec.resize((size_t)nmat * nops);
      for (int im = 0; im < nmat; ++im)
        for (int o = 0; o < nops; ++o) C.coef_spec[im * nops + o] = coef_spec[im + o * nmat];
      C.cons_diff_spec.resize((size_t)nmat * nv);
      for (int im = 0; im < nmat; ++im)
        for (int i = 0; i < nv; ++i) C.cons_diff_spec[im * nv + i] = cons_diff_spec[im + i * nmat];
    }
    C.d_op.upload(C.operation, C.stream);
    C.bcs_chan.assign(nn, 0);
    C.chain_bcs.clear();
    C.bcs_rows = 0;
    upload_hop_table(C);
  });
}

int nrg_gpu_set_bcs(int nbcs, const int32_t* coef_bcs, int nrow, const double* chain_bcs) {
  return guarded([&](Context& C) {
    need_init(C);
    NRG_REQUIRE(C.nops > 0, "set_bcs: call nrg_gpu_set_model first");
    NRG_REQUIRE(nbcs >= 0 && (nbcs == 0 || (coef_bcs && chain_bcs && nrow >= 1)), "set_bcs: bad sizes");
    const int nn = C.nops * C.nops;
    C.bcs_chan.assign(nn, 0);
    // a row matches in its forward order only, with equal references; the
    // last match sets the channel (diagonalization.F90:215-223)
    for (int ic = 0; ic < nbcs; ++ic) {
      const int a = coef_bcs[ic] - 1, b = coef_bcs[ic + nbcs] - 1, ch = coef_bcs[ic + 2 * nbcs] - 1;
      NRG_REQUIRE(a >= 0 && a < C.nops && b >= 0 && b < C.nops && ch >= 0 && ch < C.nchain,
                  "set_bcs: bad coefficient_BCS row");
      C.bcs_chan[a * C.nops + b] = ch + 1;
    }
    C.bcs_rows = nbcs ? nrow : 0;
    C.chain_bcs.assign(chain_bcs, chain_bcs + (nbcs ? (size_t)nrow * C.nchain : 0));
    upload_hop_table(C);
  });
}

int nrg_gpu_set_invariant(int nold, const double* inv) {
  return guarded([&](Context& C) {
    need_init(C);
    NRG_REQUIRE(C.ncm > 0, "set_invariant: call nrg_gpu_set_model first");
    C.d_inv.upload(inv, (size_t)nold * nold * C.ncm, C.stream);
    C.kold = nold;
    NRG_HIP(hipStreamSynchronize(C.stream));
  });
}

int nrg_gpu_set_invariant_spectrum(int nold, const double* ims) {
  return guarded([&](Context& C) {
    need_init(C);
    NRG_REQUIRE(C.nmat > 0, "set_invariant_spectrum: call nrg_gpu_set_model with spectral tables first");
    C.d_spec.upload(ims, (size_t)nold * nold * C.nmat, C.stream);
    C.kold_spec = nold;
    NRG_HIP(hipStreamSynchronize(C.stream));
  });
}

int nrg_gpu_diagonalize(int iteration, int nsub, const int32_t* subspace_info, int nfull,
                        const int32_t* basis_input, int nold, const double* eigenvalue_old,
                        const double* chain_diag, const double* chain_nondiag, double sqrt_lambda,
                        double* eigenvalue, double* eigenvector, int64_t eigenvector_size) {
  return guarded([&](Context& C) {
    need_init(C);
    if (iteration == C.fault_iteration) {
      C.fault_iteration = -1000000;
      throw Error(NRG_ERR_INJECTED, "injected fault in diagonalization at iteration " + std::to_string(iteration));
    }
    NRG_REQUIRE(C.nops > 0, "diagonalize: call nrg_gpu_set_model first");
    NRG_REQUIRE(nold == C.kold, "diagonalize: eigenvalue_old size differs from the invariant-matrix basis");
    set_blocks(C, nsub, subspace_info, nfull);
    const int rs = C.nv + 3;
    C.basis_in.assign(basis_input, basis_input + (size_t)nfull * rs);
    for (int i = 0; i < nfull; ++i) {
      const int ref = basis_input[i * rs + C.nv], op = basis_input[i * rs + C.nv + 2];
      NRG_REQUIRE(ref >= 1 && ref <= nold && op >= 1 && op <= C.nops, "diagonalize: basis_input reference/operation out of range");
    }
    int64_t n2 = 0;
    for (int b = 0; b < nsub; ++b) n2 += (int64_t)C.bn[b] * C.bn[b];
    if (eigenvector) NRG_REQUIRE(eigenvector_size == n2, "diagonalize: eigenvector_size differs from sum of dimension^2");
    C.d_basis_in.upload(C.basis_in, C.stream);
    C.d_Eold.upload(eigenvalue_old, nold, C.stream);
    C.d_H.reserve(n2);
    C.d_U.reserve(n2);
    C.d_w.reserve(nfull);
    // stage times from events, read once the eigenvalues have been waited
    // for: no synchronisation between the stages, so the host prepares the
    // eigensolver's tables while the device assembles
    NRG_HIP(hipEventRecord(C.ev_dz[0], C.stream));
    double cbcs[4] = {0, 0, 0, 0};   // chain_BCS(iteration+1, ch) (diagonalization.F90:242)
    if (C.bcs_rows) {
      NRG_REQUIRE(iteration + 1 >= 0 && iteration + 1 < C.bcs_rows, "diagonalize: iteration outside chain_BCS");
      for (int ch = 0; ch < C.nchain && ch < 4; ++ch) cbcs[ch] = C.chain_bcs[(size_t)ch * C.bcs_rows + iteration + 1];
    }
    assemble(C, chain_diag, chain_nondiag, sqrt_lambda, C.bcs_rows ? cbcs : nullptr);
    NRG_HIP(hipEventRecord(C.ev_dz[1], C.stream));
    C.u_lazy = false;
    std::string tune_key;
    TuneConfig tune_cfg;
    if (C.batch_limit <= 0 || C.batch_limit >= nsub) {
      // without an eigenvector output the back-transform waits for the
      // columns the invariant / stock calls read (ensure_columns)
      const bool defer = (eigenvector == nullptr);
      // launch configuration of the tridiagonalisation: the tuner's choice
      // for this block-size histogram (the defaults with tuning off); no
      // graph capture while a candidate is being timed
      bool explore = false;
      std::string tkey;
      TuneConfig cfg = Tuner::candidates()[0];
      if (C.tuner.mode()) {
        tkey = Tuner::key(C.bn.data(), nsub);
        cfg = C.tuner.choose(tkey, &explore);
      }
      C.eig.set_groups(C.groups_override > 0 ? C.groups_override : cfg.groups);
      C.eig.set_symv_wide(cfg.symv_wide_wg, 64.0e6);
      C.eig.set_graphs(C.graphs_on && !explore);
      C.eig.solve(nsub, C.bn.data(), C.boff.data(), C.bvoff.data(), C.d_H.p, C.d_U.p, C.d_w.p, defer,
                  C.stream, &C.eig_stats);
      // stage times and the tuner's measurement once the stream has been
      // waited for below
      tune_key = tkey;
      tune_cfg = cfg;
      C.u_lazy = defer;
    } else {
      // groups of at most batch_limit blocks, one solve each (batch_limit 1:
      // one block at a time, the reference's per-block dsyev schedule)
      EigStats sum;
      for (int b0 = 0; b0 < nsub; b0 += C.batch_limit) {
        const int nb = std::min(C.batch_limit, nsub - b0);
        // offsets relative to the group's first block (the solver sizes its
        // workspaces by the group)
        std::vector<int64_t> ao(nb), vo(nb);
        for (int q = 0; q < nb; ++q) {
          ao[q] = C.boff[b0 + q] - C.boff[b0];
          vo[q] = C.bvoff[b0 + q] - C.bvoff[b0];
        }
        EigStats st;
        C.eig.solve(nb, C.bn.data() + b0, ao.data(), vo.data(), C.d_H.p + C.boff[b0], C.d_U.p + C.boff[b0],
                    C.d_w.p + C.bvoff[b0], false, C.stream, &st);
        C.eig.stats(&st);   // waits for this group's solve
        sum.ms_tridiag += st.ms_tridiag; sum.ms_dc += st.ms_dc; sum.ms_back += st.ms_back;
        sum.ms_tridiag_issue += st.ms_tridiag_issue;
        sum.flops_alg += st.flops_alg;
        sum.max_n = std::max(sum.max_n, st.max_n);
        sum.nblk += st.nblk;
      }
      C.eig_stats = sum;
    }
    NRG_HIP(hipEventRecord(C.ev_dz[2], C.stream));
    C.d_w.download(eigenvalue, nfull, C.stream);
    if (eigenvector) C.d_U.download(eigenvector, n2, C.stream);
    NRG_HIP(hipStreamSynchronize(C.stream));
    {
      float a = 0.f, e = 0.f;
      NRG_HIP(hipEventElapsedTime(&a, C.ev_dz[0], C.ev_dz[1]));
      NRG_HIP(hipEventElapsedTime(&e, C.ev_dz[1], C.ev_dz[2]));
      C.timers.assemble = a;
      C.timers.eig = e;
    }
    if (C.batch_limit <= 0 || C.batch_limit >= nsub) {
      C.eig.stats(&C.eig_stats);
      if (C.tuner.mode() == 1) C.tuner.record(tune_key, tune_cfg, C.eig_stats.ms_tridiag);
    }
    C.timers.tridiag = C.eig_stats.ms_tridiag;
    C.timers.dc = C.eig_stats.ms_dc;
    C.timers.back = C.eig_stats.ms_back;
    C.back_timed = C.inv_timed = false;
    // global shift and chop (diagonalization.F90:351-356)
    double mn = eigenvalue[0];
    for (int i = 1; i < nfull; ++i) mn = std::min(mn, eigenvalue[i]);
    for (int i = 0; i < nfull; ++i) {
      const double v = eigenvalue[i] - mn;
      eigenvalue[i] = (v <= 1.0e-12) ? 0.0 : v;
    }
    for (int i = 0; i < nfull; ++i)
      if (!std::isfinite(eigenvalue[i])) throw Error(NRG_ERR_NUMERIC, "diagonalize: non-finite eigenvalue");
    C.E_block.assign(eigenvalue, eigenvalue + nfull);
  });
}

int nrg_gpu_set_spectrum_operators(int nmat, const int32_t* pos_op, const int32_t* neg_op) {
  return guarded([&](Context& C) {
    need_init(C);
    NRG_REQUIRE(C.nv > 0 && nmat == C.nmat && nmat > 0, "set_spectrum_operators: nmat differs from the model");
    NRG_REQUIRE(pos_op && neg_op, "set_spectrum_operators: NULL table");
    C.pos_op.resize((size_t)nmat * C.nv);
    C.neg_op.resize((size_t)nmat * C.nv);
    for (int im = 0; im < nmat; ++im)
      for (int i = 0; i < C.nv; ++i) {
        C.pos_op[im * C.nv + i] = pos_op[im + i * nmat];
        C.neg_op[im * C.nv + i] = neg_op[im + i * nmat];
      }
  });
}

int nrg_gpu_stock(int iteration, const int32_t* kept_per_block, int last) {
  return guarded([&](Context& C) {
    need_init(C);
    NRG_REQUIRE(kept_per_block != nullptr, "stock: NULL kept_per_block");
    const double t0 = wall_ms();
    ensure_columns(C, nullptr);   // kept and discarded columns enter the spectral products
    stock(C, iteration, kept_per_block, last);
    C.timers.cfs = wall_ms() - t0;
  });
}

int nrg_gpu_rho_init(int last_iteration) {
  return guarded([&](Context& C) {
    need_init(C);
    rho_init(C, last_iteration);
  });
}

int nrg_gpu_rho_shell(int iteration, double* trace) {
  return guarded([&](Context& C) {
    need_init(C);
    const double t0 = wall_ms();
    const double t = rho_shell(C, iteration);
    if (trace) *trace = t;
    C.timers.rho = wall_ms() - t0;
  });
}

int nrg_gpu_rho_get(int iteration, double* rho, int64_t size) {
  return guarded([&](Context& C) {
    need_init(C);
    HistEntry* h = cfs().at(iteration);
    NRG_REQUIRE(h && h->rho_ready, "rho_get: no reduced density for this iteration");
    int64_t tot = 0;
    for (int b = 0; b < h->nsub; ++b) tot += (int64_t)h->kept[b] * h->kept[b];
    NRG_REQUIRE(size == tot, "rho_get: size differs from the sum of kept^2");
    std::vector<double> col(tot);
    h->rho.download(col.data(), tot, C.stream);
    NRG_HIP(hipStreamSynchronize(C.stream));
    for (int b = 0; b < h->nsub; ++b) {
      const int k = h->kept[b];
      const double* s = col.data() + h->roff[b];
      double* d = rho + h->roff[b];
      for (int i = 0; i < k; ++i)
        for (int j = 0; j < k; ++j) d[(int64_t)i * k + j] = s[i + (int64_t)j * k];
    }
  });
}

int nrg_gpu_spectrum_setup(int nbin) {
  return guarded([&](Context& C) {
    need_init(C);
    spectrum_setup(C, nbin);
  });
}

int nrg_gpu_spike_cfs(int iteration, double scale) {
  return guarded([&](Context& C) {
    need_init(C);
    const double t0 = wall_ms();
    spikes(C, iteration, scale, 0);
    C.timers.cfs = wall_ms() - t0;
  });
}

int nrg_gpu_spike_zero_temperature(int iteration, double scale) {
  return guarded([&](Context& C) {
    need_init(C);
    const double t0 = wall_ms();
    spikes(C, iteration, scale, 1);
    C.timers.cfs = wall_ms() - t0;
  });
}

int nrg_gpu_save_restart(const char* path, int iteration, int nstate, const int32_t* state) {
  return guarded([&](Context& C) {
    need_init(C);
    save_restart(C, path, iteration, nstate, state);
  });
}

int nrg_gpu_load_restart(const char* path, int* iteration, int nstate, int32_t* state) {
  return guarded([&](Context& C) {
    need_init(C);
    NRG_REQUIRE(iteration != nullptr, "load_restart: null iteration");
    load_restart(C, path, iteration, nstate, state);
  });
}

int nrg_gpu_zero_temperature_weights(int gs_block, int type, int positive, double* w, int cap,
                                     int* excited_block, int* n) {
  return guarded([&](Context& C) {
    need_init(C);
    NRG_REQUIRE(w && excited_block && n, "zero_temperature_weights: null output");
    int K = -1;
    zero_temperature_weights(C, gs_block - 1, type - 1, positive ? 1 : 0, w, cap, &K, n);
    *excited_block = K + 1;
  });
}

int nrg_gpu_spectrum_bins(double* negative, double* positive, int64_t size) {
  return guarded([&](Context& C) {
    need_init(C);
    Cfs& F = cfs();
    NRG_REQUIRE(F.nbin > 0, "spectrum_bins: call nrg_gpu_spectrum_setup first");
    const int64_t n = (int64_t)F.nbin * C.nmat;
    NRG_REQUIRE(size == n, "spectrum_bins: size differs from nbin * nmat");
    if (negative) NRG_HIP(hipMemcpyAsync(negative, F.bins.p, n * sizeof(double), hipMemcpyDeviceToHost, C.stream));
    if (positive)
      NRG_HIP(hipMemcpyAsync(positive, F.bins.p + n, n * sizeof(double), hipMemcpyDeviceToHost, C.stream));
    NRG_HIP(hipStreamSynchronize(C.stream));
  });
}

int nrg_gpu_history_clear(void) {
  return guarded([&](Context& C) {
    if (C.stream) NRG_HIP(hipStreamSynchronize(C.stream));
    cfs().clear();
  });
}

int nrg_gpu_invariant(const int32_t* kept_per_block, double* out) {
  return guarded([&](Context& C) {
    need_init(C);
    NRG_REQUIRE(C.nsub > 0, "invariant: no diagonalised iteration");
    // returns once the work is queued (unless `out` is asked for): the
    // caller's host work for the next iteration overlaps the device
    ensure_columns(C, kept_per_block);
    NRG_HIP(hipEventRecord(C.ev_inv[0], C.stream));
    invariant(C, kept_per_block, out);
    NRG_HIP(hipEventRecord(C.ev_inv[1], C.stream));
    C.inv_timed = true;
  });
}

int nrg_gpu_invariant_get(double* out, int64_t size) {
  return guarded([&](Context& C) {
    need_init(C);
    const int64_t k = C.kold;
    NRG_REQUIRE(C.ncm > 0 && k > 0, "invariant_get: no invariant matrix on the device");
    NRG_REQUIRE(out && size == k * k * C.ncm, "invariant_get: size differs from k*k*numberOfConductionMatrix");
    C.d_inv.download(out, (size_t)size, C.stream);
    NRG_HIP(hipStreamSynchronize(C.stream));
  });
}

int nrg_gpu_invariant_spectrum(const int32_t* kept_per_block, double* out) {
  return guarded([&](Context& C) {
    need_init(C);
    NRG_REQUIRE(C.nsub > 0, "invariant_spectrum: no diagonalised iteration");
    const double t0 = wall_ms();
    ensure_columns(C, kept_per_block);
    invariant_spectrum(C, kept_per_block, out);
    C.timers.spectral = wall_ms() - t0;
  });
}

int nrg_gpu_stage_times(double* t, int n) {
  return guarded([&](Context& C) {
    // spans of work that returned without waiting for the device (the
    // deferred back-transform, the invariant products): read once they have
    // completed, never waited for here, so that asking for the times does
    // not drain the device between iterations; until a span completes, the
    // last completed one is reported
    auto span = [&](hipEvent_t* e, double& last) {
      const hipError_t q = hipEventQuery(e[1]);
      if (q == hipSuccess) {
        float ms = 0.f;
        NRG_HIP(hipEventElapsedTime(&ms, e[0], e[1]));
        last = ms;
      } else if (q != hipErrorNotReady) {
        NRG_HIP(q);
      }
      return last;
    };
    double back = C.timers.back;
    if (C.back_timed) back += span(C.ev_back, C.last_back_ms);
    const double inv = C.inv_timed ? span(C.ev_inv, C.last_inv_ms) : 0.0;
    const double v[13] = {C.timers.assemble, C.timers.eig, C.timers.tridiag, C.timers.dc, back,
                          inv, C.timers.spectral, C.timers.rho, C.timers.cfs,
                          C.eig_stats.flops_alg, (double)C.eig.graph_captures(), (double)C.eig.graph_replays(),
                          C.eig_stats.ms_tridiag_issue};
    for (int i = 0; i < n && i < 13; ++i) t[i] = v[i];
  });
}

int nrg_gpu_syev_batched(int nblk, const int32_t* n, const double* H, double* w, double* U) {
  return guarded([&](Context& C) {
    need_init(C);
    std::vector<int> nn(n, n + nblk);
    std::vector<int64_t> ao(nblk), vo(nblk);
    int64_t a = 0, v = 0;
    for (int b = 0; b < nblk; ++b) {
      NRG_REQUIRE(nn[b] >= 1, "syev_batched: block size < 1");
      ao[b] = a; vo[b] = v;
      a += (int64_t)nn[b] * nn[b];
      v += nn[b];
    }
    DevBuf<double> dA, dZ, dw;
    dA.upload(H, a, C.stream);
    dZ.reserve(a);
    dw.reserve(v);
    ensure_columns(C, nullptr);   // the solver's state is reused below
    C.eig.solve(nblk, nn.data(), ao.data(), vo.data(), dA.p, dZ.p, dw.p, false, C.stream, &C.eig_stats);
    dw.download(w, v, C.stream);
    dZ.download(U, a, C.stream);
    NRG_HIP(hipStreamSynchronize(C.stream));
    C.eig.stats(&C.eig_stats);
    C.timers.tridiag = C.eig_stats.ms_tridiag;
    C.timers.dc = C.eig_stats.ms_dc;
    C.timers.back = C.eig_stats.ms_back;
    C.back_timed = false;
    dA.release(); dZ.release(); dw.release();
  });
}

int nrg_gpu_dgemm(int transa, int transb, int m, int n, int k, double alpha, const double* A, int lda,
                  const double* B, int ldb, double beta, double* C_, int ldc) {
  return guarded([&](Context& C) {
    need_init(C);
    NRG_REQUIRE(m >= 0 && n >= 0 && k >= 0, "dgemm: negative size");
    if (m == 0 || n == 0) return;
    const int64_t na = (int64_t)lda * (transa ? m : k), nb = (int64_t)ldb * (transb ? k : n), nc = (int64_t)ldc * n;
    DevBuf<double> dA, dB, dC;
    dA.upload(A, std::max<int64_t>(na, 1), C.stream);
    dB.upload(B, std::max<int64_t>(nb, 1), C.stream);
    dC.upload(C_, nc, C.stream);
    GemmBatch b;
    b.add(gemm_task(m, n, k, transa, transb, 0, lda, 0, ldb, 0, ldc, alpha, beta));
    C.gemm.run(b, dA.p, dB.p, dC.p, nullptr, nullptr, nullptr, C.stream);
    dC.download(C_, nc, C.stream);
    NRG_HIP(hipStreamSynchronize(C.stream));
    dA.release(); dB.release(); dC.release();
  });
}

int nrg_gpu_dgemm_batched_noalloc(int ntask, const int32_t* mnk, int transa, int transb, const double* A,
                                  const int64_t* offA, const double* B, const int64_t* offB, double* C_,
                                  const int64_t* offC) {
  return guarded([&](Context& C) {
    need_init(C);
    dgemm_batched_noalloc(C, ntask, mnk, transa, transb, A, offA, B, offB, C_, offC);
  });
}

int nrg_gpu_dgemm_batched(int ntask, const int32_t* mnk, int transa, int transb, const double* A,
                          const int64_t* offA, const double* B, const int64_t* offB, double* C_,
                          const int64_t* offC, int slices, int repeat, double* ms) {
  return guarded([&](Context& C) {
    need_init(C);
    NRG_REQUIRE(ntask >= 0 && mnk && offA && offB && offC, "dgemm_batched: bad arguments");
    NRG_REQUIRE(slices == 0 || (slices >= 3 && slices <= 8), "dgemm_batched: slices must be 0 or 3..8");
    if (ntask == 0) return;
    GemmBatch b;
    int64_t na = 0, nb = 0, nc = 0;
    for (int i = 0; i < ntask; ++i) {
      const int m = mnk[3 * i], n = mnk[3 * i + 1], k = mnk[3 * i + 2];
      NRG_REQUIRE(m > 0 && n > 0 && k > 0, "dgemm_batched: sizes must be positive");
      const int lda = transa ? k : m, ldb = transb ? n : k;
      na = std::max(na, offA[i] + (int64_t)m * k);
      nb = std::max(nb, offB[i] + (int64_t)k * n);
      nc = std::max(nc, offC[i] + (int64_t)m * n);
      b.add(gemm_task(m, n, k, transa, transb, offA[i], lda, offB[i], ldb, offC[i], m, 1.0, 0.0));
    }
    DevBuf<double> dA, dB, dC;
    dA.upload(A, na, C.stream);
    dB.upload(B, nb, C.stream);
    dC.reserve(nc);
    auto once = [&]() {
      if (slices > 0) C.gemm.run_split(b, dA.p, dB.p, dC.p, nullptr, nullptr, nullptr, slices, C.stream);
      else {
        const int keep = C.gemm.split;
        C.gemm.split = 0;
        C.gemm.run(b, dA.p, dB.p, dC.p, nullptr, nullptr, nullptr, C.stream);
        C.gemm.split = keep;
      }
    };
    once();
    if (repeat > 0) {
      NRG_HIP(hipEventRecord(C.ev0, C.stream));
      for (int r = 0; r < repeat; ++r) once();
      NRG_HIP(hipEventRecord(C.ev1, C.stream));
      NRG_HIP(hipEventSynchronize(C.ev1));
      float t = 0.f;
      NRG_HIP(hipEventElapsedTime(&t, C.ev0, C.ev1));
      if (ms) *ms = (double)t / repeat;
    }
    dC.download(C_, nc, C.stream);
    NRG_HIP(hipStreamSynchronize(C.stream));
    dA.release(); dB.release(); dC.release();
  });
}

int nrg_gpu_set_gemm_split(int slices) {
  return guarded([&](Context& C) {
    NRG_REQUIRE(slices == 0 || (slices >= 3 && slices <= 8), "set_gemm_split: slices must be 0 or 3..8");
    C.gemm.split = slices;
    C.gemm.split_spec = slices;
    C.gemm.split_min_flops = 0.0;
  });
}

int nrg_gpu_set_gemm_split_families(int invariant_slices, int spectral_slices, double min_batch_flops) {
  return guarded([&](Context& C) {
    auto ok = [](int s) { return s == 0 || (s >= 3 && s <= 8); };
    NRG_REQUIRE(ok(invariant_slices) && ok(spectral_slices),
                "set_gemm_split_families: slices must be 0 or 3..8");
    NRG_REQUIRE(min_batch_flops >= 0.0, "set_gemm_split_families: min_batch_flops must be >= 0");
    C.gemm.split = invariant_slices;
    C.gemm.split_spec = spectral_slices;
    C.gemm.split_min_flops = min_batch_flops;
  });
}

int nrg_gpu_set_batch_limit(int max_blocks) {
  return guarded([&](Context& C) {
    NRG_REQUIRE(max_blocks >= 0, "set_batch_limit: negative limit");
    C.batch_limit = max_blocks;
  });
}

int nrg_gpu_set_graphs(int on) {
  return guarded([&](Context& C) {
    need_init(C);
    C.graphs_on = (on != 0);
  });
}

int nrg_gpu_set_tuning(int mode, const char* cache_path) {
  return guarded([&](Context& C) {
    need_init(C);
    NRG_REQUIRE(mode >= 0 && mode <= 2, "set_tuning: mode must be 0, 1 or 2");
    C.tuner.set(mode, cache_path ? std::string(cache_path) : std::string());
  });
}

int nrg_gpu_set_tridiag_symv(int kind) {
  return guarded([&](Context& C) {
    need_init(C);
    NRG_REQUIRE(kind == 0 || kind == 1, "set_tridiag_symv: kind must be 0 or 1");
    C.eig.set_symv_kind(kind);
  });
}

int nrg_gpu_set_tridiag_groups(int groups) {
  return guarded([&](Context& C) {
    need_init(C);
    NRG_REQUIRE(groups >= 0 && groups <= 4, "set_tridiag_groups: groups must be 0..4");
    C.groups_override = groups;
    C.eig.set_groups(groups > 0 ? groups : Tuner::candidates()[0].groups);
  });
}

int nrg_gpu_set_tridiag_finish(int rows) {
  return guarded([&](Context& C) {
    need_init(C);
    NRG_REQUIRE(rows >= 0 && rows <= 192, "set_tridiag_finish: rows must be 0..192");
    C.eig.set_finish(rows);
  });
}

int nrg_gpu_set_tridiag_method(int two_stage) {
  return guarded([&](Context& C) {
    need_init(C);
    C.eig.set_two_stage(two_stage != 0);
  });
}

int nrg_gpu_set_tridiag_schedule(int align_end) {
  return guarded([&](Context& C) {
    need_init(C);
    C.eig.set_align_end(align_end != 0);
  });
}

int nrg_gpu_inject_fault(int iteration) {
  return guarded([&](Context& C) { C.fault_iteration = iteration; });
}

int nrg_gpu_profile(int enable) {
  return guarded([&](Context& C) {
    if (C.stream) NRG_HIP(hipDeviceSynchronize());   // every stream of the library
    profiler().drain();
    profiler().reset();
    profiler().on = (enable != 0);
    profiler().mask = (enable == -1) ? 0xffffffffu : (uint32_t)enable;
  });
}

int nrg_gpu_profile_sample(int every) {
  return guarded([&](Context&) {
    NRG_REQUIRE(every >= 1, "profile_sample: every must be >= 1");
    profiler().every = (uint32_t)every;
  });
}

int nrg_gpu_profile_exclusive(int on) {
  return guarded([&](Context&) { profiler().exclusive = (on != 0); });
}

int nrg_gpu_work_counters(double* out, int nfam, int reset) {
  return guarded([&](Context&) {
    Profiler& P = profiler();
    for (int f = 0; f < nfam && f < KF_COUNT && out; ++f) {
      out[3 * f] = P.w_launches[f];
      out[3 * f + 1] = P.w_bytes[f];
      out[3 * f + 2] = P.w_flops[f];
    }
    if (reset) { P.w_launches = {}; P.w_bytes = {}; P.w_flops = {}; }
  });
}

int nrg_gpu_copy_calibration(int64_t nbytes, int reps, int* ok) {
  return guarded([&](Context& C) {
    need_init(C);
    const bool good = copy_calibration(nbytes, reps, C.stream);
    if (ok) *ok = good ? 1 : 0;
  });
}

int nrg_gpu_launch_log(int fam_mask) {
  return guarded([&](Context&) {
    profiler().log.clear();
    profiler().nth.clear();
    profiler().log_mask = (uint32_t)fam_mask;
  });
}

int nrg_gpu_launch_log_get(double* out, int64_t cap, int64_t* count) {
  return guarded([&](Context&) {
    const auto& L = profiler().log;
    if (count) *count = (int64_t)L.size();
    const int64_t n = std::min<int64_t>(cap, (int64_t)L.size());
    for (int64_t i = 0; i < n && out; ++i) {
      out[7 * i + 0] = L[i].fam;
      out[7 * i + 1] = L[i].threads;
      out[7 * i + 2] = L[i].kid;
      out[7 * i + 3] = (double)L[i].nth;
      out[7 * i + 4] = L[i].bytes;
      out[7 * i + 5] = L[i].flops;
      out[7 * i + 6] = L[i].timed;
    }
  });
}

int nrg_gpu_launch_log_kernel(int kid, char* name, int cap) {
  return guarded([&](Context&) {
    const auto& K = profiler().nth;
    NRG_REQUIRE(kid >= 0 && kid < (int)K.size(), "launch_log_kernel: no such kernel id");
    NRG_REQUIRE(name && cap > 0, "launch_log_kernel: no buffer");
    const std::string& nm = K[kid].name;
    const size_t len = std::min(nm.size(), (size_t)cap - 1);
    std::memcpy(name, nm.data(), len);
    name[len] = 0;
  });
}

int nrg_gpu_kernel_stats(double* out, int nfam) {
  return guarded([&](Context& C) {
    if (C.stream) NRG_HIP(hipDeviceSynchronize());
    profiler().drain();
    for (int f = 0; f < nfam && f < KF_COUNT; ++f) {
      const KStat& k = profiler().st[f];
      out[4 * f + 0] = k.ms;
      out[4 * f + 1] = (double)k.launches;
      out[4 * f + 2] = k.bytes;
      out[4 * f + 3] = k.flops;
    }
  });
}

}  // extern "C"
