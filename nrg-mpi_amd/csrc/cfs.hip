// CFS history, backward reduced-density sweep and CFS spikes on the GPU.
//
// Spectral products.  For a discarded state l of block L, the reference
// forms per type t (calculateSpectrumDGEMM.F90:392-536, module.F90:1014-1125)
//   G_ii : K = L + posOp(t),  w = U_K[:,1:kK]^T C_KL U_L[:,l]          -> negative bins
//   G_iii: K = L + negOp(t),  w = (U_L[:,l]^T C_LK U_K[:,1:kK])^T       -> positive bins
// with C_XY(a,b) = cspec(t, op_a) S(ref_b, ref_a, t) where op_a = op_b and S
// is the spectral operator of the previous iteration.  The discarded
// columns of a block are the columns after its kept prefix, so all l of one
// (K, L, t, sign) form one matrix W = U_K^T C U_L[:, kL:] (kK x dL), built in
// two gathered FP64-MFMA products while S is resident:
//   Y = U_K[rows_o, :kK]^T (c_o S_o)   per operation o   (kK x m_{L,o}, stacked)
//   W = Y U_L[stacked rows, kL:]                          (kK x dL)
// S_o is the S sub-block of the parent ranges of the two blocks' rows of
// operation o; G_ii reads it transposed.  At the last iteration the same
// builder gives the ground-state products of spikeZeroTemperature
// (calculateSpectrumDGEMM.F90:56-211): L = ground-state block, column 1
// only (the reference uses column 1 for every ground state of the block).
//
// Weights.  spikeCFS weights every spike by w(k1) sum_k2 w(k2) rho_K(k2,k1)
// (:392-536): X = rho_K W is one GEMM per group and a binning kernel forms
// W o X, omega = |E_k - E_l| scale(N) and the linear split over the two
// neighbouring log bins of binningSpike (module.F90:1210-1254), accumulated
// with FP64 atomics.
//
// Reduced density.  rho_B(N) = sum_o V_o rho_{P_o}(N+1) V_o^T with
// P_o = block of N+1 labelled label(B) - operation(o) and V_o the rows of
// U_{P_o}[:, :kP] made by operation o (reducedDensityMatrixDGEMM.F90:164-297).
// Per block the rows are gathered once into Vs = [V_o ...] (kB x sum kP),
// Z = Vs diag(rho_P) by per-operation GEMMs, rho_B = Z Vs^T by one GEMM; a
// rho_P that is exactly zero (blocks without ground states at the last
// iteration) contributes exactly zero and is skipped.
#include "cfs.h"
#include "blockrows.h"
#include <algorithm>
#include <cmath>

namespace nrg {

Cfs& cfs() {
  static Cfs c;
  return c;
}

int HistEntry::find(const int32_t* lab, int nv) const {
  for (int b = 0; b < nsub; ++b) {
    const int32_t* s = &sub[(size_t)b * (nv + 3)];
    bool eq = true;
    for (int i = 0; i < nv && eq; ++i) eq = (s[i] == lab[i]);
    if (eq) return b;
  }
  return -1;
}

namespace {

constexpr double kDegeneracy = 1.0e-8;   // module.F90:7
constexpr double kLog10Emin = -15.0;     // module.F90:970
constexpr double kMeshOfBin = 1000.0;    // module.F90:964
constexpr int64_t kYLimit = int64_t(1) << 28;   // doubles of Y per product chunk (2 GiB)

struct SpikeDev {
  int64_t start, woff, xoff, eK, eL;
  int cK, bin0;
  double factor, scale;
};

struct GatherTask {
  int64_t start, dst, src, g;
  int rows, ld_src, ld_dst, pad;
};

__device__ inline void bin_add(double* p, double v) { unsafeAtomicAdd(p, v); }

// binningSpike (module.F90:1210-1254) for one spike
__device__ inline void bin_spike(double* b, int nbin, double omega, double weight) {
  const double a = fabs(omega);
  const double pos = (a > 0.0) ? (log10(a) - kLog10Emin) * kMeshOfBin + 1.0 : -1.0;
  if (pos <= 0.0) { bin_add(b, weight); return; }
  if (pos >= (double)nbin) { bin_add(b + nbin - 1, weight); return; }
  int pref = (int)pos;
  if (pref <= 0) pref = 1;
  const double w = pos - (double)pref;
  bin_add(b + pref - 1, weight * (1.0 - w));
  bin_add(b + pref, weight * w);
}

__global__ __launch_bounds__(256) void k_bin_spikes(const SpikeDev* __restrict__ g, int ng, int64_t total,
                                                    const double* __restrict__ W, const double* __restrict__ X,
                                                    const double* __restrict__ E, double* __restrict__ bins,
                                                    int nbin) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    int lo = 0, hi = ng - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (g[mid].start <= e) lo = mid; else hi = mid - 1;
    }
    const SpikeDev s = g[lo];
    const int64_t loc = e - s.start;
    const int k = (int)(loc % s.cK);
    const int64_t l = loc / s.cK;
    const double w = W[s.woff + loc];
    const double x = (s.xoff >= 0) ? X[s.xoff + loc] : w;
    const double om = (s.eL >= 0) ? (E[s.eK + k] - E[s.eL + l]) * s.scale : E[s.eK + k] * s.scale;
    bin_spike(bins + s.bin0, nbin, om, s.factor * w * x);
  }
}

// dst[t.dst + i + c*ld_dst] = src[t.src + gidx[t.g + i] + c*ld_src]
__global__ __launch_bounds__(256) void k_gather_rows(const GatherTask* __restrict__ t, int nt, int64_t total,
                                                     const double* __restrict__ src, const int32_t* __restrict__ gidx,
                                                     double* __restrict__ dst) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    int lo = 0, hi = nt - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (t[mid].start <= e) lo = mid; else hi = mid - 1;
    }
    const GatherTask q = t[lo];
    const int64_t loc = e - q.start;
    const int i = (int)(loc % q.rows);
    const int64_t c = loc / q.rows;
    dst[q.dst + i + c * q.ld_dst] = src[q.src + gidx[q.g + i] + c * q.ld_src];
  }
}

__global__ void k_scatter_value(double* __restrict__ p, const int64_t* __restrict__ idx, int n, double v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[idx[i]] = v;
}

// per-block traces: one workgroup per block
__global__ __launch_bounds__(256) void k_block_trace(const double* __restrict__ rho, const int64_t* __restrict__ off,
                                                     const int32_t* __restrict__ k, double* __restrict__ out) {
  __shared__ double red[256];
  const int b = blockIdx.x;
  const int n = k[b];
  const double* r = rho + off[b];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) s += r[(int64_t)i * (n + 1)];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[b] = red[0];
}

unsigned grid_for(int64_t total) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 256 * 64));
}

const int32_t* label(const std::vector<int32_t>& sub, int b, int nv) { return &sub[(size_t)b * (nv + 3)]; }

// Build every spectral-product group of the current iteration into h.W.
void build_products(Context& C, HistEntry& h, int last) {
  Cfs& F = cfs();
  const int nv = C.nv, nops = C.nops, nsub = C.nsub;
  NRG_REQUIRE((int)C.pos_op.size() == C.nmat * nv && (int)C.neg_op.size() == C.nmat * nv,
              "stock: spectral operators (posOp/negOp) not set");
  NRG_REQUIRE(C.d_spec.p != nullptr && C.kold_spec > 0, "stock: no spectral operator on the device");
  std::vector<BlockRows> rows(nsub);
  std::vector<char> have(nsub, 0);
  auto R = [&](int b) -> const BlockRows& {
    if (!have[b]) { rows[b] = block_rows(C.basis_in.data(), C.bvoff[b], C.bn[b], nv, nops); have[b] = 1; }
    return rows[b];
  };
  std::vector<int32_t> lab(nv);
  auto shifted = [&](int b, const std::vector<int32_t>& op, int im) {
    const int32_t* l = label(C.sub, b, nv);
    for (int i = 0; i < nv; ++i) lab[i] = l[i] + op[im * nv + i];
    return h.find(lab.data(), nv);
  };
  std::vector<SpikeGroup>& G = h.groups;
  G.clear();
  // group form: 0 = C_KL read as S(ref_L, ref_K) (G_ii / zero-T positive),
  //             1 = C_LK^T read as S(ref_K, ref_L) (G_iii / zero-T negative)
  std::vector<int> form;
  if (!last) {
    for (int L = 0; L < nsub; ++L) {
      const int dL = C.bn[L] - h.kept[L];
      if (dL <= 0) continue;
      for (int im = 0; im < C.nmat; ++im)
        for (int sg = 0; sg < 2; ++sg) {
          const int K = shifted(L, sg == 0 ? C.pos_op : C.neg_op, im);
          if (K < 0 || h.kept[K] == 0) continue;
          G.push_back(SpikeGroup{K, L, h.kept[K], h.kept[L], dL, im, sg == 0 ? 0 : 1, 0, 1.0, 0});
          form.push_back(sg);
        }
    }
  } else {
    int g = 0;
    std::vector<int> mg(nsub, 0);
    for (int b = 0; b < nsub; ++b)
      for (int i = 0; i < C.bn[b]; ++i)
        if (h.E[h.bvoff[b] + i] < kDegeneracy) { ++mg[b]; ++g; }
    NRG_REQUIRE(g > 0, "stock: no ground state at the last iteration");
    for (int Gb = 0; Gb < nsub; ++Gb) {
      if (mg[Gb] == 0) continue;
      for (int im = 0; im < C.nmat; ++im)
        for (int sg = 0; sg < 2; ++sg) {
          const int K = shifted(Gb, sg == 0 ? C.pos_op : C.neg_op, im);
          if (K < 0) continue;
          G.push_back(SpikeGroup{K, Gb, C.bn[K], 0, 1, im, sg == 0 ? 1 : 0, 1, (double)mg[Gb] / (double)g, 0});
          form.push_back(sg);
        }
    }
  }
  // per group: the shared operations and the stacked width of Y
  const int64_t ko = C.kold_spec;
  std::vector<int64_t> ysz(G.size(), 0);
  std::vector<char> keep(G.size(), 1);
  int64_t wtot = 0;
  for (size_t q = 0; q < G.size(); ++q) {
    const BlockRows& rk = R(G[q].K);
    const BlockRows& rl = R(G[q].L);
    int64_t ma = 0;
    for (int o = 0; o < nops; ++o)
      if (!rk.by_op[o].empty() && !rl.by_op[o].empty()) ma += (int64_t)rl.by_op[o].size();
    if (ma == 0) { keep[q] = 0; continue; }
    ysz[q] = ma * G[q].cK;
    G[q].woff = wtot;
    wtot += (int64_t)G[q].cK * G[q].dL;
  }
  {
    std::vector<SpikeGroup> g2;
    std::vector<int> f2;
    std::vector<int64_t> y2;
    for (size_t q = 0; q < G.size(); ++q)
      if (keep[q]) { g2.push_back(G[q]); f2.push_back(form[q]); y2.push_back(ysz[q]); }
    G.swap(g2); form.swap(f2); ysz.swap(y2);
  }
  h.wsize = wtot;
  h.W.reserve(std::max<int64_t>(wtot, 1));
  size_t q0 = 0;
  while (q0 < G.size()) {
    // chunk of groups whose Y fits the work buffer
    size_t q1 = q0;
    int64_t ytot = 0;
    while (q1 < G.size() && (q1 == q0 || ytot + ysz[q1] <= kYLimit)) ytot += ysz[q1++];
    F.ybuf.reserve(std::max<int64_t>(ytot, 1));
    std::vector<int32_t> gidx;
    GemmBatch b1, b2;
    int64_t yoff = 0;
    for (size_t q = q0; q < q1; ++q) {
      const SpikeGroup& s = G[q];
      const BlockRows& rk = R(s.K);
      const BlockRows& rl = R(s.L);
      int64_t col = 0;
      const int64_t gl0 = (int64_t)gidx.size();
      for (int o = 0; o < nops; ++o) {
        if (rk.by_op[o].empty() || rl.by_op[o].empty()) continue;
        gidx.insert(gidx.end(), rl.by_op[o].begin(), rl.by_op[o].end());
      }
      for (int o = 0; o < nops; ++o) {
        const auto& ak = rk.by_op[o];
        const auto& al = rl.by_op[o];
        if (ak.empty() || al.empty()) continue;
        const int64_t rk0 = rk.ref_by_op[o][0] - 1, rl0 = rl.ref_by_op[o][0] - 1;
        NRG_REQUIRE(rk0 + (int64_t)ak.size() <= ko && rl0 + (int64_t)al.size() <= ko,
                    "stock: parent reference beyond the spectral operator");
        const int64_t ga = (int64_t)gidx.size();
        gidx.insert(gidx.end(), ak.begin(), ak.end());
        int64_t offS;
        int transB;
        if (form[q] == 0) { offS = rl0 + rk0 * ko; transB = 1; }   // B(a', b') = S(refL b', refK a')
        else { offS = rk0 + rl0 * ko; transB = 0; }                 // B(b', a') = S(refK b', refL a')
        GemmTask t = gemm_task(s.cK, (int)al.size(), (int)ak.size(), 1, transB, C.boff[s.K], C.bn[s.K],
                               offS + (int64_t)s.im * ko * ko, (int)ko, yoff + col * s.cK, s.cK,
                               C.coef_spec[s.im * nops + o], 0.0);
        t.offGA = ga;
        b1.add(t);
        col += (int64_t)al.size();
      }
      GemmTask t2 = gemm_task(s.cK, s.dL, (int)col, 0, 0, yoff, s.cK, C.boff[s.L] + (int64_t)s.l0 * C.bn[s.L],
                              C.bn[s.L], s.woff, s.cK, 1.0, 0.0);
      t2.offGB = gl0;
      b2.add(t2);
      yoff += ysz[q];
    }
    F.gidx.upload(gidx, C.stream);
    C.gemm.run(b1, C.d_U.p, C.d_spec.p, F.ybuf.p, F.gidx.p, nullptr, nullptr, C.stream, KF_GEMM_CFSW);
    C.gemm.run(b2, F.ybuf.p, C.d_U.p, h.W.p, nullptr, F.gidx.p, nullptr, C.stream, KF_GEMM_CFSW);
    // the gather list and Y are reused by the next chunk
    NRG_HIP(hipStreamSynchronize(C.stream));
    q0 = q1;
  }
}

}  // namespace

void stock(Context& C, int iteration, const int32_t* kept, int last) {
  Cfs& F = cfs();
  NRG_REQUIRE(C.nsub > 0, "stock: no diagonalised iteration");
  NRG_REQUIRE((int)C.E_block.size() == C.nfull, "stock: energies of the iteration are missing");
  if (!F.hist.empty() && F.hist.back()->it >= iteration) F.clear();   // a new run
  NRG_REQUIRE(F.hist.empty() || F.hist.back()->it == iteration - 1,
              "stock: truncated iterations must be stocked consecutively");
  HistEntry* h = new HistEntry;
  try {
    const int nv = C.nv;
    h->it = iteration;
    h->nsub = C.nsub;
    h->nfull = C.nfull;
    h->sub = C.sub;
    h->basis_in = C.basis_in;
    h->bn = C.bn;
    h->kept.assign(kept, kept + C.nsub);
    h->bvoff = C.bvoff;
    h->E = C.E_block;
    h->uoff.resize(C.nsub);
    int64_t u = 0;
    for (int b = 0; b < C.nsub; ++b) {
      NRG_REQUIRE(kept[b] >= 0 && kept[b] <= C.bn[b], "stock: kept count out of range");
      if (last) NRG_REQUIRE(kept[b] == C.bn[b], "stock: the last iteration keeps every state");
      h->uoff[b] = u;
      u += (int64_t)C.bn[b] * kept[b];
    }
    h->U.reserve(std::max<int64_t>(u, 1));
    for (int b = 0; b < C.nsub; ++b)
      if (kept[b] > 0)
        NRG_HIP(hipMemcpyAsync(h->U.p + h->uoff[b], C.d_U.p + C.boff[b],
                               sizeof(double) * (size_t)C.bn[b] * kept[b], hipMemcpyDeviceToDevice, C.stream));
    h->Ed.upload(h->E, C.stream);
    (void)nv;
    if (C.nmat > 0 && !C.coef_spec.empty()) build_products(C, *h, last);
    NRG_HIP(hipStreamSynchronize(C.stream));
  } catch (...) {
    h->release();
    delete h;
    throw;
  }
  F.hist.push_back(h);
}

void rho_init(Context& C, int iteration) {
  Cfs& F = cfs();
  HistEntry* h = F.at(iteration);
  NRG_REQUIRE(h != nullptr, "rho_init: iteration was not stocked");
  NRG_REQUIRE(F.hist.back() == h, "rho_init: not the last stocked iteration");
  int g = 0;
  for (double e : h->E) if (e < kDegeneracy) ++g;
  NRG_REQUIRE(g > 0, "rho_init: no ground state");
  h->roff.resize(h->nsub);
  h->rho_zero.assign(h->nsub, 1);
  int64_t r = 0;
  std::vector<int64_t> diag;
  for (int b = 0; b < h->nsub; ++b) {
    const int k = h->kept[b];
    h->roff[b] = r;
    for (int i = 0; i < k; ++i)
      if (h->E[h->bvoff[b] + i] < kDegeneracy) { diag.push_back(r + (int64_t)i * (k + 1)); h->rho_zero[b] = 0; }
    r += (int64_t)k * k;
  }
  h->rho.reserve(std::max<int64_t>(r, 1));
  fill(h->rho.p, 0.0, r, C.stream);
  DevBuf<int64_t> d;
  d.upload(diag, C.stream);
  hipLaunchKernelGGL(k_scatter_value, dim3(ceil_div((long long)diag.size(), 256)), dim3(256), 0, C.stream, h->rho.p,
                     d.p, (int)diag.size(), 1.0 / (double)g);
  NRG_HIP(hipGetLastError());
  NRG_HIP(hipStreamSynchronize(C.stream));
  d.release();
  h->rho_ready = true;
}

double rho_shell(Context& C, int iteration) {
  Cfs& F = cfs();
  HistEntry* h = F.at(iteration);
  HistEntry* p = F.at(iteration + 1);
  NRG_REQUIRE(h != nullptr && p != nullptr, "rho_shell: iteration or its successor was not stocked");
  NRG_REQUIRE(p->rho_ready, "rho_shell: the reduced density of the next iteration is not ready");
  const int nv = C.nv, nops = C.nops;
  std::vector<BlockRows> prow(p->nsub);
  std::vector<char> have(p->nsub, 0);
  h->roff.assign(h->nsub, 0);
  h->rho_zero.assign(h->nsub, 1);
  int64_t r = 0, vs = 0;
  std::vector<int32_t> lab(nv), gidx;
  std::vector<GatherTask> gt;
  GemmBatch b1, b2;
  int64_t gtot = 0;
  for (int B = 0; B < h->nsub; ++B) {
    const int kB = h->kept[B];
    h->roff[B] = r;
    r += (int64_t)kB * kB;
    if (kB == 0) continue;
    const int32_t* lb = label(h->sub, B, nv);
    struct Seg { int P, o, kP; };
    std::vector<Seg> segs;
    int64_t ktot = 0;
    for (int o = 0; o < nops; ++o) {
      for (int i = 0; i < nv; ++i) lab[i] = lb[i] - C.operation[o * (nv + C.nchain) + i];
      const int P = p->find(lab.data(), nv);
      if (P < 0 || p->kept[P] == 0) continue;
      if (!have[P]) {
        prow[P] = block_rows(p->basis_in.data(), p->bvoff[P], p->bn[P], nv, nops);
        have[P] = 1;
      }
      NRG_REQUIRE((int)prow[P].by_op[o].size() == kB,
                  "rho_shell: rows of the successor block do not match the kept states");
      if (p->rho_zero[P]) continue;
      segs.push_back(Seg{P, o, p->kept[P]});
      ktot += p->kept[P];
    }
    if (segs.empty()) continue;
    h->rho_zero[B] = 0;
    int64_t koff = 0;
    for (const Seg& s : segs) {
      const int64_t g0 = (int64_t)gidx.size();
      gidx.insert(gidx.end(), prow[s.P].by_op[s.o].begin(), prow[s.P].by_op[s.o].end());
      gt.push_back(GatherTask{gtot, vs + koff * kB, p->uoff[s.P], g0, kB, p->bn[s.P], kB, 0});
      gtot += (int64_t)kB * s.kP;
      b1.add(gemm_task(kB, s.kP, s.kP, 0, 0, vs + koff * kB, kB, p->roff[s.P], s.kP, vs + koff * kB, kB, 1.0, 0.0));
      koff += s.kP;
    }
    b2.add(gemm_task(kB, kB, (int)ktot, 0, 1, vs, kB, vs, kB, h->roff[B], kB, 1.0, 0.0));
    vs += (int64_t)kB * ktot;
  }
  h->rho.reserve(std::max<int64_t>(r, 1));
  fill(h->rho.p, 0.0, r, C.stream);
  if (!gt.empty()) {
    F.ybuf.reserve(vs);
    F.xbuf.reserve(vs);
    F.gidx.upload(gidx, C.stream);
    DevBuf<GatherTask> dt;
    dt.upload(gt, C.stream);
    plaunch(KF_RHO_GATHER, 16.0 * gtot, 0.0, k_gather_rows, dim3(grid_for(gtot)), dim3(256), 0, C.stream,
            (const GatherTask*)dt.p, (int)gt.size(), gtot, (const double*)p->U.p, (const int32_t*)F.gidx.p,
            F.ybuf.p);
    NRG_HIP(hipGetLastError());
    // Z (xbuf) = Vs (ybuf) rho_P ; rho_B = Z Vs^T
    C.gemm.run(b1, F.ybuf.p, p->rho.p, F.xbuf.p, nullptr, nullptr, nullptr, C.stream, KF_GEMM_RHO);
    C.gemm.run(b2, F.xbuf.p, F.ybuf.p, h->rho.p, nullptr, nullptr, nullptr, C.stream, KF_GEMM_RHO);
    NRG_HIP(hipStreamSynchronize(C.stream));
    dt.release();
  }
  h->rho_ready = true;
  // trace of the shell (printed by the reference, reducedDensityMatrixDGEMM.F90:364-367)
  std::vector<int32_t> kk(h->kept.begin(), h->kept.end());
  DevBuf<int64_t> doff;
  DevBuf<int32_t> dk;
  DevBuf<double> dtr;
  doff.upload(h->roff, C.stream);
  dk.upload(kk, C.stream);
  dtr.reserve(h->nsub);
  hipLaunchKernelGGL(k_block_trace, dim3(h->nsub), dim3(256), 0, C.stream, h->rho.p, doff.p, dk.p, dtr.p);
  NRG_HIP(hipGetLastError());
  std::vector<double> tr(h->nsub);
  dtr.download(tr.data(), h->nsub, C.stream);
  NRG_HIP(hipStreamSynchronize(C.stream));
  doff.release(); dk.release(); dtr.release();
  double t = 0.0;
  for (double v : tr) t += v;
  return t;
}

void spectrum_setup(Context& C, int nbin) {
  Cfs& F = cfs();
  NRG_REQUIRE(nbin >= 2 && C.nmat > 0, "spectrum_setup: bad bin count or no spectral types");
  F.nbin = nbin;
  F.bins.reserve((size_t)2 * C.nmat * nbin);
  fill(F.bins.p, 0.0, (size_t)2 * C.nmat * nbin, C.stream);
  NRG_HIP(hipStreamSynchronize(C.stream));
}

void zero_temperature_weights(Context& C, int Gb, int im, int side, double* w, int cap, int* K, int* n) {
  Cfs& F = cfs();
  NRG_REQUIRE(!F.hist.empty(), "zero_temperature_weights: nothing stocked");
  const HistEntry* h = F.hist.back();
  const int nv = C.nv;
  NRG_REQUIRE(Gb >= 0 && Gb < h->nsub && im >= 0 && im < C.nmat, "zero_temperature_weights: bad block or type");
  const std::vector<int32_t>& op = side ? C.pos_op : C.neg_op;
  std::vector<int32_t> lab(nv);
  for (int i = 0; i < nv; ++i) lab[i] = h->sub[(size_t)Gb * (nv + 3) + i] + op[im * nv + i];
  *K = h->find(lab.data(), nv);
  *n = 0;
  if (*K < 0) return;
  const int dim = h->bn[*K];
  NRG_REQUIRE(cap >= dim, "zero_temperature_weights: output too short for the excited block");
  *n = dim;
  for (int k = 0; k < dim; ++k) w[k] = 0.0;
  for (const SpikeGroup& g : h->groups)
    if (g.zero_t && g.L == Gb && g.K == *K && g.im == im && g.side == side) {
      NRG_HIP(hipMemcpyAsync(w, h->W.p + g.woff, sizeof(double) * dim, hipMemcpyDeviceToHost, C.stream));
      NRG_HIP(hipStreamSynchronize(C.stream));
      break;
    }
}

void spikes(Context& C, int iteration, double scale, int zero_temperature) {
  Cfs& F = cfs();
  HistEntry* h = F.at(iteration);
  NRG_REQUIRE(h != nullptr, "spikes: iteration was not stocked");
  NRG_REQUIRE(F.nbin > 0, "spikes: call nrg_gpu_spectrum_setup first");
  const bool zt = zero_temperature != 0;
  if (!zt) NRG_REQUIRE(h->rho_ready, "spikes: reduced density of this iteration is not ready");
  std::vector<SpikeDev> sd;
  GemmBatch bx;
  int64_t tot = 0, xo = 0;
  for (const SpikeGroup& g : h->groups) {
    NRG_REQUIRE((g.zero_t != 0) == zt, "spikes: products were stocked for the other spike routine");
    if (!zt && h->rho_zero[g.K]) continue;
    SpikeDev s;
    s.start = tot;
    s.woff = g.woff;
    s.eK = h->bvoff[g.K];
    s.eL = zt ? -1 : h->bvoff[g.L] + g.l0;
    s.cK = g.cK;
    s.bin0 = (g.side * C.nmat + g.im) * F.nbin;
    s.factor = g.factor;
    s.scale = scale;
    if (zt) {
      s.xoff = -1;
    } else {
      s.xoff = xo;
      bx.add(gemm_task(g.cK, g.dL, g.cK, 0, 0, h->roff[g.K], g.cK, g.woff, g.cK, xo, g.cK, 1.0, 0.0));
      xo += (int64_t)g.cK * g.dL;
    }
    tot += (int64_t)g.cK * g.dL;
    sd.push_back(s);
  }
  if (sd.empty()) return;
  if (!zt) {
    F.xbuf.reserve(std::max<int64_t>(xo, 1));
    C.gemm.run(bx, h->rho.p, h->W.p, F.xbuf.p, nullptr, nullptr, nullptr, C.stream, KF_GEMM_CFSX);
  }
  DevBuf<SpikeDev> d;
  d.upload(sd, C.stream);
  plaunch(KF_CFS_BIN, (zt ? 24.0 : 32.0) * tot, 0.0, k_bin_spikes, dim3(grid_for(tot)), dim3(256), 0, C.stream,
          (const SpikeDev*)d.p, (int)sd.size(), tot, (const double*)h->W.p,
          (const double*)(zt ? h->W.p : F.xbuf.p), (const double*)h->Ed.p, F.bins.p, F.nbin);
  NRG_HIP(hipGetLastError());
  NRG_HIP(hipStreamSynchronize(C.stream));
  d.release();
}

}  // namespace nrg
