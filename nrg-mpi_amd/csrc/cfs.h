// Device-resident CFS history, backward reduced-density sweep and the CFS
// spectral weights (complete-Fock-space NRG at T = 0).
//
// What the reference keeps per truncated iteration in its restore and stock
// arrays (postprocess.F90:221-374; module.F90:1345-1681) stays here in HBM:
//  * the kept columns U_b[:, 1:k_b] of every block (the column-major prefix
//    of the block's eigenvector storage), for the rho sweep;
//  * the spectral products W of every (K, L, type, sign) that a discarded
//    state l of block L contributes to (calculateSpectrumDGEMM.F90:392-536),
//    formed while the spectral operator of the previous iteration is still
//    resident, so the operator itself need not be stocked;
//  * the block table, product-basis records and energies (host copies).
// After the forward pass rho is swept backward shell by shell on the device
// and each iteration's spikes are weighted by rho and binned with atomics.
#pragma once
#include "context.h"

namespace nrg {

// One block of W: rows = the first cK columns of block K, columns = columns
// l0 .. l0+dL-1 of block L, stored column-major (ld = cK).
struct SpikeGroup {
  int K, L, cK, l0, dL, im;
  int side;          // 0: binnedSpikeNegative, 1: binnedSpikePositive
  int zero_t;        // 1: last iteration (omega = E_K, weight = factor W^2)
  double factor;
  int64_t woff;
};

struct HistEntry {
  int it = 0;
  int nsub = 0, nfull = 0;
  std::vector<int32_t> sub, basis_in;   // records, as handed to diagonalize
  std::vector<int> bn, kept;
  std::vector<int64_t> bvoff;           // block start in the block-ordered energies
  std::vector<int64_t> uoff;            // block start in U (n_b * k_b per block)
  std::vector<double> E;                // shifted energies, block order
  DevBuf<double> U, Ed, W;
  std::vector<SpikeGroup> groups;
  int64_t wsize = 0;
  // reduced density matrix, one k_b x k_b column-major block per kept block
  std::vector<int64_t> roff;
  std::vector<char> rho_zero;           // block of rho is exactly zero
  DevBuf<double> rho;
  bool rho_ready = false;
  int find(const int32_t* lab, int nv) const;
  void release() { U.release(); Ed.release(); W.release(); rho.release(); }
};

struct Cfs {
  std::vector<HistEntry*> hist;   // ascending iteration
  int nbin = 0;
  DevBuf<double> bins;            // [2][nmat][nbin]: negative then positive
  DevBuf<double> ybuf, xbuf;
  DevBuf<int32_t> gidx;
  HistEntry* at(int it) {
    for (auto* h : hist) if (h->it == it) return h;
    return nullptr;
  }
  void clear() {
    for (auto* h : hist) { h->release(); delete h; }
    hist.clear();
  }
  void release() {
    clear();
    bins.release(); ybuf.release(); xbuf.release(); gidx.release();
    nbin = 0;
  }
};

Cfs& cfs();

// postprocess.F90:221-374 stock (+ the spectral products of this iteration).
void stock(Context& C, int iteration, const int32_t* kept, int last);
// initializeReducedDensity.F90:30-78
void rho_init(Context& C, int iteration);
// reducedDensityMatrixDGEMM.F90:62-321, one shell: rho(iteration) from rho(iteration+1)
double rho_shell(Context& C, int iteration);
// calculateSpectrumDGEMM.F90:1-646 for one iteration (both spike routines)
void spikes(Context& C, int iteration, double scale, int zero_temperature);
// Spectral products of spikeZeroTemperature for one ground-state block Gb
// (0-based, last iteration), type im (0-based) and side (1 positive,
// 0 negative): W(k) over the states k of the excited block into w (zeros
// where the two blocks share no operation), its index into *K and its
// dimension into *n; *K = -1, *n = 0 when no excited block has the shifted
// labels (calculateSpectrumDGEMM.F90:114-136,166-178).
void zero_temperature_weights(Context& C, int Gb, int im, int side, double* w, int cap, int* K, int* n);
// Restart file of the CFS phase (restart.cpp): the bins, the history entries
// from `iteration` on and an opaque host state of nstate integers.
void save_restart(Context& C, const char* path, int iteration, int nstate, const int32_t* state);
void load_restart(Context& C, const char* path, int* iteration, int nstate, int32_t* state);
void spectrum_setup(Context& C, int nbin);

}  // namespace nrg
