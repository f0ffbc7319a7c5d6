#pragma once
#include "context.h"

namespace nrg {
// Launch the H-assembly of every block of the current iteration into C.d_H.
// chain_bcs: chain_BCS(iteration+1, 1:nchain) of a BCS model, else nullptr.
void assemble(Context& C, const double* chain_diag, const double* chain_nondiag, double sqrtl,
              const double* chain_bcs);
// Gathered invariant-matrix products (invariant.hip).
void invariant(Context& C, const int32_t* kept, double* out_host);
void invariant_spectrum(Context& C, const int32_t* kept, double* out_host);
}  // namespace nrg
