// Rows of one block of the product basis grouped by local operation.
//
// A block of iteration N+1 holds, for every operation o, one row per kept
// state of the parent block label + operation(o): the basis sort orders
// equal labels by reference, so those rows appear with variation (the
// parent's in-block column) 1, 2, ... and parent references forming one
// contiguous range (prepareBasis.F90:86-101; module.F90:531-602).
#pragma once
#include "common.h"

namespace nrg {

struct BlockRows {
  std::vector<std::vector<int32_t>> by_op;      // in-block row index, variation order
  std::vector<std::vector<int32_t>> ref_by_op;  // parent (old kept) index, 1-based
};

// basis_in: records of nv+3 int32 (labels, reference, variation, operation);
// s0: first record of the block, n: its dimension.
inline BlockRows block_rows(const int32_t* basis_in, int64_t s0, int n, int nv, int nops) {
  BlockRows br;
  br.by_op.assign(nops, {});
  br.ref_by_op.assign(nops, {});
  const int rs = nv + 3;
  for (int r = 0; r < n; ++r) {
    const int32_t* rec = &basis_in[(s0 + r) * rs];
    const int op = rec[nv + 2] - 1;
    const int var = rec[nv + 1];
    NRG_REQUIRE(op >= 0 && op < nops, "block rows: bad operation index");
    NRG_REQUIRE(var == (int)br.by_op[op].size() + 1,
                "block rows: rows of one operation are not in variation order");
    br.by_op[op].push_back(r);
    br.ref_by_op[op].push_back(rec[nv]);
  }
  for (int o = 0; o < nops; ++o) {
    const auto& rf = br.ref_by_op[o];
    for (size_t q = 1; q < rf.size(); ++q)
      NRG_REQUIRE(rf[q] == rf[0] + (int)q, "block rows: parent references are not contiguous");
  }
  return br;
}

}  // namespace nrg
