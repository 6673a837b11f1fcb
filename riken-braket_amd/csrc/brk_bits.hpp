// Bit insertion ("expand") masks, shared by host planners and device kernels.
//
// A gate on k operated qubits visits 2^(n-k) groups; group number g is turned
// into the base index of its group by inserting a zero at every operated bit
// position.  With the positions sorted p_0 < ... < p_{k-1} the index splits
// into k+1 bit fields; field j is shifted left by j:
//     base = sum_j (g & field_mask[j]) << j.
#pragma once

#include <algorithm>
#include <cstdint>

#include "braket_hip.h"

namespace brk {

struct expand_masks {
  uint64_t m[BRK_MAX_OPERATED_QUBITS + 1];
  int k;  // number of inserted bits
};

// `positions` need not be sorted; they must be distinct and < 64.
inline expand_masks make_expand_masks(const unsigned* positions, int k) {
  unsigned sorted[BRK_MAX_OPERATED_QUBITS];
  std::copy(positions, positions + k, sorted);
  std::sort(sorted, sorted + k);
  expand_masks e{};
  e.k = k;
  uint64_t prev = 0;  // mask of all bits below the current field
  for (int j = 0; j < k; ++j) {
    // field j holds bits [sorted[j-1]-(j-1), sorted[j]-j) of g
    const unsigned top = sorted[j] - (unsigned)j;
    const uint64_t upto = top >= 64 ? ~0ull : ((1ull << top) - 1ull);
    e.m[j] = upto ^ prev;
    prev = upto;
  }
  e.m[k] = ~0ull ^ prev;
  return e;
}

}  // namespace brk
