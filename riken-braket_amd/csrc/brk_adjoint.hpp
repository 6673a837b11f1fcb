// Launch descriptor of the adjoint-step kernel, shared by host/adjoint.cpp
// and csrc/brk_adjoint.hip.
#pragma once

#include <cstdint>

namespace brk {

constexpr int ADJ_MAX_CHAIN = 4;     // single-qubit ops on one qubit handled by one pass
constexpr int ADJ_BLOCKS = 1024;     // fixed grid: one FP64 partial per block and op

// One op of a chain, in reverse-sweep order.  M = U^dagger (un-applies the
// op), D = (dU/dtheta) U^dagger, so that dU/dtheta |psi_before> = D |psi_after>.
struct adj_op {
  double m[8];  // m00, m01, m10, m11 as (re, im)
  double d[8];
  int has_grad;
};

struct adj_chain {
  unsigned target;
  int nops;
  adj_op ops[ADJ_MAX_CHAIN];
};

// psi, lam: complex states of 2^n amplitudes (dtype BRK_C128 or BRK_C64).
// partials: nops * ADJ_BLOCKS doubles; op k's partial of block b at k * ADJ_BLOCKS + b
// (the sum over b is 2 Re <lam| D_k |psi>).
int launch_adjoint_chain(const adj_chain& c, void* psi, void* lam, uint64_t size, int dtype, double* partials,
                         void* stream);

}  // namespace brk
