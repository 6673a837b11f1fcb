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

// A run of Z-type ops (Z, CZ on two qubits) as one diagonal pass: amplitude x
// is negated when sum_t popcount(x & (x >> d[t]) & a[t]) is odd (d = 0: Z on
// each qubit of a; d > 0: CZ(q, q + d) for each qubit q of a).  Terms of equal
// d merge by XOR, so a repeated gate cancels.
constexpr int SIGN_MAX_TERMS = 16;
struct sign_terms {
  int n = 0;
  uint32_t d[SIGN_MAX_TERMS];
  uint64_t a[SIGN_MAX_TERMS];
};
int launch_sign_terms(const sign_terms& t, void* first, uint64_t size, int dtype, void* stream);

}  // namespace brk
