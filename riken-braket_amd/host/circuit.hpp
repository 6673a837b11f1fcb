// Host-side circuit representation shared by the plain runner, the fusion
// planner (precision form), the batch form and the adjoint form.
#pragma once

#include <array>
#include <cstdint>
#include <vector>

#include "braket_hip.h"

struct brk_circuit {
  struct op {
    int code;
    std::vector<unsigned> targets;
    std::vector<unsigned> controls;
    std::array<double, 8> params{};
    unsigned num_params = 0;
  };
  unsigned num_qubits = 0;
  std::vector<op> ops;
};

namespace brk {

// Logical -> physical qubit map with its inverse (the data of
// ket::mpi::qubit_permutation, ket/include/ket/mpi/qubit_permutation.hpp:40-60).
struct permutation {
  std::vector<unsigned> perm, inverse;
  explicit permutation(unsigned n) : perm(n), inverse(n) {
    for (unsigned i = 0; i < n; ++i) perm[i] = inverse[i] = i;
  }
  // qubit_permutation::permutate (qubit_permutation.hpp:242-252): swap two logical qubits' physical slots.
  void permutate(unsigned q1, unsigned q2) {
    std::swap(inverse[perm[q1]], inverse[perm[q2]]);
    std::swap(perm[q1], perm[q2]);
  }
  bool load(const unsigned* p, unsigned n);
};

// Applies one operation at physical positions (targets/controls already mapped).
int apply_physical_op(const brk_circuit::op& o, void* first, uint64_t size, int dtype, void* stream);

// The 2x2 (or 2^k x 2^k for multi-target ops) unitary of an op on its
// operated qubits, in the order targets ++ controls (bit j of the local index
// <-> qubit j of that list).  Used by the fusion planner.  Returns false for
// ops without a dense form (LOCAL_SWAP).
bool op_unitary(const brk_circuit::op& o, std::vector<unsigned>& qubits, std::vector<std::array<double, 2>>& u);

// The dependences between ops: op j depends on op i < j when they share a qubit (target or
// control) and no op between them does; a local swap depends on, and is depended on by, the
// ops before and after it (it moves qubits between slots).  mask[i] is op i's qubit mask
// (all ones for a local swap).
struct op_dependences {
  std::vector<uint64_t> mask;
  std::vector<std::vector<size_t>> succ;
  std::vector<unsigned> indeg;
};
op_dependences dependences_of(const std::vector<brk_circuit::op>& ops);

// An order of the ops for gate blocks: a topological order of their dependences (ops that
// share a qubit keep their relative order; a local swap keeps its place relative to every
// op) that a greedy scheduler packs into few blocks of num_local qubits, the line_bits
// lowest always among them, at most BLOCK_MAX_OPS ops each.  Ops on disjoint qubits
// commute, so the circuit is the same operator; its amplitudes round differently from
// those of the given order.
std::vector<size_t> block_order(const std::vector<brk_circuit::op>& ops, unsigned num_local, unsigned line_bits);

}  // namespace brk
