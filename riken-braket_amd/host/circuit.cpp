// Circuit objects of the C ABI and the plain per-gate runner.
//
// Operations address logical qubits.  The runner keeps the logical ->
// physical map of ket::mpi::qubit_permutation: every gate is applied at
// permutation[q] of its logical qubits, and BRK_OP_LOCAL_SWAP moves data
// between two physical qubits then updates the map
// (ket/include/ket/mpi/utility/detail/make_local_swap_qubit.hpp:62-66).
#include "circuit.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <new>
#include <string>

namespace brk {
int fail(int status, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
}

using cd = std::complex<double>;

namespace brk {

bool permutation::load(const unsigned* p, unsigned n) {
  std::vector<int> seen(n, 0);
  for (unsigned i = 0; i < n; ++i) {
    if (p[i] >= n || seen[p[i]]) return false;
    seen[p[i]] = 1;
    perm[i] = p[i];
    inverse[p[i]] = i;
  }
  return true;
}

static unsigned expected_params(int code) {
  switch (code) {
    case BRK_OP_U1: case BRK_OP_EX: case BRK_OP_EY: case BRK_OP_EZ:
    case BRK_OP_RX: case BRK_OP_RY: case BRK_OP_RZ: case BRK_OP_ESWAP: return 1;
    case BRK_OP_U2: case BRK_OP_PHASE_COEFF: return 2;
    case BRK_OP_U3: return 3;
    case BRK_OP_MATRIX: return 8;
    default: return 0;
  }
}

int apply_physical_op(const brk_circuit::op& o, void* first, uint64_t size, int dtype, void* stream) {
  const unsigned* t = o.targets.data();
  const unsigned nt = (unsigned)o.targets.size();
  const unsigned* c = o.controls.empty() ? nullptr : o.controls.data();
  const unsigned nc = (unsigned)o.controls.size();
  const double* p = o.params.data();
  switch (o.code) {
    case BRK_OP_H: return brk_hadamard(first, size, dtype, t[0], c, nc, stream);
    case BRK_OP_X: return brk_pauli_x(first, size, dtype, t, nt, c, nc, stream);
    case BRK_OP_Y: return brk_pauli_y(first, size, dtype, t, nt, c, nc, stream);
    case BRK_OP_Z:
    case BRK_OP_U1:
    case BRK_OP_PHASE_COEFF: {
      unsigned all[BRK_MAX_OPERATED_QUBITS];
      unsigned k = 0;
      for (unsigned i = 0; i < nt && k < BRK_MAX_OPERATED_QUBITS; ++i) all[k++] = t[i];
      for (unsigned i = 0; i < nc && k < BRK_MAX_OPERATED_QUBITS; ++i) all[k++] = c[i];
      if (o.code == BRK_OP_Z) return brk_pauli_z(first, size, dtype, all, k, stream);
      if (o.code == BRK_OP_PHASE_COEFF) return brk_phase_shift_coeff(first, size, dtype, p, all, k, stream);
      return brk_phase_shift(first, size, dtype, p[0], all, k, stream);
    }
    case BRK_OP_U2: return brk_phase_shift2(first, size, dtype, p[0], p[1], t[0], c, nc, stream);
    case BRK_OP_U3: return brk_phase_shift3(first, size, dtype, p[0], p[1], p[2], t[0], c, nc, stream);
    case BRK_OP_EX: return brk_exponential_pauli_x(first, size, dtype, p[0], t, nt, c, nc, stream);
    case BRK_OP_EY: return brk_exponential_pauli_y(first, size, dtype, p[0], t, nt, c, nc, stream);
    case BRK_OP_EZ: return brk_exponential_pauli_z(first, size, dtype, p[0], t, nt, c, nc, stream);
    case BRK_OP_RX: return brk_exponential_pauli_x(first, size, dtype, -0.5 * p[0], t, nt, c, nc, stream);
    case BRK_OP_RY: return brk_exponential_pauli_y(first, size, dtype, -0.5 * p[0], t, nt, c, nc, stream);
    case BRK_OP_RZ: return brk_exponential_pauli_z(first, size, dtype, -0.5 * p[0], t, nt, c, nc, stream);
    case BRK_OP_SWAP: return brk_swap(first, size, dtype, t[0], t[1], c, nc, stream);
    case BRK_OP_LOCAL_SWAP: return brk_swap_permutated_local_qubits(first, size, dtype, t[0], t[1], stream);
    case BRK_OP_MATRIX: return brk_gate_1q(first, size, dtype, t[0], p, c, nc, stream);
    case BRK_OP_ESWAP: return brk_exponential_swap(first, size, dtype, p[0], t[0], t[1], c, nc, stream);
    default: return fail(BRK_ERR_INVALID_ARGUMENT, "unknown op code %d", o.code);
  }
}

// --------------------------------------------------------------------------
// Dense unitaries of ops (fusion planner input)
// --------------------------------------------------------------------------

static void pauli_string_matrix(const cd p[2][2], unsigned m, std::vector<cd>& out) {
  const unsigned d = 1u << m;
  out.assign((size_t)d * d, cd{0, 0});
  for (unsigned r = 0; r < d; ++r)
    for (unsigned c = 0; c < d; ++c) {
      cd v{1, 0};
      for (unsigned j = 0; j < m; ++j) v *= p[(r >> j) & 1u][(c >> j) & 1u];
      out[(size_t)r * d + c] = v;
    }
}

bool op_unitary(const brk_circuit::op& o, std::vector<unsigned>& qubits, std::vector<std::array<double, 2>>& u) {
  if (o.code == BRK_OP_LOCAL_SWAP) return false;
  const unsigned m = (unsigned)o.targets.size();
  const cd I(0, 1);
  std::vector<cd> core;
  unsigned core_bits = m;
  const double* p = o.params.data();
  static const cd X[2][2] = {{0, 1}, {1, 0}};
  static const cd Y[2][2] = {{0, cd(0, -1)}, {cd(0, 1), 0}};
  static const cd Z[2][2] = {{1, 0}, {0, -1}};
  auto pauli_exp = [&](const cd P[2][2], double theta) {
    pauli_string_matrix(P, m, core);
    const unsigned d = 1u << m;
    for (auto& v : core) v *= I * std::sin(theta);
    for (unsigned r = 0; r < d; ++r) core[(size_t)r * d + r] += std::cos(theta);
  };
  auto two_by_two = [&](cd a, cd b, cd c, cd d) { core = {a, b, c, d}; };
  switch (o.code) {
    case BRK_OP_H: {
      const double r = 0.707106781186547524400844362104849039;
      two_by_two(r, r, r, -r);
      break;
    }
    case BRK_OP_X: pauli_string_matrix(X, m, core); break;
    case BRK_OP_Y: pauli_string_matrix(Y, m, core); break;
    case BRK_OP_Z:
    case BRK_OP_U1:
    case BRK_OP_PHASE_COEFF: {
      // acts on targets ++ controls jointly
      core_bits = m + (unsigned)o.controls.size();
      const unsigned d = 1u << core_bits;
      core.assign((size_t)d * d, cd{0, 0});
      for (unsigned r = 0; r < d; ++r) core[(size_t)r * d + r] = 1.0;
      core[(size_t)(d - 1) * d + (d - 1)] =
          o.code == BRK_OP_Z ? cd(-1, 0) : o.code == BRK_OP_U1 ? std::exp(I * p[0]) : cd(p[0], p[1]);
      break;
    }
    case BRK_OP_U2: {
      const double r = 0.707106781186547524400844362104849039;
      const cd e1 = std::exp(I * p[0]), e2 = std::exp(I * p[1]);
      two_by_two(r, -e2 * r, e1 * r, e1 * e2 * r);
      break;
    }
    case BRK_OP_U3: {
      const double s = std::sin(0.5 * p[0]), c = std::cos(0.5 * p[0]);
      const cd e2 = std::exp(I * p[1]), e3 = std::exp(I * p[2]);
      two_by_two(c, -s * e3, s * e2, c * e3 * e2);
      break;
    }
    case BRK_OP_EX: pauli_exp(X, p[0]); break;
    case BRK_OP_EY: pauli_exp(Y, p[0]); break;
    case BRK_OP_EZ: pauli_exp(Z, p[0]); break;
    case BRK_OP_RX: pauli_exp(X, -0.5 * p[0]); break;
    case BRK_OP_RY: pauli_exp(Y, -0.5 * p[0]); break;
    case BRK_OP_RZ: pauli_exp(Z, -0.5 * p[0]); break;
    case BRK_OP_SWAP: core = {1, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0, 0, 1}; break;
    case BRK_OP_ESWAP: {
      const cd e = std::exp(I * p[0]), is = I * std::sin(p[0]);
      const double c = std::cos(p[0]);
      core = {e, 0, 0, 0, 0, c, is, 0, 0, is, c, 0, 0, 0, 0, e};
      break;
    }
    case BRK_OP_MATRIX:
      two_by_two(cd(p[0], p[1]), cd(p[2], p[3]), cd(p[4], p[5]), cd(p[6], p[7]));
      break;
    default: return false;
  }
  qubits = o.targets;
  qubits.insert(qubits.end(), o.controls.begin(), o.controls.end());
  const unsigned k = (unsigned)qubits.size();
  const unsigned d = 1u << k, dc = 1u << core_bits, cmask_bits = k - core_bits;
  u.assign((size_t)d * d, {0.0, 0.0});
  const unsigned all_on = (1u << cmask_bits) - 1u;
  for (unsigned r = 0; r < d; ++r)
    for (unsigned c = 0; c < d; ++c) {
      const unsigned rc = r >> core_bits, cc = c >> core_bits;
      cd v{0, 0};
      if (rc == cc) {
        if (rc == all_on) v = core[(size_t)(r & (dc - 1)) * dc + (c & (dc - 1))];
        else if (r == c) v = 1.0;
      }
      u[(size_t)r * d + c] = {v.real(), v.imag()};
    }
  return true;
}

}  // namespace brk

using namespace brk;

extern "C" {

int brk_circuit_create(unsigned num_qubits, brk_circuit** out) {
  if (!out) return fail(BRK_ERR_INVALID_ARGUMENT, "out is null");
  if (num_qubits == 0 || num_qubits > 40) return fail(BRK_ERR_INVALID_ARGUMENT, "num_qubits %u not in [1, 40]", num_qubits);
  auto* c = new (std::nothrow) brk_circuit;
  if (!c) return fail(BRK_ERR_OUT_OF_MEMORY, "cannot allocate circuit");
  c->num_qubits = num_qubits;
  *out = c;
  return BRK_OK;
}

int brk_circuit_destroy(brk_circuit* circuit) {
  delete circuit;
  return BRK_OK;
}

int brk_circuit_add(brk_circuit* circuit, int op, const unsigned* targets, unsigned num_targets,
                    const unsigned* controls, unsigned num_controls, const double* params, unsigned num_params) {
  if (!circuit) return fail(BRK_ERR_INVALID_ARGUMENT, "circuit is null");
  if (op < BRK_OP_H || op > BRK_OP_PHASE_COEFF) return fail(BRK_ERR_INVALID_ARGUMENT, "unknown op code %d", op);
  if (num_targets == 0 || !targets) return fail(BRK_ERR_INVALID_ARGUMENT, "op %d has no targets", op);
  if (num_controls && !controls) return fail(BRK_ERR_INVALID_ARGUMENT, "controls is null");
  const unsigned need = expected_params(op);
  if (num_params != need || (need && !params))
    return fail(BRK_ERR_INVALID_ARGUMENT, "op %d takes %u parameters, got %u", op, need, num_params);
  const bool single = op == BRK_OP_H || op == BRK_OP_U2 || op == BRK_OP_U3 || op == BRK_OP_MATRIX;
  const bool pair = op == BRK_OP_SWAP || op == BRK_OP_LOCAL_SWAP || op == BRK_OP_ESWAP;
  if (single && num_targets != 1) return fail(BRK_ERR_INVALID_ARGUMENT, "op %d takes one target", op);
  if (pair && num_targets != 2) return fail(BRK_ERR_INVALID_ARGUMENT, "op %d takes two targets", op);
  if (op == BRK_OP_LOCAL_SWAP && num_controls)
    return fail(BRK_ERR_INVALID_ARGUMENT, "a local swap takes no controls");
  if (num_targets + num_controls > BRK_MAX_OPERATED_QUBITS)
    return fail(BRK_ERR_TOO_MANY_QUBITS, "more than %d operated qubits", BRK_MAX_OPERATED_QUBITS);
  brk_circuit::op o;
  o.code = op;
  uint64_t seen = 0;
  for (unsigned i = 0; i < num_targets + num_controls; ++i) {
    const unsigned q = i < num_targets ? targets[i] : controls[i - num_targets];
    if (q >= circuit->num_qubits)
      return fail(BRK_ERR_QUBIT_OUT_OF_RANGE, "qubit %u out of range for %u qubits", q, circuit->num_qubits);
    if (seen & (1ull << q)) return fail(BRK_ERR_DUPLICATE_QUBIT, "qubit %u given twice", q);
    seen |= 1ull << q;
    (i < num_targets ? o.targets : o.controls).push_back(q);
  }
  for (unsigned i = 0; i < num_params; ++i) o.params[i] = params[i];
  o.num_params = num_params;
  circuit->ops.push_back(std::move(o));
  return BRK_OK;
}

int brk_circuit_num_ops(const brk_circuit* circuit, unsigned* num_ops) {
  if (!circuit || !num_ops) return fail(BRK_ERR_INVALID_ARGUMENT, "null argument");
  *num_ops = (unsigned)circuit->ops.size();
  return BRK_OK;
}

int brk_circuit_num_qubits(const brk_circuit* circuit, unsigned* num_qubits) {
  if (!circuit || !num_qubits) return fail(BRK_ERR_INVALID_ARGUMENT, "null argument");
  *num_qubits = circuit->num_qubits;
  return BRK_OK;
}

int brk_circuit_op(const brk_circuit* circuit, unsigned i, int* op, unsigned* targets, unsigned* num_targets,
                   unsigned* controls, unsigned* num_controls, double* params, unsigned* num_params) {
  if (!circuit) return fail(BRK_ERR_INVALID_ARGUMENT, "circuit is null");
  if (i >= circuit->ops.size()) return fail(BRK_ERR_INVALID_ARGUMENT, "op %u out of range", i);
  const auto& o = circuit->ops[i];
  if (op) *op = o.code;
  if (num_targets) *num_targets = (unsigned)o.targets.size();
  if (num_controls) *num_controls = (unsigned)o.controls.size();
  if (num_params) *num_params = o.num_params;
  if (targets) std::copy(o.targets.begin(), o.targets.end(), targets);
  if (controls) std::copy(o.controls.begin(), o.controls.end(), controls);
  if (params) std::copy(o.params.begin(), o.params.begin() + o.num_params, params);
  return BRK_OK;
}

static int run_impl(const brk_circuit* circuit, void* first, int dtype, unsigned* permutation_io, float* op_ms,
                    void* stream) {
  if (!circuit) return fail(BRK_ERR_INVALID_ARGUMENT, "circuit is null");
  const unsigned n = circuit->num_qubits;
  const uint64_t size = 1ull << n;
  brk::permutation perm(n);
  if (permutation_io && !perm.load(permutation_io, n))
    return fail(BRK_ERR_INVALID_ARGUMENT, "permutation is not a permutation of 0..%u", n - 1);
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  std::vector<hipEvent_t> events;
  if (op_ms) {
    events.resize(circuit->ops.size() + 1);
    for (auto& e : events)
      if (hipEventCreate(&e) != hipSuccess) return fail(BRK_ERR_HIP, "hipEventCreate failed");
    (void)hipEventRecord(events[0], st);
  }
  int status = BRK_OK;
  brk_circuit::op mapped;
  for (size_t i = 0; i < circuit->ops.size() && status == BRK_OK; ++i) {
    const auto& o = circuit->ops[i];
    if (o.code == BRK_OP_LOCAL_SWAP) {
      status = apply_physical_op(o, first, size, dtype, stream);
      if (status == BRK_OK) perm.permutate(perm.inverse[o.targets[0]], perm.inverse[o.targets[1]]);
    } else {
      mapped.code = o.code;
      mapped.params = o.params;
      mapped.num_params = o.num_params;
      mapped.targets.resize(o.targets.size());
      mapped.controls.resize(o.controls.size());
      for (size_t j = 0; j < o.targets.size(); ++j) mapped.targets[j] = perm.perm[o.targets[j]];
      for (size_t j = 0; j < o.controls.size(); ++j) mapped.controls[j] = perm.perm[o.controls[j]];
      status = apply_physical_op(mapped, first, size, dtype, stream);
    }
    if (op_ms && status == BRK_OK) (void)hipEventRecord(events[i + 1], st);
  }
  if (op_ms) {
    if (status == BRK_OK) {
      if (hipEventSynchronize(events.back()) != hipSuccess) status = fail(BRK_ERR_HIP, "hipEventSynchronize failed");
      for (size_t i = 0; status == BRK_OK && i < circuit->ops.size(); ++i)
        if (hipEventElapsedTime(&op_ms[i], events[i], events[i + 1]) != hipSuccess)
          status = fail(BRK_ERR_HIP, "hipEventElapsedTime failed");
    }
    for (auto& e : events) (void)hipEventDestroy(e);
  }
  if (status == BRK_OK && permutation_io)
    for (unsigned i = 0; i < n; ++i) permutation_io[i] = perm.perm[i];
  return status;
}

int brk_circuit_run(const brk_circuit* circuit, void* first, int dtype, unsigned* permutation, void* stream) {
  return run_impl(circuit, first, dtype, permutation, nullptr, stream);
}

int brk_circuit_run_timed(const brk_circuit* circuit, void* first, int dtype, unsigned* permutation, float* op_ms,
                          void* stream) {
  if (!op_ms) return fail(BRK_ERR_INVALID_ARGUMENT, "op_ms is null");
  return run_impl(circuit, first, dtype, permutation, op_ms, stream);
}

}  // extern "C"
