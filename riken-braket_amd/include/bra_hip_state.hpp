// bra_hip_state.hpp -- a bra::state-shaped backend over the C ABI.
//
// bra dispatches every gate of a circuit to a virtual do_* hook of
// bra::state (bra/include/bra/state.hpp:1021-1317); bra::nompi_state is the
// CPU implementation (bra/src/nompi_state.cpp).  hip_state offers the same
// hooks for the gates of the plain path, with the same qubit conventions
// (Z-type gates take "control" qubits), over a device state owned here.  A
// maintainer wiring it into bra derives from bra::state and forwards each
// override to the member of the same name; the brk-bra front end
// (riken-braket_amd/bra/brk_bra.cpp) uses it directly.
#pragma once

#include <array>
#include <complex>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "braket_hip.h"

namespace brk {
namespace bra {

class error : public std::runtime_error {
  int status_;

 public:
  error(int status, const std::string& what) : std::runtime_error(what), status_(status) {}
  int status() const noexcept { return status_; }
};

inline void check(int status) {
  if (status != BRK_OK) throw error(status, std::string(brk_status_string(status)) + ": " + brk_last_error());
}

class hip_state {
 public:
  using complex_type = std::complex<double>;

  // nompi_state(initial_integer, total_num_qubits, ...) (bra/src/nompi_state.cpp:102-117)
  hip_state(unsigned num_qubits, std::uint64_t initial_integer, void* stream = nullptr)
      : n_(num_qubits), size_(std::uint64_t{1} << num_qubits), stream_(stream) {
    check(brk_malloc(&data_, size_ * sizeof(complex_type)));
    check(brk_set_basis_state(data_, size_, BRK_C128, initial_integer, stream_));
  }
  hip_state(const hip_state&) = delete;
  hip_state& operator=(const hip_state&) = delete;
  ~hip_state() {
    if (data_) (void)brk_free(data_);
  }

  unsigned num_qubits() const { return n_; }
  void* data() const { return data_; }
  std::uint64_t size() const { return size_; }

  // --- gate hooks (bra/src/nompi_state.cpp:326-1054) ---------------------
  void do_hadamard(unsigned q) { check(brk_hadamard(data_, size_, BRK_C128, q, nullptr, 0, stream_)); }
  void do_not_(unsigned q) { do_pauli_x(q); }
  void do_pauli_x(unsigned q) { do_pauli_xn({q}); }
  void do_pauli_xx(unsigned q1, unsigned q2) { do_pauli_xn({q1, q2}); }
  void do_pauli_xn(const std::vector<unsigned>& qs) {
    check(brk_pauli_x(data_, size_, BRK_C128, qs.data(), (unsigned)qs.size(), nullptr, 0, stream_));
  }
  void do_pauli_y(unsigned q) { do_pauli_yn({q}); }
  void do_pauli_yn(const std::vector<unsigned>& qs) {
    check(brk_pauli_y(data_, size_, BRK_C128, qs.data(), (unsigned)qs.size(), nullptr, 0, stream_));
  }
  void do_pauli_z(unsigned control) { check(brk_pauli_z(data_, size_, BRK_C128, &control, 1, stream_)); }
  void do_pauli_zz(unsigned q1, unsigned q2) {
    do_pauli_z(q1);
    do_pauli_z(q2);
  }
  void do_swap(unsigned q1, unsigned q2) { check(brk_swap(data_, size_, BRK_C128, q1, q2, nullptr, 0, stream_)); }
  void do_u1(double phase, unsigned control) {
    check(brk_phase_shift(data_, size_, BRK_C128, phase, &control, 1, stream_));
  }
  void do_u2(double phase1, double phase2, unsigned q) {
    check(brk_phase_shift2(data_, size_, BRK_C128, phase1, phase2, q, nullptr, 0, stream_));
  }
  void do_u3(double phase1, double phase2, double phase3, unsigned q) {
    check(brk_phase_shift3(data_, size_, BRK_C128, phase1, phase2, phase3, q, nullptr, 0, stream_));
  }
  void do_phase_shift(complex_type coeff, unsigned control) {
    const double c[2] = {coeff.real(), coeff.imag()};
    check(brk_phase_shift_coeff(data_, size_, BRK_C128, c, &control, 1, stream_));
  }
  void do_exponential_pauli_x(double phase, unsigned q) {
    check(brk_exponential_pauli_x(data_, size_, BRK_C128, phase, &q, 1, nullptr, 0, stream_));
  }
  void do_exponential_pauli_y(double phase, unsigned q) {
    check(brk_exponential_pauli_y(data_, size_, BRK_C128, phase, &q, 1, nullptr, 0, stream_));
  }
  void do_exponential_pauli_z(double phase, unsigned q) {
    check(brk_exponential_pauli_z(data_, size_, BRK_C128, phase, &q, 1, nullptr, 0, stream_));
  }
  void do_controlled_not(unsigned target, unsigned control) {
    check(brk_controlled_not(data_, size_, BRK_C128, target, &control, 1, stream_));
  }
  void do_controlled_pauli_z(unsigned c1, unsigned c2) {
    const unsigned q[2] = {c1, c2};
    check(brk_pauli_z(data_, size_, BRK_C128, q, 2, stream_));
  }
  void do_controlled_phase_shift(complex_type coeff, unsigned c1, unsigned c2) {
    const double c[2] = {coeff.real(), coeff.imag()};
    const unsigned q[2] = {c1, c2};
    check(brk_phase_shift_coeff(data_, size_, BRK_C128, c, q, 2, stream_));
  }
  void do_toffoli(unsigned target, unsigned c1, unsigned c2) {
    const unsigned c[2] = {c1, c2};
    check(brk_controlled_not(data_, size_, BRK_C128, target, c, 2, stream_));
  }
  // every gate of a parsed circuit, in order (what interpreter::apply_circuit does,
  // bra/src/interpreter.cpp:851-872); nompi mode applies gates at their own qubits
  void apply(const brk_circuit* circuit) { check(brk_circuit_run(circuit, data_, BRK_C128, nullptr, stream_)); }

  // --- observables ----------------------------------------------------------
  // ket::all_spin_expectation_values (ket/include/ket/all_spin_expectation_values.hpp:22-42)
  std::vector<std::array<double, 3>> do_expectation_values() const {
    std::vector<std::array<double, 3>> out(n_);
    for (unsigned q = 0; q < n_; ++q)
      check(brk_spin_expectation_value(data_, size_, BRK_C128, q, out[q].data(), stream_));
    return out;
  }
  std::vector<complex_type> amplitudes(const std::vector<std::uint64_t>& indices) const {
    std::vector<complex_type> out;
    if (indices.empty()) {
      out.resize(size_);
      check(brk_memcpy_dtoh(out.data(), data_, size_ * sizeof(complex_type), stream_));
      return out;
    }
    out.resize(indices.size());
    for (size_t i = 0; i < indices.size(); ++i)
      check(brk_memcpy_dtoh(&out[i], static_cast<const complex_type*>(data_) + indices[i], sizeof(complex_type),
                            stream_));
    return out;
  }

 private:
  unsigned n_;
  std::uint64_t size_;
  void* stream_;
  void* data_ = nullptr;
};

}  // namespace bra
}  // namespace brk
