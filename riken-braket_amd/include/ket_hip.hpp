// ket_hip.hpp -- ket-shaped C++ facade over the C ABI (include/braket_hip.h).
//
// Call sites written for the reference's ket library,
//     ket::gate::hadamard(policy, first, last, 3_q, 0_cq);
//     ket::gate::ranges::phase_shift3(policy, state, theta, phi, lambda, 2_q);
// compile unchanged against device states when `policy` is a ket::hip::policy
// and `first`/`last`/`state` are ket::hip device iterators / states.  The
// functions mirror the reference's names, argument order (iterator pair or
// range, real parameters, target qubits, control qubits) and meaning
// (ket/include/ket/gate/*.hpp); they enqueue HIP kernels on the policy's stream
// and throw ket::hip::error for an invalid request (the reference asserts).
//
// ket::qubit and ket::control: when the reference's ket/qubit.hpp and
// ket/control.hpp are included first their types are used; otherwise this
// header provides minimal types with the same interface subset.
#pragma once

#include <array>
#include <cmath>
#include <complex>
#include <limits>
#include <random>
#include <cstdint>
#include <iterator>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "braket_hip.h"

#ifndef KET_QUBIT_HPP
#define KET_QUBIT_HPP
namespace ket {
template <typename StateInteger = std::uint64_t, typename BitInteger = unsigned int>
class qubit {
  BitInteger bit_;

 public:
  using state_integer_type = StateInteger;
  using bit_integer_type = BitInteger;
  constexpr qubit() noexcept : bit_{0u} {}
  explicit constexpr qubit(BitInteger bit) noexcept : bit_{bit} {}
  explicit constexpr operator BitInteger() const noexcept { return bit_; }
  friend constexpr bool operator==(qubit a, qubit b) noexcept { return a.bit_ == b.bit_; }
  friend constexpr bool operator!=(qubit a, qubit b) noexcept { return a.bit_ != b.bit_; }
  friend constexpr bool operator<(qubit a, qubit b) noexcept { return a.bit_ < b.bit_; }
};
namespace literals {
inline namespace qubit_literals {
inline constexpr qubit<> operator"" _q(unsigned long long bit) noexcept { return qubit<>{static_cast<unsigned>(bit)}; }
}  // namespace qubit_literals
}  // namespace literals
}  // namespace ket
#endif

#ifndef KET_CONTROL_HPP
#define KET_CONTROL_HPP
namespace ket {
template <typename Qubit>
class control {
  Qubit qubit_;

 public:
  using qubit_type = Qubit;
  constexpr control() noexcept : qubit_{} {}
  explicit constexpr control(Qubit q) noexcept : qubit_{q} {}
  constexpr Qubit qubit() const noexcept { return qubit_; }
};
template <typename Qubit>
inline constexpr control<Qubit> make_control(Qubit q) { return control<Qubit>{q}; }
namespace literals {
inline namespace control_literals {
inline constexpr control<qubit<>> operator"" _cq(unsigned long long bit) noexcept {
  return control<qubit<>>{qubit<>{static_cast<unsigned>(bit)}};
}
}  // namespace control_literals
}  // namespace literals
}  // namespace ket
#endif

namespace ket {
namespace hip {

class error : public std::runtime_error {
  int status_;

 public:
  error(int status, const std::string& what) : std::runtime_error(what), status_(status) {}
  int status() const noexcept { return status_; }
};

inline void check(int status) {
  if (status != BRK_OK) throw error(status, std::string(brk_status_string(status)) + ": " + brk_last_error());
}

// Execution policy: the HIP stream the gates are enqueued on (the role of
// ket::utility::policy::parallel in the reference).
struct policy {
  void* stream = nullptr;
};
inline policy make_policy(void* stream = nullptr) { return policy{stream}; }

template <typename Complex> struct dtype_of;
template <> struct dtype_of<std::complex<double>> { static constexpr int value = BRK_C128; };
template <> struct dtype_of<std::complex<float>> { static constexpr int value = BRK_C64; };

// Random-access position in a device state (device pointer; never dereferenced on the host).
template <typename Complex>
class device_iterator {
  Complex* p_ = nullptr;

 public:
  using value_type = Complex;
  using difference_type = std::ptrdiff_t;
  constexpr device_iterator() = default;
  explicit constexpr device_iterator(Complex* p) : p_(p) {}
  constexpr Complex* get() const { return p_; }
  constexpr device_iterator operator+(difference_type n) const { return device_iterator(p_ + n); }
  constexpr difference_type operator-(device_iterator o) const { return p_ - o.p_; }
  constexpr bool operator==(device_iterator o) const { return p_ == o.p_; }
  constexpr bool operator!=(device_iterator o) const { return p_ != o.p_; }
};

// Owning state vector in device memory (the range the ranges:: functions take).
template <typename Complex>
class state {
  Complex* data_ = nullptr;
  std::uint64_t size_ = 0;

 public:
  using value_type = Complex;
  explicit state(unsigned num_qubits) : size_(std::uint64_t{1} << num_qubits) {
    void* p = nullptr;
    check(brk_malloc(&p, size_ * sizeof(Complex)));
    data_ = static_cast<Complex*>(p);
    check(brk_set_basis_state(data_, size_, dtype_of<Complex>::value, 0, nullptr));
  }
  explicit state(const std::vector<Complex>& host) : size_(host.size()) {
    void* p = nullptr;
    check(brk_malloc(&p, size_ * sizeof(Complex)));
    data_ = static_cast<Complex*>(p);
    upload(host);
  }
  state(const state&) = delete;
  state& operator=(const state&) = delete;
  state(state&& o) noexcept : data_(o.data_), size_(o.size_) { o.data_ = nullptr; }
  ~state() {
    if (data_) (void)brk_free(data_);
  }
  std::uint64_t size() const { return size_; }
  Complex* data() const { return data_; }
  device_iterator<Complex> begin() const { return device_iterator<Complex>(data_); }
  device_iterator<Complex> end() const { return device_iterator<Complex>(data_ + size_); }
  void upload(const std::vector<Complex>& host, void* stream = nullptr) {
    if (host.size() != size_) throw error(BRK_ERR_INVALID_ARGUMENT, "size mismatch");
    check(brk_memcpy_htod(data_, host.data(), size_ * sizeof(Complex), stream));
  }
  std::vector<Complex> download(void* stream = nullptr) const {
    std::vector<Complex> out(size_);
    check(brk_memcpy_dtoh(out.data(), data_, size_ * sizeof(Complex), stream));
    return out;
  }
};

namespace detail {

// Splits a ket argument pack into target qubits and control qubits.
struct operands {
  unsigned targets[BRK_MAX_OPERATED_QUBITS];
  unsigned controls[BRK_MAX_OPERATED_QUBITS];
  unsigned nt = 0, nc = 0;
};
template <typename S, typename B>
inline void collect(operands& o, ::ket::qubit<S, B> q) {
  if (o.nt >= BRK_MAX_OPERATED_QUBITS) throw error(BRK_ERR_TOO_MANY_QUBITS, "too many target qubits");
  o.targets[o.nt++] = static_cast<unsigned>(static_cast<B>(q));
}
template <typename Q>
inline void collect(operands& o, ::ket::control<Q> c) {
  if (o.nc >= BRK_MAX_OPERATED_QUBITS) throw error(BRK_ERR_TOO_MANY_QUBITS, "too many control qubits");
  o.controls[o.nc++] = static_cast<unsigned>(static_cast<typename Q::bit_integer_type>(c.qubit()));
}
template <typename... Qs>
inline operands split(Qs... qs) {
  operands o;
  (collect(o, qs), ...);
  return o;
}
// all listed qubits, targets first (for the "control-qubit" forms: Z, U1)
inline void all_qubits(const operands& o, unsigned* out, unsigned* n) {
  *n = 0;
  for (unsigned i = 0; i < o.nt; ++i) out[(*n)++] = o.targets[i];
  for (unsigned i = 0; i < o.nc; ++i) out[(*n)++] = o.controls[i];
}
template <typename Complex>
inline std::uint64_t span(device_iterator<Complex> first, device_iterator<Complex> last) {
  return static_cast<std::uint64_t>(last - first);
}
inline const unsigned* ptr_or_null(const unsigned* p, unsigned n) { return n ? p : nullptr; }

// runtime forms: targets are one ket::qubit or a range of them, controls a range of ket::control
template <typename S, typename B>
inline void add_all(operands& o, const ::ket::qubit<S, B>& q) {
  collect(o, q);
}
template <typename Range>
inline auto add_all(operands& o, const Range& r) -> decltype(std::begin(r), std::end(r), void()) {
  for (const auto& q : r) collect(o, q);
}
template <typename Targets>
inline operands from_ranges(const Targets& targets) {
  operands o;
  add_all(o, targets);
  return o;
}
template <typename Targets, typename Controls>
inline operands from_ranges(const Targets& targets, const Controls& controls) {
  operands o;
  add_all(o, targets);
  add_all(o, controls);
  return o;
}
// iterator-pair forms: [first, last) of ket::qubit or ket::control values
template <typename It>
inline void add_all(operands& o, It first, It last) {
  for (; first != last; ++first) collect(o, *first);
}
template <typename It>
using if_iterator = decltype(*std::declval<It&>(), ++std::declval<It&>(), std::declval<It>() != std::declval<It>(), 0);

}  // namespace detail
}  // namespace hip

namespace gate {

// Core of every gate: (policy, first, last, real parameters, operands).  The public
// forms below -- variadic ket::qubit / ket::control arguments, ranges::, runtime::
// qubit ranges -- all build the operand list and call these.
namespace hip_core {
using hip::detail::operands;
using hip::detail::ptr_or_null;
using hip::detail::span;
template <typename Complex>
using iter = hip::device_iterator<Complex>;
template <typename Complex>
constexpr int dtype() { return hip::dtype_of<Complex>::value; }

inline void one_target(const operands& o, const char* name) {
  if (o.nt != 1) throw hip::error(BRK_ERR_INVALID_ARGUMENT, std::string(name) + " takes one target qubit");
}
inline std::complex<double> exp_i(double phase) { return std::exp(std::complex<double>{0.0, phase}); }

// H -- ket/include/ket/gate/hadamard.hpp:34-223
template <typename Complex>
inline void hadamard(hip::policy p, iter<Complex> first, iter<Complex> last, const operands& o) {
  one_target(o, "hadamard");
  hip::check(brk_hadamard(first.get(), span(first, last), dtype<Complex>(), o.targets[0],
                          ptr_or_null(o.controls, o.nc), o.nc, p.stream));
}
// X...X under controls -- pauli_x.hpp; NOT / CNOT / TOFFOLI -- not_.hpp, controlled_not.hpp, toffoli.hpp
template <typename Complex>
inline void pauli_x(hip::policy p, iter<Complex> first, iter<Complex> last, const operands& o) {
  hip::check(brk_pauli_x(first.get(), span(first, last), dtype<Complex>(), o.targets, o.nt,
                         ptr_or_null(o.controls, o.nc), o.nc, p.stream));
}
template <typename Complex>
inline void not_(hip::policy p, iter<Complex> first, iter<Complex> last, const operands& o) {
  pauli_x(p, first, last, o);
}
template <typename Complex>
inline void controlled_not(hip::policy p, iter<Complex> first, iter<Complex> last, const operands& o) {
  if (o.nt != 1 || o.nc == 0) throw hip::error(BRK_ERR_INVALID_ARGUMENT, "controlled_not takes a target and controls");
  hip::check(brk_controlled_not(first.get(), span(first, last), dtype<Complex>(), o.targets[0], o.controls, o.nc,
                                p.stream));
}
template <typename Complex>
inline void toffoli(hip::policy p, iter<Complex> first, iter<Complex> last, const operands& o) {
  if (o.nt != 1 || o.nc != 2) throw hip::error(BRK_ERR_INVALID_ARGUMENT, "toffoli takes a target and two controls");
  controlled_not(p, first, last, o);
}
// Y...Y under controls -- pauli_y.hpp
template <typename Complex>
inline void pauli_y(hip::policy p, iter<Complex> first, iter<Complex> last, const operands& o) {
  hip::check(brk_pauli_y(first.get(), span(first, last), dtype<Complex>(), o.targets, o.nt,
                         ptr_or_null(o.controls, o.nc), o.nc, p.stream));
}
// Z on control qubits (Z, CZ, CnZ) -- pauli_z.hpp:55-210 -- and with one ket::qubit target
// (CZ_{tc}, :305-390): -1 where every listed bit is 1.  Two or more targets (ZZ, C...CZ...Z,
// :265-303, :392-470) are a product of Z's: -1 on odd parity of the targets under the controls.
template <typename Complex>
inline void pauli_z(hip::policy p, iter<Complex> first, iter<Complex> last, const operands& o) {
  if (o.nt >= 2) {
    hip::check(brk_pauli_zn(first.get(), span(first, last), dtype<Complex>(), o.targets, o.nt,
                            ptr_or_null(o.controls, o.nc), o.nc, p.stream));
    return;
  }
  unsigned all[2 * BRK_MAX_OPERATED_QUBITS], n = 0;
  hip::detail::all_qubits(o, all, &n);
  hip::check(brk_pauli_z(first.get(), span(first, last), dtype<Complex>(), all, n, p.stream));
}
// a coefficient where every listed qubit is 1: U1, controlled U1, sqrt(Z) -- phase_shift.hpp:37-93,
// controlled_phase_shift.hpp:30-110, sqrt_pauli_z.hpp:35-230
template <typename Complex>
inline void phase_shift_coeff(hip::policy p, iter<Complex> first, iter<Complex> last, std::complex<double> coeff,
                              const operands& o) {
  unsigned all[2 * BRK_MAX_OPERATED_QUBITS], n = 0;
  hip::detail::all_qubits(o, all, &n);
  const double c[2] = {coeff.real(), coeff.imag()};
  hip::check(brk_phase_shift_coeff(first.get(), span(first, last), dtype<Complex>(), c, all, n, p.stream));
}
template <typename Complex>
inline void phase_shift(hip::policy p, iter<Complex> first, iter<Complex> last, double phase, const operands& o) {
  unsigned all[2 * BRK_MAX_OPERATED_QUBITS], n = 0;
  hip::detail::all_qubits(o, all, &n);
  hip::check(brk_phase_shift(first.get(), span(first, last), dtype<Complex>(), phase, all, n, p.stream));
}
template <typename Complex>
inline void controlled_phase_shift_coeff(hip::policy p, iter<Complex> first, iter<Complex> last,
                                         std::complex<double> coeff, const operands& o) {
  phase_shift_coeff(p, first, last, coeff, o);
}
template <typename Complex>
inline void controlled_phase_shift(hip::policy p, iter<Complex> first, iter<Complex> last, double phase,
                                   const operands& o) {
  phase_shift(p, first, last, phase, o);
}
// sqrt(Z): i where every listed bit is 1 (sqrt_pauli_z.hpp:35-420, one target :428-460);
// two or more targets (sZZ, C...CsZ...Z, :462-700): i^f, f = number of set target bits.
template <typename Complex>
inline void sqrt_pauli_z(hip::policy p, iter<Complex> first, iter<Complex> last, const operands& o) {
  if (o.nt >= 2) {
    hip::check(brk_sqrt_pauli_zn(first.get(), span(first, last), dtype<Complex>(), o.targets, o.nt,
                                 ptr_or_null(o.controls, o.nc), o.nc, p.stream));
    return;
  }
  phase_shift_coeff(p, first, last, {0.0, 1.0}, o);
}
// U2, U3 -- phase_shift.hpp:653-709, 1195-1254
template <typename Complex>
inline void phase_shift2(hip::policy p, iter<Complex> first, iter<Complex> last, double phase1, double phase2,
                         const operands& o) {
  one_target(o, "phase_shift2");
  hip::check(brk_phase_shift2(first.get(), span(first, last), dtype<Complex>(), phase1, phase2, o.targets[0],
                              ptr_or_null(o.controls, o.nc), o.nc, p.stream));
}
template <typename Complex>
inline void phase_shift3(hip::policy p, iter<Complex> first, iter<Complex> last, double phase1, double phase2,
                         double phase3, const operands& o) {
  one_target(o, "phase_shift3");
  hip::check(brk_phase_shift3(first.get(), span(first, last), dtype<Complex>(), phase1, phase2, phase3,
                              o.targets[0], ptr_or_null(o.controls, o.nc), o.nc, p.stream));
}
// exp(i phase P...P) -- exponential_pauli_{x,y,z}.hpp; the _coeff forms take e^{i phase}
#define KET_HIP_EXP_PAULI(NAME, FN)                                                                         \
  template <typename Complex>                                                                               \
  inline void NAME(hip::policy p, iter<Complex> first, iter<Complex> last, double phase, const operands& o) { \
    hip::check(FN(first.get(), span(first, last), dtype<Complex>(), phase, o.targets, o.nt,                  \
                  ptr_or_null(o.controls, o.nc), o.nc, p.stream));                                          \
  }                                                                                                         \
  template <typename Complex>                                                                               \
  inline void NAME##_coeff(hip::policy p, iter<Complex> first, iter<Complex> last, std::complex<double> coeff, \
                           const operands& o) {                                                             \
    const double c[2] = {coeff.real(), coeff.imag()};                                                       \
    hip::check(FN##_coeff(first.get(), span(first, last), dtype<Complex>(), c, o.targets, o.nt,             \
                          ptr_or_null(o.controls, o.nc), o.nc, p.stream));                                  \
  }
KET_HIP_EXP_PAULI(exponential_pauli_x, brk_exponential_pauli_x)
KET_HIP_EXP_PAULI(exponential_pauli_y, brk_exponential_pauli_y)
KET_HIP_EXP_PAULI(exponential_pauli_z, brk_exponential_pauli_z)
#undef KET_HIP_EXP_PAULI
// SWAP and exp(i phase SWAP) under controls -- swap.hpp:30-75, exponential_swap.hpp:31-350
template <typename Complex>
inline void swap(hip::policy p, iter<Complex> first, iter<Complex> last, const operands& o) {
  if (o.nt != 2) throw hip::error(BRK_ERR_INVALID_ARGUMENT, "swap takes two target qubits");
  hip::check(brk_swap(first.get(), span(first, last), dtype<Complex>(), o.targets[0], o.targets[1],
                      ptr_or_null(o.controls, o.nc), o.nc, p.stream));
}
template <typename Complex>
inline void exponential_swap_coeff(hip::policy p, iter<Complex> first, iter<Complex> last,
                                   std::complex<double> coeff, const operands& o) {
  if (o.nt != 2) throw hip::error(BRK_ERR_INVALID_ARGUMENT, "exponential_swap takes two target qubits");
  const double c[2] = {coeff.real(), coeff.imag()};
  hip::check(brk_exponential_swap_coeff(first.get(), span(first, last), dtype<Complex>(), c, o.targets[0],
                                        o.targets[1], ptr_or_null(o.controls, o.nc), o.nc, p.stream));
}
template <typename Complex>
inline void exponential_swap(hip::policy p, iter<Complex> first, iter<Complex> last, double phase,
                             const operands& o) {
  if (o.nt != 2) throw hip::error(BRK_ERR_INVALID_ARGUMENT, "exponential_swap takes two target qubits");
  hip::check(brk_exponential_swap(first.get(), span(first, last), dtype<Complex>(), phase, o.targets[0],
                                  o.targets[1], ptr_or_null(o.controls, o.nc), o.nc, p.stream));
}
// fixed 2x2 gates on one target under controls, rows (m00, m01; m10, m11)
template <typename Complex>
inline void matrix_1q(hip::policy p, iter<Complex> first, iter<Complex> last, std::complex<double> m00,
                      std::complex<double> m01, std::complex<double> m10, std::complex<double> m11,
                      const operands& o, const char* name) {
  one_target(o, name);
  const double m[8] = {m00.real(), m00.imag(), m01.real(), m01.imag(), m10.real(), m10.imag(), m11.real(), m11.imag()};
  hip::check(brk_gate_1q(first.get(), span(first, last), dtype<Complex>(), o.targets[0], m,
                         ptr_or_null(o.controls, o.nc), o.nc, p.stream));
}
// sqrt(X), sqrt(Y) and adjoints -- sqrt_pauli_x.hpp:31-32, 258-259; sqrt_pauli_y.hpp:32-33, 255-256
template <typename Complex>
inline void sqrt_pauli_x(hip::policy p, iter<Complex> first, iter<Complex> last, const operands& o) {
  const std::complex<double> a{0.5, 0.5}, b{0.5, -0.5};
  matrix_1q(p, first, last, a, b, b, a, o, "sqrt_pauli_x");
}
template <typename Complex>
inline void adj_sqrt_pauli_x(hip::policy p, iter<Complex> first, iter<Complex> last, const operands& o) {
  const std::complex<double> a{0.5, -0.5}, b{0.5, 0.5};
  matrix_1q(p, first, last, a, b, b, a, o, "adj_sqrt_pauli_x");
}
template <typename Complex>
inline void sqrt_pauli_y(hip::policy p, iter<Complex> first, iter<Complex> last, const operands& o) {
  const std::complex<double> a{0.5, 0.5};
  matrix_1q(p, first, last, a, -a, a, a, o, "sqrt_pauli_y");
}
template <typename Complex>
inline void adj_sqrt_pauli_y(hip::policy p, iter<Complex> first, iter<Complex> last, const operands& o) {
  const std::complex<double> a{0.5, -0.5};
  matrix_1q(p, first, last, a, a, -a, a, o, "adj_sqrt_pauli_y");
}
// +X / -X and +Y / -Y half-pi rotations -- x_rotation_half_pi.hpp:33-34, 253-254; y_rotation_half_pi.hpp:32-33, 250-251
template <typename Complex>
inline void x_rotation_half_pi(hip::policy p, iter<Complex> first, iter<Complex> last, const operands& o) {
  const double r = 0.707106781186547524400844362104849039;
  matrix_1q(p, first, last, {r, 0.0}, {0.0, r}, {0.0, r}, {r, 0.0}, o, "x_rotation_half_pi");
}
template <typename Complex>
inline void adj_x_rotation_half_pi(hip::policy p, iter<Complex> first, iter<Complex> last, const operands& o) {
  const double r = 0.707106781186547524400844362104849039;
  matrix_1q(p, first, last, {r, 0.0}, {0.0, -r}, {0.0, -r}, {r, 0.0}, o, "adj_x_rotation_half_pi");
}
template <typename Complex>
inline void y_rotation_half_pi(hip::policy p, iter<Complex> first, iter<Complex> last, const operands& o) {
  const double r = 0.707106781186547524400844362104849039;
  matrix_1q(p, first, last, {r, 0.0}, {r, 0.0}, {-r, 0.0}, {r, 0.0}, o, "y_rotation_half_pi");
}
template <typename Complex>
inline void adj_y_rotation_half_pi(hip::policy p, iter<Complex> first, iter<Complex> last, const operands& o) {
  const double r = 0.707106781186547524400844362104849039;
  matrix_1q(p, first, last, {r, 0.0}, {-r, 0.0}, {r, 0.0}, {r, 0.0}, o, "adj_y_rotation_half_pi");
}
// controlled V(theta) = [[1+e, 1-e], [1-e, 1+e]] / 2 on the target under controls -- controlled_v.hpp:30-90
template <typename Complex>
inline void controlled_v_coeff(hip::policy p, iter<Complex> first, iter<Complex> last, std::complex<double> coeff,
                               const operands& o) {
  if (o.nc == 0) throw hip::error(BRK_ERR_INVALID_ARGUMENT, "controlled_v takes at least one control qubit");
  const std::complex<double> plus = 0.5 * (1.0 + coeff), minus = 0.5 * (1.0 - coeff);
  matrix_1q(p, first, last, plus, minus, minus, plus, o, "controlled_v");
}
template <typename Complex>
inline void controlled_v(hip::policy p, iter<Complex> first, iter<Complex> last, double phase, const operands& o) {
  controlled_v_coeff(p, first, last, exp_i(phase), o);
}

// adjoints (adj_* of each gate header): self-inverse gates map to themselves, the
// others to the same gate with negated phase / conjugated coefficient; U2(phi, lambda)^+ =
// U3(-pi/2, -lambda, -phi), U3(theta, phi, lambda)^+ = U3(-theta, -lambda, -phi)
#define KET_HIP_SELF_ADJOINT(NAME)                                                                       \
  template <typename Complex>                                                                            \
  inline void adj_##NAME(hip::policy p, iter<Complex> first, iter<Complex> last, const operands& o) {    \
    NAME(p, first, last, o);                                                                             \
  }
KET_HIP_SELF_ADJOINT(hadamard)
KET_HIP_SELF_ADJOINT(pauli_x)
KET_HIP_SELF_ADJOINT(not_)
KET_HIP_SELF_ADJOINT(controlled_not)
KET_HIP_SELF_ADJOINT(toffoli)
KET_HIP_SELF_ADJOINT(pauli_y)
KET_HIP_SELF_ADJOINT(pauli_z)
KET_HIP_SELF_ADJOINT(swap)
#undef KET_HIP_SELF_ADJOINT
#define KET_HIP_PHASE_ADJOINT(NAME)                                                                      \
  template <typename Complex>                                                                            \
  inline void adj_##NAME(hip::policy p, iter<Complex> first, iter<Complex> last, double phase,           \
                         const operands& o) {                                                            \
    NAME(p, first, last, -phase, o);                                                                     \
  }
#define KET_HIP_COEFF_ADJOINT(NAME)                                                                      \
  template <typename Complex>                                                                            \
  inline void adj_##NAME(hip::policy p, iter<Complex> first, iter<Complex> last, std::complex<double> c, \
                         const operands& o) {                                                            \
    NAME(p, first, last, std::conj(c), o);                                                               \
  }
KET_HIP_PHASE_ADJOINT(phase_shift)
KET_HIP_PHASE_ADJOINT(controlled_phase_shift)
KET_HIP_PHASE_ADJOINT(exponential_pauli_x)
KET_HIP_PHASE_ADJOINT(exponential_pauli_y)
KET_HIP_PHASE_ADJOINT(exponential_pauli_z)
KET_HIP_PHASE_ADJOINT(exponential_swap)
KET_HIP_PHASE_ADJOINT(controlled_v)
KET_HIP_COEFF_ADJOINT(phase_shift_coeff)
KET_HIP_COEFF_ADJOINT(controlled_phase_shift_coeff)
KET_HIP_COEFF_ADJOINT(exponential_pauli_x_coeff)
KET_HIP_COEFF_ADJOINT(exponential_pauli_y_coeff)
KET_HIP_COEFF_ADJOINT(exponential_pauli_z_coeff)
KET_HIP_COEFF_ADJOINT(exponential_swap_coeff)
KET_HIP_COEFF_ADJOINT(controlled_v_coeff)
#undef KET_HIP_PHASE_ADJOINT
#undef KET_HIP_COEFF_ADJOINT
template <typename Complex>
inline void adj_sqrt_pauli_z(hip::policy p, iter<Complex> first, iter<Complex> last, const operands& o) {
  if (o.nt >= 2) {
    hip::check(brk_adj_sqrt_pauli_zn(first.get(), span(first, last), dtype<Complex>(), o.targets, o.nt,
                                     ptr_or_null(o.controls, o.nc), o.nc, p.stream));
    return;
  }
  phase_shift_coeff(p, first, last, {0.0, -1.0}, o);
}
template <typename Complex>
inline void adj_phase_shift2(hip::policy p, iter<Complex> first, iter<Complex> last, double phase1, double phase2,
                             const operands& o) {
  phase_shift3(p, first, last, -1.57079632679489661923132169163975144, -phase2, -phase1, o);
}
template <typename Complex>
inline void adj_phase_shift3(hip::policy p, iter<Complex> first, iter<Complex> last, double phase1, double phase2,
                             double phase3, const operands& o) {
  phase_shift3(p, first, last, -phase1, -phase3, -phase2, o);
}
}  // namespace hip_core

// Public forms of gate NAME with real parameters PARAMS (declarations) / ARGS (names):
//   NAME(policy, first, last, params..., qubit, control...)          -- ket::gate::NAME
//   ranges::NAME(policy, state, params..., qubit, control...)        -- ket::gate::ranges::NAME
//   runtime::qubit_ranges::NAME(policy, first, last, params..., targets[, controls])
//   runtime::ranges::NAME(policy, state, params..., targets[, controls])
// where targets is one ket::qubit or a range of them and controls a range of ket::control
// (the runtime forms of the reference, e.g. hadamard.hpp:281-470, pauli_x.hpp:288-572), and
// the iterator-pair runtime forms
//   runtime::NAME(policy, first, last, params..., target)
//   runtime::NAME(policy, first, last, params..., target, control_first, control_last)
//   runtime::NAME(policy, first, last, params..., target_first, target_last)
//   runtime::NAME(policy, first, last, params..., target_first, target_last, control_first, control_last)
// (hadamard.hpp:407-437, pauli_x.hpp:395-440).
#define KET_HIP_UNPAREN(...) __VA_ARGS__
#define KET_HIP_GATE(NAME, PARAMS, ARGS)                                                                      \
  template <typename Complex, typename... Qubits>                                                             \
  inline void NAME(hip::policy p, hip::device_iterator<Complex> first, hip::device_iterator<Complex> last     \
                       KET_HIP_UNPAREN PARAMS,                                                                \
                   Qubits... qubits) {                                                                        \
    hip_core::NAME(p, first, last KET_HIP_UNPAREN ARGS, hip::detail::split(qubits...));                       \
  }                                                                                                           \
  namespace ranges {                                                                                          \
  template <typename Complex, typename... Qubits>                                                             \
  inline hip::state<Complex>& NAME(hip::policy p, hip::state<Complex>& s KET_HIP_UNPAREN PARAMS,              \
                                   Qubits... qubits) {                                                        \
    hip_core::NAME(p, s.begin(), s.end() KET_HIP_UNPAREN ARGS, hip::detail::split(qubits...));               \
    return s;                                                                                                 \
  }                                                                                                           \
  }                                                                                                           \
  namespace runtime {                                                                                         \
  template <typename Complex, typename Target>                                                                \
  inline void NAME(hip::policy p, hip::device_iterator<Complex> first, hip::device_iterator<Complex> last     \
                       KET_HIP_UNPAREN PARAMS,                                                                \
                   const Target& target) {                                                                    \
    hip_core::NAME(p, first, last KET_HIP_UNPAREN ARGS, hip::detail::from_ranges(target));                    \
  }                                                                                                           \
  template <typename Complex, typename S, typename B, typename ControlIt,                                      \
            hip::detail::if_iterator<ControlIt> = 0>                                                          \
  inline void NAME(hip::policy p, hip::device_iterator<Complex> first, hip::device_iterator<Complex> last     \
                       KET_HIP_UNPAREN PARAMS,                                                                \
                   ::ket::qubit<S, B> target, ControlIt control_first, ControlIt control_last) {              \
    hip::detail::operands o;                                                                                  \
    hip::detail::collect(o, target);                                                                          \
    hip::detail::add_all(o, control_first, control_last);                                                     \
    hip_core::NAME(p, first, last KET_HIP_UNPAREN ARGS, o);                                                   \
  }                                                                                                           \
  template <typename Complex, typename TargetIt, hip::detail::if_iterator<TargetIt> = 0>                      \
  inline void NAME(hip::policy p, hip::device_iterator<Complex> first, hip::device_iterator<Complex> last     \
                       KET_HIP_UNPAREN PARAMS,                                                                \
                   TargetIt target_first, TargetIt target_last) {                                             \
    hip::detail::operands o;                                                                                  \
    hip::detail::add_all(o, target_first, target_last);                                                       \
    hip_core::NAME(p, first, last KET_HIP_UNPAREN ARGS, o);                                                   \
  }                                                                                                           \
  template <typename Complex, typename TargetIt, typename ControlIt, hip::detail::if_iterator<TargetIt> = 0,  \
            hip::detail::if_iterator<ControlIt> = 0>                                                          \
  inline void NAME(hip::policy p, hip::device_iterator<Complex> first, hip::device_iterator<Complex> last     \
                       KET_HIP_UNPAREN PARAMS,                                                                \
                   TargetIt target_first, TargetIt target_last, ControlIt control_first,                      \
                   ControlIt control_last) {                                                                  \
    hip::detail::operands o;                                                                                  \
    hip::detail::add_all(o, target_first, target_last);                                                       \
    hip::detail::add_all(o, control_first, control_last);                                                     \
    hip_core::NAME(p, first, last KET_HIP_UNPAREN ARGS, o);                                                   \
  }                                                                                                           \
  namespace qubit_ranges {                                                                                    \
  template <typename Complex, typename Targets, typename Controls>                                            \
  inline void NAME(hip::policy p, hip::device_iterator<Complex> first, hip::device_iterator<Complex> last     \
                       KET_HIP_UNPAREN PARAMS,                                                                \
                   const Targets& targets, const Controls& controls) {                                        \
    hip_core::NAME(p, first, last KET_HIP_UNPAREN ARGS, hip::detail::from_ranges(targets, controls));         \
  }                                                                                                           \
  template <typename Complex, typename Targets>                                                               \
  inline void NAME(hip::policy p, hip::device_iterator<Complex> first, hip::device_iterator<Complex> last     \
                       KET_HIP_UNPAREN PARAMS,                                                                \
                   const Targets& targets) {                                                                  \
    hip_core::NAME(p, first, last KET_HIP_UNPAREN ARGS, hip::detail::from_ranges(targets));                   \
  }                                                                                                           \
  }                                                                                                           \
  namespace ranges {                                                                                          \
  template <typename Complex, typename Targets, typename Controls>                                            \
  inline hip::state<Complex>& NAME(hip::policy p, hip::state<Complex>& s KET_HIP_UNPAREN PARAMS,              \
                                   const Targets& targets, const Controls& controls) {                        \
    hip_core::NAME(p, s.begin(), s.end() KET_HIP_UNPAREN ARGS, hip::detail::from_ranges(targets, controls));  \
    return s;                                                                                                 \
  }                                                                                                           \
  template <typename Complex, typename Targets>                                                               \
  inline hip::state<Complex>& NAME(hip::policy p, hip::state<Complex>& s KET_HIP_UNPAREN PARAMS,              \
                                   const Targets& targets) {                                                  \
    hip_core::NAME(p, s.begin(), s.end() KET_HIP_UNPAREN ARGS, hip::detail::from_ranges(targets));            \
    return s;                                                                                                 \
  }                                                                                                           \
  }                                                                                                           \
  }

#define KET_HIP_GATE_AND_ADJOINT(NAME, PARAMS, ARGS) \
  KET_HIP_GATE(NAME, PARAMS, ARGS)                   \
  KET_HIP_GATE(adj_##NAME, PARAMS, ARGS)

KET_HIP_GATE_AND_ADJOINT(hadamard, (), ())
KET_HIP_GATE_AND_ADJOINT(pauli_x, (), ())
KET_HIP_GATE_AND_ADJOINT(not_, (), ())
KET_HIP_GATE_AND_ADJOINT(controlled_not, (), ())
KET_HIP_GATE_AND_ADJOINT(toffoli, (), ())
KET_HIP_GATE_AND_ADJOINT(pauli_y, (), ())
KET_HIP_GATE_AND_ADJOINT(pauli_z, (), ())
KET_HIP_GATE_AND_ADJOINT(swap, (), ())
KET_HIP_GATE_AND_ADJOINT(sqrt_pauli_x, (), ())
KET_HIP_GATE_AND_ADJOINT(sqrt_pauli_y, (), ())
KET_HIP_GATE_AND_ADJOINT(sqrt_pauli_z, (), ())
KET_HIP_GATE_AND_ADJOINT(x_rotation_half_pi, (), ())
KET_HIP_GATE_AND_ADJOINT(y_rotation_half_pi, (), ())
KET_HIP_GATE_AND_ADJOINT(phase_shift_coeff, (, std::complex<double> coeff), (, coeff))
KET_HIP_GATE_AND_ADJOINT(phase_shift, (, double phase), (, phase))
KET_HIP_GATE_AND_ADJOINT(controlled_phase_shift_coeff, (, std::complex<double> coeff), (, coeff))
KET_HIP_GATE_AND_ADJOINT(controlled_phase_shift, (, double phase), (, phase))
KET_HIP_GATE_AND_ADJOINT(phase_shift2, (, double phase1, double phase2), (, phase1, phase2))
KET_HIP_GATE_AND_ADJOINT(phase_shift3, (, double phase1, double phase2, double phase3), (, phase1, phase2, phase3))
KET_HIP_GATE_AND_ADJOINT(exponential_pauli_x, (, double phase), (, phase))
KET_HIP_GATE_AND_ADJOINT(exponential_pauli_y, (, double phase), (, phase))
KET_HIP_GATE_AND_ADJOINT(exponential_pauli_z, (, double phase), (, phase))
KET_HIP_GATE_AND_ADJOINT(exponential_pauli_x_coeff, (, std::complex<double> coeff), (, coeff))
KET_HIP_GATE_AND_ADJOINT(exponential_pauli_y_coeff, (, std::complex<double> coeff), (, coeff))
KET_HIP_GATE_AND_ADJOINT(exponential_pauli_z_coeff, (, std::complex<double> coeff), (, coeff))
KET_HIP_GATE_AND_ADJOINT(exponential_swap, (, double phase), (, phase))
KET_HIP_GATE_AND_ADJOINT(exponential_swap_coeff, (, std::complex<double> coeff), (, coeff))
KET_HIP_GATE_AND_ADJOINT(controlled_v, (, double phase), (, phase))
KET_HIP_GATE_AND_ADJOINT(controlled_v_coeff, (, std::complex<double> coeff), (, coeff))
#undef KET_HIP_GATE_AND_ADJOINT
#undef KET_HIP_GATE
#undef KET_HIP_UNPAREN

}  // namespace gate

// <psi2|psi1> -- ket/include/ket/inner_product.hpp:32-56
template <typename Complex1, typename Complex2>
inline std::complex<double> inner_product(hip::policy p, hip::device_iterator<Complex1> first1,
                                          hip::device_iterator<Complex1> last1, hip::device_iterator<Complex2> first2) {
  double r[2];
  hip::check(brk_inner_product(first1.get(), hip::dtype_of<Complex1>::value, first2.get(),
                               hip::dtype_of<Complex2>::value, hip::detail::span(first1, last1), r, p.stream));
  return {r[0], r[1]};
}

// |<psi2|psi1>|^2 -- ket/include/ket/fidelity.hpp:18-25
template <typename Complex1, typename Complex2>
inline double fidelity(hip::policy p, hip::device_iterator<Complex1> first1, hip::device_iterator<Complex1> last1,
                       hip::device_iterator<Complex2> first2) {
  double r = 0.0;
  hip::check(brk_fidelity(first1.get(), hip::dtype_of<Complex1>::value, first2.get(), hip::dtype_of<Complex2>::value,
                          hip::detail::span(first1, last1), &r, p.stream));
  return r;
}

// {Re, Im of sum conj(a0) a1, (sum |a0|^2 - |a1|^2) / 2} -- spin_expectation_value.hpp:20-79
template <typename Complex, typename S, typename B>
inline std::array<double, 3> spin_expectation_value(hip::policy p, hip::device_iterator<Complex> first,
                                                    hip::device_iterator<Complex> last, ::ket::qubit<S, B> q) {
  std::array<double, 3> r{};
  hip::check(brk_spin_expectation_value(first.get(), hip::detail::span(first, last), hip::dtype_of<Complex>::value,
                                        static_cast<unsigned>(static_cast<B>(q)), r.data(), p.stream));
  return r;
}

// every qubit's spin_expectation_value -- ket/include/ket/all_spin_expectation_values.hpp:22-42
template <typename Complex>
inline std::vector<std::array<double, 3>> all_spin_expectation_values(hip::policy p, hip::device_iterator<Complex> first,
                                                                      hip::device_iterator<Complex> last) {
  const std::uint64_t size = hip::detail::span(first, last);
  unsigned n = 0;
  while ((std::uint64_t{1} << n) < size) ++n;
  std::vector<std::array<double, 3>> out(n);
  for (unsigned q = 0; q < n; ++q)
    hip::check(brk_spin_expectation_value(first.get(), size, hip::dtype_of<Complex>::value, q, out[q].data(),
                                          p.stream));
  return out;
}

namespace hip {
namespace detail {
// std::uniform_real_distribution<double>{0, m}(g) draws generate_canonical(g) * m
// (positive_random_value_upto.hpp:12-18); the device applies m, so only the canonical part is drawn.
template <typename RandomNumberGenerator>
inline double canonical(RandomNumberGenerator& g) {
  return std::generate_canonical<double, std::numeric_limits<double>::digits>(g);
}
template <typename Complex>
inline unsigned qubit_count(hip::device_iterator<Complex> first, hip::device_iterator<Complex> last) {
  unsigned n = 0;
  while ((std::uint64_t{1} << n) < span(first, last)) ++n;
  return n;
}
}  // namespace detail
}  // namespace hip

// ket::generate_events (generate_events.hpp:15-66): num_events draws; the state is left
// unchanged (the reference overwrites it with its cumulative probabilities)
template <typename Complex, typename StateInteger, typename Allocator, typename RandomNumberGenerator>
inline void generate_events(hip::policy p, std::vector<StateInteger, Allocator>& result,
                            hip::device_iterator<Complex> first, hip::device_iterator<Complex> last, int num_events,
                            RandomNumberGenerator& g) {
  result.clear();
  if (num_events <= 0) return;
  std::vector<double> c(static_cast<std::size_t>(num_events));
  for (auto& u : c) u = hip::detail::canonical(g);
  std::vector<std::uint64_t> events(c.size());
  hip::check(brk_generate_events(first.get(), hip::detail::span(first, last), hip::dtype_of<Complex>::value, c.data(),
                                 static_cast<unsigned>(c.size()), events.data(), p.stream));
  result.assign(events.begin(), events.end());
}
template <typename Complex, typename StateInteger, typename Allocator, typename RandomNumberGenerator>
inline void generate_events(hip::policy p, std::vector<StateInteger, Allocator>& result,
                            hip::device_iterator<Complex> first, hip::device_iterator<Complex> last, int num_events,
                            const RandomNumberGenerator&, typename RandomNumberGenerator::result_type seed) {
  RandomNumberGenerator g(seed);
  generate_events(p, result, first, last, num_events, g);
}
// ket::measure (measure.hpp:14-40): one draw, then the state becomes that basis state
template <typename Complex, typename RandomNumberGenerator>
inline std::uint64_t measure(hip::policy p, hip::device_iterator<Complex> first, hip::device_iterator<Complex> last,
                             RandomNumberGenerator& g) {
  const double c = hip::detail::canonical(g);
  std::uint64_t index = 0;
  const std::uint64_t size = hip::detail::span(first, last);
  hip::check(brk_generate_events(first.get(), size, hip::dtype_of<Complex>::value, &c, 1u, &index, p.stream));
  hip::check(brk_set_basis_state(first.get(), size, hip::dtype_of<Complex>::value, index, p.stream));
  return index;
}
// ket::shor_box (shor_box.hpp:60-98) with the layout bra uses (bra/src/state.cpp:2173-2191): the
// exponent qubits are the top ones in ascending order, the modular-exponentiation qubits 0, 1, ...
template <typename Complex, typename StateInteger, typename Qubits>
inline void shor_box(hip::policy p, hip::device_iterator<Complex> first, hip::device_iterator<Complex> last,
                     StateInteger base, StateInteger divisor, const Qubits& exponent_qubits,
                     const Qubits& modular_exponentiation_qubits) {
  const unsigned n = hip::detail::qubit_count(first, last);
  const unsigned m = static_cast<unsigned>(std::distance(std::begin(exponent_qubits), std::end(exponent_qubits)));
  unsigned k = 0;
  bool layout = true;
  for (const auto& q : modular_exponentiation_qubits) layout = layout && static_cast<unsigned>(q) == k++;
  for (const auto& q : exponent_qubits) layout = layout && static_cast<unsigned>(q) == k++;
  if (!layout || k != n || m > n)
    throw hip::error(BRK_ERR_UNSUPPORTED, "shor_box supports bra's layout only (exponent qubits on top, ascending)");
  hip::check(brk_shor_box(first.get(), hip::detail::span(first, last), hip::dtype_of<Complex>::value, m,
                          static_cast<std::uint64_t>(divisor), static_cast<std::uint64_t>(base), p.stream));
}

namespace ranges {
template <typename Complex>
inline std::vector<std::array<double, 3>> all_spin_expectation_values(hip::policy p, hip::state<Complex>& s) {
  return ::ket::all_spin_expectation_values(p, s.begin(), s.end());
}
template <typename Complex, typename StateInteger, typename Allocator, typename RandomNumberGenerator>
inline void generate_events(hip::policy p, std::vector<StateInteger, Allocator>& result, hip::state<Complex>& s,
                            int num_events, RandomNumberGenerator& g) {
  ::ket::generate_events(p, result, s.begin(), s.end(), num_events, g);
}
template <typename Complex, typename StateInteger, typename Allocator, typename RandomNumberGenerator>
inline void generate_events(hip::policy p, std::vector<StateInteger, Allocator>& result, hip::state<Complex>& s,
                            int num_events, const RandomNumberGenerator& g,
                            typename RandomNumberGenerator::result_type seed) {
  ::ket::generate_events(p, result, s.begin(), s.end(), num_events, g, seed);
}
template <typename Complex, typename RandomNumberGenerator>
inline std::uint64_t measure(hip::policy p, hip::state<Complex>& s, RandomNumberGenerator& g) {
  return ::ket::measure(p, s.begin(), s.end(), g);
}
template <typename Complex, typename StateInteger, typename Qubits>
inline hip::state<Complex>& shor_box(hip::policy p, hip::state<Complex>& s, StateInteger base, StateInteger divisor,
                                     const Qubits& exponent_qubits, const Qubits& modular_exponentiation_qubits) {
  ::ket::shor_box(p, s.begin(), s.end(), base, divisor, exponent_qubits, modular_exponentiation_qubits);
  return s;
}
}  // namespace ranges

namespace gate {
#ifndef KET_GATE_PROJECTIVE_MEASUREMENT_HPP
enum class outcome : int { unspecified = -1, zero = 0, one = 1 };  // projective_measurement.hpp:30
#endif
// P(qubit = 0 / 1), then the kept half times P^(-1/2) and the other half +0:
// projective_measurement.hpp:137-176, clear.hpp:85-108, set.hpp:88-111
template <typename Complex, typename RandomNumberGenerator, typename S, typename B>
inline outcome projective_measurement(hip::policy p, hip::device_iterator<Complex> first,
                                      hip::device_iterator<Complex> last, RandomNumberGenerator& g,
                                      ::ket::qubit<S, B> qubit) {
  const unsigned q = static_cast<unsigned>(static_cast<B>(qubit));
  const std::uint64_t size = hip::detail::span(first, last);
  double pr[2];
  hip::check(brk_zero_one_probabilities(first.get(), size, hip::dtype_of<Complex>::value, q, pr, p.stream));
  const double r = std::uniform_real_distribution<double>{0.0, pr[0] + pr[1]}(g);
  const int v = r < pr[0] ? 0 : 1;
  hip::check(brk_project_qubit(first.get(), size, hip::dtype_of<Complex>::value, q, v, std::pow(pr[v], -0.5),
                               p.stream));
  return v ? outcome::one : outcome::zero;
}
template <typename Complex, typename S, typename B>
inline void clear(hip::policy p, hip::device_iterator<Complex> first, hip::device_iterator<Complex> last,
                  ::ket::qubit<S, B> qubit) {
  const unsigned q = static_cast<unsigned>(static_cast<B>(qubit));
  const std::uint64_t size = hip::detail::span(first, last);
  double pr[2];
  hip::check(brk_zero_one_probabilities(first.get(), size, hip::dtype_of<Complex>::value, q, pr, p.stream));
  hip::check(brk_project_qubit(first.get(), size, hip::dtype_of<Complex>::value, q, 0, std::pow(pr[0], -0.5),
                               p.stream));
}
template <typename Complex, typename S, typename B>
inline void set(hip::policy p, hip::device_iterator<Complex> first, hip::device_iterator<Complex> last,
                ::ket::qubit<S, B> qubit) {
  const unsigned q = static_cast<unsigned>(static_cast<B>(qubit));
  const std::uint64_t size = hip::detail::span(first, last);
  double pr[2];
  hip::check(brk_zero_one_probabilities(first.get(), size, hip::dtype_of<Complex>::value, q, pr, p.stream));
  hip::check(brk_project_qubit(first.get(), size, hip::dtype_of<Complex>::value, q, 1, std::pow(pr[1], -0.5),
                               p.stream));
}
namespace ranges {
template <typename Complex, typename RandomNumberGenerator, typename S, typename B>
inline outcome projective_measurement(hip::policy p, hip::state<Complex>& s, RandomNumberGenerator& g,
                                      ::ket::qubit<S, B> qubit) {
  return ::ket::gate::projective_measurement(p, s.begin(), s.end(), g, qubit);
}
template <typename Complex, typename S, typename B>
inline hip::state<Complex>& clear(hip::policy p, hip::state<Complex>& s, ::ket::qubit<S, B> qubit) {
  ::ket::gate::clear(p, s.begin(), s.end(), qubit);
  return s;
}
template <typename Complex, typename S, typename B>
inline hip::state<Complex>& set(hip::policy p, hip::state<Complex>& s, ::ket::qubit<S, B> qubit) {
  ::ket::gate::set(p, s.begin(), s.end(), qubit);
  return s;
}
}  // namespace ranges
}  // namespace gate

namespace mpi {

// Logical -> physical qubit map -- ket/include/ket/mpi/qubit_permutation.hpp:40-302
template <typename StateInteger = std::uint64_t, typename BitInteger = unsigned int>
class qubit_permutation {
  std::vector<BitInteger> data_, inverse_;

 public:
  explicit qubit_permutation(std::size_t n) : data_(n), inverse_(n) {
    for (std::size_t i = 0; i < n; ++i) data_[i] = inverse_[i] = static_cast<BitInteger>(i);
  }
  std::size_t size() const { return data_.size(); }
  BitInteger operator[](::ket::qubit<StateInteger, BitInteger> q) const { return data_[static_cast<BitInteger>(q)]; }
  BitInteger inverse(BitInteger physical) const { return inverse_[physical]; }
  const std::vector<BitInteger>& data() const { return data_; }
  // qubit_permutation::permutate (qubit_permutation.hpp:242-252)
  void permutate(::ket::qubit<StateInteger, BitInteger> q1, ::ket::qubit<StateInteger, BitInteger> q2) {
    const auto a = static_cast<BitInteger>(q1), b = static_cast<BitInteger>(q2);
    std::swap(inverse_[data_[a]], inverse_[data_[b]]);
    std::swap(data_[a], data_[b]);
  }
};

// permutate_bits / inverse_permutate_bits (qubit_permutation.hpp:348-484): bit b of a logical
// index moves to bit permutation[b] of the physical index.
template <typename S, typename B, typename U>
inline U permutate_bits(const qubit_permutation<S, B>& perm, U logical) {
  U out = 0;
  for (std::size_t b = 0; b < perm.size(); ++b)
    if ((logical >> b) & U{1}) out |= U{1} << perm.data()[b];
  return out;
}
template <typename S, typename B, typename U>
inline U inverse_permutate_bits(const qubit_permutation<S, B>& perm, U physical) {
  U out = 0;
  for (std::size_t b = 0; b < perm.size(); ++b)
    if ((physical >> perm.data()[b]) & U{1}) out |= U{1} << b;
  return out;
}

namespace utility {
// Local qubit swap: exchanges the data of two physical qubits and updates the
// permutation (swap_permutated_local_qubits.hpp:26-81 + make_local_swap_qubit.hpp:62-66).
template <typename Complex, typename S, typename B>
inline void swap_permutated_local_qubits(hip::policy p, hip::state<Complex>& s, qubit_permutation<S, B>& perm,
                                         B physical1, B physical2) {
  hip::check(brk_swap_permutated_local_qubits(s.data(), s.size(), hip::dtype_of<Complex>::value, physical1,
                                              physical2, p.stream));
  perm.permutate(::ket::qubit<S, B>{perm.inverse(physical1)}, ::ket::qubit<S, B>{perm.inverse(physical2)});
}
}  // namespace utility
}  // namespace mpi
}  // namespace ket
