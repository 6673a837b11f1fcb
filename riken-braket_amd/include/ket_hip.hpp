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
#include <complex>
#include <cstdint>
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

}  // namespace detail
}  // namespace hip

namespace gate {

// H -- ket/include/ket/gate/hadamard.hpp:34-223
template <typename Complex, typename... Qubits>
inline void hadamard(hip::policy p, hip::device_iterator<Complex> first, hip::device_iterator<Complex> last,
                     Qubits... qubits) {
  auto o = hip::detail::split(qubits...);
  if (o.nt != 1) throw hip::error(BRK_ERR_INVALID_ARGUMENT, "hadamard takes one target qubit");
  hip::check(brk_hadamard(first.get(), hip::detail::span(first, last), hip::dtype_of<Complex>::value, o.targets[0],
                          hip::detail::ptr_or_null(o.controls, o.nc), o.nc, p.stream));
}

// X...X under controls -- pauli_x.hpp
template <typename Complex, typename... Qubits>
inline void pauli_x(hip::policy p, hip::device_iterator<Complex> first, hip::device_iterator<Complex> last,
                    Qubits... qubits) {
  auto o = hip::detail::split(qubits...);
  hip::check(brk_pauli_x(first.get(), hip::detail::span(first, last), hip::dtype_of<Complex>::value, o.targets, o.nt,
                         hip::detail::ptr_or_null(o.controls, o.nc), o.nc, p.stream));
}

// NOT / CNOT -- not_.hpp:33-51, controlled_not.hpp:30-37
template <typename Complex, typename... Qubits>
inline void not_(hip::policy p, hip::device_iterator<Complex> first, hip::device_iterator<Complex> last,
                 Qubits... qubits) {
  pauli_x(p, first, last, qubits...);
}
template <typename Complex, typename... Qubits>
inline void controlled_not(hip::policy p, hip::device_iterator<Complex> first, hip::device_iterator<Complex> last,
                           Qubits... qubits) {
  auto o = hip::detail::split(qubits...);
  if (o.nt != 1 || o.nc == 0) throw hip::error(BRK_ERR_INVALID_ARGUMENT, "controlled_not takes a target and controls");
  hip::check(brk_controlled_not(first.get(), hip::detail::span(first, last), hip::dtype_of<Complex>::value,
                                o.targets[0], o.controls, o.nc, p.stream));
}

// Y...Y under controls -- pauli_y.hpp
template <typename Complex, typename... Qubits>
inline void pauli_y(hip::policy p, hip::device_iterator<Complex> first, hip::device_iterator<Complex> last,
                    Qubits... qubits) {
  auto o = hip::detail::split(qubits...);
  hip::check(brk_pauli_y(first.get(), hip::detail::span(first, last), hip::dtype_of<Complex>::value, o.targets, o.nt,
                         hip::detail::ptr_or_null(o.controls, o.nc), o.nc, p.stream));
}

// Z on control qubits (Z, CZ, CnZ) -- pauli_z.hpp:55-210; targets and controls are treated alike
template <typename Complex, typename... Qubits>
inline void pauli_z(hip::policy p, hip::device_iterator<Complex> first, hip::device_iterator<Complex> last,
                    Qubits... qubits) {
  auto o = hip::detail::split(qubits...);
  unsigned all[2 * BRK_MAX_OPERATED_QUBITS], n = 0;
  hip::detail::all_qubits(o, all, &n);
  hip::check(brk_pauli_z(first.get(), hip::detail::span(first, last), hip::dtype_of<Complex>::value, all, n, p.stream));
}

// U1 from its coefficient e^{i theta} -- phase_shift.hpp:37-93
template <typename Complex, typename... Qubits>
inline void phase_shift_coeff(hip::policy p, hip::device_iterator<Complex> first, hip::device_iterator<Complex> last,
                              std::complex<double> coeff, Qubits... qubits) {
  auto o = hip::detail::split(qubits...);
  unsigned all[2 * BRK_MAX_OPERATED_QUBITS], n = 0;
  hip::detail::all_qubits(o, all, &n);
  const double c[2] = {coeff.real(), coeff.imag()};
  hip::check(brk_phase_shift_coeff(first.get(), hip::detail::span(first, last), hip::dtype_of<Complex>::value, c, all,
                                   n, p.stream));
}

// U1(theta) -- phase_shift.hpp:510-521
template <typename Complex, typename... Qubits>
inline void phase_shift(hip::policy p, hip::device_iterator<Complex> first, hip::device_iterator<Complex> last,
                        double phase, Qubits... qubits) {
  auto o = hip::detail::split(qubits...);
  unsigned all[2 * BRK_MAX_OPERATED_QUBITS], n = 0;
  hip::detail::all_qubits(o, all, &n);
  hip::check(brk_phase_shift(first.get(), hip::detail::span(first, last), hip::dtype_of<Complex>::value, phase, all, n,
                             p.stream));
}

// U2(phase1, phase2) -- phase_shift.hpp:653-709
template <typename Complex, typename... Qubits>
inline void phase_shift2(hip::policy p, hip::device_iterator<Complex> first, hip::device_iterator<Complex> last,
                         double phase1, double phase2, Qubits... qubits) {
  auto o = hip::detail::split(qubits...);
  if (o.nt != 1) throw hip::error(BRK_ERR_INVALID_ARGUMENT, "phase_shift2 takes one target qubit");
  hip::check(brk_phase_shift2(first.get(), hip::detail::span(first, last), hip::dtype_of<Complex>::value, phase1,
                              phase2, o.targets[0], hip::detail::ptr_or_null(o.controls, o.nc), o.nc, p.stream));
}

// U3(phase1, phase2, phase3) -- phase_shift.hpp:1195-1254
template <typename Complex, typename... Qubits>
inline void phase_shift3(hip::policy p, hip::device_iterator<Complex> first, hip::device_iterator<Complex> last,
                         double phase1, double phase2, double phase3, Qubits... qubits) {
  auto o = hip::detail::split(qubits...);
  if (o.nt != 1) throw hip::error(BRK_ERR_INVALID_ARGUMENT, "phase_shift3 takes one target qubit");
  hip::check(brk_phase_shift3(first.get(), hip::detail::span(first, last), hip::dtype_of<Complex>::value, phase1,
                              phase2, phase3, o.targets[0], hip::detail::ptr_or_null(o.controls, o.nc), o.nc,
                              p.stream));
}

#define KET_HIP_EXPONENTIAL_PAULI(NAME, FN)                                                                      \
  template <typename Complex, typename... Qubits>                                                                 \
  inline void NAME(hip::policy p, hip::device_iterator<Complex> first, hip::device_iterator<Complex> last,        \
                   double phase, Qubits... qubits) {                                                              \
    auto o = hip::detail::split(qubits...);                                                                       \
    hip::check(FN(first.get(), hip::detail::span(first, last), hip::dtype_of<Complex>::value, phase, o.targets,  \
                  o.nt, hip::detail::ptr_or_null(o.controls, o.nc), o.nc, p.stream));                             \
  }
// exp(i phase P...P) -- exponential_pauli_{x,y,z}.hpp
KET_HIP_EXPONENTIAL_PAULI(exponential_pauli_x, brk_exponential_pauli_x)
KET_HIP_EXPONENTIAL_PAULI(exponential_pauli_y, brk_exponential_pauli_y)
KET_HIP_EXPONENTIAL_PAULI(exponential_pauli_z, brk_exponential_pauli_z)
#undef KET_HIP_EXPONENTIAL_PAULI

// SWAP under controls -- swap.hpp:30-75
template <typename Complex, typename... Qubits>
inline void swap(hip::policy p, hip::device_iterator<Complex> first, hip::device_iterator<Complex> last,
                 Qubits... qubits) {
  auto o = hip::detail::split(qubits...);
  if (o.nt != 2) throw hip::error(BRK_ERR_INVALID_ARGUMENT, "swap takes two target qubits");
  hip::check(brk_swap(first.get(), hip::detail::span(first, last), hip::dtype_of<Complex>::value, o.targets[0],
                      o.targets[1], hip::detail::ptr_or_null(o.controls, o.nc), o.nc, p.stream));
}

// exp(i phase SWAP) under controls -- exponential_swap.hpp:31-90, 300-350
template <typename Complex, typename... Qubits>
inline void exponential_swap(hip::policy p, hip::device_iterator<Complex> first, hip::device_iterator<Complex> last,
                             double phase, Qubits... qubits) {
  auto o = hip::detail::split(qubits...);
  if (o.nt != 2) throw hip::error(BRK_ERR_INVALID_ARGUMENT, "exponential_swap takes two target qubits");
  hip::check(brk_exponential_swap(first.get(), hip::detail::span(first, last), hip::dtype_of<Complex>::value, phase,
                                  o.targets[0], o.targets[1], hip::detail::ptr_or_null(o.controls, o.nc), o.nc,
                                  p.stream));
}

namespace ranges {
// range forms: (policy, state, ...) -> state&, as ket::gate::ranges::*
#define KET_HIP_RANGE(NAME)                                                        \
  template <typename Complex, typename... Args>                                    \
  inline hip::state<Complex>& NAME(hip::policy p, hip::state<Complex>& s, Args... args) { \
    ::ket::gate::NAME(p, s.begin(), s.end(), args...);                             \
    return s;                                                                      \
  }
KET_HIP_RANGE(hadamard)
KET_HIP_RANGE(pauli_x)
KET_HIP_RANGE(not_)
KET_HIP_RANGE(controlled_not)
KET_HIP_RANGE(pauli_y)
KET_HIP_RANGE(pauli_z)
KET_HIP_RANGE(phase_shift_coeff)
KET_HIP_RANGE(phase_shift)
KET_HIP_RANGE(phase_shift2)
KET_HIP_RANGE(phase_shift3)
KET_HIP_RANGE(exponential_pauli_x)
KET_HIP_RANGE(exponential_pauli_y)
KET_HIP_RANGE(exponential_pauli_z)
KET_HIP_RANGE(swap)
KET_HIP_RANGE(exponential_swap)
#undef KET_HIP_RANGE
}  // namespace ranges
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
