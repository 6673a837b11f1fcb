/*
 * ket_oracle.h -- CPU oracle for the gate-application path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * ket library's algorithm (jhpc-quantum/RIKEN-braket, ket/include/ket/...),
 * used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
 * the checker and as the CPU baseline.  The engine (libbraket_hip.so) never
 * links, loads or calls it.
 *
 * States are arrays of interleaved doubles (re, im), i.e. the memory layout
 * of std::complex<double>[size]; size is a power of two; basis-index bit q is
 * qubit q.  Loops are OpenMP "parallel for" with a static schedule over the
 * 2^(n-k) groups of a k-qubit gate, like ket::utility::loop_n with
 * policy::parallel (ket/include/ket/utility/parallel/loop_n.hpp:155-201).
 */
#ifndef KET_ORACLE_H
#define KET_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

void or_set_num_threads(int n);
int or_get_num_threads(void);

/* index masks of ket's bit-mask driver (gate_bitmasks.hpp:50-130) and the
   address of local value `v` of group `g` (index_with_qubits.hpp:91-112).
   qubits are in argument order; index_masks gets k+1 entries. */
void or_make_index_masks(const unsigned* qubits, unsigned k, uint64_t* index_masks);
uint64_t or_index_with_qubits(uint64_t g, uint64_t v, const unsigned* qubits, unsigned k,
                              const uint64_t* index_masks);

/* gates (hadamard.hpp, pauli_x.hpp, pauli_y.hpp, pauli_z.hpp, phase_shift.hpp,
   exponential_pauli_{x,y,z}.hpp, swap.hpp) */
void or_hadamard(double* s, uint64_t size, unsigned target, const unsigned* controls, unsigned nc);
void or_pauli_x(double* s, uint64_t size, const unsigned* targets, unsigned nt, const unsigned* controls,
                unsigned nc);
void or_pauli_y(double* s, uint64_t size, const unsigned* targets, unsigned nt, const unsigned* controls,
                unsigned nc);
void or_pauli_z(double* s, uint64_t size, const unsigned* qubits, unsigned nq);
void or_pauli_zn(double* s, uint64_t size, const unsigned* targets, unsigned nt, const unsigned* controls,
                 unsigned nc);
void or_sqrt_pauli_zn(double* s, uint64_t size, const unsigned* targets, unsigned nt, const unsigned* controls,
                      unsigned nc, int adjoint);
void or_phase_shift_coeff(double* s, uint64_t size, const double coeff[2], const unsigned* qubits, unsigned nq);
void or_phase_shift2(double* s, uint64_t size, double phase1, double phase2, unsigned target,
                     const unsigned* controls, unsigned nc);
void or_phase_shift3(double* s, uint64_t size, double phase1, double phase2, double phase3, unsigned target,
                     const unsigned* controls, unsigned nc);
void or_exponential_pauli_x(double* s, uint64_t size, double phase, const unsigned* targets, unsigned nt,
                            const unsigned* controls, unsigned nc);
void or_exponential_pauli_y(double* s, uint64_t size, double phase, const unsigned* targets, unsigned nt,
                            const unsigned* controls, unsigned nc);
void or_exponential_pauli_z(double* s, uint64_t size, double phase, const unsigned* targets, unsigned nt,
                            const unsigned* controls, unsigned nc);
void or_swap(double* s, uint64_t size, unsigned q1, unsigned q2, const unsigned* controls, unsigned nc);
void or_exponential_swap(double* s, uint64_t size, double phase, unsigned q1, unsigned q2, const unsigned* controls,
                         unsigned nc);
void or_gate_1q(double* s, uint64_t size, unsigned target, const double m[8], const unsigned* controls,
                unsigned nc);

/* local qubit swap and permutation bookkeeping
   (mpi/utility/detail/swap_permutated_local_qubits.hpp:26-81,
    mpi/qubit_permutation.hpp:242-252, :348-484) */
void or_swap_permutated_local_qubits(double* s, uint64_t size, unsigned p1, unsigned p2);
void or_permutate(unsigned* perm, unsigned* inverse, unsigned n, unsigned q1, unsigned q2);
uint64_t or_permutate_bits(const unsigned* perm, unsigned n, uint64_t x);
uint64_t or_inverse_permutate_bits(const unsigned* perm, const unsigned* inverse, unsigned n, uint64_t x);

/* reductions (inner_product.hpp:32-56, spin_expectation_value.hpp:20-79,
   expectation_value.hpp:53-131 with pauli_index_coeff.hpp:32-59) */
void or_inner_product(const double* a, const double* b, uint64_t size, double out[2]);
double or_norm2(const double* a, uint64_t size);
void or_spin_expectation_value(const double* s, uint64_t size, unsigned qubit, double out[3]);
void or_expectation_pauli_string(const double* s, uint64_t size, const char* pauli, const unsigned* qubits,
                                 unsigned nq, double out[2]);
double or_expectation_ising(const double* s, uint64_t size, const double* J, const double* h);
void or_apply_ising(double* out, const double* psi, uint64_t size, const double* J, const double* h);

/* measurement and sampling.  Random numbers come from std::mt19937_64
   (bra/include/bra/state.hpp:459) drawn through std::uniform_real_distribution
   (ket/utility/positive_random_value_upto.hpp:12-18), whose libstdc++ form is
   generate_canonical: one 64-bit draw g, canonical = (double)g / 2^64, and the
   value canonical * maximum. */
typedef struct or_mt64 {
  uint64_t mt[312];
  unsigned index;
} or_mt64;
void or_mt64_seed(or_mt64* g, uint64_t seed);
uint64_t or_mt64_next(or_mt64* g);
double or_mt64_canonical(or_mt64* g);
/* P(0), P(1) of a qubit, summed in long double (projective_measurement.hpp:28-62) */
void or_zero_one_probabilities(const double* s, uint64_t size, unsigned qubit, double out[2]);
/* keep the half where the qubit equals value, times multiplier; +0 elsewhere
   (projective_measurement.hpp:64-135, clear.hpp:60-83, set.hpp:61-86) */
void or_project(double* s, uint64_t size, unsigned qubit, int value, double multiplier);
/* ket::gate::projective_measurement with the canonical draw given (projective_measurement.hpp:137-176) */
int or_projective_measurement(double* s, uint64_t size, unsigned qubit, double canonical);
/* ket::generate_events without the in-place scan: a sequential FP64 inclusive
   scan of |a|^2 and, per event, upper_bound of canonical[e] * total
   (generate_events.hpp:15-48); the state is left unchanged. */
void or_generate_events(const double* s, uint64_t size, const double* canonical, unsigned num_events,
                        uint64_t* events);

/* ket::shor_box (shor_box.hpp:60-98): exponent qubits are the top m */
void or_shor_box(double* s, uint64_t size, unsigned m, uint64_t divisor, uint64_t base);

#ifdef __cplusplus
}
#endif

#endif
