/*
 * braket_hip.h -- C ABI of the MI355X (gfx950) state-vector engine.
 *
 * Every entry point takes plain pointers and sizes.  A state vector is a
 * device array of 2^n amplitudes ("first" + "size", exactly like the
 * [first, last) iterator pair of the reference's ket gate functions), in one
 * of two element types:
 *
 *   BRK_C128  std::complex<double>  (16 B: re, im)   -- the plain path
 *   BRK_C64   std::complex<float>   ( 8 B: re, im)   -- precision / batch forms
 *
 * Basis index bit q is qubit q (qubit 0 = least significant bit).  All gate
 * functions update the state in place, allocate nothing, and are enqueued on
 * the HIP stream passed as `stream` (NULL = the null stream); they return
 * before the kernel finishes.  Reductions synchronise the stream and return
 * their result on the host.  Functions never throw; each returns a brk_status
 * and brk_last_error() gives a message for the calling thread.
 *
 * Two calls on the same state must not run concurrently on different streams
 * (the same rule the reference states for its parallel policies).  Calls on
 * different states may come from several host threads: the library's
 * per-device buffers (reduction partials, adjoint-gradient partials) are held
 * by one call at a time, from its launch until it has read its result back.
 */
#ifndef BRAKET_HIP_H
#define BRAKET_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Status codes and element types                                            */
/* ------------------------------------------------------------------------ */

typedef enum brk_status {
  BRK_OK = 0,
  BRK_ERR_INVALID_ARGUMENT = 1,   /* null pointer, size not a power of two, bad count */
  BRK_ERR_QUBIT_OUT_OF_RANGE = 2, /* 2^qubit >= size                                  */
  BRK_ERR_DUPLICATE_QUBIT = 3,    /* the same qubit given twice                       */
  BRK_ERR_TOO_MANY_QUBITS = 4,    /* more operated qubits than the kernel supports    */
  BRK_ERR_UNSUPPORTED_DTYPE = 5,
  BRK_ERR_OUT_OF_MEMORY = 6,
  BRK_ERR_HIP = 7,                /* HIP runtime or launch failure                     */
  BRK_ERR_UNSUPPORTED = 8         /* valid request that this build does not implement */
} brk_status;

typedef enum brk_dtype { BRK_C128 = 0, BRK_C64 = 1 } brk_dtype;

/* Maximum number of operated qubits (targets + controls) of one plain gate. */
#define BRK_MAX_OPERATED_QUBITS 16

const char* brk_status_string(int status);
/* Message describing the last failure on the calling thread ("" if none). */
const char* brk_last_error(void);
/* 1 when the library was built for gfx950 and a device is visible. */
int brk_device_count(int* count);
const char* brk_version(void);
/* "<marketing name> (<gfx arch>, <CUs> CUs)" of the current device. */
int brk_device_name(char* buffer, size_t size);

/* Launch tuning of the streaming gate kernels (no reference counterpart; the
   reference's analogue is its compile-time cache blocking,
   KET_DEFAULT_NUM_ON_CACHE_QUBITS, gate_bitmasks.hpp:614-616).  Keys:
   "pair_lds", "diag_lds" (bytes of LDS reserved per block, 0..65536: caps
   resident blocks per CU), "pair_remap", "diag_remap" (0/1: XCD-contiguous
   tile order), "nt_min_bytes" (states at least this large are streamed with
   non-temporal loads/stores), "max_blocks" (grid cap of the gate kernels),
   "line_bytes" (write granularity for control bits), "fused_groups",
   "fused_staged_groups" (grid caps of the fused kernels), "fused_staged"
   (0/1: LDS-staged or register-direct fused tiles), "fused_occupancy" (2 or
   4), "fused_chunked" (0/1), "block_lds" (65536, 32768 or 16384: gate-block
   chunk bytes), "block_run_bytes" (power of two in [16, 4096]: the lowest
   qubits covering this many bytes stay local in every gate block),
   "adjoint_blocks" (0/1: blocked or per-run adjoint sweep), "fused_chunks"
   (1/0: width-4/5 fused plans run as chunk passes of several blocks, or one
   launch per block).  Each key can
   also be set by the environment variable BRK_TUNE_<KEY> read at first use.
   Process-wide and not synchronised with launches: set before launching, not
   while another thread launches.  brk_get_tune reads a key back. */
int brk_tune(const char* key, int64_t value);
int brk_get_tune(const char* key, int64_t* value);

/* ------------------------------------------------------------------------ */
/* Device memory and state preparation (no reference counterpart: the       */
/* reference owns a std::vector; bra::nompi_state's constructor builds the  */
/* initial basis state, bra/include/bra/nompi_state.hpp:155-165).            */
/* ------------------------------------------------------------------------ */

int brk_malloc(void** device_ptr, size_t bytes);
int brk_free(void* device_ptr);
int brk_memcpy_htod(void* device_dst, const void* host_src, size_t bytes, void* stream);
int brk_memcpy_dtoh(void* host_dst, const void* device_src, size_t bytes, void* stream);
int brk_memcpy_dtod(void* device_dst, const void* device_src, size_t bytes, void* stream);
int brk_stream_synchronize(void* stream);
/* Runtime helpers for callers that reach HIP only through this library (one
   HIP runtime per process): free/total device memory, default-priority
   streams, and events for timing work on a stream. */
int brk_mem_info(size_t* free_bytes, size_t* total_bytes);
int brk_stream_create(void** stream);
int brk_stream_destroy(void* stream);
int brk_event_create(void** event);
int brk_event_destroy(void* event);
int brk_event_record(void* event, void* stream);
int brk_event_synchronize(void* event);
int brk_event_elapsed_ms(void* start, void* end, float* ms);

/* state := |basis_index> */
int brk_set_basis_state(void* first, uint64_t size, int dtype, uint64_t basis_index, void* stream);

/* A device state with its stream: the owned counterpart of the range the
   reference's gate functions take (ket's std::vector, bra::nompi_state's
   data_, bra/include/bra/nompi_state.hpp:155-165).  create allocates 2^n
   amplitudes of dtype and sets |0..0>; brk_state_data gives the
   (first, size, dtype) triple and the stream every other call takes.
   upload / download copy the whole state to or from host memory (2^n
   amplitudes of the state's dtype) and synchronise the stream; norm2 is
   brk_norm2 on the state. */
typedef struct brk_state brk_state;
int brk_state_create(unsigned num_qubits, int dtype, void* stream, brk_state** out);
int brk_state_destroy(brk_state* state);
int brk_state_data(const brk_state* state, void** first, uint64_t* size, int* dtype, void** stream);
int brk_state_set_basis(brk_state* state, uint64_t basis_index);
int brk_state_upload(brk_state* state, const void* host);
int brk_state_download(const brk_state* state, void* host);
int brk_state_norm2(const brk_state* state, double* result);
/* Deterministic, index-only fills used to build states of any size on the
   device (test states of the reference's ket/test/ *.cpp are of this form):
     re[i] = scales[0] * ((params[0]*i + params[1]) mod params[2])
     im[i] = scales[1] * ((params[3]*i + params[4]) mod params[5])
   with uint64 wrap-around arithmetic and "mod 0" meaning no reduction. */
int brk_fill_modular(void* first, uint64_t size, int dtype, const double scales[2],
                     const uint64_t params[6], void* stream);
/* re[i] = scale * u(seed, 2i), im[i] = scale * u(seed, 2i+1) where u is the
   splitmix64 finaliser of seed + (counter+1)*0x9E3779B97F4A7C15 mapped to
   [-1, 1) with 53 bits: (z >> 11) * 2^-53 * 2 - 1. */
int brk_fill_hash(void* first, uint64_t size, int dtype, uint64_t seed, double scale, void* stream);
/* dst (dtype_dst) := src (dtype_src), elementwise conversion c128 <-> c64. */
int brk_convert(void* dst, int dtype_dst, const void* src, int dtype_src, uint64_t size, void* stream);
/* Streaming copy kernel (bandwidth calibration and state copies). */
int brk_copy(void* dst, const void* src, uint64_t size, int dtype, void* stream);

/* ------------------------------------------------------------------------ */
/* Plain per-gate form.  Argument order follows the reference's iterator    */
/* form: (first, last) -> (first, size, dtype), then real parameters, then  */
/* the target qubit(s), then the control qubits.                            */
/* ------------------------------------------------------------------------ */

/* H on `target`, applied where every control bit is 1.
   Replaces ket::gate::hadamard (ket/include/ket/gate/hadamard.hpp:34-72 H,
   :77-130 CH, :133-223 CnH). */
int brk_hadamard(void* first, uint64_t size, int dtype, unsigned target,
                 const unsigned* controls, unsigned num_controls, void* stream);

/* X on every qubit of `targets` at once (X, XX, Xn), under controls.
   Replaces ket::gate::pauli_x (ket/include/ket/gate/pauli_x.hpp:31-61 X,
   :116-165 CX, multi-target/multi-control forms after :166). */
int brk_pauli_x(void* first, uint64_t size, int dtype, const unsigned* targets, unsigned num_targets,
                const unsigned* controls, unsigned num_controls, void* stream);

/* NOT / CNOT / CnNOT on one target.  Replaces ket::gate::not_
   (ket/include/ket/gate/not_.hpp:33-51) and ket::gate::controlled_not
   (ket/include/ket/gate/controlled_not.hpp:30-37). */
int brk_controlled_not(void* first, uint64_t size, int dtype, unsigned target,
                       const unsigned* controls, unsigned num_controls, void* stream);

/* Y on every qubit of `targets` at once (Y, YY, Yn), under controls.
   Replaces ket::gate::pauli_y (ket/include/ket/gate/pauli_y.hpp). */
int brk_pauli_y(void* first, uint64_t size, int dtype, const unsigned* targets, unsigned num_targets,
                const unsigned* controls, unsigned num_controls, void* stream);

/* Z acting on "control" qubits: the amplitude is negated where every listed
   bit is 1 (Z, CZ, CnZ).  Replaces ket::gate::pauli_z, case 1
   (ket/include/ket/gate/pauli_z.hpp:55-88 Z, :90-137 CZ, :139-210 CnZ). */
int brk_pauli_z(void* first, uint64_t size, int dtype, const unsigned* qubits, unsigned num_qubits,
                void* stream);

/* Z on every qubit of `targets` (Z...Z strings) where every control bit is 1:
   the amplitude is negated where an odd number of the target bits is 1.
   Replaces ket::gate::pauli_z, case 2 (ket/include/ket/gate/pauli_z.hpp:265-303
   Z and ZZ on ket::qubit targets, :305-350 CZ with a target, :392-470 C...CZ...Z). */
int brk_pauli_zn(void* first, uint64_t size, int dtype, const unsigned* targets, unsigned num_targets,
                 const unsigned* controls, unsigned num_controls, void* stream);

/* sqrt(Z) on every qubit of `targets` (sZ...sZ strings) where every control bit
   is 1: the amplitude is multiplied by i^f, f = number of target bits that are 1;
   the adjoint by (-i)^f.  Replaces ket::gate::sqrt_pauli_z and
   adj_sqrt_pauli_z, case 2 (ket/include/ket/gate/sqrt_pauli_z.hpp:428-700,
   :710-1000).  The control-qubit-only forms (case 1, :35-420) are
   brk_phase_shift_coeff with coeff = i or -i. */
int brk_sqrt_pauli_zn(void* first, uint64_t size, int dtype, const unsigned* targets, unsigned num_targets,
                      const unsigned* controls, unsigned num_controls, void* stream);
int brk_adj_sqrt_pauli_zn(void* first, uint64_t size, int dtype, const unsigned* targets, unsigned num_targets,
                          const unsigned* controls, unsigned num_controls, void* stream);

/* Multiply the amplitudes whose listed bits are all 1 by `coeff`
   (U1 / CU1 / CnU1 given exp(i theta)).  Replaces ket::gate::phase_shift_coeff
   (ket/include/ket/gate/phase_shift.hpp:37-93 and the controlled forms). */
int brk_phase_shift_coeff(void* first, uint64_t size, int dtype, const double coeff[2],
                          const unsigned* qubits, unsigned num_qubits, void* stream);
/* Same with coeff = exp(i phase).  Replaces ket::gate::phase_shift
   (ket/include/ket/gate/phase_shift.hpp:510-521). */
int brk_phase_shift(void* first, uint64_t size, int dtype, double phase,
                    const unsigned* qubits, unsigned num_qubits, void* stream);

/* U2(phase1, phase2) on `target` under controls.  Replaces
   ket::gate::phase_shift2 (ket/include/ket/gate/phase_shift.hpp:653-709). */
int brk_phase_shift2(void* first, uint64_t size, int dtype, double phase1, double phase2,
                     unsigned target, const unsigned* controls, unsigned num_controls, void* stream);

/* U3(phase1, phase2, phase3) on `target` under controls.  Replaces
   ket::gate::phase_shift3 (ket/include/ket/gate/phase_shift.hpp:1195-1254). */
int brk_phase_shift3(void* first, uint64_t size, int dtype, double phase1, double phase2, double phase3,
                     unsigned target, const unsigned* controls, unsigned num_controls, void* stream);

/* exp(i phase X...X) on `targets` under controls.  Replaces
   ket::gate::exponential_pauli_x (ket/include/ket/gate/exponential_pauli_x.hpp). */
int brk_exponential_pauli_x(void* first, uint64_t size, int dtype, double phase,
                            const unsigned* targets, unsigned num_targets,
                            const unsigned* controls, unsigned num_controls, void* stream);

/* exp(i phase Y...Y) on `targets` under controls.  Replaces
   ket::gate::exponential_pauli_y (ket/include/ket/gate/exponential_pauli_y.hpp:37-90,
   wrapper :473-495).  RY(theta) == exponential_pauli_y(-theta/2). */
int brk_exponential_pauli_y(void* first, uint64_t size, int dtype, double phase,
                            const unsigned* targets, unsigned num_targets,
                            const unsigned* controls, unsigned num_controls, void* stream);

/* exp(i phase Z...Z) on `targets` under controls.  Replaces
   ket::gate::exponential_pauli_z (ket/include/ket/gate/exponential_pauli_z.hpp:36-78,
   wrapper :367-390).  RZ(theta) == exponential_pauli_z(-theta/2). */
int brk_exponential_pauli_z(void* first, uint64_t size, int dtype, double phase,
                            const unsigned* targets, unsigned num_targets,
                            const unsigned* controls, unsigned num_controls, void* stream);

/* The same three gates from the coefficient coeff = {cos, sin} = e^{i phase}
   instead of the phase: the _coeff forms of ket
   (exponential_pauli_{x,y,z}.hpp: exponential_pauli_*_coeff). */
int brk_exponential_pauli_x_coeff(void* first, uint64_t size, int dtype, const double coeff[2],
                                  const unsigned* targets, unsigned num_targets,
                                  const unsigned* controls, unsigned num_controls, void* stream);
int brk_exponential_pauli_y_coeff(void* first, uint64_t size, int dtype, const double coeff[2],
                                  const unsigned* targets, unsigned num_targets,
                                  const unsigned* controls, unsigned num_controls, void* stream);
int brk_exponential_pauli_z_coeff(void* first, uint64_t size, int dtype, const double coeff[2],
                                  const unsigned* targets, unsigned num_targets,
                                  const unsigned* controls, unsigned num_controls, void* stream);

/* SWAP of qubit1 and qubit2 under controls.  Replaces ket::gate::swap
   (ket/include/ket/gate/swap.hpp:30-75 and the controlled form after :76). */
int brk_swap(void* first, uint64_t size, int dtype, unsigned qubit1, unsigned qubit2,
             const unsigned* controls, unsigned num_controls, void* stream);

/* exp(i phase SWAP) on qubit1, qubit2 under controls: |00> and |11> times
   e^{i phase}; (|01>, |10>) mixed by [[cos, i sin], [i sin, cos]].  Replaces
   ket::gate::exponential_swap (ket/include/ket/gate/exponential_swap.hpp:31-90,
   controlled form :92-…, wrappers :300-350). */
int brk_exponential_swap(void* first, uint64_t size, int dtype, double phase, unsigned qubit1, unsigned qubit2,
                         const unsigned* controls, unsigned num_controls, void* stream);
/* The same from coeff = e^{i phase} (exponential_swap_coeff, exponential_swap.hpp:31-90). */
int brk_exponential_swap_coeff(void* first, uint64_t size, int dtype, const double coeff[2], unsigned qubit1,
                               unsigned qubit2, const unsigned* controls, unsigned num_controls, void* stream);

/* General single-qubit gate: matrix = {m00.re, m00.im, m01.re, m01.im,
   m10.re, m10.im, m11.re, m11.im}, (a0', a1') = M (a0, a1), under controls.
   Replaces the generic driver ket::gate::gate applied with a 2x2 functor
   (ket/include/ket/gate/detail/gate_bitmasks.hpp:470-586). */
int brk_gate_1q(void* first, uint64_t size, int dtype, unsigned target, const double matrix[8],
                const unsigned* controls, unsigned num_controls, void* stream);

/* Data movement of a local qubit swap: exchanges the amplitudes whose
   physical bits (qubit1, qubit2) are (1,0) and (0,1).  The logical->physical
   map is kept on the host (ket::mpi::qubit_permutation::permutate).
   Replaces ket::mpi::utility::detail::swap_permutated_local_qubits
   (ket/include/ket/mpi/utility/detail/swap_permutated_local_qubits.hpp:26-81)
   for the single-device case (one data block). */
int brk_swap_permutated_local_qubits(void* first, uint64_t size, int dtype,
                                     unsigned permutated_qubit1, unsigned permutated_qubit2, void* stream);

/* ------------------------------------------------------------------------ */
/* FP64 reductions (partial sums in double, deterministic two-pass tree).   */
/* ------------------------------------------------------------------------ */

/* result = <psi2|psi1> = sum conj(psi2[i]) psi1[i].  Replaces ket::inner_product
   (ket/include/ket/inner_product.hpp:32-56).  dtype1/dtype2 may differ
   (c64 vs c128 fidelity). */
int brk_inner_product(const void* first1, int dtype1, const void* first2, int dtype2, uint64_t size,
                      double result[2], void* stream);
/* result = |<psi2|psi1>|^2.  Replaces ket::fidelity (ket/include/ket/fidelity.hpp:18-25). */
int brk_fidelity(const void* first1, int dtype1, const void* first2, int dtype2, uint64_t size,
                 double* result, void* stream);
/* result = <psi|psi>. */
int brk_norm2(const void* first, uint64_t size, int dtype, double* result, void* stream);
/* result = {sum Re(conj(a0) a1), sum Im(conj(a0) a1), (sum |a0|^2 - |a1|^2) / 2}
   over the pairs of `qubit`.  Replaces ket::spin_expectation_value
   (ket/include/ket/spin_expectation_value.hpp:20-79). */
int brk_spin_expectation_value(const void* first, uint64_t size, int dtype, unsigned qubit,
                               double result[3], void* stream);
/* <psi| P |psi> for a Pauli string P (characters I, X, Y, Z; character k acts on
   qubits[k]).  Replaces ket::expectation_value with the Pauli functor of
   bra::nompi_state::do_expectation_value (bra/src/nompi_state.cpp:1155-1202,
   ket/include/ket/gate/utility/pauli_index_coeff.hpp:32-59). */
int brk_expectation_pauli_string(const void* first, uint64_t size, int dtype, const char* pauli_string,
                                 const unsigned* qubits, unsigned num_qubits, double result[2], void* stream);
/* <psi| H |psi> for the open-chain Ising Hamiltonian
   H = sum_{i<n-1} J[i] Z_i Z_{i+1} + sum_i h[i] X_i, n = log2(size).
   (Composition of spin_expectation_value terms; no single reference function.) */
int brk_expectation_ising(const void* first, uint64_t size, int dtype, const double* J, const double* h,
                          double* result, void* stream);
/* out := H |psi> for the same Hamiltonian (the adjoint method's starting vector). */
int brk_apply_ising(void* out, const void* psi, uint64_t size, int dtype, const double* J, const double* h,
                    void* stream);

/* Overwrites the state with Shor's modular-exponentiation superposition:
   2^{-m/2} sum_{x < 2^m} |reverse_m(x)> on the top m qubits (m =
   num_exponent_qubits) times |v_x> on the n - m low qubits, where v_0 = 1,
   v_x = base^x mod divisor and only the low n - m bits of v_x are kept.
   Replaces ket::shor_box (ket/include/ket/shor_box.hpp:60-98) as
   bra::state::shor_box calls it (bra/src/state.cpp:2173-2191).  Needs
   (divisor - 1) * base < 2^64 (the reference's 64-bit recurrence does not
   wrap there; BRK_ERR_UNSUPPORTED otherwise).  Allocates a 2^m-entry table
   for the call and synchronises the stream. */
int brk_shor_box(void* first, uint64_t size, int dtype, unsigned num_exponent_qubits, uint64_t divisor,
                 uint64_t base, void* stream);

/* ------------------------------------------------------------------------ */
/* Measurement and sampling.  The random draws stay with the caller (bra    */
/* keeps a std::mt19937_64 per state, bra/include/bra/state.hpp:459); these */
/* calls take the canonical numbers u in [0, 1) that                        */
/* std::uniform_real_distribution turns into u * maximum.                  */
/* ------------------------------------------------------------------------ */
/* result = {P(qubit = 0), P(qubit = 1)}, FP64 sums in a fixed order.  Replaces
   projective_measurement_detail::zero_one_probabilities
   (ket/include/ket/gate/projective_measurement.hpp:28-62) and the zero / one
   probabilities of clear.hpp:26-55 and set.hpp:26-55. */
int brk_zero_one_probabilities(const void* first, uint64_t size, int dtype, unsigned qubit, double result[2],
                               void* stream);
/* Amplitudes whose `qubit` bit equals `value` are multiplied by `multiplier`;
   the others become +0.  Replaces change_state_after_measuring_zero/one
   (projective_measurement.hpp:64-135) and the state updates of
   ket::gate::clear (clear.hpp:57-83) and ket::gate::set (set.hpp:57-86), which
   pass multiplier = p^(-1/2). */
int brk_project_qubit(void* first, uint64_t size, int dtype, unsigned qubit, int value, double multiplier,
                      void* stream);
/* events[e] = the smallest basis index i whose cumulative probability
   sum_{j<=i} |a_j|^2 exceeds canonical[e] * total (upper_bound over the
   inclusive scan).  Replaces ket::generate_events
   (ket/include/ket/generate_events.hpp:15-48) and the sampling step of
   ket::measure (ket/include/ket/measure.hpp:14-40).  The reference overwrites
   the state with the scan; this call leaves it unchanged.  The scan is summed
   per chunk of 2^13 amplitudes and then across chunks, so an event can differ
   from a sequential scan only where canonical[e] * total lies within rounding
   of a cumulative boundary.  Synchronises the stream. */
int brk_generate_events(const void* first, uint64_t size, int dtype, const double* canonical, unsigned num_events,
                        uint64_t* events, void* stream);
/* brk_generate_events without any device allocation: no counterpart in the
   reference.  For callers that cannot allocate device memory (tight serving
   loops, code that forbids hipMalloc).  Same arguments, checks and events
   (bit for bit) as brk_generate_events; its buffers live in a static device
   array of 6.5 MiB that each device holds until the process ends.  Limits:
   states of at most 2^32 amplitudes (32 qubits), at most 65536 events per
   call (BRK_ERR_INVALID_ARGUMENT beyond either; the draws of a larger sample
   can be passed in consecutive calls).  Calls on one device share that array
   and take turns under a per-device lock held until the events are on the
   host.  Synchronises the stream; not for use under stream capture. */
int brk_generate_events_noalloc(const void* first, uint64_t size, int dtype, const double* canonical,
                                unsigned num_events, uint64_t* events, void* stream);

/* ------------------------------------------------------------------------ */
/* Circuits: an ordered list of operations on logical qubits, run with the  */
/* plain per-gate kernels.  Stands in for the gate list bra's interpreter   */
/* builds (bra/src/interpreter.cpp:371-825) and executes through            */
/* bra::nompi_state (interpreter::apply_circuit, interpreter.cpp:851-872).  */
/* ------------------------------------------------------------------------ */

typedef struct brk_circuit brk_circuit;

typedef enum brk_op {
  BRK_OP_H = 0,          /* targets[0], controls                                   */
  BRK_OP_X = 1,          /* targets (X..X), controls                               */
  BRK_OP_Y = 2,          /* targets (Y..Y), controls                               */
  BRK_OP_Z = 3,          /* -1 where every bit of targets and controls is 1         */
  BRK_OP_U1 = 4,         /* exp(i p0) where every bit of targets and controls is 1  */
  BRK_OP_U2 = 5,         /* U2(p0, p1) on targets[0], controls                     */
  BRK_OP_U3 = 6,         /* U3(p0, p1, p2) on targets[0], controls                 */
  BRK_OP_EX = 7,         /* exp(i p0 X..X) on targets, controls                    */
  BRK_OP_EY = 8,         /* exp(i p0 Y..Y)                                         */
  BRK_OP_EZ = 9,         /* exp(i p0 Z..Z)                                         */
  BRK_OP_RX = 10,        /* RX(p0) = EX(-p0/2)                                     */
  BRK_OP_RY = 11,        /* RY(p0) = EY(-p0/2)                                     */
  BRK_OP_RZ = 12,        /* RZ(p0) = EZ(-p0/2)                                     */
  BRK_OP_SWAP = 13,      /* SWAP(targets[0], targets[1]), controls                 */
  BRK_OP_LOCAL_SWAP = 14,/* data swap of PHYSICAL qubits targets[0], targets[1] and
                            permutate(inverse[t0], inverse[t1]) of the permutation  */
  BRK_OP_MATRIX = 15,    /* 2x2 matrix p0..p7 (as brk_gate_1q) on targets[0]       */
  BRK_OP_ESWAP = 16,     /* exp(i p0 SWAP) on targets[0], targets[1], controls      */
  BRK_OP_PHASE_COEFF = 17/* p0 + i p1 where every bit of targets and controls is 1
                            (ket::gate::phase_shift_coeff: S = sqrt(Z) with i, ...)  */
} brk_op;

int brk_circuit_create(unsigned num_qubits, brk_circuit** out);
int brk_circuit_destroy(brk_circuit* circuit);
/* Appends one operation; qubits are validated against num_qubits here. */
int brk_circuit_add(brk_circuit* circuit, int op, const unsigned* targets, unsigned num_targets,
                    const unsigned* controls, unsigned num_controls, const double* params, unsigned num_params);
int brk_circuit_num_ops(const brk_circuit* circuit, unsigned* num_ops);
int brk_circuit_num_qubits(const brk_circuit* circuit, unsigned* num_qubits);
/* Reads operation i back: its code, targets, controls and parameters (arrays of
   BRK_MAX_OPERATED_QUBITS unsigned / 8 doubles; any pointer may be NULL). */
int brk_circuit_op(const brk_circuit* circuit, unsigned i, int* op, unsigned* targets, unsigned* num_targets,
                   unsigned* controls, unsigned* num_controls, double* params, unsigned* num_params);
/* Runs the circuit on a state of 2^num_qubits amplitudes.  `permutation`
   (num_qubits entries, logical -> physical, in/out) maps each logical qubit
   to its bit in the state, as ket::mpi::qubit_permutation does
   (ket/include/ket/mpi/qubit_permutation.hpp:40-302); NULL = identity,
   local swaps then update an internal copy only. */
int brk_circuit_run(const brk_circuit* circuit, void* first, int dtype, unsigned* permutation, void* stream);
/* Same, recording HIP events around every operation; op_ms[i] receives the
   device time of operation i (synchronises the stream). */
int brk_circuit_run_timed(const brk_circuit* circuit, void* first, int dtype, unsigned* permutation,
                          float* op_ms, void* stream);

/* ------------------------------------------------------------------------ */
/* Gate blocks: runs of consecutive gates applied per chunk of the state    */
/* held in LDS (2^12 amplitudes for complex<double>, 2^13 for               */
/* complex<float>), each gate with the per-gate kernels' arithmetic, so the */
/* amplitudes are bit-identical to brk_circuit_run's.  One HBM read and     */
/* write per block instead of per gate.  Replaces the gate application of   */
/* bra's BEGIN/END FUSION (bra/src/state.cpp:2193-2217,                     */
/* bra/src/nompi_state.cpp:1282-1357) and of ket's cache-aware gate         */
/* functions (ket/include/ket/gate/detail/gate_bitmasks.hpp:587-2134).      */
/* ------------------------------------------------------------------------ */

typedef struct brk_block_plan brk_block_plan;

/* Groups the circuit's ops (greedily, in order) into blocks whose qubits,
   with the 3 (c128) or 4 (c64) lowest qubits, fit in the local qubits of a
   chunk; a local swap joins a block as the SWAP of its two physical qubits,
   a single op stays a plain step.  `permutation` (may be
   NULL = identity) is the logical->physical map at the start, as for
   brk_circuit_run.  Planning needs no device. */
int brk_block_plan_create(const brk_circuit* circuit, int dtype, const unsigned* permutation,
                          brk_block_plan** out);
/* Call after the work of brk_block_plan_run on its stream has completed. */
int brk_block_plan_destroy(brk_block_plan* plan);
int brk_block_plan_num_steps(const brk_block_plan* plan, unsigned* num_steps);
/* Step i covers ops [first_op, first_op + num_ops) of the circuit; num_local
   is 0 for a plain step, else the block's local (physical) qubits, ascending,
   are written to local_qubits (up to 13 entries). */
int brk_block_plan_step(const brk_block_plan* plan, unsigned i, unsigned* first_op, unsigned* num_ops,
                        unsigned* num_local, unsigned* local_qubits);
int brk_block_plan_final_permutation(const brk_block_plan* plan, unsigned* permutation);
/* Runs the plan on a state of 2^num_qubits amplitudes of the plan's dtype;
   the first run uploads the block programs to the current device. */
int brk_block_plan_run(brk_block_plan* plan, void* first, void* stream);
/* Plans, runs and destroys in one call (synchronises the stream); the
   permutation is updated as by brk_circuit_run. */
int brk_circuit_run_blocked(const brk_circuit* circuit, void* first, int dtype, unsigned* permutation,
                            void* stream);
/* An order of the circuit's ops that packs into fewer blocks: a topological
   order of their dependences (ops sharing a qubit, and every op and a local
   swap, keep their relative order; ops on disjoint qubits commute exactly)
   built block by block -- ops within the block's qubits first, else the op
   whose qubits let the most ops run per added qubit -- for blocks of
   num_local qubits that always hold the line_bits lowest ones.  order[k] is
   the index of the k-th op (circuit-size entries).  The reordered circuit is
   the same operator; its amplitudes round differently.  The adjoint gradient
   sweeps its circuit in this order.  No counterpart in the reference: bra and
   ket apply gates in the given order. */
int brk_circuit_block_order(const brk_circuit* circuit, unsigned num_local, unsigned line_bits, unsigned* order);

/* ------------------------------------------------------------------------ */
/* Precision form: greedy gate fusion into dense 4..6-qubit unitaries,      */
/* applied to a complex64 state as batched complex GEMMs on MFMA (FP32      */
/* multiply, FP32 accumulate).  Semantics follow bra's BEGIN/END FUSION     */
/* (bra/src/state.cpp:2193-2217, bra/src/nompi_state.cpp:1285-1357): a      */
/* fused block is the in-order product of its gates; the block matrix is    */
/* formed in FP64 on the host and rounded to FP32 once.                     */
/* ------------------------------------------------------------------------ */

typedef struct brk_fused_plan brk_fused_plan;

/* Plans `circuit` with blocks of exactly `width` qubits (4, 5 or 6; smaller
   blocks are widened with identity qubits).  Gates move into an earlier block
   only past gates on disjoint qubits (they commute).  Local swaps become
   SWAP gates on physical qubits, so the planned state has the same physical
   layout as brk_circuit_run's.  `permutation` (may be NULL = identity) is the
   logical->physical map at the start.  Needs num_qubits >= width + 5.
   The block matrices are uploaded to the current device. */
int brk_fused_plan_create(const brk_circuit* circuit, unsigned width, const unsigned* permutation,
                          brk_fused_plan** out);
int brk_fused_plan_destroy(brk_fused_plan* plan);
int brk_fused_plan_num_blocks(const brk_fused_plan* plan, unsigned* num_blocks);
/* Block i: its qubits (sorted, `width` entries), the number of gates it holds,
   and (if matrix != NULL) its 2^width x 2^width matrix as (re, im) pairs,
   row-major, local index bit j <-> qubits[j]. */
int brk_fused_plan_block(const brk_fused_plan* plan, unsigned i, unsigned* qubits, unsigned* num_gates,
                         double* matrix);
/* Logical->physical permutation after the circuit (num_qubits entries). */
int brk_fused_plan_final_permutation(const brk_fused_plan* plan, unsigned* permutation);
/* Number of chunk passes of a width-4 or width-5 plan on at least 13 qubits:
   consecutive blocks (the plan's order puts blocks on few qubits next to each
   other; blocks move only past blocks on disjoint qubits) whose qubits, with
   the 4 lowest, fit in 13 local qubits -- at most 8 blocks at width 4, 2 at
   width 5 -- run on one chunk of 2^13 amplitudes held in LDS, one HBM read
   and write per pass instead of per block.  A block there is Gauss's
   three-product complex GEMM in FP32 (the per-block launches use the real
   embedding): the same amplitudes within FP32 rounding.  0 when the plan has
   no passes. */
int brk_fused_plan_num_passes(const brk_fused_plan* plan, unsigned* num_passes);
/* Applies every block in order to a BRK_C64 state of 2^num_qubits amplitudes
   (in chunk passes when the plan has them and tuning fused_chunks is 1). */
int brk_fused_apply(const brk_fused_plan* plan, void* first, int dtype, void* stream);
/* Same, with the device time of launch i in block_ms[i] (synchronises): one
   entry per pass when the plan runs in chunk passes, else one per block;
   block_ms has room for num_blocks entries. */
int brk_fused_apply_timed(const brk_fused_plan* plan, void* first, int dtype, float* block_ms, void* stream);

/* ------------------------------------------------------------------------ */
/* Differentiation form (no reference counterpart): adjoint (reverse-sweep) */
/* gradient of E(theta) = <psi(theta)| H |psi(theta)>, psi = U(theta)|psi0>, */
/* H the open-chain Ising Hamiltonian of brk_expectation_ising.             */
/* ------------------------------------------------------------------------ */

/* Number of parameters of a circuit: one per RX/RY/RZ/EX/EY/EZ/U1 op, in op order. */
int brk_circuit_num_parameters(const brk_circuit* circuit, unsigned* num_parameters);
/* `psi` holds |psi0> on entry; it is overwritten (on return it holds |psi0>
   again up to rounding).  `work` is a second state of the same size and dtype.
   energy receives E; gradient[k] = dE/dp_k for the k-th parameter
   (num_gradient must equal brk_circuit_num_parameters).  FP64 accumulation.
   Circuits with BRK_OP_LOCAL_SWAP are not supported. */
int brk_adjoint_ising_gradient(const brk_circuit* circuit, void* psi, void* work, int dtype, const double* J,
                               const double* h, double* energy, double* gradient, unsigned num_gradient,
                               void* stream);

/* ------------------------------------------------------------------------ */
/* Batching form: many small independent circuits in one persistent launch, */
/* captured once as a HIP graph and replayed.  Each circuit runs on its own */
/* complex64 state from |0...0> and yields its FP64 open-chain Ising energy */
/* sum_i J_i <Z_i Z_i+1> + sum_i h_i <X_i>.  Reference ancestor: bra's      */
/* CIRCUITS mode, which advances several states in turn                     */
/* (bra/src/bra.cpp:315-336, bra/src/interpreter.cpp:411-419).              */
/* ------------------------------------------------------------------------ */

typedef struct brk_batch brk_batch;

#define BRK_BATCH_MAX_QUBITS 24
#define BRK_BATCH_MAX_LOCAL 13

/* planner op kinds (brk_batch_stage_op) */
enum {
  BRK_BATCH_GATE = 0,       /* 2x2 matrix on one qubit                             */
  BRK_BATCH_SIGN = 1,       /* negate amplitudes whose index has every mask bit set */
  BRK_BATCH_MEASURE_X = 2,  /* add h_q <X_q>                                        */
  BRK_BATCH_MEASURE_ZZ = 3  /* add sum_i J_i <Z_i Z_i+1>                            */
};

/* Compiles `num_circuits` circuits (1..BRK_BATCH_MAX_QUBITS qubits each) with
   their Hamiltonians: J[i] has num_qubits-1 entries, h[i] num_qubits (J, h,
   J[i] or h[i] may be NULL = zeros).  Circuits may hold uncontrolled
   single-qubit ops (consecutive ones on one qubit are multiplied into one
   2x2 matrix in FP64) and BRK_OP_Z on any qubits (Z, CZ, CnZ); anything else
   is BRK_ERR_UNSUPPORTED.  Planning needs no device; the first run uploads
   the program to the current device and captures the graph. */
int brk_batch_create(const brk_circuit* const* circuits, unsigned num_circuits, const double* const* J,
                     const double* const* h, brk_batch** out);
int brk_batch_destroy(brk_batch* batch);
/* Plan introspection: a circuit runs as stages; a stage has up to
   BRK_BATCH_MAX_LOCAL local qubits (ascending) and applies its ops, in the
   order listed, to the chunks of the state over those qubits. */
int brk_batch_num_stages(const brk_batch* batch, unsigned circuit, unsigned* num_stages);
int brk_batch_stage(const brk_batch* batch, unsigned circuit, unsigned stage, unsigned* local_qubits,
                    unsigned* num_local, unsigned* num_ops);
/* Op k of a stage: its kind, its qubits as a mask and, for BRK_BATCH_GATE,
   the FP64 matrix (row-major (re, im) pairs; the device uses it rounded to FP32). */
int brk_batch_stage_op(const brk_batch* batch, unsigned circuit, unsigned stage, unsigned k, int* kind,
                       uint64_t* qubit_mask, double* matrix);
/* Runs every circuit; energies[i] receives circuit i's energy (host array of
   num_circuits).  Ordered after the work already on `stream`; returns when done. */
int brk_batch_run(brk_batch* batch, double* energies, void* stream);
/* Same, with the device time of the graph launch and, per circuit (may be
   NULL), the wall-clock times in ms when a workgroup took it from the queue
   and when its energy was written, both relative to the first take. */
int brk_batch_run_timed(brk_batch* batch, double* energies, double* elapsed_ms, double* circuit_start_ms,
                        double* circuit_end_ms, void* stream);

/* ------------------------------------------------------------------------ */
/* bra front end (nompi mode): quantum-assembler (qcx) text -> statements.  */
/* Replaces the parsing half of bra::interpreter::invoke                    */
/* (bra/src/interpreter.cpp:371-872, mnemonics as in docs/bra.md:109-188);  */
/* the gates between two directives form one brk_circuit.  Subset:          */
/*   QUBITS, CIRCUITS 1, INITIAL STATE, BIT ASSIGNMENT (accepted: nompi     */
/*   mode does not permute), comments after '!', case-insensitive mnemonics;*/
/*   gates I IC In, H, NOT, X, Y, Z, Xn/XX.., Yn/YY.., Zn/ZZ.., SWAP, ESWAP  */
/*   ESWAP+, S S+,                                                          */
/*   SZ SZ+, SZZ.. SZn (+: sqrt(Z) strings), T T+, U1 U1+, U2 U2+, U3 U3+,  */
/*   R R+, U U+, +X -X +Y -Y, SX SX+,                                       */
/*   SY SY+, EX EY EZ (+ adjoints, EXX../EXn forms), TOFFOLI, and every one */
/*   of them behind a C..C / Cn control prefix (controls come first);       */
/*   DEPOLARIZING CHANNEL P_X=..,P_Y=..,P_Z=..,SEED=.. (brk_qcx_depolarizing */
/*   and brk_qcx_gate give what a runner needs to insert the noise);        */
/*   DO MEASUREMENT [p], BEGIN MEASUREMENT, DO AMPLITUDES [i..],            */
/*   BEGIN FUSION / END FUSION (gates in between run in order), M, CLEAR,   */
/*   SET, SHORBOX m divisor base, GENERATE EVENTS n seed and EXIT (reading */
/*   stops after the last two, as bra's does).                              */
/* Anything else fails with BRK_ERR_UNSUPPORTED and a message naming the    */
/* line.                                                                    */
/* ------------------------------------------------------------------------ */

typedef struct brk_qcx brk_qcx;

typedef enum brk_qcx_kind {
  BRK_QCX_GATES = 0,        /* a brk_circuit of consecutive gates               */
  BRK_QCX_MEASUREMENT = 1,  /* spin expectation values of every qubit (precision) */
  BRK_QCX_AMPLITUDES = 2,   /* amplitudes (all, or the listed indices)          */
  BRK_QCX_BEGIN_FUSION = 3,
  BRK_QCX_END_FUSION = 4,
  BRK_QCX_PROJECTIVE_MEASUREMENT = 5, /* M q: indices = {q}                  */
  BRK_QCX_CLEAR = 6,                  /* CLEAR q: indices = {q}              */
  BRK_QCX_SET = 7,                    /* SET q: indices = {q}                */
  BRK_QCX_GENERATE_EVENTS = 8,        /* indices = {num_events}, precision = seed
                                         (negative: continue the state's generator) */
  BRK_QCX_EXIT = 9,                   /* sample one basis state and collapse onto it */
  BRK_QCX_SHOR_BOX = 10               /* indices = {num_exponent_qubits, divisor, base} */
} brk_qcx_kind;

int brk_qcx_parse(const char* text, brk_qcx** out);
/* The program's DEPOLARIZING CHANNEL statement (bra/src/interpreter.cpp:431-445,
   :1890-1975): enabled 0/1, probabilities {P_X, P_Y, P_Z} (0 when not given),
   seed (-1 when not given; bra seeds a generator of its own only when seed > 0,
   bra/src/bra.cpp:317-320, else the noise draws come from the state's
   generator). */
int brk_qcx_depolarizing(const brk_qcx* program, int* enabled, double probabilities[3], int64_t* seed);
/* The program's gates inside gate statement `statement` (BRK_QCX_GATES): their
   number, and for gate `gate` the ops [first_op, first_op + num_ops) of the
   statement's circuit it became (0 ops for identity gates) and the qubits after
   which bra::state draws a depolarizing noise (BRK_MAX_OPERATED_QUBITS entries
   at most), in bra's order (bra/src/state.cpp:1000-3330: targets then controls;
   operand order for the gates bra reads on "control" qubits only). */
int brk_qcx_num_gates(const brk_qcx* program, unsigned statement, unsigned* num_gates);
int brk_qcx_gate(const brk_qcx* program, unsigned statement, unsigned gate, unsigned* first_op, unsigned* num_ops,
                 unsigned* noise_qubits, unsigned* num_noise_qubits);
int brk_qcx_destroy(brk_qcx* program);
int brk_qcx_num_qubits(const brk_qcx* program, unsigned* num_qubits);
/* the INITIAL STATE value (0 if absent): the run starts from that basis state */
int brk_qcx_initial_state(const brk_qcx* program, uint64_t* basis_index);
int brk_qcx_num_statements(const brk_qcx* program, unsigned* num_statements);
/* Statement i: kind; for BRK_QCX_GATES its circuit (owned by the program);
   for BRK_QCX_MEASUREMENT the precision; for BRK_QCX_AMPLITUDES the indices
   (num_indices 0 = all); for the other kinds the operands listed with the
   kind.  Output pointers may be NULL. */
int brk_qcx_statement(const brk_qcx* program, unsigned i, int* kind, const brk_circuit** gates, int* precision,
                      unsigned* num_indices, const uint64_t** indices);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* BRAKET_HIP_H */
