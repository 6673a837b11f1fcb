// Plain per-gate form: in-place gate kernels on a 2^n-amplitude state.
//
// Two kernel families cover every gate of the reference's gate set that the
// engine exposes:
//
//  * k_pair  -- visits the 2^(n-k) groups of a gate with k inserted bit
//    positions; each group yields one amplitude pair (i0, i1 = i0 ^ flip),
//    read once and written once.  Modes: Hadamard, swap (X/CNOT/SWAP/local
//    swap), general 2x2 matrix (U2, U3, EX, EY, Y, user matrices).
//  * k_diag  -- visits only the amplitudes whose control bits are all 1 and
//    multiplies each by one of two coefficients chosen by the parity of its
//    target bits (EZ...Z, U1, Z, CZ, CnZ).  Untouched amplitudes are never
//    read, so CZ moves a quarter of the state.
//
// Thread t of block b handles groups b*1024 + t + 256*u, u = 0..3: for every
// u a wavefront touches 64 consecutive groups, i.e. contiguous amplitudes in
// each half of the pair whenever the lowest operated qubit is >= 6.
// Group -> index mapping: brk::expand_index (bit insertion, the same
// decomposition as ket's index masks, gate_bitmasks.hpp:50-130).
#include "brk_internal.hpp"

#include <cmath>
#include <complex>

namespace brk {

enum pair_mode { MODE_H = 0, MODE_SWAP = 1, MODE_MATRIX = 2 };

struct pair_args {
  uint64_t m[BRK_MAX_OPERATED_QUBITS + 1];
  int k;
  uint64_t setmask;   // OR'ed into the expanded index (controls, SWAP's first qubit)
  uint64_t flipmask;  // i1 = i0 ^ flipmask
  uint64_t signmask;  // MODE_MATRIX: off-diagonal entries change sign when parity(i0 & signmask) is odd
  double u[8];        // MODE_MATRIX: m00, m01, m10, m11 as (re, im)
};

constexpr int GATE_THREADS = 256;

// Dynamic LDS is reserved only to cap the number of resident blocks per CU
// (see tuning::pair_lds): on MI355X a far-pair butterfly streams best with
// about 16-32 KiB of loads in flight per CU; more concurrent waves cost
// bandwidth.  The array is never read or written.
extern __shared__ char occupancy_lds[];

// One amplitude pair: (a[i0], a[i0 ^ flipmask]) := M (a[i0], a[i0 ^ flipmask]).
template <typename R, int MODE, int K, bool NT>
__device__ __forceinline__ void pair_update(cpx<R>* __restrict__ s, uint64_t g, const pair_args& a) {
  const uint64_t i0 = expand_index<K>(g, a.m, a.k) | a.setmask;
  const uint64_t i1 = i0 ^ a.flipmask;
  const cpx<R> x0 = load_amp<NT>(s + i0);
  const cpx<R> x1 = load_amp<NT>(s + i1);
  cpx<R> y0, y1;
  if constexpr (MODE == MODE_H) {
    // (a0 + a1)/sqrt2, (a0 - a1)/sqrt2 -- ket/include/ket/gate/hadamard.hpp:66-70
    const R r = (R)0.707106781186547524400844362104849039;
    y0 = cscale(cadd(x0, x1), r);
    y1 = cscale(csub(x0, x1), r);
  } else if constexpr (MODE == MODE_SWAP) {
    y0 = x1;
    y1 = x0;
  } else {
    const cpx<R> m00 = {(R)a.u[0], (R)a.u[1]};
    cpx<R> m01 = {(R)a.u[2], (R)a.u[3]};
    cpx<R> m10 = {(R)a.u[4], (R)a.u[5]};
    const cpx<R> m11 = {(R)a.u[6], (R)a.u[7]};
    if (__popcll(i0 & a.signmask) & 1) {
      m01 = {-m01.re, -m01.im};
      m10 = {-m10.re, -m10.im};
    }
    y0 = cadd(cmul(m00, x0), cmul(m01, x1));
    y1 = cadd(cmul(m10, x0), cmul(m11, x1));
  }
  store_amp<NT>(s + i0, y0);
  store_amp<NT>(s + i1, y1);
}

template <typename R, int MODE, int K, bool NT>
__global__ __launch_bounds__(GATE_THREADS) void k_pair(cpx<R>* __restrict__ s, uint64_t ngroups, pair_args a,
                                                      int remap) {
  if (a.k < 0) occupancy_lds[threadIdx.x] = 0;  // never taken; keeps the reservation
  const uint64_t stride = (uint64_t)gridDim.x * GATE_THREADS;
  for (uint64_t g = xcd_tile(blockIdx.x, gridDim.x, remap) * GATE_THREADS + threadIdx.x; g < ngroups; g += stride)
    pair_update<R, MODE, K, NT>(s, g, a);
}

struct diag_args {
  uint64_t m[BRK_MAX_OPERATED_QUBITS + 1];
  int k;
  uint64_t setmask;  // bits that must be 1 (controls), OR'ed into the expanded index
  uint64_t zmask;    // parity mask selecting d1
  double d0[2], d1[2];
};

template <typename R, int K, bool PARITY, bool NT>
__global__ __launch_bounds__(GATE_THREADS) void k_diag(cpx<R>* __restrict__ s, uint64_t nitems, diag_args a,
                                                      int remap) {
  if (a.k < 0) occupancy_lds[threadIdx.x] = 0;  // never taken; keeps the reservation
  const cpx<R> c0 = {(R)a.d0[0], (R)a.d0[1]};
  const cpx<R> c1 = {(R)a.d1[0], (R)a.d1[1]};
  for (uint64_t g = xcd_tile(blockIdx.x, gridDim.x, remap) * GATE_THREADS + threadIdx.x; g < nitems;
       g += (uint64_t)gridDim.x * GATE_THREADS) {
    const uint64_t i = expand_index<K>(g, a.m, a.k) | a.setmask;
    cpx<R> c = c0;
    if constexpr (PARITY) {
      if (__popcll(i & a.zmask) & 1) c = c1;
    }
    store_amp<NT>(s + i, cmul(load_amp<NT>(s + i), c));
  }
}

// ---------------------------------------------------------------------------
// Launch helpers
// ---------------------------------------------------------------------------

// One tile of GATE_THREADS items per block; above MAX_GATE_BLOCKS tiles the
// blocks loop (grid stride), keeping gridDim.x * blockDim.x < 2^32.
constexpr uint64_t MAX_GATE_BLOCKS = 1ull << 22;

static int grid_blocks(uint64_t items, unsigned* blocks) {
  uint64_t b = (items + GATE_THREADS - 1) / GATE_THREADS;
  if (b == 0) b = 1;
  if (b > MAX_GATE_BLOCKS) b = MAX_GATE_BLOCKS;
  *blocks = (unsigned)b;
  return BRK_OK;
}

template <typename R, int MODE, bool NT>
static void launch_pair_k(cpx<R>* s, uint64_t ngroups, const pair_args& a, unsigned blocks, hipStream_t st) {
  const tuning& t = get_tuning();
  const dim3 g(blocks), b(GATE_THREADS);
  const size_t lds = (size_t)t.pair_lds;
  switch (a.k) {
    case 1: hipLaunchKernelGGL((k_pair<R, MODE, 1, NT>), g, b, lds, st, s, ngroups, a, t.pair_remap); break;
    case 2: hipLaunchKernelGGL((k_pair<R, MODE, 2, NT>), g, b, lds, st, s, ngroups, a, t.pair_remap); break;
    case 3: hipLaunchKernelGGL((k_pair<R, MODE, 3, NT>), g, b, lds, st, s, ngroups, a, t.pair_remap); break;
    default: hipLaunchKernelGGL((k_pair<R, MODE, -1, NT>), g, b, lds, st, s, ngroups, a, t.pair_remap); break;
  }
}

template <typename R, bool NT>
static void launch_pair_mode(int mode, cpx<R>* s, uint64_t ngroups, const pair_args& a, unsigned blocks,
                             hipStream_t st) {
  if (mode == MODE_H) launch_pair_k<R, MODE_H, NT>(s, ngroups, a, blocks, st);
  else if (mode == MODE_SWAP) launch_pair_k<R, MODE_SWAP, NT>(s, ngroups, a, blocks, st);
  else launch_pair_k<R, MODE_MATRIX, NT>(s, ngroups, a, blocks, st);
}

static int launch_pair(int dtype, int mode, void* first, uint64_t ngroups, const pair_args& a, void* stream) {
  unsigned blocks;
  if (int st = grid_blocks(ngroups, &blocks)) return st;
  hipStream_t st = as_stream(stream);
  const uint64_t state_bytes = (ngroups << a.k) * amp_bytes(dtype);
  const bool nt = state_bytes >= get_tuning().nt_min_bytes;
  if (dtype == BRK_C128) {
    auto* s = (cpx<double>*)first;
    if (nt) launch_pair_mode<double, true>(mode, s, ngroups, a, blocks, st);
    else launch_pair_mode<double, false>(mode, s, ngroups, a, blocks, st);
  } else {
    auto* s = (cpx<float>*)first;
    if (nt) launch_pair_mode<float, true>(mode, s, ngroups, a, blocks, st);
    else launch_pair_mode<float, false>(mode, s, ngroups, a, blocks, st);
  }
  return check_launch("pair gate");
}

template <typename R, bool PARITY, bool NT>
static void launch_diag_k(cpx<R>* s, uint64_t n, const diag_args& a, unsigned blocks, hipStream_t st) {
  const tuning& t = get_tuning();
  const dim3 g(blocks), b(GATE_THREADS);
  const size_t lds = (size_t)t.diag_lds;
  switch (a.k) {
    case 0: hipLaunchKernelGGL((k_diag<R, 0, PARITY, NT>), g, b, lds, st, s, n, a, t.diag_remap); break;
    case 1: hipLaunchKernelGGL((k_diag<R, 1, PARITY, NT>), g, b, lds, st, s, n, a, t.diag_remap); break;
    case 2: hipLaunchKernelGGL((k_diag<R, 2, PARITY, NT>), g, b, lds, st, s, n, a, t.diag_remap); break;
    case 3: hipLaunchKernelGGL((k_diag<R, 3, PARITY, NT>), g, b, lds, st, s, n, a, t.diag_remap); break;
    default: hipLaunchKernelGGL((k_diag<R, -1, PARITY, NT>), g, b, lds, st, s, n, a, t.diag_remap); break;
  }
}

template <typename R, bool NT>
static void launch_diag_p(bool parity, cpx<R>* s, uint64_t n, const diag_args& a, unsigned blocks,
                          hipStream_t st) {
  if (parity) launch_diag_k<R, true, NT>(s, n, a, blocks, st);
  else launch_diag_k<R, false, NT>(s, n, a, blocks, st);
}

static int launch_diag(int dtype, void* first, uint64_t nitems, const diag_args& a, void* stream) {
  unsigned blocks;
  if (int st = grid_blocks(nitems, &blocks)) return st;
  hipStream_t st = as_stream(stream);
  const bool parity = a.zmask != 0;
  const uint64_t state_bytes = (nitems << a.k) * amp_bytes(dtype);
  const bool nt = state_bytes >= get_tuning().nt_min_bytes;
  if (dtype == BRK_C128) {
    if (nt) launch_diag_p<double, true>(parity, (cpx<double>*)first, nitems, a, blocks, st);
    else launch_diag_p<double, false>(parity, (cpx<double>*)first, nitems, a, blocks, st);
  } else {
    if (nt) launch_diag_p<float, true>(parity, (cpx<float>*)first, nitems, a, blocks, st);
    else launch_diag_p<float, false>(parity, (cpx<float>*)first, nitems, a, blocks, st);
  }
  return check_launch("diagonal gate");
}

static pair_args make_pair_args(const unsigned* positions, int k, uint64_t setmask, uint64_t flipmask) {
  pair_args a{};
  expand_masks e = make_expand_masks(positions, k);
  std::copy(e.m, e.m + BRK_MAX_OPERATED_QUBITS + 1, a.m);
  a.k = k;
  a.setmask = setmask;
  a.flipmask = flipmask;
  a.signmask = 0;
  return a;
}

static diag_args make_diag_args(const unsigned* positions, int k, uint64_t setmask, uint64_t zmask,
                                std::complex<double> d0, std::complex<double> d1) {
  diag_args a{};
  expand_masks e = make_expand_masks(positions, k);
  std::copy(e.m, e.m + BRK_MAX_OPERATED_QUBITS + 1, a.m);
  a.k = k;
  a.setmask = setmask;
  a.zmask = zmask;
  a.d0[0] = d0.real();
  a.d0[1] = d0.imag();
  a.d1[0] = d1.real();
  a.d1[1] = d1.imag();
  return a;
}

// Single target + controls, general 2x2 matrix or Hadamard mode.
static int single_target(void* first, uint64_t size, int dtype, unsigned target, const unsigned* controls,
                         unsigned nc, int mode, const double* u, void* stream) {
  unsigned nq;
  if (int st = check_state(first, size, dtype, &nq)) return st;
  unsigned ops[BRK_MAX_OPERATED_QUBITS];
  unsigned cnt = 0;
  if (int st = collect_qubits(&target, 1, nq, ops, &cnt, "target")) return st;
  if (int st = collect_qubits(controls, nc, nq, ops, &cnt, "control")) return st;
  pair_args a = make_pair_args(ops, (int)cnt, mask_of(controls, nc), 1ull << target);
  if (u) std::copy(u, u + 8, a.u);
  return launch_pair(dtype, mode, first, size >> cnt, a, stream);
}

// Targets flipped together (X...X strings) + controls; pairs (i0, i0 ^ T) with the
// lowest target bit of i0 equal to 0.
static int multi_target_flip(void* first, uint64_t size, int dtype, const unsigned* targets, unsigned nt,
                             const unsigned* controls, unsigned nc, int mode, const double* u,
                             uint64_t signmask_is_targets, void* stream) {
  unsigned nq;
  if (int st = check_state(first, size, dtype, &nq)) return st;
  if (nt == 0) return fail(BRK_ERR_INVALID_ARGUMENT, "no target qubits");
  unsigned ops[BRK_MAX_OPERATED_QUBITS];
  unsigned cnt = 0;
  if (int st = collect_qubits(targets, nt, nq, ops, &cnt, "target")) return st;
  if (int st = collect_qubits(controls, nc, nq, ops, &cnt, "control")) return st;
  const uint64_t tmask = mask_of(targets, nt);
  unsigned lowest = targets[0];
  for (unsigned i = 1; i < nt; ++i) lowest = std::min(lowest, targets[i]);
  unsigned pos[BRK_MAX_OPERATED_QUBITS];
  unsigned np = 0;
  pos[np++] = lowest;
  for (unsigned i = 0; i < nc; ++i) pos[np++] = controls[i];
  pair_args a = make_pair_args(pos, (int)np, mask_of(controls, nc), tmask);
  if (u) std::copy(u, u + 8, a.u);
  if (signmask_is_targets) a.signmask = tmask;
  return launch_pair(dtype, mode, first, size >> np, a, stream);
}

// (-i)^m as a complex number.
static std::complex<double> minus_i_pow(unsigned m) {
  switch (m & 3u) {
    case 0: return {1.0, 0.0};
    case 1: return {0.0, -1.0};
    case 2: return {-1.0, 0.0};
    default: return {0.0, 1.0};
  }
}

static void put(double* u, int slot, std::complex<double> v) {
  u[2 * slot] = v.real();
  u[2 * slot + 1] = v.imag();
}

}  // namespace brk

using namespace brk;

extern "C" {

int brk_hadamard(void* first, uint64_t size, int dtype, unsigned target, const unsigned* controls,
                 unsigned num_controls, void* stream) {
  return single_target(first, size, dtype, target, controls, num_controls, MODE_H, nullptr, stream);
}

int brk_gate_1q(void* first, uint64_t size, int dtype, unsigned target, const double matrix[8],
                const unsigned* controls, unsigned num_controls, void* stream) {
  if (!matrix) return fail(BRK_ERR_INVALID_ARGUMENT, "matrix is null");
  return single_target(first, size, dtype, target, controls, num_controls, MODE_MATRIX, matrix, stream);
}

int brk_pauli_x(void* first, uint64_t size, int dtype, const unsigned* targets, unsigned num_targets,
                const unsigned* controls, unsigned num_controls, void* stream) {
  return multi_target_flip(first, size, dtype, targets, num_targets, controls, num_controls, MODE_SWAP, nullptr,
                           0, stream);
}

int brk_controlled_not(void* first, uint64_t size, int dtype, unsigned target, const unsigned* controls,
                       unsigned num_controls, void* stream) {
  return multi_target_flip(first, size, dtype, &target, 1, controls, num_controls, MODE_SWAP, nullptr, 0, stream);
}

int brk_pauli_y(void* first, uint64_t size, int dtype, const unsigned* targets, unsigned num_targets,
                const unsigned* controls, unsigned num_controls, void* stream) {
  // Y...Y: (Y^m a)[i] = (-i)^m (-1)^{popcount(i & T)} a[i ^ T]; for one target
  // Y = [[0, -i], [i, 0]] (docs/bra.md "Y t").
  const std::complex<double> ph = minus_i_pow(num_targets);
  const double sgn_m = (num_targets & 1u) ? -1.0 : 1.0;
  double u[8];
  put(u, 0, {0.0, 0.0});
  put(u, 1, ph);
  put(u, 2, ph * sgn_m);
  put(u, 3, {0.0, 0.0});
  return multi_target_flip(first, size, dtype, targets, num_targets, controls, num_controls, MODE_MATRIX, u, 1,
                           stream);
}

int brk_pauli_z(void* first, uint64_t size, int dtype, const unsigned* qubits, unsigned num_qubits, void* stream) {
  unsigned nq;
  if (int st = check_state(first, size, dtype, &nq)) return st;
  if (num_qubits == 0) return fail(BRK_ERR_INVALID_ARGUMENT, "no qubits");
  unsigned ops[BRK_MAX_OPERATED_QUBITS];
  unsigned cnt = 0;
  if (int st = collect_qubits(qubits, num_qubits, nq, ops, &cnt, "control")) return st;
  diag_args a = make_diag_args(ops, (int)cnt, mask_of(qubits, num_qubits), 0, {-1.0, 0.0}, {-1.0, 0.0});
  return launch_diag(dtype, first, size >> cnt, a, stream);
}

int brk_phase_shift_coeff(void* first, uint64_t size, int dtype, const double coeff[2], const unsigned* qubits,
                          unsigned num_qubits, void* stream) {
  unsigned nq;
  if (int st = check_state(first, size, dtype, &nq)) return st;
  if (!coeff) return fail(BRK_ERR_INVALID_ARGUMENT, "coeff is null");
  unsigned ops[BRK_MAX_OPERATED_QUBITS];
  unsigned cnt = 0;
  if (int st = collect_qubits(qubits, num_qubits, nq, ops, &cnt, "control")) return st;
  std::complex<double> c{coeff[0], coeff[1]};
  diag_args a = make_diag_args(ops, (int)cnt, mask_of(qubits, num_qubits), 0, c, c);
  return launch_diag(dtype, first, size >> cnt, a, stream);
}

int brk_phase_shift(void* first, uint64_t size, int dtype, double phase, const unsigned* qubits,
                    unsigned num_qubits, void* stream) {
  // exp(i phase) computed once on the host with std::exp, as ket::utility::exp_i
  // does (ket/include/ket/utility/exp_i.hpp:24-29).
  std::complex<double> c = std::exp(std::complex<double>{0.0, phase});
  double cc[2] = {c.real(), c.imag()};
  return brk_phase_shift_coeff(first, size, dtype, cc, qubits, num_qubits, stream);
}

int brk_phase_shift2(void* first, uint64_t size, int dtype, double phase1, double phase2, unsigned target,
                     const unsigned* controls, unsigned num_controls, void* stream) {
  // U2(p1, p2) = 1/sqrt2 [[1, -e^{i p2}], [e^{i p1}, e^{i(p1+p2)}]]
  // (ket/include/ket/gate/phase_shift.hpp:649-652)
  const double r = 0.707106781186547524400844362104849039;
  std::complex<double> e1 = std::exp(std::complex<double>{0.0, phase1});
  std::complex<double> e2 = std::exp(std::complex<double>{0.0, phase2});
  double u[8];
  put(u, 0, {r, 0.0});
  put(u, 1, -e2 * r);
  put(u, 2, e1 * r);
  put(u, 3, (e2 * e1) * r);
  return single_target(first, size, dtype, target, controls, num_controls, MODE_MATRIX, u, stream);
}

int brk_phase_shift3(void* first, uint64_t size, int dtype, double phase1, double phase2, double phase3,
                     unsigned target, const unsigned* controls, unsigned num_controls, void* stream) {
  // U3(t, p, l) = [[c, -e^{il} s], [e^{ip} s, e^{i(p+l)} c]], c = cos(t/2), s = sin(t/2)
  // (ket/include/ket/gate/phase_shift.hpp:1195-1198, coefficients :1219-1229)
  const double sine = std::sin(0.5 * phase1);
  const double cosine = std::cos(0.5 * phase1);
  std::complex<double> e2 = std::exp(std::complex<double>{0.0, phase2});
  std::complex<double> e3 = std::exp(std::complex<double>{0.0, phase3});
  double u[8];
  put(u, 0, {cosine, 0.0});
  put(u, 1, -(sine * e3));
  put(u, 2, sine * e2);
  put(u, 3, (cosine * e3) * e2);
  return single_target(first, size, dtype, target, controls, num_controls, MODE_MATRIX, u, stream);
}

int brk_exponential_pauli_x(void* first, uint64_t size, int dtype, double phase, const unsigned* targets,
                            unsigned num_targets, const unsigned* controls, unsigned num_controls, void* stream) {
  // exp(i phase X..X) = cos I + i sin X..X; pairs (i, i ^ T).
  std::complex<double> e = std::exp(std::complex<double>{0.0, phase});
  const double c = e.real(), s = e.imag();
  double u[8];
  put(u, 0, {c, 0.0});
  put(u, 1, {0.0, s});
  put(u, 2, {0.0, s});
  put(u, 3, {c, 0.0});
  return multi_target_flip(first, size, dtype, targets, num_targets, controls, num_controls, MODE_MATRIX, u, 0,
                           stream);
}

int brk_exponential_pauli_y(void* first, uint64_t size, int dtype, double phase, const unsigned* targets,
                            unsigned num_targets, const unsigned* controls, unsigned num_controls, void* stream) {
  // exp(i phase Y..Y) = cos I + i sin Y..Y, with (Y^m a)[i] = (-i)^m (-1)^{popcount(i & T)} a[i ^ T].
  // For one target: a0' = c a0 + s a1, a1' = -s a0 + c a1
  // (ket/include/ket/gate/exponential_pauli_y.hpp:36-38, :84-87).
  if (!targets && num_targets) return fail(BRK_ERR_INVALID_ARGUMENT, "targets is null");
  std::complex<double> e = std::exp(std::complex<double>{0.0, phase});
  const double c = e.real(), s = e.imag();
  const std::complex<double> ph = minus_i_pow(num_targets);
  const double sgn_m = (num_targets & 1u) ? -1.0 : 1.0;
  double u[8];
  put(u, 0, {c, 0.0});
  put(u, 1, std::complex<double>{0.0, s} * ph);
  put(u, 2, std::complex<double>{0.0, s} * ph * sgn_m);
  put(u, 3, {c, 0.0});
  return multi_target_flip(first, size, dtype, targets, num_targets, controls, num_controls, MODE_MATRIX, u, 1,
                           stream);
}

int brk_exponential_pauli_z(void* first, uint64_t size, int dtype, double phase, const unsigned* targets,
                            unsigned num_targets, const unsigned* controls, unsigned num_controls, void* stream) {
  // exp(i phase Z..Z): e^{i phase} on even parity of the target bits, e^{-i phase} on odd
  // (ket/include/ket/gate/exponential_pauli_z.hpp:36-38, :73-74).
  unsigned nq;
  if (int st = check_state(first, size, dtype, &nq)) return st;
  if (num_targets == 0) return fail(BRK_ERR_INVALID_ARGUMENT, "no target qubits");
  unsigned ops[BRK_MAX_OPERATED_QUBITS];
  unsigned cnt = 0;
  if (int st = collect_qubits(targets, num_targets, nq, ops, &cnt, "target")) return st;
  unsigned nt = cnt;
  if (int st = collect_qubits(controls, num_controls, nq, ops, &cnt, "control")) return st;
  std::complex<double> e = std::exp(std::complex<double>{0.0, phase});
  diag_args a = make_diag_args(ops + nt, (int)(cnt - nt), mask_of(controls, num_controls),
                               mask_of(targets, num_targets), e, std::conj(e));
  return launch_diag(dtype, first, size >> (cnt - nt), a, stream);
}

int brk_swap(void* first, uint64_t size, int dtype, unsigned qubit1, unsigned qubit2, const unsigned* controls,
             unsigned num_controls, void* stream) {
  unsigned nq;
  if (int st = check_state(first, size, dtype, &nq)) return st;
  unsigned ops[BRK_MAX_OPERATED_QUBITS];
  unsigned cnt = 0;
  if (int st = collect_qubits(&qubit1, 1, nq, ops, &cnt, "target")) return st;
  if (int st = collect_qubits(&qubit2, 1, nq, ops, &cnt, "target")) return st;
  if (int st = collect_qubits(controls, num_controls, nq, ops, &cnt, "control")) return st;
  // pairs ...1(q1)...0(q2)... <-> ...0(q1)...1(q2)... (ket/include/ket/gate/swap.hpp:30-32)
  pair_args a = make_pair_args(ops, (int)cnt, mask_of(controls, num_controls) | (1ull << qubit1),
                               (1ull << qubit1) | (1ull << qubit2));
  return launch_pair(dtype, MODE_SWAP, first, size >> cnt, a, stream);
}

int brk_swap_permutated_local_qubits(void* first, uint64_t size, int dtype, unsigned permutated_qubit1,
                                     unsigned permutated_qubit2, void* stream) {
  // With one data block the run-wise swap_ranges of the reference
  // (swap_permutated_local_qubits.hpp:64-77) exchanges exactly the amplitudes a SWAP
  // of the two physical qubits exchanges.
  return brk_swap(first, size, dtype, permutated_qubit1, permutated_qubit2, nullptr, 0, stream);
}

}  // extern "C"
