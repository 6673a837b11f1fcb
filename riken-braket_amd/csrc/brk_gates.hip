// Plain per-gate form: in-place gate kernels on a 2^n-amplitude state.
//
// Kernel families (all one work item per thread, 256-thread blocks, every
// state access a whole amplitude -- one 16-byte access for c128):
//
//  * k_pair  -- the 2^(n-k) groups of a gate with k inserted bit positions
//    each yield one amplitude pair (i0, i1 = i0 ^ flip), read once and written
//    once.  Modes: Hadamard, swap (X/CNOT/SWAP/local swap), 2x2 matrix (U2,
//    U3, EX, EY, Y, user matrices).  For targets >= 6 a wavefront touches 64
//    consecutive amplitudes on each side of the pair.
//  * k_lane  -- single-target gates on a low qubit t < 6: lane l of a
//    wavefront owns amplitude base + l and gets its partner base + (l ^ 2^t)
//    with a cross-lane shuffle, so every load and store of a wavefront is one
//    contiguous KiB (no partial cache lines).
//  * k_lane_swap -- SWAP / local swap of a low qubit (< 6) with a high one:
//    the two halves of the high qubit stream contiguously and the low-qubit
//    exchange happens across lanes.
//  * k_diag  -- multiplies the amplitudes of a subspace by a coefficient
//    chosen by the number of set target bits: one coefficient (U1, Z, CZ),
//    two by its parity (EZ...Z, Z...Z strings), or four by its value mod 4
//    (sqrt(Z)...sqrt(Z) strings).
//
// Controls: a control bit at or above the line boundary (the bits addressing
// amplitudes inside one 64-byte line) is removed from the iteration space, so
// untouched amplitudes are never read.  A control bit below it stays in the
// iteration and is tested per amplitude ("cond"), so that every line the
// kernel writes is written whole -- a partial-line write costs a
// read-modify-write in the memory system.
//
// Group -> index mapping: brk::expand_index (bit insertion, the same
// decomposition as ket's index masks, gate_bitmasks.hpp:50-130).
#include "brk_internal.hpp"
#include "brk_block.hpp"

#include <cmath>
#include <complex>

namespace brk {

enum pair_mode { MODE_H = 0, MODE_SWAP = 1, MODE_MATRIX = 2, MODE_ESWAP = 3 };

struct gate_args {
  uint64_t m[BRK_MAX_OPERATED_QUBITS + 1];  // expand masks
  int k;                                    // inserted bit positions
  uint64_t setmask;   // OR'ed into the expanded index (controls above the line boundary, SWAP's first qubit)
  uint64_t flipmask;  // pair kernels: i1 = i0 ^ flipmask
  uint64_t signmask;  // matrix mode: off-diagonal entries change sign when parity(i0 & signmask) is odd
  uint64_t condmask;  // the op acts only where all these bits are 1 (controls kept in the iteration)
  uint64_t zmask;     // diagonal kernel: target mask whose set-bit count selects the coefficient
  uint64_t swapmask;  // k_lane swap mode on two low qubits: acts where parity(i & swapmask) is odd
  int lane_bit;       // lane kernels: the low qubit
  double u[8];        // matrix (m00, m01, m10, m11) or diagonal coefficients (d0, d1[, d2, d3]) as (re, im)
};

// How k_diag picks the coefficient from popcount(i & zmask).
enum diag_select { DIAG_ONE = 0, DIAG_PARITY = 1, DIAG_MOD4 = 2 };

constexpr int GATE_THREADS = 256;
constexpr unsigned LANE_BITS = 6;  // log2(wavefront size)

// Dynamic LDS is reserved only to cap the number of resident blocks per CU
// (tuning::pair_lds): on MI355X a far-pair butterfly streams best with about
// 16-32 KiB of loads in flight per CU; more concurrent waves cost bandwidth.
// The array is never read or written.
extern __shared__ char occupancy_lds[];

template <typename R>
__device__ __forceinline__ cpx<R> shfl_xor_amp(cpx<R> v, int mask) {
  return cpx<R>{__shfl_xor(v.re, mask, 64), __shfl_xor(v.im, mask, 64)};
}

// Component-wise select (a struct-valued ?: can make the compiler go through scratch memory).
template <typename R>
__device__ __forceinline__ cpx<R> csel(bool c, cpx<R> a, cpx<R> b) {
  return cpx<R>{c ? a.re : b.re, c ? a.im : b.im};
}

// u: the gate's 8 coefficients (any indexable type: kernel argument or program memory)
template <typename R, typename U>
__device__ __forceinline__ void load_matrix(const U& u, cpx<R>& m00, cpx<R>& m01, cpx<R>& m10, cpx<R>& m11) {
  m00 = {(R)u[0], (R)u[1]};
  m01 = {(R)u[2], (R)u[3]};
  m10 = {(R)u[4], (R)u[5]};
  m11 = {(R)u[6], (R)u[7]};
}

// The arithmetic of one amplitude pair (i0, i1), shared by every kernel that
// applies a pair gate: the plain families and the LDS-resident gate blocks.
template <typename R, int MODE, typename U>
__device__ __forceinline__ void pair_math(const U& u, uint64_t signmask, uint64_t i0, cpx<R> x0, cpx<R> x1,
                                          cpx<R>& y0, cpx<R>& y1) {
  if constexpr (MODE == MODE_H) {
    // (a0 + a1)/sqrt2, (a0 - a1)/sqrt2 -- ket/include/ket/gate/hadamard.hpp:66-70
    const R r = (R)0.707106781186547524400844362104849039;
    y0 = cscale(cadd(x0, x1), r);
    y1 = cscale(csub(x0, x1), r);
  } else if constexpr (MODE == MODE_SWAP) {
    y0 = x1;
    y1 = x0;
  } else if constexpr (MODE == MODE_ESWAP) {
    // exp(i theta SWAP) (ket/include/ket/gate/exponential_swap.hpp:31-90): the pair
    // (i0, i0 ^ flipmask) is (00, 11) when the higher target bit of i0 is 0 -- both
    // get e^{i theta} -- and (10, 01) otherwise: a' = cos a + i sin b, b' = cos b + i sin a.
    const cpx<R> e = {(R)u[0], (R)u[1]};
    const R c = (R)u[2], sn = (R)u[3];
    const bool mix = (i0 & signmask) != 0;
    const cpx<R> m0 = {x0.re * c - x1.im * sn, x0.im * c + x1.re * sn};
    const cpx<R> m1 = {x1.re * c - x0.im * sn, x1.im * c + x0.re * sn};
    y0 = csel(mix, m0, cmul(x0, e));
    y1 = csel(mix, m1, cmul(x1, e));
  } else {
    cpx<R> m00, m01, m10, m11;
    load_matrix(u, m00, m01, m10, m11);
    if (signmask != 0) {  // uniform: most gates have no sign mask
      const bool flip = (__popcll(i0 & signmask) & 1) != 0;
      m01 = csel(flip, cpx<R>{-m01.re, -m01.im}, m01);
      m10 = csel(flip, cpx<R>{-m10.re, -m10.im}, m10);
    }
    // uniform branches on the coefficients' exact zeros
    if (u[1] == 0.0 && u[3] == 0.0 && u[5] == 0.0 && u[7] == 0.0) {
      // a real 2x2 (RY, EY): real coefficients times amplitudes, the form of
      // c a0 + s a1 in exponential_pauli_y.hpp:36-38, :84-87
      y0 = {m00.re * x0.re + m01.re * x1.re, m00.re * x0.im + m01.re * x1.im};
      y1 = {m10.re * x0.re + m11.re * x1.re, m10.re * x0.im + m11.re * x1.im};
    } else if (u[1] == 0.0 && u[2] == 0.0 && u[4] == 0.0 && u[7] == 0.0) {
      // [[r, i s], [i s', r']] (RX, EX): cos a0 + i sin a1 (exponential_pauli_x.hpp:37-38)
      y0 = {m00.re * x0.re - m01.im * x1.im, m00.re * x0.im + m01.im * x1.re};
      y1 = {m11.re * x1.re - m10.im * x0.im, m11.re * x1.im + m10.im * x0.re};
    } else {
      y0 = cadd(cmul(m00, x0), cmul(m01, x1));
      y1 = cadd(cmul(m10, x0), cmul(m11, x1));
    }
  }
}

template <typename R, int MODE, int K, bool NT, bool COND>
__global__ __launch_bounds__(GATE_THREADS) void k_pair(cpx<R>* __restrict__ s, uint64_t ngroups, gate_args a,
                                                      int remap) {
  if (a.k < 0) occupancy_lds[threadIdx.x] = 0;  // never taken; keeps the reservation
  const uint64_t stride = (uint64_t)gridDim.x * GATE_THREADS;
  for (uint64_t g = xcd_tile(blockIdx.x, gridDim.x, remap) * GATE_THREADS + threadIdx.x; g < ngroups;
       g += stride) {
    const uint64_t i0 = expand_index<K>(g, a.m, a.k) | a.setmask;
    const uint64_t i1 = i0 ^ a.flipmask;
    const cpx<R> x0 = load_amp<NT>(s + i0);
    const cpx<R> x1 = load_amp<NT>(s + i1);
    cpx<R> y0, y1;
    pair_math<R, MODE>(a.u, a.signmask, i0, x0, x1, y0, y1);
    if (COND) {
      const bool on = (i0 & a.condmask) == a.condmask;
      y0 = csel(on, y0, x0);
      y1 = csel(on, y1, x1);
    }
    store_amp<NT>(s + i0, y0);
    store_amp<NT>(s + i1, y1);
  }
}

// Single target t = a.lane_bit < 6.  Item g -> amplitude i; whole wavefronts
// are in or out of range (the item count is a multiple of 64), so every lane
// takes part in the shuffle.
template <typename R, int MODE, int K, bool NT, bool COND, bool SWAP2>
__global__ __launch_bounds__(GATE_THREADS) void k_lane(cpx<R>* __restrict__ s, uint64_t nitems, gate_args a,
                                                      int remap) {
  if (a.k < 0) occupancy_lds[threadIdx.x] = 0;
  const uint64_t stride = (uint64_t)gridDim.x * GATE_THREADS;
  const int lane_mask = SWAP2 ? (int)a.swapmask : (1 << a.lane_bit);
  for (uint64_t g = xcd_tile(blockIdx.x, gridDim.x, remap) * GATE_THREADS + threadIdx.x; g < nitems;
       g += stride) {
    const uint64_t i = expand_index<K>(g, a.m, a.k) | a.setmask;
    const cpx<R> x = load_amp<NT>(s + i);
    const cpx<R> p = shfl_xor_amp(x, lane_mask);
    cpx<R> y;
    if constexpr (SWAP2) {
      y = csel((__popcll(i & a.swapmask) & 1) != 0, p, x);  // SWAP of two low qubits
    } else {
      const bool one = (i >> a.lane_bit) & 1ull;
      cpx<R> y0, y1;
      // pair (a0, a1) = (this, partner) on the 0 side, (partner, this) on the 1 side
      pair_math<R, MODE>(a.u, a.signmask, i & ~(1ull << a.lane_bit), csel(one, p, x), csel(one, x, p), y0, y1);
      y = csel(one, y1, y0);
    }
    if (COND) y = csel((i & a.condmask) == a.condmask, y, x);
    store_amp<NT>(s + i, y);
  }
}

// SWAP of the low qubit a.lane_bit (< 6) with a high qubit h (>= 6): pairs
// i0 (h = 0) and i1 = i0 | 2^h, exchange across lanes l <-> l ^ 2^lane_bit.
template <typename R, int K, bool NT, bool COND>
__global__ __launch_bounds__(GATE_THREADS) void k_lane_swap(cpx<R>* __restrict__ s, uint64_t ngroups, gate_args a,
                                                           int remap) {
  if (a.k < 0) occupancy_lds[threadIdx.x] = 0;
  const uint64_t stride = (uint64_t)gridDim.x * GATE_THREADS;
  const int lane_mask = 1 << a.lane_bit;
  for (uint64_t g = xcd_tile(blockIdx.x, gridDim.x, remap) * GATE_THREADS + threadIdx.x; g < ngroups;
       g += stride) {
    const uint64_t i0 = expand_index<K>(g, a.m, a.k) | a.setmask;
    const uint64_t i1 = i0 ^ a.flipmask;
    const cpx<R> x0 = load_amp<NT>(s + i0);
    const cpx<R> x1 = load_amp<NT>(s + i1);
    const cpx<R> p1 = shfl_xor_amp(x1, lane_mask);  // value at i1 ^ 2^lane_bit
    const cpx<R> p0 = shfl_xor_amp(x0, lane_mask);  // value at i0 ^ 2^lane_bit
    const bool low_one = (i0 >> a.lane_bit) & 1ull;
    cpx<R> y0 = csel(low_one, p1, x0);  // (h=0, l=1) <- (h=1, l=0)
    cpx<R> y1 = csel(low_one, x1, p0);  // (h=1, l=0) <- (h=0, l=1)
    if (COND) {
      const bool on = (i0 & a.condmask) == a.condmask;
      y0 = csel(on, y0, x0);
      y1 = csel(on, y1, x1);
    }
    store_amp<NT>(s + i0, y0);
    store_amp<NT>(s + i1, y1);
  }
}

template <typename R, int K, int SELECT, bool NT, bool COND>
__global__ __launch_bounds__(GATE_THREADS) void k_diag(cpx<R>* __restrict__ s, uint64_t nitems, gate_args a,
                                                      int remap) {
  if (a.k < 0) occupancy_lds[threadIdx.x] = 0;
  const cpx<R> c0 = {(R)a.u[0], (R)a.u[1]};
  const cpx<R> c1 = {(R)a.u[2], (R)a.u[3]};
  const cpx<R> c2 = {(R)a.u[4], (R)a.u[5]};
  const cpx<R> c3 = {(R)a.u[6], (R)a.u[7]};
  // Z and CZ: both coefficients -1, an exact negation
  const bool negate = SELECT != DIAG_MOD4 && a.u[0] == -1.0 && a.u[1] == 0.0 && a.u[2] == -1.0 && a.u[3] == 0.0;
  const uint64_t stride = (uint64_t)gridDim.x * GATE_THREADS;
  for (uint64_t g = xcd_tile(blockIdx.x, gridDim.x, remap) * GATE_THREADS + threadIdx.x; g < nitems;
       g += stride) {
    const uint64_t i = expand_index<K>(g, a.m, a.k) | a.setmask;
    const cpx<R> x = load_amp<NT>(s + i);
    cpx<R> c = c0;
    if constexpr (SELECT == DIAG_PARITY) c = csel((__popcll(i & a.zmask) & 1) != 0, c1, c0);
    if constexpr (SELECT == DIAG_MOD4) {
      const unsigned n = (unsigned)__popcll(i & a.zmask);
      c = csel((n & 2u) != 0, csel((n & 1u) != 0, c3, c2), csel((n & 1u) != 0, c1, c0));
    }
    cpx<R> y = negate ? cneg(x) : cmul(x, c);
    if (COND) y = csel((i & a.condmask) == a.condmask, y, x);
    store_amp<NT>(s + i, y);
  }
}

// ---------------------------------------------------------------------------
// Gate blocks: one workgroup per chunk of 2^L amplitudes (the local bits of a
// block_program), held in LDS while the block's gates run in order with the
// arithmetic above; one HBM read and one write per amplitude per block.
// ---------------------------------------------------------------------------

typedef const __attribute__((address_space(4))) block_program* cprogram;

// Chunk index -> global bits outside the local positions.
__device__ __forceinline__ uint64_t chunk_base(uint64_t c, cprogram prog, uint32_t L) {
  for (uint32_t r = 0; r < L; ++r) {
    const uint32_t q = prog->pos[r];
    c = ((c >> q) << (q + 1)) | (c & ((1ull << q) - 1ull));
  }
  return c;
}

// Local index of item v of a gate: zeros inserted at op.ins[0 .. nins), then
// the cond bits set.  The common cases (no insertion: uncontrolled diagonal
// gates; one: uncontrolled pair gates) avoid the loop.
__device__ __forceinline__ uint32_t op_index(uint32_t v, const __attribute__((address_space(4))) block_op& op,
                                             uint32_t nins, uint32_t set, uint32_t lower0) {
  if (nins == 0) return v;
  if (nins == 1) return ((v & ~lower0) << 1) | (v & lower0) | set;
  for (uint32_t r = 0; r < nins; ++r) {
    const uint32_t q = op.ins[r];
    v = ((v >> q) << (q + 1)) | (v & ((1u << q) - 1u));
  }
  return v | set;
}

typedef const __attribute__((address_space(4))) block_op& cop;

// One gate over the whole chunk: a strided loop over exactly the local indices it changes.
template <typename R, int PER>
__device__ __forceinline__ void single_pass(cpx<R>* lds, cop op, uint32_t t) {
  constexpr uint32_t CAP = PER * BLOCK_THREADS;
  const int mode = op.mode;
  const uint32_t nins = op.nins, set = op.cond;
  const uint32_t items = CAP >> nins;
  const uint32_t lower0 = (1u << op.ins[0]) - 1u;  // below the first inserted bit
  if (mode == BLOCK_DIAG) {
    const cpx<R> c0 = {(R)op.u[0], (R)op.u[1]}, c1 = {(R)op.u[2], (R)op.u[3]};
    const uint32_t zmask = op.zmask;
    if (op.negate) {  // the visited indices are exactly the cond subspace
      for (uint32_t v = t; v < items; v += BLOCK_THREADS) {
        const uint32_t j = op_index(v, op, nins, set, lower0);
        lds[j] = cneg(lds[j]);
      }
      return;
    }
    for (uint32_t v = t; v < items; v += BLOCK_THREADS) {
      const uint32_t j = op_index(v, op, nins, set, lower0);
      lds[j] = cmul(lds[j], csel((__popc(j & zmask) & 1) != 0, c1, c0));
    }
  } else {
    const uint32_t flip = op.flip, sign = op.sign;
    double u[8];
    for (int k = 0; k < 8; ++k) u[k] = op.u[k];
    for (uint32_t v = t; v < items; v += BLOCK_THREADS) {
      const uint32_t j0 = op_index(v, op, nins, set, lower0), j1 = j0 ^ flip;
      const cpx<R> x0 = lds[j0], x1 = lds[j1];
      cpx<R> y0, y1;
      switch (mode) {
        case BLOCK_H: pair_math<R, MODE_H>(u, sign, j0, x0, x1, y0, y1); break;
        case BLOCK_SWAP: pair_math<R, MODE_SWAP>(u, sign, j0, x0, x1, y0, y1); break;
        case BLOCK_ESWAP: pair_math<R, MODE_ESWAP>(u, sign, j0, x0, x1, y0, y1); break;
        default: pair_math<R, MODE_MATRIX>(u, sign, j0, x0, x1, y0, y1); break;
      }
      lds[j0] = y0;
      lds[j1] = y1;
    }
  }
}

// A pair gate on a thread's group in registers: group index m pairs with m ^ F when the
// gate's pair bit is clear in m and its condition bits are set in the local index j[m].
template <int F, typename R, int PER>
__device__ __forceinline__ void group_pairs(cpx<R> (&x)[PER], const uint32_t (&j)[PER], cop op) {
  if constexpr (F < PER) {
    const int mode = op.mode;
    const uint32_t pair = op.group_pair, cond = op.cond, sign = op.sign;
    double u[8];
    for (int k = 0; k < 8; ++k) u[k] = op.u[k];
    if (pair == (uint32_t)(F & -F)) {
      // the common case: i0 is the member with the lowest flipped bit clear
#pragma unroll
      for (int m = 0; m < PER; ++m) {
        if ((m & (F & -F)) != 0) continue;
        const int m1 = m ^ F;
        cpx<R> y0, y1;
        switch (mode) {
          case BLOCK_H: pair_math<R, MODE_H>(u, sign, j[m], x[m], x[m1], y0, y1); break;
          case BLOCK_SWAP: pair_math<R, MODE_SWAP>(u, sign, j[m], x[m], x[m1], y0, y1); break;
          case BLOCK_ESWAP: pair_math<R, MODE_ESWAP>(u, sign, j[m], x[m], x[m1], y0, y1); break;
          default: pair_math<R, MODE_MATRIX>(u, sign, j[m], x[m], x[m1], y0, y1); break;
        }
        if (cond == 0) {
          x[m] = y0;
          x[m1] = y1;
        } else {
          const bool on = (j[m] & cond) == cond;
          x[m] = csel(on, y0, x[m]);
          x[m1] = csel(on, y1, x[m1]);
        }
      }
      return;
    }
#pragma unroll
    for (int m = 0; m < PER; ++m) {
      // each pair once, from its member with the lowest flipped bit clear; i0 is the member
      // with the gate's pair bit clear (another flipped bit for SWAP)
      if ((m & (F & -F)) != 0) continue;
      const int m1 = m ^ F;
      const bool first_is_i0 = ((uint32_t)m & pair) == 0;
      const uint32_t i0 = first_is_i0 ? j[m] : j[m1];
      const bool on = (i0 & cond) == cond;
      const cpx<R> x0 = csel(first_is_i0, x[m], x[m1]), x1 = csel(first_is_i0, x[m1], x[m]);
      cpx<R> y0, y1;
      switch (mode) {
        case BLOCK_H: pair_math<R, MODE_H>(u, sign, i0, x0, x1, y0, y1); break;
        case BLOCK_SWAP: pair_math<R, MODE_SWAP>(u, sign, i0, x0, x1, y0, y1); break;
        case BLOCK_ESWAP: pair_math<R, MODE_ESWAP>(u, sign, i0, x0, x1, y0, y1); break;
        default: pair_math<R, MODE_MATRIX>(u, sign, i0, x0, x1, y0, y1); break;
      }
      const cpx<R> nm = csel(first_is_i0, y0, y1), nm1 = csel(first_is_i0, y1, y0);
      x[m] = csel(on, nm, x[m]);
      x[m1] = csel(on, nm1, x[m1]);
    }
  }
}

// An unconditioned diagonal gate whose parity bits are group bits ZG (over the group index):
// each member's coefficient is fixed at compile time.
template <int ZG, typename R, int PER>
__device__ __forceinline__ void group_diag(cpx<R> (&x)[PER], cpx<R> c0, cpx<R> c1) {
  if constexpr (ZG < PER) {
#pragma unroll
    for (int m = 0; m < PER; ++m) x[m] = cmul(x[m], (__builtin_popcount(m & ZG) & 1) ? c1 : c0);
  }
}

// A segment: each thread loads one group of PER amplitudes (local indices differing in
// the group bits only), applies the segment's gates in registers and stores it back.
template <typename R, int PER>
__device__ __forceinline__ void segment_pass(cpx<R>* lds, cprogram prog, uint32_t first_op, uint32_t num_ops,
                                             const __attribute__((address_space(4))) uint32_t* group, uint32_t t) {
  constexpr int GB = PER == 1 ? 0 : PER == 2 ? 1 : PER == 4 ? 2 : PER == 8 ? 3 : 4;
  uint32_t base = t;
#pragma unroll
  for (int i = 0; i < GB; ++i) {
    const uint32_t q = group[i];
    base = ((base >> q) << (q + 1)) | (base & ((1u << q) - 1u));
  }
  uint32_t j[PER];
  cpx<R> x[PER];
#pragma unroll
  for (int m = 0; m < PER; ++m) {
    uint32_t dep = 0;
#pragma unroll
    for (int i = 0; i < GB; ++i)
      if ((m >> i) & 1) dep |= 1u << group[i];
    j[m] = base | dep;
    x[m] = lds[j[m]];
  }
  for (uint32_t o = 0; o < num_ops; ++o) {
    cop op = prog->ops[first_op + o];
    if (op.mode == BLOCK_DIAG) {
      // in a segment: negations under conditions, and unconditioned diagonal gates (the
      // planner gives the others a pass)
      if (op.negate) {
        const uint32_t cond = op.cond;
#pragma unroll
        for (int m = 0; m < PER; ++m) x[m] = cneg_if(x[m], (j[m] & cond) == cond);
        continue;
      }
      const cpx<R> c0 = {(R)op.u[0], (R)op.u[1]}, c1 = {(R)op.u[2], (R)op.u[3]};
      switch (op.group_zmask) {  // parity bits among the group bits: no per-amplitude select
        case 0: group_diag<0>(x, c0, c1); continue;
        case 1: group_diag<1>(x, c0, c1); continue;
        case 2: group_diag<2>(x, c0, c1); continue;
        case 3: group_diag<3>(x, c0, c1); continue;
        case 4: group_diag<4>(x, c0, c1); continue;
        case 5: group_diag<5>(x, c0, c1); continue;
        case 6: group_diag<6>(x, c0, c1); continue;
        case 7: group_diag<7>(x, c0, c1); continue;
        case 8: group_diag<8>(x, c0, c1); continue;
        case 9: group_diag<9>(x, c0, c1); continue;
        case 10: group_diag<10>(x, c0, c1); continue;
        case 11: group_diag<11>(x, c0, c1); continue;
        case 12: group_diag<12>(x, c0, c1); continue;
        case 13: group_diag<13>(x, c0, c1); continue;
        case 14: group_diag<14>(x, c0, c1); continue;
        case 15: group_diag<15>(x, c0, c1); continue;
        default: break;
      }
      const uint32_t zmask = op.zmask;
#pragma unroll
      for (int m = 0; m < PER; ++m) x[m] = cmul(x[m], csel((__popc(j[m] & zmask) & 1) != 0, c1, c0));
    } else {
      switch (op.group_flip) {
        case 1: group_pairs<1>(x, j, op); break;
        case 2: group_pairs<2>(x, j, op); break;
        case 3: group_pairs<3>(x, j, op); break;
        case 4: group_pairs<4>(x, j, op); break;
        case 5: group_pairs<5>(x, j, op); break;
        case 6: group_pairs<6>(x, j, op); break;
        case 8: group_pairs<8>(x, j, op); break;
        case 9: group_pairs<9>(x, j, op); break;
        case 10: group_pairs<10>(x, j, op); break;
        default: group_pairs<12>(x, j, op); break;
      }
    }
  }
#pragma unroll
  for (int m = 0; m < PER; ++m) lds[j[m]] = x[m];
}

// One workgroup per chunk of exactly PER * BLOCK_THREADS amplitudes
// (L = 9 + log2 PER), so every global access of the load and store phases is
// unconditional.  The gates run as passes over the chunk in LDS: segments
// (several gates on register-resident groups) and single-gate passes.
template <typename R, bool NT, int PER>
__global__ __launch_bounds__(BLOCK_THREADS) void k_block(cpx<R>* __restrict__ s, const block_program* program) {
  constexpr uint32_t CAP = PER * BLOCK_THREADS;
  __shared__ cpx<R> lds[CAP];
  const cprogram prog = (cprogram)program;  // uniform: scalar loads
  const uint32_t L = prog->num_local;
  // XCD-contiguous chunk order: consecutive chunks (neighbouring runs of memory) on one XCD
  const uint64_t base = chunk_base(xcd_tile(blockIdx.x, gridDim.x, 1), prog, L);
  // local index j = t + BLOCK_THREADS * k: the bits of t go to pos[0 .. 8]
  const uint32_t t = threadIdx.x;
  uint64_t low = 0;
  for (uint32_t r = 0; r < (uint32_t)BLOCK_THREAD_BITS; ++r) low |= (uint64_t)((t >> r) & 1u) << prog->pos[r];
#pragma unroll
  for (uint32_t k = 0; k < PER; ++k)
    lds[t + BLOCK_THREADS * k] = load_amp<NT>(s + (base | low | prog->deposit_high[k]));
  __syncthreads();
  const uint32_t npasses = prog->num_passes;
  for (uint32_t pi = 0; pi < npasses; ++pi) {
    const auto& pass = prog->passes[pi];
    if (pass.kind == BLOCK_PASS_SINGLE) single_pass<R, PER>(lds, prog->ops[pass.first_op], t);
    else segment_pass<R, PER>(lds, prog, pass.first_op, pass.num_ops, pass.group, t);
    __syncthreads();
  }
#pragma unroll
  for (uint32_t k = 0; k < PER; ++k)
    store_amp<NT>(s + (base | low | prog->deposit_high[k]), lds[t + BLOCK_THREADS * k]);
}

// Reverse sweep of the adjoint method over one chunk of psi and lambda (both in
// LDS).  The program's passes (the circuit's inverse ops in reverse order) are
// register segments or single ops.  A parametric op reads each amplitude pair of
// both states once, adds its gradient term (summed over the chunk in a fixed
// order) and un-applies the op to both.  Non-parametric ops are un-applied with
// the plain arithmetic; the RX / RY / RZ-type parametric forms use fused
// multiply-adds (the sweep's results are sums, not amplitudes compared bitwise).
typedef const __attribute__((address_space(4))) adjoint_block_program* cadjoint;

typedef const __attribute__((address_space(4))) double* cdouble;

// A lane's value from another lane of its row by a DPP control (both 32-bit halves).  Every
// control used below selects a lane inside the row, so all lanes read a valid source.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  return __hiloint2double(__builtin_amdgcn_update_dpp(0, hi, CTRL, 0xF, 0xF, false),
                          __builtin_amdgcn_update_dpp(0, lo, CTRL, 0xF, 0xF, false));
}

// The sum of v over the 64 lanes of the wave, in a fixed order, without LDS traffic: pairs and
// quads (quad permutations), 8 and 16 lanes (half-row and row mirrors; every lane then holds
// its row's sum), then the four row sums added in row order.  Uniform result.
__device__ __forceinline__ double wave_sum_dpp(double v) {
  v += dpp_f64<0xB1>(v);   // quad_perm [1, 0, 3, 2]
  v += dpp_f64<0x4E>(v);   // quad_perm [2, 3, 0, 1]
  v += dpp_f64<0x141>(v);  // row_half_mirror
  v += dpp_f64<0x140>(v);  // row_mirror
  auto lane = [&](int l) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                            __builtin_amdgcn_readlane(__double2loint(v), l));
  };
  return ((lane(0) + lane(16)) + lane(32)) + lane(48);
}

__device__ __forceinline__ float fma_r(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_r(double a, double b, double c) { return __builtin_fma(a, b, c); }

// The gradient term of one pair -- psi (x0, x1), lambda (y0, y1) -- added to two FP64 sums.
// For the Z, Y and X forms Re <lam| D |psi> = w0 s0 + w1 s1 with the weights of
// gradient_term:
//   Z: D = diag(i a, i b):        s0 = Im(conj(l0) p0), s1 = Im(conj(l1) p1);
//   Y: D = [[0, a], [b, 0]]:      s0 = Re(conj(l0) p1), s1 = Re(conj(l1) p0);
//   X: D = [[0, i a], [i b, 0]]:  s0 = Im(conj(l0) p1), s1 = Im(conj(l1) p0).
// The general form adds Re <lam| D |psi> to s0.
template <int KIND, typename R>
__device__ __forceinline__ void pair_gradient(cdouble d, cpx<R> x0, cpx<R> x1, cpx<R> y0, cpx<R> y1, double& s0,
                                              double& s1) {
  const double p0r = x0.re, p0i = x0.im, p1r = x1.re, p1i = x1.im;
  const double l0r = y0.re, l0i = y0.im, l1r = y1.re, l1i = y1.im;
  if constexpr (KIND == ADJ_GRAD_Z) {
    s0 = __builtin_fma(l0i, p0r, s0);
    s0 = __builtin_fma(-l0r, p0i, s0);
    s1 = __builtin_fma(l1i, p1r, s1);
    s1 = __builtin_fma(-l1r, p1i, s1);
  } else if constexpr (KIND == ADJ_GRAD_Y) {
    s0 = __builtin_fma(l0r, p1r, s0);
    s0 = __builtin_fma(l0i, p1i, s0);
    s1 = __builtin_fma(l1r, p0r, s1);
    s1 = __builtin_fma(l1i, p0i, s1);
  } else if constexpr (KIND == ADJ_GRAD_X) {
    s0 = __builtin_fma(l0i, p1r, s0);
    s0 = __builtin_fma(-l0r, p1i, s0);
    s1 = __builtin_fma(l1i, p0r, s1);
    s1 = __builtin_fma(-l1r, p0i, s1);
  } else {
    const double d00r = d[0], d00i = d[1], d01r = d[2], d01i = d[3], d10r = d[4], d10i = d[5], d11r = d[6],
                 d11i = d[7];
    const double q0r = d00r * p0r - d00i * p0i + d01r * p1r - d01i * p1i;
    const double q0i = d00r * p0i + d00i * p0r + d01r * p1i + d01i * p1r;
    const double q1r = d10r * p0r - d10i * p0i + d11r * p1r - d11i * p1i;
    const double q1i = d10r * p0i + d10i * p0r + d11r * p1i + d11i * p1r;
    s0 += l0r * q0r + l0i * q0i + l1r * q1r + l1i * q1i;
  }
}

// Re <lam| D |psi> from the two sums of pair_gradient: the weights are a and b of D.
template <int KIND>
__device__ __forceinline__ double gradient_term(cdouble d, double s0, double s1) {
  if constexpr (KIND == ADJ_GRAD_Z) return d[1] * s0 + d[7] * s1;
  else if constexpr (KIND == ADJ_GRAD_Y) return d[2] * s0 + d[4] * s1;
  else if constexpr (KIND == ADJ_GRAD_X) return d[3] * s0 + d[5] * s1;
  else return s0;
}

// The inverse of a one-target parametric op in matrix mode on the pair (x0, x1), i0 = j0:
// the matrix-mode arithmetic in general; the Y form (a real 2x2) and the X form
// ([[r, i s], [i s', r']]) with fused multiply-adds.
template <int KIND, typename R>
__device__ __forceinline__ void unapply_pair(const double (&u)[8], uint64_t sign, uint32_t j0, cpx<R> x0,
                                             cpx<R> x1, cpx<R>& y0, cpx<R>& y1) {
  if constexpr (KIND == ADJ_GRAD_Y) {
    const R m00 = (R)u[0], m01 = (R)u[2], m10 = (R)u[4], m11 = (R)u[6];
    y0 = {fma_r(m01, x1.re, m00 * x0.re), fma_r(m01, x1.im, m00 * x0.im)};
    y1 = {fma_r(m11, x1.re, m10 * x0.re), fma_r(m11, x1.im, m10 * x0.im)};
  } else if constexpr (KIND == ADJ_GRAD_X) {
    const R m00 = (R)u[0], m01 = (R)u[3], m10 = (R)u[5], m11 = (R)u[6];
    y0 = {fma_r(-m01, x1.im, m00 * x0.re), fma_r(m01, x1.re, m00 * x0.im)};
    y1 = {fma_r(m11, x1.re, -(m10 * x0.im)), fma_r(m10, x0.re, m11 * x1.im)};
  } else {
    pair_math<R, MODE_MATRIX>(u, sign, j0, x0, x1, y0, y1);
  }
}

template <typename R>
__device__ __forceinline__ cpx<R> cmul_fma(cpx<R> x, cpx<R> k) {
  return {fma_r(x.re, k.re, -(x.im * k.im)), fma_r(x.re, k.im, x.im * k.re)};
}

// The inverse of a diagonal parametric op (RZ / EZ: parity of the target; U1: condition
// on the target) on one amplitude of each state.
template <typename R>
__device__ __forceinline__ void unapply_diag(cop op, uint32_t j, cpx<R>& x, cpx<R>& y) {
  const cpx<R> c0 = {(R)op.u[0], (R)op.u[1]}, c1 = {(R)op.u[2], (R)op.u[3]};
  const bool on = (j & op.cond) == op.cond;
  const cpx<R> k = csel((__popc(j & op.zmask) & 1) != 0, c1, c0);
  x = csel(on, cmul_fma(x, k), x);
  y = csel(on, cmul_fma(y, k), y);
}

// The Z form's inverse on a pair (j0 with the target clear, j1 with it set): RZ / EZ
// (no condition) multiply the pair's members by c0 and c1; U1 multiplies j1 by c0.
template <typename R>
__device__ __forceinline__ void unapply_z_pair(cop op, cpx<R>& x0, cpx<R>& x1, cpx<R>& y0, cpx<R>& y1) {
  const cpx<R> c0 = {(R)op.u[0], (R)op.u[1]}, c1 = {(R)op.u[2], (R)op.u[3]};
  if (op.cond == 0) {
    x0 = cmul_fma(x0, c0);
    y0 = cmul_fma(y0, c0);
    x1 = cmul_fma(x1, c1);
    y1 = cmul_fma(y1, c1);
  } else {
    x1 = cmul_fma(x1, c0);
    y1 = cmul_fma(y1, c0);
  }
}

// A parametric op over the whole chunk: per pair (j0, j1) of its target bit, read psi and
// lambda once, add Re <lam| D |psi>, un-apply the op to both, write both back.  Returns the
// thread's share of the gradient term.
template <int KIND, typename R, int PER>
__device__ __forceinline__ double param_pass(cpx<R>* lp, cpx<R>* ll, cop op, cdouble d, uint32_t t) {
  constexpr uint32_t CAP = PER * BLOCK_THREADS;
  const uint32_t b = op.pair_bit, lower = (1u << b) - 1u;
  const bool diag = op.mode == BLOCK_DIAG;
  const uint64_t sign = op.sign;
  double u[8];
  for (int k = 0; k < 8; ++k) u[k] = op.u[k];
  double s0 = 0.0, s1 = 0.0;
  for (uint32_t v = t; v < CAP / 2; v += BLOCK_THREADS) {
    const uint32_t j0 = ((v & ~lower) << 1) | (v & lower), j1 = j0 | (1u << b);
    cpx<R> x0 = lp[j0], x1 = lp[j1], y0 = ll[j0], y1 = ll[j1];
    pair_gradient<KIND>(d, x0, x1, y0, y1, s0, s1);
    if (KIND == ADJ_GRAD_Z) {
      unapply_z_pair(op, x0, x1, y0, y1);
    } else if (KIND == ADJ_GRAD_GENERAL && diag) {
      unapply_diag(op, j0, x0, y0);
      unapply_diag(op, j1, x1, y1);
    } else {
      cpx<R> a0, a1, e0, e1;
      unapply_pair<KIND>(u, sign, j0, x0, x1, a0, a1);
      unapply_pair<KIND>(u, sign, j0, y0, y1, e0, e1);
      x0 = a0;
      x1 = a1;
      y0 = e0;
      y1 = e1;
    }
    lp[j0] = x0;
    lp[j1] = x1;
    ll[j0] = y0;
    ll[j1] = y1;
  }
  return gradient_term<KIND>(d, s0, s1);
}

// A parametric op on a thread's group in registers: pairs (m, m | P) of the op's target
// group bit P.  Adds the gradient term, then un-applies the op to both states.  Every
// instance reads and writes all PER amplitudes of both groups, so that the instances for
// different P, merged by the compiler, still address the group arrays statically.
template <int KIND, int P, typename R, int PER>
__device__ __forceinline__ double group_param(cpx<R> (&xp)[PER], cpx<R> (&xl)[PER], const uint32_t (&j)[PER],
                                              cop op, cdouble d) {
  double s0 = 0.0, s1 = 0.0;
  if constexpr (P < PER) {
    cpx<R> a[PER], b[PER];
#pragma unroll
    for (int m = 0; m < PER; ++m) {
      a[m] = xp[m];
      b[m] = xl[m];
    }
    if (KIND == ADJ_GRAD_Z) {
#pragma unroll
      for (int m = 0; m < PER; ++m) {
        if ((m & P) != 0) continue;
        pair_gradient<KIND>(d, a[m], a[m | P], b[m], b[m | P], s0, s1);
        unapply_z_pair(op, a[m], a[m | P], b[m], b[m | P]);
      }
    } else if (KIND == ADJ_GRAD_GENERAL && op.mode == BLOCK_DIAG) {
#pragma unroll
      for (int m = 0; m < PER; ++m) {
        if ((m & P) != 0) continue;
        pair_gradient<KIND>(d, a[m], a[m | P], b[m], b[m | P], s0, s1);
      }
#pragma unroll
      for (int m = 0; m < PER; ++m) unapply_diag(op, j[m], a[m], b[m]);
    } else {
      const uint64_t sign = op.sign;
      double u[8];
      for (int k = 0; k < 8; ++k) u[k] = op.u[k];
#pragma unroll
      for (int m = 0; m < PER; ++m) {
        if ((m & P) != 0) continue;
        const int m1 = m | P;
        pair_gradient<KIND>(d, a[m], a[m1], b[m], b[m1], s0, s1);
        cpx<R> a0, a1, e0, e1;
        unapply_pair<KIND>(u, sign, j[m], a[m], a[m1], a0, a1);
        unapply_pair<KIND>(u, sign, j[m], b[m], b[m1], e0, e1);
        a[m] = a0;
        a[m1] = a1;
        b[m] = e0;
        b[m1] = e1;
      }
    }
#pragma unroll
    for (int m = 0; m < PER; ++m) {
      xp[m] = a[m];
      xl[m] = b[m];
    }
  }
  return gradient_term<KIND>(d, s0, s1);
}

template <int KIND, typename R, int PER>
__device__ __forceinline__ double group_param_any(cpx<R> (&xp)[PER], cpx<R> (&xl)[PER], const uint32_t (&j)[PER],
                                                  cop op, cdouble d) {
  switch (op.group_pair) {
    case 1: return group_param<KIND, 1>(xp, xl, j, op, d);
    case 2: return group_param<KIND, 2>(xp, xl, j, op, d);
    case 4: return group_param<KIND, 4>(xp, xl, j, op, d);
    default: return group_param<KIND, 8>(xp, xl, j, op, d);
  }
}

// A register segment of the reverse sweep: each thread loads one group of psi and one of
// lambda (local indices differing in the group bits only), runs the segment's ops on both --
// the gradient term of a parametric op first, then its inverse -- and stores them; the
// segment's gradient terms are then reduced over the workgroup in a fixed order.
template <typename R, int PER>
__device__ __forceinline__ void adjoint_segment(cpx<R>* lp, cpx<R>* ll, cadjoint aprog, uint32_t first_op,
                                                uint32_t num_ops,
                                                const __attribute__((address_space(4))) uint32_t* group, uint32_t t,
                                                double (*wave_sums)[BLOCK_THREADS / 64], double* partials) {
  constexpr int GB = PER == 1 ? 0 : PER == 2 ? 1 : PER == 4 ? 2 : PER == 8 ? 3 : 4;
  const cprogram prog = (cprogram)&aprog->base;
  uint32_t base = t;
#pragma unroll
  for (int i = 0; i < GB; ++i) {
    const uint32_t q = group[i];
    base = ((base >> q) << (q + 1)) | (base & ((1u << q) - 1u));
  }
  uint32_t j[PER];
  cpx<R> xp[PER], xl[PER];
#pragma unroll
  for (int m = 0; m < PER; ++m) {
    uint32_t dep = 0;
#pragma unroll
    for (int i = 0; i < GB; ++i)
      if ((m >> i) & 1) dep |= 1u << group[i];
    j[m] = base | dep;
    xp[m] = lp[j[m]];
    xl[m] = ll[j[m]];
  }
  uint32_t ngrad = 0;
  for (uint32_t o = first_op; o < first_op + num_ops; ++o) {
    cop op = prog->ops[o];
    if (aprog->grad_slot[o] >= 0) {
      double g;
      switch (aprog->grad_kind[o]) {
        case ADJ_GRAD_Z: g = group_param_any<ADJ_GRAD_Z>(xp, xl, j, op, aprog->d[o]); break;
        case ADJ_GRAD_Y: g = group_param_any<ADJ_GRAD_Y>(xp, xl, j, op, aprog->d[o]); break;
        case ADJ_GRAD_X: g = group_param_any<ADJ_GRAD_X>(xp, xl, j, op, aprog->d[o]); break;
        default: g = group_param_any<ADJ_GRAD_GENERAL>(xp, xl, j, op, aprog->d[o]); break;
      }
      // the term's wave sum now (fixed order), kept per wave until the segment ends
      g = wave_sum_dpp(g);
      if ((t & 63) == 0) wave_sums[ngrad][t >> 6] = g;
      ++ngrad;
      continue;
    }
    if (op.mode == BLOCK_DIAG) {
      const uint32_t cond = op.cond, zmask = op.zmask;
      if (op.negate) {
#pragma unroll
        for (int m = 0; m < PER; ++m) {
          const bool on = (j[m] & cond) == cond;
          xp[m] = cneg_if(xp[m], on);
          xl[m] = cneg_if(xl[m], on);
        }
        continue;
      }
      const cpx<R> c0 = {(R)op.u[0], (R)op.u[1]}, c1 = {(R)op.u[2], (R)op.u[3]};
#pragma unroll
      for (int m = 0; m < PER; ++m) {
        const bool on = (j[m] & cond) == cond;
        const cpx<R> k = csel((__popc(j[m] & zmask) & 1) != 0, c1, c0);
        xp[m] = csel(on, cmul(xp[m], k), xp[m]);
        xl[m] = csel(on, cmul(xl[m], k), xl[m]);
      }
    } else {
      switch (op.group_flip) {
        case 1: group_pairs<1>(xp, j, op); group_pairs<1>(xl, j, op); break;
        case 2: group_pairs<2>(xp, j, op); group_pairs<2>(xl, j, op); break;
        case 3: group_pairs<3>(xp, j, op); group_pairs<3>(xl, j, op); break;
        case 4: group_pairs<4>(xp, j, op); group_pairs<4>(xl, j, op); break;
        case 5: group_pairs<5>(xp, j, op); group_pairs<5>(xl, j, op); break;
        case 6: group_pairs<6>(xp, j, op); group_pairs<6>(xl, j, op); break;
        case 8: group_pairs<8>(xp, j, op); group_pairs<8>(xl, j, op); break;
        case 9: group_pairs<9>(xp, j, op); group_pairs<9>(xl, j, op); break;
        case 10: group_pairs<10>(xp, j, op); group_pairs<10>(xl, j, op); break;
        default: group_pairs<12>(xp, j, op); group_pairs<12>(xl, j, op); break;
      }
    }
  }
#pragma unroll
  for (int m = 0; m < PER; ++m) {
    lp[j[m]] = xp[m];
    ll[j[m]] = xl[m];
  }
  __syncthreads();
  if (t < ngrad) {
    // the t-th parametric op of the segment
    uint32_t seen = 0, slot = 0;
    for (uint32_t o = first_op; o < first_op + num_ops; ++o)
      if (aprog->grad_slot[o] >= 0) {
        if (seen == t) slot = (uint32_t)aprog->grad_slot[o];
        ++seen;
      }
    double sum = 0.0;
    for (int w = 0; w < BLOCK_THREADS / 64; ++w) sum += wave_sums[t][w];
    partials[(uint64_t)slot * gridDim.x + blockIdx.x] = sum;
  }
}

template <typename R, int PER>
__global__ __launch_bounds__(BLOCK_THREADS) void k_adjoint_block(cpx<R>* __restrict__ psi, cpx<R>* __restrict__ lam,
                                                                const adjoint_block_program* program,
                                                                double* __restrict__ partials) {
  constexpr uint32_t CAP = PER * BLOCK_THREADS;
  __shared__ cpx<R> lp[CAP], ll[CAP];
  // Per-wave gradient sums, two sets used in turn by the passes that make gradient terms: a pass
  // writes its set, waits at a barrier and a few threads add it up; the next such pass writes the
  // other set, so those threads finish while the workgroup goes on, and the pass after that starts
  // writing the first set only after the next pass's barrier, which they have passed.
  __shared__ double wave_sums[2][ADJ_SEGMENT_GRADIENTS][BLOCK_THREADS / 64];
  uint32_t wbuf = 0;
  const cadjoint aprog = (cadjoint)program;
  const cprogram prog = (cprogram)&program->base;
  const uint32_t L = prog->num_local;
  const uint32_t chunk = (uint32_t)xcd_tile(blockIdx.x, gridDim.x, 1);  // XCD-contiguous chunk order
  const uint64_t base = chunk_base(chunk, prog, L);
  const uint32_t t = threadIdx.x;
  uint64_t low = 0;
  for (uint32_t r = 0; r < (uint32_t)BLOCK_THREAD_BITS; ++r) low |= (uint64_t)((t >> r) & 1u) << prog->pos[r];
#pragma unroll
  for (uint32_t k = 0; k < PER; ++k) {
    const uint64_t i = base | low | prog->deposit_high[k];
    lp[t + BLOCK_THREADS * k] = load_amp<true>(psi + i);
    ll[t + BLOCK_THREADS * k] = load_amp<true>(lam + i);
  }
  __syncthreads();
  const uint32_t npasses = prog->num_passes;
  for (uint32_t pi = 0; pi < npasses; ++pi) {
    const auto& pass = prog->passes[pi];
    if (pass.kind == BLOCK_PASS_SEGMENT) {
      // (ends with its barrier between the LDS writes and the gradient sums)
      adjoint_segment<R, PER>(lp, ll, aprog, pass.first_op, pass.num_ops, pass.group, t, wave_sums[wbuf], partials);
      wbuf ^= 1u;
      continue;
    }
    const uint32_t o = pass.first_op;
    const auto& op = prog->ops[o];
    const int slot = aprog->grad_slot[o];
    if (slot < 0) {
      single_pass<R, PER>(lp, op, t);
      single_pass<R, PER>(ll, op, t);
      __syncthreads();
      continue;
    }
    // a parametric op (one target, no controls) with the strided loop over its pairs
    double acc;
    switch (aprog->grad_kind[o]) {
      case ADJ_GRAD_Z: acc = param_pass<ADJ_GRAD_Z, R, PER>(lp, ll, op, aprog->d[o], t); break;
      case ADJ_GRAD_Y: acc = param_pass<ADJ_GRAD_Y, R, PER>(lp, ll, op, aprog->d[o], t); break;
      case ADJ_GRAD_X: acc = param_pass<ADJ_GRAD_X, R, PER>(lp, ll, op, aprog->d[o], t); break;
      default: acc = param_pass<ADJ_GRAD_GENERAL, R, PER>(lp, ll, op, aprog->d[o], t); break;
    }
    acc = wave_sum_dpp(acc);
    if ((t & 63) == 0) wave_sums[wbuf][0][t >> 6] = acc;
    __syncthreads();
    if (t == 0) {
      double sum = 0.0;
      for (int w = 0; w < BLOCK_THREADS / 64; ++w) sum += wave_sums[wbuf][0][w];
      partials[(uint64_t)slot * gridDim.x + blockIdx.x] = sum;
    }
    wbuf ^= 1u;
  }
#pragma unroll
  for (uint32_t k = 0; k < PER; ++k) {
    const uint64_t i = base | low | prog->deposit_high[k];
    store_amp<true>(psi + i, lp[t + BLOCK_THREADS * k]);
    store_amp<true>(lam + i, ll[t + BLOCK_THREADS * k]);
  }
}

int launch_adjoint_block(const adjoint_block_program* device_program, const adjoint_block_program& program,
                         void* psi, void* lam, uint64_t size, int dtype, double* partials, void* stream) {
  unsigned nq;
  if (int st = check_state(psi, size, dtype, &nq)) return st;
  if (int st = check_state(lam, size, dtype, &nq)) return st;
  const unsigned L = program.base.num_local;
  const unsigned want = dtype == BRK_C128 ? ADJ_BLOCK_LOCAL_C128 : ADJ_BLOCK_LOCAL_C64;
  if (L != want || L > nq || program.base.num_ops > (uint32_t)BLOCK_MAX_OPS)
    return fail(BRK_ERR_INVALID_ARGUMENT, "adjoint block does not fit (%u local qubits)", L);
  const uint64_t chunks = size >> L;
  if (chunks > 0x7fffffffull) return fail(BRK_ERR_INVALID_ARGUMENT, "too many chunks");
  hipStream_t st = as_stream(stream);
  if (dtype == BRK_C128)
    hipLaunchKernelGGL((k_adjoint_block<double, 4>), dim3((unsigned)chunks), dim3(BLOCK_THREADS), 0, st,
                       (cpx<double>*)psi, (cpx<double>*)lam, device_program, partials);
  else
    hipLaunchKernelGGL((k_adjoint_block<float, 8>), dim3((unsigned)chunks), dim3(BLOCK_THREADS), 0, st,
                       (cpx<float>*)psi, (cpx<float>*)lam, device_program, partials);
  return check_launch("adjoint block");
}

// ---------------------------------------------------------------------------
// Launch helpers
// ---------------------------------------------------------------------------

unsigned block_max_local(int dtype) {
  const unsigned amp_log2 = dtype == BRK_C128 ? 4 : 3;
  return (unsigned)__builtin_ctz((unsigned)get_tuning().block_lds) - amp_log2;
}

bool adjoint_blocks_enabled() { return get_tuning().adjoint_blocks != 0; }

unsigned block_run_qubits(int dtype) {
  const unsigned amp_log2 = dtype == BRK_C128 ? 4 : 3;
  const unsigned run_log2 = (unsigned)__builtin_ctz((unsigned)get_tuning().block_run_bytes);
  return run_log2 > amp_log2 ? run_log2 - amp_log2 : 0;
}

template <typename R, bool NT, int PER>
static int launch_block_inst(uint64_t chunks, hipStream_t st, cpx<R>* s, const block_program* p) {
  hipLaunchKernelGGL((k_block<R, NT, PER>), dim3((unsigned)chunks), dim3(BLOCK_THREADS), 0, st, s, p);
  return BRK_OK;
}

// Picks the instantiation whose chunk is exactly 2^L amplitudes (L = 9 .. 13).
template <typename R, bool NT>
static int launch_block_per(uint64_t chunks, hipStream_t st, cpx<R>* s, const block_program* p, unsigned L) {
  switch (L - BLOCK_THREAD_BITS) {
    case 0: return launch_block_inst<R, NT, 1>(chunks, st, s, p);
    case 1: return launch_block_inst<R, NT, 2>(chunks, st, s, p);
    case 2: return launch_block_inst<R, NT, 4>(chunks, st, s, p);
    case 3: return launch_block_inst<R, NT, 8>(chunks, st, s, p);
    default:
      if constexpr (sizeof(R) == 4) return launch_block_inst<R, NT, 16>(chunks, st, s, p);
      return fail(BRK_ERR_INVALID_ARGUMENT, "no gate-block kernel for %u local qubits", L);
  }
}

int launch_block(const block_program* device_program, const block_program& program, void* first, uint64_t size,
                 int dtype, void* stream) {
  unsigned nq;
  if (int st = check_state(first, size, dtype, &nq)) return st;
  const uint32_t cap = dtype == BRK_C128 ? BLOCK_MAX_LOCAL_C128 : BLOCK_MAX_LOCAL_C64;
  if (program.num_local > cap || program.num_local < (uint32_t)BLOCK_THREAD_BITS || program.num_local > nq ||
      program.num_ops > (uint32_t)BLOCK_MAX_OPS || program.num_passes > (uint32_t)BLOCK_MAX_OPS)
    return fail(BRK_ERR_INVALID_ARGUMENT, "block program does not fit (%u local qubits, %u ops, %u passes)",
                program.num_local, program.num_ops, program.num_passes);
  const uint64_t chunks = size >> program.num_local;
  if (chunks > 0x7fffffffull) return fail(BRK_ERR_INVALID_ARGUMENT, "too many chunks");
  const bool nt = size * amp_bytes(dtype) >= get_tuning().nt_min_bytes;
  hipStream_t st = as_stream(stream);
  const unsigned L = program.num_local;
  int rc;
  if (dtype == BRK_C128)
    rc = nt ? launch_block_per<double, true>(chunks, st, (cpx<double>*)first, device_program, L)
            : launch_block_per<double, false>(chunks, st, (cpx<double>*)first, device_program, L);
  else
    rc = nt ? launch_block_per<float, true>(chunks, st, (cpx<float>*)first, device_program, L)
            : launch_block_per<float, false>(chunks, st, (cpx<float>*)first, device_program, L);
  if (rc) return rc;
  return check_launch("block");
}

static unsigned grid_blocks(uint64_t items) {
  uint64_t b = (items + GATE_THREADS - 1) / GATE_THREADS;
  if (b == 0) b = 1;
  const uint64_t cap = std::min<uint64_t>(get_tuning().max_blocks, (1ull << 24) - 1);
  return (unsigned)std::min(b, cap);
}

// Expands a runtime K into the template instantiations K = 0..3 or generic.
#define BRK_DISPATCH_K(k, KERNEL, ...)                               \
  switch (k) {                                                       \
    case 0: KERNEL(0, __VA_ARGS__); break;                           \
    case 1: KERNEL(1, __VA_ARGS__); break;                           \
    case 2: KERNEL(2, __VA_ARGS__); break;                           \
    case 3: KERNEL(3, __VA_ARGS__); break;                           \
    default: KERNEL(-1, __VA_ARGS__); break;                         \
  }

struct launch_ctx {
  dim3 grid, block;
  size_t lds;
  hipStream_t st;
  int remap;
};

template <typename R, int MODE, bool NT, bool COND>
static void pair_k(const launch_ctx& c, cpx<R>* s, uint64_t n, const gate_args& a) {
#define L(K, ...) hipLaunchKernelGGL((k_pair<R, MODE, K, NT, COND>), c.grid, c.block, c.lds, c.st, s, n, a, c.remap)
  BRK_DISPATCH_K(a.k, L, 0)
#undef L
}

template <typename R, int MODE, bool NT, bool COND, bool SWAP2>
static void lane_k(const launch_ctx& c, cpx<R>* s, uint64_t n, const gate_args& a) {
#define L(K, ...) \
  hipLaunchKernelGGL((k_lane<R, MODE, K, NT, COND, SWAP2>), c.grid, c.block, c.lds, c.st, s, n, a, c.remap)
  BRK_DISPATCH_K(a.k, L, 0)
#undef L
}

template <typename R, bool NT, bool COND>
static void lane_swap_k(const launch_ctx& c, cpx<R>* s, uint64_t n, const gate_args& a) {
#define L(K, ...) hipLaunchKernelGGL((k_lane_swap<R, K, NT, COND>), c.grid, c.block, c.lds, c.st, s, n, a, c.remap)
  BRK_DISPATCH_K(a.k, L, 0)
#undef L
}

template <typename R, int SELECT, bool NT, bool COND>
static void diag_k(const launch_ctx& c, cpx<R>* s, uint64_t n, const gate_args& a) {
#define L(K, ...) hipLaunchKernelGGL((k_diag<R, K, SELECT, NT, COND>), c.grid, c.block, c.lds, c.st, s, n, a, c.remap)
  BRK_DISPATCH_K(a.k, L, 0)
#undef L
}

enum kernel_family { FAM_PAIR, FAM_LANE, FAM_LANE_SWAP2, FAM_LANE_SWAP, FAM_DIAG };

// Picks the instantiation from (dtype, family, mode, nt, cond, parity) and launches.
template <typename R, bool NT, bool COND>
static void launch_family(int family, int mode, const launch_ctx& c, cpx<R>* s, uint64_t n, const gate_args& a) {
  switch (family) {
    case FAM_PAIR:
      if (mode == MODE_H) pair_k<R, MODE_H, NT, COND>(c, s, n, a);
      else if (mode == MODE_SWAP) pair_k<R, MODE_SWAP, NT, COND>(c, s, n, a);
      else if (mode == MODE_ESWAP) pair_k<R, MODE_ESWAP, NT, COND>(c, s, n, a);
      else pair_k<R, MODE_MATRIX, NT, COND>(c, s, n, a);
      break;
    case FAM_LANE:
      if (mode == MODE_H) lane_k<R, MODE_H, NT, COND, false>(c, s, n, a);
      else if (mode == MODE_SWAP) lane_k<R, MODE_SWAP, NT, COND, false>(c, s, n, a);
      else lane_k<R, MODE_MATRIX, NT, COND, false>(c, s, n, a);
      break;
    case FAM_LANE_SWAP2: lane_k<R, MODE_SWAP, NT, COND, true>(c, s, n, a); break;
    case FAM_LANE_SWAP: lane_swap_k<R, NT, COND>(c, s, n, a); break;
    default:
      if (mode == DIAG_MOD4) diag_k<R, DIAG_MOD4, NT, COND>(c, s, n, a);
      else if (a.zmask) diag_k<R, DIAG_PARITY, NT, COND>(c, s, n, a);
      else diag_k<R, DIAG_ONE, NT, COND>(c, s, n, a);
      break;
  }
}

static int launch(int dtype, int family, int mode, void* first, uint64_t size, uint64_t items, const gate_args& a,
                  void* stream) {
  const tuning& t = get_tuning();
  // one load per thread (diagonal, lane kernels) vs two (pair kernels): different occupancy caps
  const bool single = family == FAM_DIAG || family == FAM_LANE || family == FAM_LANE_SWAP2;
  launch_ctx c{dim3(grid_blocks(items)), dim3(GATE_THREADS), (size_t)(single ? t.diag_lds : t.pair_lds),
               as_stream(stream), single ? t.diag_remap : t.pair_remap};
  const bool nt = size * amp_bytes(dtype) >= t.nt_min_bytes;
  const bool cond = a.condmask != 0;
  if (dtype == BRK_C128) {
    auto* s = (cpx<double>*)first;
    if (nt) cond ? launch_family<double, true, true>(family, mode, c, s, items, a)
               : launch_family<double, true, false>(family, mode, c, s, items, a);
    else cond ? launch_family<double, false, true>(family, mode, c, s, items, a)
              : launch_family<double, false, false>(family, mode, c, s, items, a);
  } else {
    auto* s = (cpx<float>*)first;
    if (nt) cond ? launch_family<float, true, true>(family, mode, c, s, items, a)
               : launch_family<float, true, false>(family, mode, c, s, items, a);
    else cond ? launch_family<float, false, true>(family, mode, c, s, items, a)
              : launch_family<float, false, false>(family, mode, c, s, items, a);
  }
  return check_launch("gate");
}

// ---------------------------------------------------------------------------
// Host-side planning: which kernel, which bits are inserted, which tested
// ---------------------------------------------------------------------------

// Number of low bits addressing amplitudes inside one line (tuning::line_bytes).
static unsigned line_bits(int dtype) {
  const uint64_t per_line = std::max<uint64_t>(1, get_tuning().line_bytes / amp_bytes(dtype));
  return (unsigned)__builtin_ctzll(per_line);
}

// Low qubits whose pair partner the lane kernels fetch across the wavefront: inside a line,
// and below the wavefront width (a shuffle mask of 2^6 or more would return the lane's own value).
static unsigned lane_target_bits(int dtype) { return std::min(line_bits(dtype), LANE_BITS); }

struct plan {
  unsigned pos[BRK_MAX_OPERATED_QUBITS];  // inserted positions
  unsigned np = 0;
  uint64_t setmask = 0, condmask = 0;
  void insert(unsigned q) { pos[np++] = q; }
  void control(unsigned q, unsigned threshold) {
    if (q < threshold) condmask |= 1ull << q;
    else {
      insert(q);
      setmask |= 1ull << q;
    }
  }
};

static gate_args make_args(const plan& p) {
  gate_args a{};
  expand_masks e = make_expand_masks(p.pos, (int)p.np);
  std::copy(e.m, e.m + BRK_MAX_OPERATED_QUBITS + 1, a.m);
  a.k = (int)p.np;
  a.setmask = p.setmask;
  a.condmask = p.condmask;
  return a;
}

struct validated {
  unsigned nq;
};

static int validate(void* first, uint64_t size, int dtype, const unsigned* t, unsigned nt, const unsigned* c,
                    unsigned nc, validated* v) {
  if (int st = check_state(first, size, dtype, &v->nq)) return st;
  unsigned ops[BRK_MAX_OPERATED_QUBITS];
  unsigned cnt = 0;
  if (int st = collect_qubits(t, nt, v->nq, ops, &cnt, "target")) return st;
  if (int st = collect_qubits(c, nc, v->nq, ops, &cnt, "control")) return st;
  return BRK_OK;
}

// Single target under controls (H, U2, U3, EX/EY/Y on one qubit, X/CNOT, matrices).
static int single_target(void* first, uint64_t size, int dtype, unsigned target, const unsigned* controls,
                         unsigned nc, int mode, const double* u, void* stream) {
  validated v;
  if (int st = validate(first, size, dtype, &target, 1, controls, nc, &v)) return st;
  plan p;
  gate_args a;
  // Below the line boundary the two amplitudes of a pair share a line: the lane
  // kernel keeps every access a whole contiguous wavefront.
  if (v.nq >= LANE_BITS && target < lane_target_bits(dtype)) {
    for (unsigned i = 0; i < nc; ++i) p.control(controls[i], LANE_BITS);
    a = make_args(p);
    a.lane_bit = (int)target;
    if (u) std::copy(u, u + 8, a.u);
    return launch(dtype, FAM_LANE, mode, first, size, size >> p.np, a, stream);
  }
  p.insert(target);
  for (unsigned i = 0; i < nc; ++i) p.control(controls[i], line_bits(dtype));
  a = make_args(p);
  a.flipmask = 1ull << target;
  if (u) std::copy(u, u + 8, a.u);
  return launch(dtype, FAM_PAIR, mode, first, size, size >> p.np, a, stream);
}

// Targets flipped together (X...X, Y...Y strings) under controls; pairs (i0, i0 ^ T)
// with the lowest target bit of i0 equal to 0.
static int multi_target_flip(void* first, uint64_t size, int dtype, const unsigned* targets, unsigned nt,
                             const unsigned* controls, unsigned nc, int mode, const double* u, bool sign_by_targets,
                             void* stream) {
  if (nt == 0) return fail(BRK_ERR_INVALID_ARGUMENT, "no target qubits");
  if (!targets) return fail(BRK_ERR_INVALID_ARGUMENT, "target list is null");
  if (nt == 1) return single_target(first, size, dtype, targets[0], controls, nc, mode, u, stream);
  validated v;
  if (int st = validate(first, size, dtype, targets, nt, controls, nc, &v)) return st;
  unsigned lowest = targets[0];
  for (unsigned i = 1; i < nt; ++i) lowest = std::min(lowest, targets[i]);
  plan p;
  p.insert(lowest);
  for (unsigned i = 0; i < nc; ++i) p.control(controls[i], line_bits(dtype));
  gate_args a = make_args(p);
  a.flipmask = mask_of(targets, nt);
  if (sign_by_targets) a.signmask = a.flipmask;
  if (u) std::copy(u, u + 8, a.u);
  return launch(dtype, FAM_PAIR, mode, first, size, size >> p.np, a, stream);
}

static int diagonal(void* first, uint64_t size, int dtype, const unsigned* controls, unsigned nc, uint64_t zmask,
                    std::complex<double> d0, std::complex<double> d1, void* stream) {
  plan p;
  for (unsigned i = 0; i < nc; ++i) p.control(controls[i], line_bits(dtype));
  gate_args a = make_args(p);
  a.zmask = zmask;
  a.u[0] = d0.real();
  a.u[1] = d0.imag();
  a.u[2] = d1.real();
  a.u[3] = d1.imag();
  return launch(dtype, FAM_DIAG, zmask ? DIAG_PARITY : DIAG_ONE, first, size, size >> p.np, a, stream);
}

// d[popcount(i & zmask) mod 4] where every control bit of i is 1.
static int diagonal_mod4(void* first, uint64_t size, int dtype, const unsigned* controls, unsigned nc,
                         uint64_t zmask, const std::complex<double> (&d)[4], void* stream) {
  plan p;
  for (unsigned i = 0; i < nc; ++i) p.control(controls[i], line_bits(dtype));
  gate_args a = make_args(p);
  a.zmask = zmask;
  for (int k = 0; k < 4; ++k) {
    a.u[2 * k] = d[k].real();
    a.u[2 * k + 1] = d[k].imag();
  }
  return launch(dtype, FAM_DIAG, DIAG_MOD4, first, size, size >> p.np, a, stream);
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

// ---------------------------------------------------------------------------
// Gate coefficients, computed once on the host.  Every path that applies a
// gate (the entry points below, gate blocks) takes them from these helpers.
// ---------------------------------------------------------------------------

// exp(i phase) with std::exp, as ket::utility::exp_i does (ket/include/ket/utility/exp_i.hpp:24-29).
static std::complex<double> exp_i(double phase) { return std::exp(std::complex<double>{0.0, phase}); }

// Y...Y: (Y^m a)[i] = (-i)^m (-1)^{popcount(i & T)} a[i ^ T]; for one target
// Y = [[0, -i], [i, 0]] (docs/bra.md "Y t").  Sign by the targets.
static void coeffs_pauli_y(unsigned num_targets, double* u) {
  const std::complex<double> ph = minus_i_pow(num_targets);
  const double sgn_m = (num_targets & 1u) ? -1.0 : 1.0;
  put(u, 0, {0.0, 0.0});
  put(u, 1, ph);
  put(u, 2, ph * sgn_m);
  put(u, 3, {0.0, 0.0});
}

// U2(p1, p2) = 1/sqrt2 [[1, -e^{i p2}], [e^{i p1}, e^{i(p1+p2)}]] (ket/include/ket/gate/phase_shift.hpp:649-652)
static void coeffs_u2(double phase1, double phase2, double* u) {
  const double r = 0.707106781186547524400844362104849039;
  const std::complex<double> e1 = exp_i(phase1), e2 = exp_i(phase2);
  put(u, 0, {r, 0.0});
  put(u, 1, -e2 * r);
  put(u, 2, e1 * r);
  put(u, 3, (e2 * e1) * r);
}

// U3(t, p, l) = [[c, -e^{il} s], [e^{ip} s, e^{i(p+l)} c]], c = cos(t/2), s = sin(t/2)
// (ket/include/ket/gate/phase_shift.hpp:1195-1198, coefficients :1219-1229)
static void coeffs_u3(double phase1, double phase2, double phase3, double* u) {
  const double sine = std::sin(0.5 * phase1);
  const double cosine = std::cos(0.5 * phase1);
  const std::complex<double> e2 = exp_i(phase2), e3 = exp_i(phase3);
  put(u, 0, {cosine, 0.0});
  put(u, 1, -(sine * e3));
  put(u, 2, sine * e2);
  put(u, 3, (cosine * e3) * e2);
}

// exp(i phase X..X) = cos I + i sin X..X; pairs (i, i ^ T).  e = e^{i phase}.
static void coeffs_exp_x(std::complex<double> e, double* u) {
  const double c = e.real(), s = e.imag();
  put(u, 0, {c, 0.0});
  put(u, 1, {0.0, s});
  put(u, 2, {0.0, s});
  put(u, 3, {c, 0.0});
}

// exp(i phase Y..Y) = cos I + i sin Y..Y, with (Y^m a)[i] = (-i)^m (-1)^{popcount(i & T)} a[i ^ T].
// For one target: a0' = c a0 + s a1, a1' = -s a0 + c a1
// (ket/include/ket/gate/exponential_pauli_y.hpp:36-38, :84-87).  Sign by the targets.
static void coeffs_exp_y(std::complex<double> e, unsigned num_targets, double* u) {
  const double c = e.real(), s = e.imag();
  const std::complex<double> ph = minus_i_pow(num_targets);
  const double sgn_m = (num_targets & 1u) ? -1.0 : 1.0;
  put(u, 0, {c, 0.0});
  put(u, 1, std::complex<double>{0.0, s} * ph);
  put(u, 2, std::complex<double>{0.0, s} * ph * sgn_m);
  put(u, 3, {c, 0.0});
}

// exp(i phase SWAP): e^{i phase} for (00, 11), cos and sin of the same coefficient for (01, 10)
// (ket/include/ket/gate/exponential_swap.hpp:31-90).
static void coeffs_exp_swap(std::complex<double> e, double* u) {
  put(u, 0, e);
  u[2] = e.real();
  u[3] = e.imag();
}

int describe_physical_op(int code, const unsigned* t, unsigned nt, const unsigned* c, unsigned nc,
                         const double* p, gate_desc* d) {
  *d = gate_desc{};
  const uint64_t tmask = nt ? mask_of(t, nt) : 0, cmask = nc ? mask_of(c, nc) : 0;
  d->qubits = tmask | cmask;
  d->cond = cmask;
  unsigned lowest = 64;
  for (unsigned i = 0; i < nt; ++i) lowest = std::min(lowest, t[i]);
  auto pair = [&](int mode) {
    d->mode = mode;
    d->pair_bit = lowest;
    d->flip = tmask;
  };
  auto diag = [&](uint64_t cond, uint64_t zmask, std::complex<double> d0, std::complex<double> d1) {
    d->mode = BLOCK_DIAG;
    d->cond = cond;
    d->zmask = zmask;
    put(d->u, 0, d0);
    put(d->u, 1, d1);
  };
  switch (code) {
    case BRK_OP_H: pair(BLOCK_H); break;
    case BRK_OP_X: pair(BLOCK_SWAP); break;
    case BRK_OP_Y:
      pair(BLOCK_MATRIX);
      coeffs_pauli_y(nt, d->u);
      d->sign = tmask;
      break;
    case BRK_OP_Z: diag(tmask | cmask, 0, {-1.0, 0.0}, {-1.0, 0.0}); break;
    case BRK_OP_U1: {
      const std::complex<double> e = exp_i(p[0]);
      diag(tmask | cmask, 0, e, e);
      break;
    }
    case BRK_OP_PHASE_COEFF: diag(tmask | cmask, 0, {p[0], p[1]}, {p[0], p[1]}); break;
    case BRK_OP_U2: pair(BLOCK_MATRIX); coeffs_u2(p[0], p[1], d->u); break;
    case BRK_OP_U3: pair(BLOCK_MATRIX); coeffs_u3(p[0], p[1], p[2], d->u); break;
    case BRK_OP_MATRIX: pair(BLOCK_MATRIX); std::copy(p, p + 8, d->u); break;
    case BRK_OP_EX: pair(BLOCK_MATRIX); coeffs_exp_x(exp_i(p[0]), d->u); break;
    case BRK_OP_RX: pair(BLOCK_MATRIX); coeffs_exp_x(exp_i(-0.5 * p[0]), d->u); break;
    case BRK_OP_EY:
    case BRK_OP_RY:
      pair(BLOCK_MATRIX);
      coeffs_exp_y(exp_i(code == BRK_OP_EY ? p[0] : -0.5 * p[0]), nt, d->u);
      d->sign = tmask;
      break;
    case BRK_OP_EZ:
    case BRK_OP_RZ: {
      const std::complex<double> e = exp_i(code == BRK_OP_EZ ? p[0] : -0.5 * p[0]);
      diag(cmask, tmask, e, std::conj(e));
      break;
    }
    case BRK_OP_SWAP:
      // pairs ...1(t0)...0(t1)... <-> ...0(t0)...1(t1)...: t1 is the inserted bit, t0 must be set
      d->mode = BLOCK_SWAP;
      d->pair_bit = t[1];
      d->flip = tmask;
      d->cond = cmask | (1ull << t[0]);
      break;
    case BRK_OP_ESWAP:
      pair(BLOCK_ESWAP);
      d->sign = 1ull << std::max(t[0], t[1]);
      coeffs_exp_swap(exp_i(p[0]), d->u);
      break;
    default: return fail(BRK_ERR_UNSUPPORTED, "op %d has no gate description", code);
  }
  return BRK_OK;
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
                           false, stream);
}

int brk_controlled_not(void* first, uint64_t size, int dtype, unsigned target, const unsigned* controls,
                       unsigned num_controls, void* stream) {
  return single_target(first, size, dtype, target, controls, num_controls, MODE_SWAP, nullptr, stream);
}

int brk_pauli_y(void* first, uint64_t size, int dtype, const unsigned* targets, unsigned num_targets,
                const unsigned* controls, unsigned num_controls, void* stream) {
  double u[8];
  coeffs_pauli_y(num_targets, u);
  return multi_target_flip(first, size, dtype, targets, num_targets, controls, num_controls, MODE_MATRIX, u, true,
                           stream);
}

int brk_pauli_z(void* first, uint64_t size, int dtype, const unsigned* qubits, unsigned num_qubits, void* stream) {
  if (num_qubits == 0) return fail(BRK_ERR_INVALID_ARGUMENT, "no qubits");
  validated v;
  if (int st = validate(first, size, dtype, nullptr, 0, qubits, num_qubits, &v)) return st;
  return diagonal(first, size, dtype, qubits, num_qubits, 0, {-1.0, 0.0}, {-1.0, 0.0}, stream);
}

}  // extern "C"

namespace brk {

// Targets then controls in one list (the qubits of a diagonal acting where all are 1).
static unsigned join_qubits(const unsigned* t, unsigned nt, const unsigned* c, unsigned nc,
                            unsigned (&all)[2 * BRK_MAX_OPERATED_QUBITS]) {
  unsigned k = 0;
  for (unsigned i = 0; i < nt; ++i) all[k++] = t[i];
  for (unsigned i = 0; i < nc; ++i) all[k++] = c[i];
  return k;
}

// C..C(sZ_t1 ... sZ_tm) and its adjoint: i^f (or (-i)^f), f = number of set target bits, where
// every control is 1 (ket/include/ket/gate/sqrt_pauli_z.hpp:565-700 and :846-1000: the
// coefficient by f mod 4; one target: i on the all-ones subspace, :428-460).
static int sqrt_z_string(void* first, uint64_t size, int dtype, const unsigned* targets, unsigned nt,
                         const unsigned* controls, unsigned nc, bool adjoint, void* stream) {
  if (nt == 0) return fail(BRK_ERR_INVALID_ARGUMENT, "no target qubits");
  if (!targets) return fail(BRK_ERR_INVALID_ARGUMENT, "target list is null");
  validated v;
  if (int st = validate(first, size, dtype, targets, nt, controls, nc, &v)) return st;
  const std::complex<double> i{0.0, adjoint ? -1.0 : 1.0};
  if (nt == 1) {
    unsigned all[2 * BRK_MAX_OPERATED_QUBITS];
    return diagonal(first, size, dtype, all, join_qubits(targets, nt, controls, nc, all), 0, i, i, stream);
  }
  const std::complex<double> d[4] = {{1.0, 0.0}, i, {-1.0, 0.0}, -i};
  return diagonal_mod4(first, size, dtype, controls, nc, mask_of(targets, nt), d, stream);
}

}  // namespace brk

extern "C" {

int brk_pauli_zn(void* first, uint64_t size, int dtype, const unsigned* targets, unsigned num_targets,
                 const unsigned* controls, unsigned num_controls, void* stream) {
  if (num_targets == 0) return fail(BRK_ERR_INVALID_ARGUMENT, "no target qubits");
  if (!targets) return fail(BRK_ERR_INVALID_ARGUMENT, "target list is null");
  validated v;
  if (int st = validate(first, size, dtype, targets, num_targets, controls, num_controls, &v)) return st;
  if (num_targets == 1) {  // CnZ: -1 where the target and every control are 1 (pauli_z.hpp:352-390)
    unsigned all[2 * BRK_MAX_OPERATED_QUBITS];
    const unsigned k = join_qubits(targets, num_targets, controls, num_controls, all);
    return diagonal(first, size, dtype, all, k, 0, {-1.0, 0.0}, {-1.0, 0.0}, stream);
  }
  // C..C(Z_t1 ... Z_tm): -1 on odd parity of the targets (pauli_z.hpp:392-470)
  return diagonal(first, size, dtype, controls, num_controls, mask_of(targets, num_targets), {1.0, 0.0},
                  {-1.0, 0.0}, stream);
}

int brk_sqrt_pauli_zn(void* first, uint64_t size, int dtype, const unsigned* targets, unsigned num_targets,
                      const unsigned* controls, unsigned num_controls, void* stream) {
  return sqrt_z_string(first, size, dtype, targets, num_targets, controls, num_controls, false, stream);
}

int brk_adj_sqrt_pauli_zn(void* first, uint64_t size, int dtype, const unsigned* targets, unsigned num_targets,
                          const unsigned* controls, unsigned num_controls, void* stream) {
  return sqrt_z_string(first, size, dtype, targets, num_targets, controls, num_controls, true, stream);
}

int brk_phase_shift_coeff(void* first, uint64_t size, int dtype, const double coeff[2], const unsigned* qubits,
                          unsigned num_qubits, void* stream) {
  if (!coeff) return fail(BRK_ERR_INVALID_ARGUMENT, "coeff is null");
  validated v;
  if (int st = validate(first, size, dtype, nullptr, 0, qubits, num_qubits, &v)) return st;
  const std::complex<double> c{coeff[0], coeff[1]};
  return diagonal(first, size, dtype, qubits, num_qubits, 0, c, c, stream);
}

int brk_phase_shift(void* first, uint64_t size, int dtype, double phase, const unsigned* qubits,
                    unsigned num_qubits, void* stream) {
  const std::complex<double> c = exp_i(phase);
  const double cc[2] = {c.real(), c.imag()};
  return brk_phase_shift_coeff(first, size, dtype, cc, qubits, num_qubits, stream);
}

int brk_phase_shift2(void* first, uint64_t size, int dtype, double phase1, double phase2, unsigned target,
                     const unsigned* controls, unsigned num_controls, void* stream) {
  double u[8];
  coeffs_u2(phase1, phase2, u);
  return single_target(first, size, dtype, target, controls, num_controls, MODE_MATRIX, u, stream);
}

int brk_phase_shift3(void* first, uint64_t size, int dtype, double phase1, double phase2, double phase3,
                     unsigned target, const unsigned* controls, unsigned num_controls, void* stream) {
  double u[8];
  coeffs_u3(phase1, phase2, phase3, u);
  return single_target(first, size, dtype, target, controls, num_controls, MODE_MATRIX, u, stream);
}

}  // extern "C"

namespace brk {

static int exp_x(void* first, uint64_t size, int dtype, std::complex<double> e, const unsigned* targets,
                 unsigned num_targets, const unsigned* controls, unsigned num_controls, void* stream) {
  double u[8];
  coeffs_exp_x(e, u);
  return multi_target_flip(first, size, dtype, targets, num_targets, controls, num_controls, MODE_MATRIX, u, false,
                           stream);
}

static int exp_y(void* first, uint64_t size, int dtype, std::complex<double> e, const unsigned* targets,
                 unsigned num_targets, const unsigned* controls, unsigned num_controls, void* stream) {
  double u[8];
  coeffs_exp_y(e, num_targets, u);
  return multi_target_flip(first, size, dtype, targets, num_targets, controls, num_controls, MODE_MATRIX, u, true,
                           stream);
}

// exp(i phase Z..Z): e^{i phase} on even parity of the target bits, e^{-i phase} on odd
// (ket/include/ket/gate/exponential_pauli_z.hpp:36-38, :73-74).
static int exp_z(void* first, uint64_t size, int dtype, std::complex<double> e, const unsigned* targets,
                 unsigned num_targets, const unsigned* controls, unsigned num_controls, void* stream) {
  if (num_targets == 0) return fail(BRK_ERR_INVALID_ARGUMENT, "no target qubits");
  validated v;
  if (int st = validate(first, size, dtype, targets, num_targets, controls, num_controls, &v)) return st;
  return diagonal(first, size, dtype, controls, num_controls, mask_of(targets, num_targets), e, std::conj(e),
                  stream);
}

static int exp_swap(void* first, uint64_t size, int dtype, std::complex<double> e, unsigned qubit1, unsigned qubit2,
                    const unsigned* controls, unsigned num_controls, void* stream) {
  const unsigned t[2] = {qubit1, qubit2};
  validated v;
  if (int st = validate(first, size, dtype, t, 2, controls, num_controls, &v)) return st;
  const unsigned lo = std::min(qubit1, qubit2), hi = std::max(qubit1, qubit2);
  plan p;
  p.insert(lo);
  for (unsigned i = 0; i < num_controls; ++i) p.control(controls[i], line_bits(dtype));
  gate_args a = make_args(p);
  a.flipmask = (1ull << lo) | (1ull << hi);
  a.signmask = 1ull << hi;
  coeffs_exp_swap(e, a.u);
  return launch(dtype, FAM_PAIR, MODE_ESWAP, first, size, size >> p.np, a, stream);
}

static std::complex<double> coefficient(const double* c) { return {c[0], c[1]}; }

}  // namespace brk

extern "C" {

int brk_exponential_pauli_x(void* first, uint64_t size, int dtype, double phase, const unsigned* targets,
                            unsigned num_targets, const unsigned* controls, unsigned num_controls, void* stream) {
  return exp_x(first, size, dtype, exp_i(phase), targets, num_targets, controls, num_controls, stream);
}

int brk_exponential_pauli_y(void* first, uint64_t size, int dtype, double phase, const unsigned* targets,
                            unsigned num_targets, const unsigned* controls, unsigned num_controls, void* stream) {
  return exp_y(first, size, dtype, exp_i(phase), targets, num_targets, controls, num_controls, stream);
}

int brk_exponential_pauli_z(void* first, uint64_t size, int dtype, double phase, const unsigned* targets,
                            unsigned num_targets, const unsigned* controls, unsigned num_controls, void* stream) {
  return exp_z(first, size, dtype, exp_i(phase), targets, num_targets, controls, num_controls, stream);
}

int brk_exponential_pauli_x_coeff(void* first, uint64_t size, int dtype, const double coeff[2],
                                  const unsigned* targets, unsigned num_targets, const unsigned* controls,
                                  unsigned num_controls, void* stream) {
  if (!coeff) return fail(BRK_ERR_INVALID_ARGUMENT, "coeff is null");
  return exp_x(first, size, dtype, coefficient(coeff), targets, num_targets, controls, num_controls, stream);
}

int brk_exponential_pauli_y_coeff(void* first, uint64_t size, int dtype, const double coeff[2],
                                  const unsigned* targets, unsigned num_targets, const unsigned* controls,
                                  unsigned num_controls, void* stream) {
  if (!coeff) return fail(BRK_ERR_INVALID_ARGUMENT, "coeff is null");
  return exp_y(first, size, dtype, coefficient(coeff), targets, num_targets, controls, num_controls, stream);
}

int brk_exponential_pauli_z_coeff(void* first, uint64_t size, int dtype, const double coeff[2],
                                  const unsigned* targets, unsigned num_targets, const unsigned* controls,
                                  unsigned num_controls, void* stream) {
  if (!coeff) return fail(BRK_ERR_INVALID_ARGUMENT, "coeff is null");
  return exp_z(first, size, dtype, coefficient(coeff), targets, num_targets, controls, num_controls, stream);
}

int brk_swap(void* first, uint64_t size, int dtype, unsigned qubit1, unsigned qubit2, const unsigned* controls,
             unsigned num_controls, void* stream) {
  const unsigned t[2] = {qubit1, qubit2};
  validated v;
  if (int st = validate(first, size, dtype, t, 2, controls, num_controls, &v)) return st;
  // pairs ...1(q1)...0(q2)... <-> ...0(q1)...1(q2)... (ket/include/ket/gate/swap.hpp:30-32)
  const unsigned lo = std::min(qubit1, qubit2), hi = std::max(qubit1, qubit2);
  plan p;
  gate_args a;
  if (v.nq >= LANE_BITS && hi < LANE_BITS && lo < lane_target_bits(dtype)) {
    for (unsigned i = 0; i < num_controls; ++i) p.control(controls[i], LANE_BITS);
    a = make_args(p);
    a.swapmask = (1ull << lo) | (1ull << hi);
    return launch(dtype, FAM_LANE_SWAP2, MODE_SWAP, first, size, size >> p.np, a, stream);
  }
  if (v.nq > LANE_BITS && lo < lane_target_bits(dtype) && hi >= LANE_BITS) {
    p.insert(hi);
    for (unsigned i = 0; i < num_controls; ++i) p.control(controls[i], LANE_BITS);
    a = make_args(p);
    a.flipmask = 1ull << hi;
    a.lane_bit = (int)lo;
    return launch(dtype, FAM_LANE_SWAP, MODE_SWAP, first, size, size >> p.np, a, stream);
  }
  p.insert(qubit1);
  p.insert(qubit2);
  p.setmask |= 1ull << qubit1;
  for (unsigned i = 0; i < num_controls; ++i) p.control(controls[i], line_bits(dtype));
  a = make_args(p);
  a.flipmask = (1ull << qubit1) | (1ull << qubit2);
  return launch(dtype, FAM_PAIR, MODE_SWAP, first, size, size >> p.np, a, stream);
}

int brk_exponential_swap(void* first, uint64_t size, int dtype, double phase, unsigned qubit1, unsigned qubit2,
                         const unsigned* controls, unsigned num_controls, void* stream) {
  return exp_swap(first, size, dtype, exp_i(phase), qubit1, qubit2, controls, num_controls, stream);
}

int brk_exponential_swap_coeff(void* first, uint64_t size, int dtype, const double coeff[2], unsigned qubit1,
                               unsigned qubit2, const unsigned* controls, unsigned num_controls, void* stream) {
  if (!coeff) return fail(BRK_ERR_INVALID_ARGUMENT, "coeff is null");
  return exp_swap(first, size, dtype, coefficient(coeff), qubit1, qubit2, controls, num_controls, stream);
}

int brk_swap_permutated_local_qubits(void* first, uint64_t size, int dtype, unsigned permutated_qubit1,
                                     unsigned permutated_qubit2, void* stream) {
  // With one data block the run-wise swap_ranges of the reference
  // (swap_permutated_local_qubits.hpp:64-77) exchanges exactly the amplitudes a SWAP
  // of the two physical qubits exchanges.
  return brk_swap(first, size, dtype, permutated_qubit1, permutated_qubit2, nullptr, 0, stream);
}

}  // extern "C"
