// Gate blocks: runs of consecutive gates applied to one LDS-resident chunk of
// the state per workgroup.
//
// A gate is first described in physical-qubit terms (gate_desc) by the same
// code that computes the plain kernels' coefficients, so a block applies each
// gate with the plain path's arithmetic and gives bit-identical amplitudes.
// The planner (host/block.cpp) turns the descriptions of one block into a
// block_program over local bit positions; the kernel (brk_gates.hip) loads
// the 2^L amplitudes of a chunk, applies the program in LDS and stores them.
// Plain C++ only: included by host code compiled without HIP.
#pragma once

#include <cstdint>

namespace brk {

constexpr int BLOCK_LDS_BYTES = 65536;            // largest chunk: 64 KiB of LDS (two workgroups per CU)
constexpr int BLOCK_THREADS = 512;
constexpr int BLOCK_MAX_LOCAL_C128 = 12;          // 2^12 x 16 B = 64 KiB
constexpr int BLOCK_MAX_LOCAL_C64 = 13;           // 2^13 x 8 B  = 64 KiB
constexpr int BLOCK_MAX_OPS = 64;
constexpr int BLOCK_THREAD_BITS = 9;              // log2(BLOCK_THREADS)

// modes of a described gate: the pair modes of the plain kernels, or diagonal
enum block_mode { BLOCK_H = 0, BLOCK_SWAP = 1, BLOCK_MATRIX = 2, BLOCK_ESWAP = 3, BLOCK_DIAG = 4 };

// One gate in physical-qubit terms.  Pair modes act on (i0, i0 ^ flip) for
// every i0 with bit pair_bit clear and all `cond` bits set; diagonal mode
// multiplies every amplitude with all `cond` bits set by u[0..1] (even parity
// of i & zmask) or u[2..3] (odd).  `sign` is the matrix mode's sign-flip
// parity mask and ESWAP's higher-qubit bit.
struct gate_desc {
  int mode;
  unsigned pair_bit;
  uint64_t flip, cond, sign, zmask;
  uint64_t qubits;  // every qubit the gate reads or writes
  double u[8];
};

// Describes an op (physical qubits; op codes of brk_op_code).  Returns
// BRK_OK, or BRK_ERR_UNSUPPORTED for ops without a description (local swap).
int describe_physical_op(int code, const unsigned* targets, unsigned num_targets, const unsigned* controls,
                         unsigned num_controls, const double* params, gate_desc* out);

// A gate of a block with its masks over local bit positions (bit r of a
// chunk index <-> physical qubit pos[r]).  The gate visits only the local
// indices it acts on: item v -> j = v with zeros inserted at ins[0..nins)
// (ascending: the pair bit and the cond bits), then the cond bits set.
struct block_op {
  int32_t mode;
  uint32_t pair_bit, flip, cond, sign, zmask;
  uint32_t nins;
  uint32_t ins[15];
  uint32_t group_flip, group_pair;  // register segments: flip and pair bit over the group index
  uint32_t negate;                   // diagonal with u = (-1, -1) (Z, CZ): an exact negation on cond
  uint32_t group_zmask;              // register segments: zmask over the group index, ~0 if not inside
  double u[8];
};

// A pass over the chunk in LDS.  A segment pass gives every thread one group of
// 2^(L - 9) amplitudes -- local indices that differ only in the bits group[0..]
// -- in registers and applies several gates to it: diagonal gates, and pair
// gates whose flipped bits are group bits.  A single pass applies one gate
// with the strided loop.
enum block_pass_kind { BLOCK_PASS_SEGMENT = 0, BLOCK_PASS_SINGLE = 1 };
struct block_pass {
  int32_t kind;
  uint32_t first_op, num_ops;
  uint32_t group[4];  // ascending local bits of the group (L - 9 of them)
};

struct block_program {
  uint32_t num_ops;
  uint32_t num_local;                   // L
  uint32_t num_passes;
  uint32_t pos[16];                     // ascending physical qubits of the local bits
  uint64_t deposit_high[16];            // physical offset of local index k << BLOCK_THREAD_BITS
  block_pass passes[BLOCK_MAX_OPS];
  block_op ops[BLOCK_MAX_OPS];
};

// Reverse sweep of the adjoint method in blocks: chunks of psi and lambda are
// held in LDS together (2^11 complex<double> or 2^12 complex<float>
// amplitudes each, 32 KiB apiece) while the block's ops -- the inverses of
// the circuit's ops, in reverse order -- are un-applied to both.  Before a
// parametric op (one target, no controls) is un-applied, its term
// 2 Re <lambda| D |psi> is summed over the chunk into
// partials[grad_slot * num_chunks + chunk] (FP64, fixed order).
constexpr int ADJ_BLOCK_LOCAL_C128 = 11;
constexpr int ADJ_BLOCK_LOCAL_C64 = 12;
constexpr int ADJ_SEGMENT_GRADIENTS = 8;  // parametric ops per register segment
// Forms of a parametric op's arithmetic, from the exact zeros of D and of the inverse op's
// 2x2: the general formulas with their zero terms left out (the same values).
//   Z: D = diag(i a, i b), the inverse op diagonal on the parity of the target
//      (RZ, EZ) or on its 1 side (U1);
//   Y: D = [[0, a], [b, 0]] real, the inverse op a real 2x2 (RY, EY);
//   X: D = [[0, i a], [i b, 0]], the inverse op [[r, i s], [i s', r']] (RX, EX).
enum adjoint_grad_kind { ADJ_GRAD_GENERAL = 0, ADJ_GRAD_Z = 1, ADJ_GRAD_Y = 2, ADJ_GRAD_X = 3 };
struct adjoint_block_program {
  block_program base;                  // the inverse ops; their passes
  int32_t grad_slot[BLOCK_MAX_OPS];    // -1: no gradient term
  int32_t grad_kind[BLOCK_MAX_OPS];    // adjoint_grad_kind of a parametric op
  double d[BLOCK_MAX_OPS][8];          // D = (dU/dtheta) U^dagger, (re, im) of d00, d01, d10, d11
};
int launch_adjoint_block(const adjoint_block_program* device_program, const adjoint_block_program& program,
                         void* psi, void* lam, uint64_t size, int dtype, double* partials, void* stream);
// out[p] = scale * (sum over c of partials[p * chunks + c]) for p < rows, on the device in a
// fixed order (one workgroup per row: strided per-thread sums, then a tree).
int sum_gradient_partials(const double* partials, unsigned rows, uint64_t chunks, double scale, double* out,
                          void* stream);

// Whether the adjoint gradient runs in blocks (tuning adjoint_blocks).
bool adjoint_blocks_enabled();

// A block_program for the described gates (physical qubits of an n-qubit state):
// local qubits are their qubits, the line_bits lowest qubits, then the lowest others,
// num_local in all.
block_program make_block_program(const gate_desc* descs, unsigned count, unsigned n, unsigned num_local,
                                 unsigned line_bits);

// Local qubits of a chunk for this dtype under the current tuning (block_lds).
unsigned block_max_local(int dtype);
// Number of lowest qubits every chunk keeps local (contiguous runs of block_run_bytes).
unsigned block_run_qubits(int dtype);

// Launches one block over a state of `size` amplitudes: 2^(n - L) workgroups.
// `device_program` points to a block_program in device memory whose copy on
// the host is `program`.
int launch_block(const block_program* device_program, const block_program& program, void* first, uint64_t size,
                 int dtype, void* stream);

}  // namespace brk
