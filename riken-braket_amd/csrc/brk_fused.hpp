// Launch descriptor of the fused-block kernel (precision form), shared by the
// host planner (host/fusion.cpp) and the kernel translation unit.
#pragma once

#include <cstdint>

namespace brk {

constexpr int FUSED_COLUMNS = 32;  // tile columns: the 5 lowest qubits outside the block
constexpr int FUSED_MAX_ROWS = 64; // 2^6 block-row values

// A tile is the 2^k x 32 complex matrix X[row][col]: row = value of the k
// block qubits, col = value of the 5 lowest qubits outside the block, for one
// value of every other qubit (the tile number, "outer" index).  Element
// (row, col) of tile t lives at
//     expand(t) | row_bits[row] | col_bits[col]
// where expand() inserts zeros at the k + 5 tile positions and row_bits /
// col_bits deposit the row / column value into the block / column positions.
struct fused_launch {
  int width;                           // block qubits k (4..6)
  uint64_t outer_masks[17];            // expand masks inserting zeros at the k + 5 tile bits
  int outer_k;                         // = k + 5
  uint64_t ntiles;                     // 2^(n - k - 5)
  uint64_t row_bits[FUSED_MAX_ROWS];   // deposit of each block-row value
  uint64_t col_bits[FUSED_COLUMNS];    // deposit of each column value
  const float* matrix;                 // device: real embedding [[Ur, -Ui], [Ui, Ur]], transposed, swizzled
  int chunked;                         // tile order (set at launch from the tuning)

  // Staged form (k <= 5): a workgroup tile is 2^k rows x 128 columns (the 7
  // lowest qubits outside the block), moved between HBM and LDS in address
  // order.  Tile bit i (ascending qubit order) is qubit stile_qubit[i]; in the
  // LDS tile, stored as [row][128 columns], it is index bit stile_lds_bit[i]
  // (column bit c -> c, block-row bit r -> 7 + r).
  uint64_t souter_masks[17];           // expand masks inserting zeros at the k + 7 tile bits
  int souter_k;                        // = k + 7
  uint64_t sntiles;                    // 2^(n - k - 7); 0 when n < k + 7
  uint8_t stile_qubit[16];
  uint8_t stile_lds_bit[16];
};

constexpr int FUSED_STAGED_COLUMNS = 128;

// Launches one block (state is complex64).  Defined in csrc/brk_fused.hip.
int launch_fused_block(const fused_launch& L, void* state, void* stream);

// Chunk passes: several blocks applied to one chunk of 2^13 complex64 amplitudes
// (64 KiB) held in LDS, one HBM read and one write per pass.  Local bit r of a
// chunk index is physical qubit pos[r] (ascending); the chunk number, expanded
// with zeros at those positions (outer_masks), gives the other bits.  Op o of
// the pass applies block matrix `matrix` (index into the plan's embeddings) on
// the local bits row_bits[0..k) (block-row bit j: the block's j-th lowest
// qubit); the remaining local bits, ascending, carry the column index.
constexpr int FUSED_CHUNK_LOCAL = 13;
constexpr int FUSED_CHUNK_THREADS = 512;
constexpr int FUSED_CHUNK_MAX_OPS = 64;
struct fused_chunk_op {
  uint32_t matrix;
  uint32_t row_bits[6];
  uint32_t col_bits[9];  // 13 - k column bits (k = 4: 9, k = 5: 8)
};
struct fused_chunk_program {
  uint32_t num_ops;
  uint32_t pos[FUSED_CHUNK_LOCAL];
  uint64_t outer_masks[17];           // expand masks inserting zeros at the local positions
  uint64_t deposit_high[16];          // physical offset of local index k << 9 (k < 16)
  fused_chunk_op ops[FUSED_CHUNK_MAX_OPS];
};
// The blocks of one pass at width k: the pass's three-product forms (3 x 2^k x 2^k FP32) stay
// in 32 KiB of LDS beside two chunk buffers: 8 at k = 4, 2 at k = 5.
constexpr int fused_chunk_max_blocks(int k) { return k == 4 ? 8 : k == 5 ? 2 : 0; }
// Launches one pass of width-k blocks over a complex64 state of `size` amplitudes:
// one persistent workgroup per CU.  `device_program` is the pass in device memory,
// `program` its host copy, `matrices` the plan's device three-product forms
// (fused_split_floats(width) per block).
int launch_fused_chunk(const fused_chunk_program* device_program, const fused_chunk_program& program, int width,
                       const float* matrices, void* state, uint64_t size, void* stream);

// Whether plans of width 4 and 5 run in chunk passes (tuning fused_chunks).
bool fused_chunks_enabled();

// Size in floats of a block's device matrix for width k.
inline uint64_t fused_matrix_floats(int k) { return (uint64_t)(2 << k) * (uint64_t)(2 << k); }
// Size in floats of a block's three-product form for chunk passes: Ur, Ui and Ur + Ui, each
// 2^k x 2^k, stored transposed (element (i, c) at c * 2^k + i).
inline uint64_t fused_split_floats(int k) { return 3ull << (2 * k); }

}  // namespace brk
