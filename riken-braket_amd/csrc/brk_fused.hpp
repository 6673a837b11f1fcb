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

// Size in floats of a block's device matrix for width k.
inline uint64_t fused_matrix_floats(int k) { return (uint64_t)(2 << k) * (uint64_t)(2 << k); }

}  // namespace brk
