// Precision form: one fused k-qubit block (k = 4, 5, 6) applied to a
// complex64 state as a batched complex GEMM on MFMA.
//
// For every outer index the 2^k amplitudes of the block qubits form a column
// vector x; the block computes y = U x.  A tile holds 32 such columns (the 32
// values of the 5 lowest qubits outside the block), so it is a 2^k x 32
// complex matrix X and the tile update is Y = U X.  The complex product runs
// as one real GEMM with the real embedding
//     [Yr]   [Ur  -Ui] [Xr]
//     [Yi] = [Ui   Ur] [Xi]        (2^(k+1) x 2^(k+1) times 2^(k+1) x 32)
// on v_mfma_f32_32x32x2_f32 (FP32 products, FP32 accumulation): 2^(k+1)/32
// output tiles of 32 x 32, 2^k MFMA steps of K = 2 each.
//
// Register-direct data flow (one wavefront = one tile at a time, no LDS for
// the state, no barriers in the tile loop):
//  * B operand.  MFMA step st needs B[kk][col] = real row 2*st + kk of
//    [Xr; Xi] in lane col + 32*kk.  Lane (col, kk) therefore loads the complex
//    elements X[2j + kk][col], j < 2^(k-1), straight from HBM into registers:
//    its real parts feed steps j, its imaginary parts steps j + 2^(k-1).  The
//    two half-waves read the two values of the lowest block qubit, so a load
//    instruction covers both halves of a row pair.
//  * A operand.  The block's embedded matrix sits in LDS (loaded once per
//    workgroup), transposed, odd columns XOR-swizzled by 32 so the two
//    half-waves read disjoint banks.
//  * D.  Lane (col, kk) ends with D[8(v/4) + 4kk + v%4 + 32t][col]; the real
//    and imaginary rows of the same complex row are in the same lane (tiles t
//    and t + DT/2, or registers v and v + 8 when DT = 1), so each result is
//    stored as one complex element, in place.
//
// MFMA 32x32x2 f32 operand layout (lane l, accumulator register v):
//   A[i][k] = lane (i + 32k), B[k][j] = lane (j + 32k),
//   D[8(v/4) + 4(l/32) + v%4][l%32] = acc[v] of lane l.
#include "brk_internal.hpp"
#include "brk_fused.hpp"

namespace brk {

typedef float f16v __attribute__((ext_vector_type(16)));

constexpr int FUSED_THREADS = 256;  // 4 wavefronts per workgroup, one tile each at a time

template <int KB>
__global__ __launch_bounds__(FUSED_THREADS, 2) void k_fused(cpx<float>* __restrict__ s, fused_launch L) {
  constexpr int R = 1 << KB;   // complex rows
  constexpr int W = 2 * R;     // real rows of the embedding
  constexpr int HALF = R / 2;  // complex elements per lane
  constexpr int DT = W / 32;   // 32-row output tiles
  extern __shared__ float A[];  // W*W floats: element (r, c) at c*W + (r ^ swizzle(c))
  for (int i = threadIdx.x; i < W * W; i += FUSED_THREADS) A[i] = L.matrix[i];
  __syncthreads();

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i32 = lane & 31, kk = lane >> 5;
  const int swz = (W >= 64 && kk) ? 32 : 0;
  const uint64_t col_bits = L.col_bits[i32];
  const uint64_t load_half = kk ? L.row_bits[1] : 0;   // row bit 0 = kk for the loads
  const uint64_t store_half = kk ? L.row_bits[4] : 0;  // row bit 2 = kk for the stores
  // Tile order: interleaved (wave g takes tiles g, g + G, g + 2G, ...; the
  // waves of the device sweep the state together) or chunked (wave g takes a
  // contiguous run of tiles).
  const uint64_t nwaves = (uint64_t)gridDim.x * (FUSED_THREADS / 64);
  const uint64_t gwave = (uint64_t)blockIdx.x * (FUSED_THREADS / 64) + wave;
  uint64_t stride = nwaves, tile = gwave, tile_end = L.ntiles;
  if (L.chunked) {
    const uint64_t chunk = (L.ntiles + nwaves - 1) / nwaves;
    stride = 1;
    tile = gwave * chunk;
    tile_end = std::min<uint64_t>(L.ntiles, tile + chunk);
  }

  // k <= 5: the next tile's loads are issued before this tile's MFMAs and
  // stores (two tiles in registers); k = 6 has no registers to spare for that.
  constexpr bool PIPELINE = KB <= 5;
  cpx<float> xn[HALF];
  if (PIPELINE && tile < tile_end) {
    const uint64_t b0 = expand_index<KB + 5>(tile, L.outer_masks, L.outer_k) | col_bits;
#pragma unroll
    for (int j = 0; j < HALF; ++j) xn[j] = load_amp<true>(s + (b0 | L.row_bits[2 * j] | load_half));
  }
  for (; tile < tile_end; tile += stride) {
    const uint64_t base = expand_index<KB + 5>(tile, L.outer_masks, L.outer_k) | col_bits;
    cpx<float> x[HALF];
    if constexpr (PIPELINE) {
#pragma unroll
      for (int j = 0; j < HALF; ++j) x[j] = xn[j];
      const uint64_t next = tile + stride;
      if (next < tile_end) {
        const uint64_t bn = expand_index<KB + 5>(next, L.outer_masks, L.outer_k) | col_bits;
#pragma unroll
        for (int j = 0; j < HALF; ++j) xn[j] = load_amp<true>(s + (bn | L.row_bits[2 * j] | load_half));
      }
    } else {
#pragma unroll
      for (int j = 0; j < HALF; ++j) x[j] = load_amp<true>(s + (base | L.row_bits[2 * j] | load_half));
    }

    f16v acc[DT];
#pragma unroll
    for (int t = 0; t < DT; ++t) acc[t] = f16v{};
#pragma unroll
    for (int st = 0; st < R; ++st) {
      const float b = st < HALF ? x[st].re : x[st - HALF].im;
#pragma unroll
      for (int t = 0; t < DT; ++t) {
        const float a = A[(2 * st + kk) * W + ((32 * t + i32) ^ swz)];
        acc[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc[t], 0, 0, 0);
      }
      // k = 6: keep the A-fragment reads of a step with their MFMAs (hoisting all
      // 256 of them would exceed the register budget)
      if constexpr (KB == 6) __builtin_amdgcn_sched_barrier(0);
    }

    if constexpr (DT == 1) {
      // rows 0..15 real parts (v < 8), rows 16..31 imaginary parts (v + 8)
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        const int row = 8 * (v >> 2) + (v & 3);  // + 4*kk through store_half
        store_amp<true>(s + (base | L.row_bits[row] | store_half), cpx<float>{acc[0][v], acc[0][v + 8]});
      }
    } else {
#pragma unroll
      for (int t = 0; t < DT / 2; ++t) {
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const int row = 32 * t + 8 * (v >> 2) + (v & 3);
          store_amp<true>(s + (base | L.row_bits[row] | store_half), cpx<float>{acc[t][v], acc[t + DT / 2][v]});
        }
      }
    }
  }
}

template <int KB>
static int launch_k(const fused_launch& L, void* state, void* stream) {
  constexpr int W = 2 << KB;
  const size_t lds = sizeof(float) * (size_t)W * W;
  static bool configured = false;
  if (!configured) {
    // above 64 KiB of dynamic LDS a kernel has to opt in
    BRK_HIP_CHECK(hipFuncSetAttribute((const void*)k_fused<KB>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds));
    configured = true;
  }
  const uint64_t groups = (L.ntiles + 3) / 4;
  const unsigned grid = (unsigned)std::min<uint64_t>(groups, (uint64_t)get_tuning().fused_groups);
  fused_launch Lc = L;
  Lc.chunked = get_tuning().fused_chunked;
  hipLaunchKernelGGL(k_fused<KB>, dim3(grid), dim3(FUSED_THREADS), lds, as_stream(stream), (cpx<float>*)state, Lc);
  return check_launch("fused block");
}

int launch_fused_block(const fused_launch& L, void* state, void* stream) {
  switch (L.width) {
    case 4: return launch_k<4>(L, state, stream);
    case 5: return launch_k<5>(L, state, stream);
    case 6: return launch_k<6>(L, state, stream);
    default: return fail(BRK_ERR_UNSUPPORTED, "fused block width %d not in 4..6", L.width);
  }
}

}  // namespace brk
