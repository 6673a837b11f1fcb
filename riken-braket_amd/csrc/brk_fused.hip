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
// Loads and stores are ordinary (temporal): when the block holds some of qubits
// 0..2 each store instruction covers only part of every 64-byte line, and the
// L2 merges those pieces before write-back (non-temporal stores went to memory
// as partial lines and cost up to 2.7x the write traffic).
//
// MFMA 32x32x2 f32 operand layout (lane l, accumulator register v):
//   A[i][k] = lane (i + 32k), B[k][j] = lane (j + 32k),
//   D[8(v/4) + 4(l/32) + v%4][l%32] = acc[v] of lane l.
#include <mutex>
#include <set>
#include <utility>

#include "brk_internal.hpp"
#include "brk_fused.hpp"

namespace brk {

typedef float f16v __attribute__((ext_vector_type(16)));

constexpr int FUSED_THREADS = 256;  // 4 wavefronts per workgroup, one tile each at a time

// OCC = workgroups per CU the register budget is cut for: 2 (software-pipelined,
// the next tile's loads in flight during this tile's MFMAs) or 4 (no pipelining,
// more wavefronts per SIMD hide the memory latency instead).
template <int KB, int OCC>
__global__ __launch_bounds__(FUSED_THREADS, OCC) void k_fused(cpx<float>* __restrict__ s, fused_launch L) {
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
  constexpr bool PIPELINE = KB <= 5 && OCC == 2;
  cpx<float> xn[HALF];
  if (PIPELINE && tile < tile_end) {
    const uint64_t b0 = expand_index<KB + 5>(tile, L.outer_masks, L.outer_k) | col_bits;
#pragma unroll
    for (int j = 0; j < HALF; ++j) xn[j] = load_amp<false>(s + (b0 | L.row_bits[2 * j] | load_half));
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
        for (int j = 0; j < HALF; ++j) xn[j] = load_amp<false>(s + (bn | L.row_bits[2 * j] | load_half));
      }
    } else {
#pragma unroll
      for (int j = 0; j < HALF; ++j) x[j] = load_amp<false>(s + (base | L.row_bits[2 * j] | load_half));
    }

    f16v acc[DT];
#pragma unroll
    for (int t = 0; t < DT; ++t) acc[t] = f16v{};
    // OCC = 4: read the matrix from LDS in every tile instead of holding all of it in
    // registers (an opaque row offset keeps the reads inside the loop)
    int arow = 0;
    if constexpr (OCC == 4) asm volatile("" : "+v"(arow));
#pragma unroll
    for (int st = 0; st < R; ++st) {
      const float b = st < HALF ? x[st].re : x[st - HALF].im;
#pragma unroll
      for (int t = 0; t < DT; ++t) {
        const float a = A[(2 * st + kk) * W + arow + ((32 * t + i32) ^ swz)];
        acc[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc[t], 0, 0, 0);
      }
      // k = 6: keep the A-fragment reads of a step with their MFMAs (hoisting all
      // 256 of them would exceed the register budget)
      if constexpr (KB == 6 || OCC == 4) __builtin_amdgcn_sched_barrier(0);
    }

    if constexpr (DT == 1) {
      // rows 0..15 real parts (v < 8), rows 16..31 imaginary parts (v + 8)
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        const int row = 8 * (v >> 2) + (v & 3);  // + 4*kk through store_half
        store_amp<false>(s + (base | L.row_bits[row] | store_half), cpx<float>{acc[0][v], acc[0][v + 8]});
      }
    } else {
#pragma unroll
      for (int t = 0; t < DT / 2; ++t) {
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const int row = 32 * t + 8 * (v >> 2) + (v & 3);
          store_amp<false>(s + (base | L.row_bits[row] | store_half), cpx<float>{acc[t][v], acc[t + DT / 2][v]});
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Staged form (k = 4, 5): one 256-thread workgroup owns a tile of 2^k rows x
// 128 columns (the 7 lowest qubits outside the block).  HBM <-> LDS copies
// run in address order, 16 bytes per thread per access, so every instruction
// covers whole 64-byte lines (qubits 0..6 are always tile bits); wave w reads
// the B operands of columns 32w..32w+31 from LDS, runs the same MFMA chain as
// the register-direct kernel, and writes its D rows back over its own
// columns.  The next tile's loads are in flight during the MFMAs.  Address
// order <-> LDS position: both are sums over the set bits of the address-order
// index e, looked up in two 64-entry tables each (bits 0..5 and 6..11 of e).
// ---------------------------------------------------------------------------
constexpr int STAGED_THREADS = 256;

struct f4 {
  float x, y, z, w;
} __attribute__((aligned(16)));
typedef float v4f __attribute__((ext_vector_type(4)));

// Whole-line traffic streams past the caches (non-temporal, 16 bytes per lane).
__device__ __forceinline__ f4 load_f4(const cpx<float>* p) {
  const v4f v = __builtin_nontemporal_load(reinterpret_cast<const v4f*>(p));
  return f4{v.x, v.y, v.z, v.w};
}
__device__ __forceinline__ void store_f4(cpx<float>* p, f4 a) {
  const v4f v = {a.x, a.y, a.z, a.w};
  __builtin_nontemporal_store(v, reinterpret_cast<v4f*>(p));
}

// Two consecutive complex64 amplitudes as one 16-byte access (non-temporal for large states).
template <bool NT>
__device__ __forceinline__ v4f load_pair(const cpx<float>* p) {
  if constexpr (NT) return __builtin_nontemporal_load(reinterpret_cast<const v4f*>(p));
  else return *reinterpret_cast<const v4f*>(p);
}
template <bool NT>
__device__ __forceinline__ void store_pair(cpx<float>* p, v4f v) {
  if constexpr (NT) __builtin_nontemporal_store(v, reinterpret_cast<v4f*>(p));
  else *reinterpret_cast<v4f*>(p) = v;
}

template <int KB>
__global__ __launch_bounds__(STAGED_THREADS, 3) void k_fused_staged(cpx<float>* __restrict__ s, fused_launch L) {
  constexpr int R = 1 << KB, W = 2 * R, HALF = R / 2, DT = W / 32;
  constexpr int COLS = FUSED_STAGED_COLUMNS;
  constexpr int ELEMS = R * COLS;
  constexpr int PIECES = ELEMS / 2 / STAGED_THREADS;  // 16-byte accesses per thread per tile
  extern __shared__ float smem[];
  float* A = smem;                                            // W*W floats (layout of the register-direct kernel)
  cpx<float>* buf = reinterpret_cast<cpx<float>*>(smem + W * W);  // the tile, [row][COLS]
  uint32_t* l_lo = reinterpret_cast<uint32_t*>(buf + ELEMS);  // LDS index of bits 0..5 of e
  uint32_t* l_hi = l_lo + 64;                                 // ... of bits 6..11
  uint64_t* g_lo = reinterpret_cast<uint64_t*>(l_hi + 64);    // address bits of bits 0..5 of e
  uint64_t* g_hi = g_lo + 64;                                 // ... of bits 6..11

  const int tid = threadIdx.x;
  for (int i = tid; i < W * W; i += STAGED_THREADS) A[i] = L.matrix[i];
  for (int t = tid; t < 128; t += STAGED_THREADS) {
    const int v = t & 63, first = t < 64 ? 0 : 6, count = t < 64 ? 6 : L.souter_k - 6;
    uint32_t li = 0;
    uint64_t gi = 0;
    for (int b = 0; b < count; ++b)
      if ((v >> b) & 1) {
        li |= 1u << L.stile_lds_bit[first + b];
        gi |= 1ull << L.stile_qubit[first + b];
      }
    if (t < 64) {
      l_lo[v] = li;
      g_lo[v] = gi;
    } else {
      l_hi[v] = li;
      g_hi[v] = gi;
    }
  }
  __syncthreads();

  const int lane = tid & 63, wave = tid >> 6;
  const int i32 = lane & 31, kk = lane >> 5;
  const int swz = (W >= 64 && kk) ? 32 : 0;
  const uint32_t one = l_lo[1];  // LDS index bit of e's bit 0 (qubit 0)
  const uint64_t ntiles = L.sntiles;

  // Workgroup b runs on XCD b % 8; with the grid a multiple of 8 and at least 8
  // tiles per XCD, XCD x walks the contiguous tile range [x T/8, (x+1) T/8)
  // with its G/8 workgroups; otherwise the workgroups interleave over all tiles.
  const bool by_xcd = (gridDim.x % 8) == 0 && ntiles >= 8 * (uint64_t)(gridDim.x / 8) && (ntiles % 8) == 0;
  const uint64_t g8 = gridDim.x / 8, per_xcd = ntiles / 8, xcd = blockIdx.x % 8, lw = blockIdx.x / 8;
  const uint64_t iters = by_xcd ? (per_xcd + g8 - 1 - lw) / g8 : (ntiles + gridDim.x - 1 - blockIdx.x) / gridDim.x;
  auto tile_of = [&](uint64_t it) { return by_xcd ? xcd * per_xcd + lw + it * g8 : blockIdx.x + it * gridDim.x; };

  f4 pre[PIECES];
  uint64_t tile = tile_of(0);
  if (iters > 0) {
    const uint64_t base = expand_index<KB + 7>(tile, L.souter_masks, L.souter_k);
#pragma unroll
    for (int p = 0; p < PIECES; ++p) {
      const int e = 2 * (tid + STAGED_THREADS * p);
      pre[p] = load_f4(s + (base | g_lo[e & 63] | g_hi[e >> 6]));
    }
  }
  for (uint64_t it = 0; it < iters; ++it) {
    tile = tile_of(it);
    const uint64_t base = expand_index<KB + 7>(tile, L.souter_masks, L.souter_k);
    // 1. the tile into LDS
#pragma unroll
    for (int p = 0; p < PIECES; ++p) {
      const int e = 2 * (tid + STAGED_THREADS * p);
      const uint32_t li = l_lo[e & 63] | l_hi[e >> 6];
      buf[li] = cpx<float>{pre[p].x, pre[p].y};
      buf[li | one] = cpx<float>{pre[p].z, pre[p].w};
    }
    __syncthreads();
    // 2. this wave's B operands: lane (col, kk) holds X[2j + kk][col]
    cpx<float> x[HALF];
#pragma unroll
    for (int j = 0; j < HALF; ++j) x[j] = buf[(2 * j + kk) * COLS + 32 * wave + i32];
    // 3. the next tile's loads
    const uint64_t next = tile_of(it + 1);
    if (it + 1 < iters) {
      const uint64_t nbase = expand_index<KB + 7>(next, L.souter_masks, L.souter_k);
#pragma unroll
      for (int p = 0; p < PIECES; ++p) {
        const int e = 2 * (tid + STAGED_THREADS * p);
        pre[p] = load_f4(s + (nbase | g_lo[e & 63] | g_hi[e >> 6]));
      }
    }
    // 4. Y = U X on the wave's 32 columns
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
    }
    // 5. D over the wave's own columns (no other wave reads or writes them before the barrier)
    if constexpr (DT == 1) {
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        const int row = 8 * (v >> 2) + (v & 3) + 4 * kk;
        buf[row * COLS + 32 * wave + i32] = cpx<float>{acc[0][v], acc[0][v + 8]};
      }
    } else {
#pragma unroll
      for (int t = 0; t < DT / 2; ++t) {
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const int row = 32 * t + 8 * (v >> 2) + (v & 3) + 4 * kk;
          buf[row * COLS + 32 * wave + i32] = cpx<float>{acc[t][v], acc[t + DT / 2][v]};
        }
      }
    }
    __syncthreads();
    // 6. the tile back to HBM in address order
#pragma unroll
    for (int p = 0; p < PIECES; ++p) {
      const int e = 2 * (tid + STAGED_THREADS * p);
      const uint32_t li = l_lo[e & 63] | l_hi[e >> 6];
      const cpx<float> a0 = buf[li], a1 = buf[li | one];
      store_f4(s + (base | g_lo[e & 63] | g_hi[e >> 6]), f4{a0.re, a0.im, a1.re, a1.im});
    }
    __syncthreads();
  }
}

template <int KB>
static size_t staged_lds_bytes() {
  constexpr int W = 2 << KB;
  return sizeof(float) * W * W + sizeof(cpx<float>) * (size_t)(1 << KB) * FUSED_STAGED_COLUMNS +
         128 * sizeof(uint32_t) + 128 * sizeof(uint64_t);
}

template <int KB>
static int launch_staged(const fused_launch& L, void* state, void* stream) {
  const size_t lds = staged_lds_bytes<KB>();
  if (const int rc = allow_dynamic_lds((const void*)k_fused_staged<KB>, lds)) return rc;
  const unsigned grid = (unsigned)std::min<uint64_t>(L.sntiles, (uint64_t)get_tuning().fused_staged_groups);
  hipLaunchKernelGGL(k_fused_staged<KB>, dim3(grid), dim3(STAGED_THREADS), lds, as_stream(stream),
                     (cpx<float>*)state, L);
  return check_launch("fused block (staged)");
}

template <int KB, int OCC>
static int launch_k(const fused_launch& L, void* state, void* stream) {
  constexpr int W = 2 << KB;
  const size_t lds = sizeof(float) * (size_t)W * W;
  if (const int rc = allow_dynamic_lds((const void*)k_fused<KB, OCC>, lds)) return rc;
  const uint64_t groups = (L.ntiles + 3) / 4;
  const unsigned grid = (unsigned)std::min<uint64_t>(groups, (uint64_t)get_tuning().fused_groups);
  fused_launch Lc = L;
  Lc.chunked = get_tuning().fused_chunked;
  hipLaunchKernelGGL((k_fused<KB, OCC>), dim3(grid), dim3(FUSED_THREADS), lds, as_stream(stream),
                     (cpx<float>*)state, Lc);
  return check_launch("fused block");
}

// ---------------------------------------------------------------------------
// Chunk passes (k = 4, 5): several blocks applied to one chunk of 2^13 complex64
// amplitudes held in LDS, one HBM read and one write per pass.  A persistent
// 1024-thread workgroup per CU walks a contiguous run of chunks; its 160 KiB of LDS
// hold two chunk buffers (2 x 64 KiB) and the pass's blocks in three-product form
// (up to 8 at k = 4, 2 at k = 5, loaded once).  For chunk i every thread first stores
// its 4 amplitude pairs of chunk i - 1 from the other buffer and issues the loads of its
// 4 pairs of chunk i + 1 into registers (16-byte accesses); then the 16 waves run the
// pass's blocks on chunk i (one barrier per block) and, before the last block's barrier,
// each thread writes its loaded pairs into the other buffer.  A barrier waits only
// for LDS operations and the blocks issue no global memory accesses, so the chunk
// transfers stay in flight while the blocks run.  A thread touches the same 8 slots
// of either buffer for the transfers, so its reads (for the stores) and later writes
// (of the next chunk) of one buffer are ordered by program order.
//
// A block on local bits row_bits[] sees the chunk as 2^k rows x 2^(13-k) columns
// (the other local bits), 2^(8-k) tiles of 32 columns.  Gauss's complex product
//   T1 = Ur Xr, T2 = Ui Xi, T3 = (Ur + Ui)(Xr + Xi);  Yr = T1 - T2, Yi = (T3 - T1) - T2
// runs on v_mfma_f32_16x16x4_f32 (FP32 products and accumulation): three real
// products of 2^k x 2^k by 2^k x 32 instead of the 2^(k+1)-square embedding, 3/4 of the
// MFMA cycles, and six independent accumulation chains per wave.  Wave w takes column
// tile w % 2^(8-k) and 16-row output tile w / 2^(8-k) (k = 5 has two).  MFMA 16x16x4
// operand layout (lane l, accumulator register v):
//   A[i][kk] = lane (i + 16 kk), B[kk][j] = lane (j + 16 kk),
//   D[4 (l / 16) + v][l % 16] = acc[v] of lane l.
// ---------------------------------------------------------------------------
typedef const __attribute__((address_space(4))) fused_chunk_program* cchunk;

constexpr int CHUNK_THREADS = 1024;
constexpr size_t CHUNK_LDS_BYTES = 160 * 1024;  // the whole LDS of a CU
constexpr size_t CHUNK_BUFFER_BYTES = sizeof(cpx<float>) << FUSED_CHUNK_LOCAL;

typedef float f4v __attribute__((ext_vector_type(4)));

template <int KB>
struct chunk_shape {
  static constexpr int R = 1 << KB;
  static constexpr int NCOL = FUSED_CHUNK_LOCAL - KB;  // column bits
  static constexpr int NTILES = 1 << (NCOL - 5);       // 32-column tiles per block
  static constexpr int MT = R / 16;                    // 16-row output tiles per column tile
  static constexpr int STEPS = R / 4;                  // MFMA steps of K = 4
  static constexpr int WAVES = CHUNK_THREADS / 64;
  static constexpr int AFLOATS = 3 * R * R;            // one block's three-product form
  static constexpr int SLOTS = fused_chunk_max_blocks(KB);
  static_assert(NTILES * MT == WAVES, "one (column tile, row tile) per wave");
  static_assert(2 * CHUNK_BUFFER_BYTES + sizeof(float) * AFLOATS * SLOTS <= CHUNK_LDS_BYTES, "LDS");
};

template <int KB, bool NT>
__global__ __launch_bounds__(CHUNK_THREADS, 1) void k_fused_chunk(cpx<float>* __restrict__ s,
                                                                   const fused_chunk_program* program,
                                                                   const float* __restrict__ matrices,
                                                                   uint64_t nchunks) {
  using S = chunk_shape<KB>;
  constexpr int R = S::R, STEPS = S::STEPS;
  constexpr int CHUNK = 1 << FUSED_CHUNK_LOCAL;
  constexpr int PAIRS = CHUNK / (2 * CHUNK_THREADS);  // 4 amplitude pairs (16 bytes each) per thread
  extern __shared__ float smem[];
  cpx<float>* bufs = reinterpret_cast<cpx<float>*>(smem);  // two chunk buffers
  float* slots = smem + 4 * CHUNK;                          // the pass's three-product forms
  const cchunk prog = (cchunk)program;
  const uint32_t t = threadIdx.x;
  const uint32_t nops = prog->num_ops;  // 1 .. SLOTS (checked by the host)

  // This workgroup's chunks.  Workgroup b runs on XCD b % 8; with the grid a multiple of 8,
  // XCD x walks the contiguous chunk range [x N/8, (x+1) N/8) with its G/8 workgroups, else
  // the workgroups interleave over all chunks.
  const uint64_t G = gridDim.x;
  const bool by_xcd = (G % 8) == 0 && (nchunks % 8) == 0 && nchunks >= G;
  const uint64_t g8 = G / 8, per_xcd = nchunks / 8, xcd = blockIdx.x % 8, lw = blockIdx.x / 8;
  const uint64_t iters = by_xcd ? (per_xcd + g8 - 1 - lw) / g8 : (nchunks + G - 1 - blockIdx.x) / G;
  auto chunk_of = [&](uint64_t it) { return by_xcd ? xcd * per_xcd + lw + it * g8 : blockIdx.x + it * G; };
  // chunk number -> the bits outside the local positions
  auto base_of = [&](uint64_t c) {
    uint64_t b = c & prog->outer_masks[0];
    for (int j = 1; j <= FUSED_CHUNK_LOCAL; ++j) b |= (c & prog->outer_masks[j]) << j;
    return b;
  };
  // transfers in 16-byte pairs (local bit 0 is qubit 0, so a pair is two consecutive amplitudes):
  // pair k of thread t holds local indices 2 (t + 1024 k) and the one after, the bits of t at
  // pos[1 .. 10], k at pos[11 .. 12] (deposit_high[4k] places local index 2048 k = (4k) << 9).
  // Half as many memory instructions as 8-byte amplitudes, which shortens the transfer issue at the
  // start of every chunk.
  uint64_t low = 0;
  for (int r = 0; r < 10; ++r) low |= (uint64_t)((t >> r) & 1u) << prog->pos[r + 1];
  auto pair_at = [&](uint64_t b, int k) { return s + (b | low | prog->deposit_high[4 * k]); };

  // prologue: the pass's three-product forms into their slots, chunk 0 into buffer 0
  for (uint32_t m = 0; m < nops; ++m) {
    const v4f* A4 = reinterpret_cast<const v4f*>(matrices + (uint64_t)prog->ops[m].matrix * (uint64_t)S::AFLOATS);
    for (uint32_t i = t; i < (uint32_t)(S::AFLOATS / 4); i += CHUNK_THREADS)
      reinterpret_cast<v4f*>(slots + m * S::AFLOATS)[i] = A4[i];
  }
  v4f pre[PAIRS];
  if (iters > 0) {
    const uint64_t b = base_of(chunk_of(0));
#pragma unroll
    for (int k = 0; k < PAIRS; ++k) pre[k] = load_pair<NT>(pair_at(b, k));
#pragma unroll
    for (int k = 0; k < PAIRS; ++k) reinterpret_cast<v4f*>(bufs)[t + CHUNK_THREADS * k] = pre[k];
  }
  __syncthreads();

  const int lane = t & 63, wave = t >> 6;
  const int li = lane & 15, lk = lane >> 4;  // A row / B and D column within a 16-tile; K row within a step
  const int tile = wave % S::NTILES;         // column tile
  const int mt = wave / S::NTILES;           // 16-row output tile
  for (uint64_t it = 0; it < iters; ++it) {
    cpx<float>* lds = bufs + (it & 1) * CHUNK;
    cpx<float>* other = bufs + ((it + 1) & 1) * CHUNK;
    if (it > 0) {  // chunk it - 1 back to HBM
      const uint64_t b = base_of(chunk_of(it - 1));
#pragma unroll
      for (int k = 0; k < PAIRS; ++k)
        store_pair<NT>(pair_at(b, k), reinterpret_cast<const v4f*>(other)[t + CHUNK_THREADS * k]);
    }
    // chunk it + 1 into registers -- on the last iteration the current chunk again: its loads and
    // the write into the free buffer below are then harmless, and on every path the loads are
    // waited for before the chunk's last barrier (so the compiler does not wait for them, and for
    // the stores issued after them, when the next iteration reuses their registers)
    {
      const uint64_t b = base_of(chunk_of(it + 1 < iters ? it + 1 : it));
#pragma unroll
      for (int k = 0; k < PAIRS; ++k) pre[k] = load_pair<NT>(pair_at(b, k));
    }
    for (uint32_t o = 0; o < nops; ++o) {
      const auto& op = prog->ops[o];
      const float* Ar = slots + o * S::AFLOATS;  // Ur, Ui, Ur + Ui, each transposed
      // local offset of block row r (bit b of r at local bit row_bits[b])
      auto rowoff = [&](int r) {
        uint32_t d = 0;
#pragma unroll
        for (int b = 0; b < KB; ++b) d |= ((uint32_t)(r >> b) & 1u) << op.row_bits[b];
        return d;
      };
      // column 16 nt + li of this wave's tile: bits 0..3 from li, bit 4 from nt, bits 5.. from the tile
      uint32_t colbase = 0;
#pragma unroll
      for (int b = 0; b < 4; ++b) colbase |= ((uint32_t)(li >> b) & 1u) << op.col_bits[b];
#pragma unroll
      for (int b = 5; b < S::NCOL; ++b) colbase |= ((uint32_t)(tile >> (b - 5)) & 1u) << op.col_bits[b];
      const uint32_t col[2] = {colbase, colbase | (1u << op.col_bits[4])};
      const uint32_t krow = rowoff(lk);  // the lane's K row within a step
      // B operands: X[4 st + lk][16 nt + li]
      cpx<float> x[STEPS][2];
#pragma unroll
      for (int st = 0; st < STEPS; ++st)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) x[st][nt] = lds[col[nt] | rowoff(4 * st) | krow];
      f4v acc[3][2];
#pragma unroll
      for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) acc[p][nt] = f4v{};
#pragma unroll
      for (int st = 0; st < STEPS; ++st) {
        const int ai = (4 * st + lk) * R + 16 * mt + li;  // A[16 mt + li][4 st + lk], transposed
        const float a0 = Ar[ai], a1 = Ar[R * R + ai], a2 = Ar[2 * R * R + ai];
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
          const cpx<float> v = x[st][nt];
          acc[0][nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, v.re, acc[0][nt], 0, 0, 0);
          acc[1][nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, v.im, acc[1][nt], 0, 0, 0);
          acc[2][nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a2, v.re + v.im, acc[2][nt], 0, 0, 0);
        }
      }
      // k = 5: both row tiles of a column tile gather all its rows; every gather ends before a
      // wave writes its rows
      if constexpr (S::MT > 1) __syncthreads();
      // D: rows 16 mt + 4 lk + v, column 16 nt + li
      const uint32_t drow = rowoff(16 * mt + 4 * lk);
#pragma unroll
      for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const float t1 = acc[0][nt][v], t2 = acc[1][nt][v], t3 = acc[2][nt][v];
          lds[col[nt] | drow | rowoff(v)] = cpx<float>{t1 - t2, (t3 - t1) - t2};
        }
      if (o + 1 < nops) __syncthreads();  // the block done: its D rows in LDS
    }
    // chunk it + 1 into the other buffer (this thread's slots), then the chunk's last barrier
#pragma unroll
    for (int k = 0; k < PAIRS; ++k) reinterpret_cast<v4f*>(other)[t + CHUNK_THREADS * k] = pre[k];
    __syncthreads();
  }
  if (iters > 0) {  // the last chunk back to HBM
    const cpx<float>* last = bufs + ((iters - 1) & 1) * CHUNK;
    const uint64_t b = base_of(chunk_of(iters - 1));
#pragma unroll
    for (int k = 0; k < PAIRS; ++k)
      store_pair<NT>(pair_at(b, k), reinterpret_cast<const v4f*>(last)[t + CHUNK_THREADS * k]);
  }
}

template <int KB, bool NT>
static int launch_chunk(const fused_chunk_program* device_program, const float* matrices, cpx<float>* s,
                        uint64_t chunks, hipStream_t st) {
  constexpr size_t lds = CHUNK_LDS_BYTES;
  if (const int rc = allow_dynamic_lds((const void*)k_fused_chunk<KB, NT>, lds)) return rc;
  // one persistent workgroup per CU (a multiple of 8 when the chunks allow: XCD-contiguous runs)
  const unsigned grid = (unsigned)std::min<uint64_t>(chunks, (uint64_t)device_cus());
  hipLaunchKernelGGL((k_fused_chunk<KB, NT>), dim3(grid), dim3(CHUNK_THREADS), lds, st, s, device_program, matrices,
                     chunks);
  return check_launch("fused chunk pass");
}

bool fused_chunks_enabled() { return get_tuning().fused_chunks != 0; }

int launch_fused_chunk(const fused_chunk_program* device_program, const fused_chunk_program& program, int width,
                       const float* matrices, void* state, uint64_t size, void* stream) {
  unsigned nq;
  if (int st = check_state(state, size, BRK_C64, &nq)) return st;
  // (a pass holds 1 .. fused_chunk_max_blocks(width) blocks: the embeddings it keeps in LDS)
  if (nq < (unsigned)FUSED_CHUNK_LOCAL || (width != 4 && width != 5) || program.num_ops == 0 ||
      program.num_ops > (uint32_t)fused_chunk_max_blocks(width))
    return fail(BRK_ERR_INVALID_ARGUMENT, "fused chunk pass does not fit (%u qubits, %u ops, width %d)", nq,
                program.num_ops, width);
  const uint64_t chunks = size >> FUSED_CHUNK_LOCAL;
  if (chunks > 0x7fffffffull) return fail(BRK_ERR_INVALID_ARGUMENT, "too many chunks");
  const bool nt = size * sizeof(cpx<float>) >= get_tuning().nt_min_bytes;
  hipStream_t st = as_stream(stream);
  auto* s = (cpx<float>*)state;
  if (width == 4)
    return nt ? launch_chunk<4, true>(device_program, matrices, s, chunks, st)
              : launch_chunk<4, false>(device_program, matrices, s, chunks, st);
  return nt ? launch_chunk<5, true>(device_program, matrices, s, chunks, st)
            : launch_chunk<5, false>(device_program, matrices, s, chunks, st);
}

int launch_fused_block(const fused_launch& L, void* state, void* stream) {
  if (get_tuning().fused_staged && L.sntiles) {
    if (L.width == 4) return launch_staged<4>(L, state, stream);
    if (L.width == 5) return launch_staged<5>(L, state, stream);
  }
  switch (L.width) {
    case 4: return get_tuning().fused_occupancy == 4 ? launch_k<4, 4>(L, state, stream) : launch_k<4, 2>(L, state, stream);
    case 5: return get_tuning().fused_occupancy == 4 ? launch_k<5, 4>(L, state, stream) : launch_k<5, 2>(L, state, stream);
    case 6: return launch_k<6, 2>(L, state, stream);  // 64 KiB of LDS per workgroup: two per CU at most
    default: return fail(BRK_ERR_UNSUPPORTED, "fused block width %d not in 4..6", L.width);
  }
}

}  // namespace brk
