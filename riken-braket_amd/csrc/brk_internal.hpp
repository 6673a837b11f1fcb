// Internal helpers shared by the HIP translation units of the engine:
// error reporting, argument validation, the bit-insertion ("expand") index
// map, and the device complex type.
#pragma once

#include <mutex>

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdarg>
#include <string>
#include <algorithm>

#include "braket_hip.h"
#include "brk_bits.hpp"

namespace brk {

// ---------------------------------------------------------------------------
// Errors
// ---------------------------------------------------------------------------

// Records a message for brk_last_error() on this thread and returns `status`.
int fail(int status, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

#define BRK_HIP_CHECK(expr)                                                     \
  do {                                                                          \
    hipError_t brk_e_ = (expr);                                                 \
    if (brk_e_ != hipSuccess)                                                   \
      return ::brk::fail(BRK_ERR_HIP, "%s failed: %s", #expr,                   \
                         hipGetErrorString(brk_e_));                            \
  } while (0)

// Checks the launch that was just enqueued.
inline int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(BRK_ERR_HIP, "%s launch failed: %s", what, hipGetErrorString(e));
  return BRK_OK;
}

// ---------------------------------------------------------------------------
// Device complex type: re then im, aligned to its size so that a c128 element
// is one 16-byte load/store and a c64 element one 8-byte load/store.
// ---------------------------------------------------------------------------

template <typename R>
struct alignas(2 * sizeof(R)) cpx {
  R re, im;
};

template <typename R>
__device__ __forceinline__ cpx<R> cmul(cpx<R> a, cpx<R> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename R>
__device__ __forceinline__ cpx<R> cadd(cpx<R> a, cpx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <typename R>
__device__ __forceinline__ cpx<R> csub(cpx<R> a, cpx<R> b) { return {a.re - b.re, a.im - b.im}; }
template <typename R>
__device__ __forceinline__ cpx<R> cscale(cpx<R> a, R s) { return {a.re * s, a.im * s}; }
template <typename R>
__device__ __forceinline__ cpx<R> cneg(cpx<R> a) { return {-a.re, -a.im}; }
// -a when `on`, else a: the sign bits flipped by an integer XOR (the same values as cneg)
__device__ __forceinline__ cpx<double> cneg_if(cpx<double> a, bool on) {
  const uint64_t m = (uint64_t)on << 63;
  return {__longlong_as_double(__double_as_longlong(a.re) ^ (long long)m),
          __longlong_as_double(__double_as_longlong(a.im) ^ (long long)m)};
}
__device__ __forceinline__ cpx<float> cneg_if(cpx<float> a, bool on) {
  const uint32_t m = (uint32_t)on << 31;
  return {__uint_as_float(__float_as_uint(a.re) ^ m), __uint_as_float(__float_as_uint(a.im) ^ m)};
}

inline size_t amp_bytes(int dtype) { return dtype == BRK_C128 ? 16 : 8; }

// One amplitude as a native 2-vector, so that it can carry the non-temporal
// (streaming) hint: a c128 element is one 16-byte access, a c64 one 8-byte.
template <typename R> struct vec2;
template <> struct vec2<double> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct vec2<float> { typedef float type __attribute__((ext_vector_type(2))); };

template <bool NT, typename R>
__device__ __forceinline__ cpx<R> load_amp(const cpx<R>* p) {
  using V = typename vec2<R>::type;
  V v;
  if constexpr (NT) v = __builtin_nontemporal_load(reinterpret_cast<const V*>(p));
  else v = *reinterpret_cast<const V*>(p);
  return cpx<R>{v.x, v.y};
}

template <bool NT, typename R>
__device__ __forceinline__ void store_amp(cpx<R>* p, cpx<R> c) {
  using V = typename vec2<R>::type;
  V v = {c.re, c.im};
  if constexpr (NT) __builtin_nontemporal_store(v, reinterpret_cast<V*>(p));
  else *reinterpret_cast<V*>(p) = v;
}

// Block number -> tile number.  The dispatcher hands blocks to the 8 XCDs
// round robin (block b -> XCD b % 8); with remap on, the blocks of one XCD
// walk one contiguous eighth of the tiles instead of every eighth tile.  The
// map is a bijection on [0, nblocks) for any nblocks.
__device__ __forceinline__ uint64_t xcd_tile(uint32_t b, uint32_t nblocks, int remap) {
  if (!remap) return b;
  const uint32_t x = b & 7u, k = b >> 3;
  const uint32_t q = nblocks >> 3, r = nblocks & 7u;
  return (uint64_t)(x * q + (x < r ? x : r) + k);
}

// Launch tuning of the streaming gate kernels.  Defaults come from sweeps on
// MI355X; every field can be overridden with brk_tune() or the environment
// variable BRK_TUNE_<KEY> (upper case) read at first use.
struct tuning {
  int pair_lds;        // bytes of LDS reserved per 256-thread block of the pair kernel (occupancy cap)
  int diag_lds;        // same for the diagonal kernel
  int pair_remap;      // XCD-contiguous tile order for the pair kernel
  int diag_remap;      // same for the diagonal kernel
  uint64_t nt_min_bytes;  // states at least this large use non-temporal loads/stores
  uint64_t max_blocks;    // grid cap of the gate kernels (larger states loop)
  uint64_t line_bytes;    // write granularity: control bits addressing inside it are tested, not skipped
  uint64_t fused_groups;  // grid cap of the fused-block kernel (workgroups loop over the tiles)
  int fused_occupancy;    // 2 or 4: workgroups per CU the fused kernel (k <= 5) is compiled for
  int fused_staged;       // k <= 5: 1 = LDS-staged tiles in address order, 0 = register-direct tiles
  uint64_t fused_staged_groups;  // grid of the staged kernel (workgroups loop over the tiles)
  int fused_chunked;      // fused tiles: 0 = interleaved over all waves, 1 = contiguous run per wave
  int block_lds;          // gate blocks: LDS bytes per chunk (65536, 32768 or 16384) -> local qubits
  int block_run_bytes;    // gate blocks: the lowest qubits addressing this many bytes are always local
  int adjoint_blocks;     // adjoint gradient: 1 = forward pass and reverse sweep in gate blocks, 0 = per-run passes
  int fused_chunks;       // precision form (k = 4, 5): 1 = several blocks per LDS-resident chunk pass, 0 = a launch per block
};
tuning& get_tuning();

// Compute units of the current device (cached per device).
unsigned device_cus();

// Raises a kernel's dynamic-LDS limit to `bytes` (above 64 KiB a kernel has to opt in) once per
// device: the attribute belongs to the device the call is made on, and launches may come from
// several host threads.
int allow_dynamic_lds(const void* kernel, size_t bytes);

// Bit insertion: see brk_bits.hpp.

template <int K>
__device__ __forceinline__ uint64_t expand_index(uint64_t g, const uint64_t* m, int k_runtime) {
  if constexpr (K >= 0) {
    uint64_t r = g & m[0];
#pragma unroll
    for (int j = 1; j <= K; ++j) r |= (g & m[j]) << j;
    return r;
  } else {
    uint64_t r = g & m[0];
    for (int j = 1; j <= k_runtime; ++j) r |= (g & m[j]) << j;
    return r;
  }
}

// ---------------------------------------------------------------------------
// Host-side validation
// ---------------------------------------------------------------------------

int check_state(const void* first, uint64_t size, int dtype, unsigned* num_qubits);

// Collects `n` qubits into `out` (appending at *count), checking the range
// against `num_qubits` and distinctness against everything already in `out`.
int collect_qubits(const unsigned* qubits, unsigned n, unsigned num_qubits, unsigned* out, unsigned* count,
                   const char* role);

inline uint64_t mask_of(const unsigned* q, unsigned n) {
  uint64_t m = 0;
  for (unsigned i = 0; i < n; ++i) m |= 1ull << q[i];
  return m;
}

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Per-device scratch used by the reductions (partial sums).  Allocated on the
// first reduction on a device and kept until the process ends.  A reduction
// holds `busy` from its launch until it has read the partials back, so calls
// from several host threads on one device take turns.
struct reduce_scratch {
  double* device = nullptr;   // RED_MAX_BLOCKS * RED_MAX_VALUES doubles
  double* host = nullptr;     // pinned mirror
  std::mutex busy;
};
int get_reduce_scratch(reduce_scratch** out);

constexpr int RED_BLOCKS = 1024;      // fixed grid -> fixed summation order
constexpr int RED_THREADS = 256;
constexpr int RED_MAX_VALUES = 4;

}  // namespace brk
