// Internal helpers shared by the HIP translation units of the engine:
// error reporting, argument validation, the bit-insertion ("expand") index
// map, and the device complex type.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdarg>
#include <string>
#include <algorithm>

#include "braket_hip.h"

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

inline size_t amp_bytes(int dtype) { return dtype == BRK_C128 ? 16 : 8; }

// ---------------------------------------------------------------------------
// Bit insertion.  A gate on k operated qubits visits 2^(n-k) groups; group
// number g is turned into the base index of its group by inserting a zero at
// every operated bit position.  With the positions sorted p_0 < ... < p_{k-1}
// the index splits into k+1 bit fields; field j is shifted left by j:
//     base = sum_j (g & field_mask[j]) << j.
// ---------------------------------------------------------------------------

struct expand_masks {
  uint64_t m[BRK_MAX_OPERATED_QUBITS + 1];
  int k;  // number of inserted bits
};

// `positions` need not be sorted; they must be distinct and < 64.
inline expand_masks make_expand_masks(const unsigned* positions, int k) {
  unsigned sorted[BRK_MAX_OPERATED_QUBITS];
  std::copy(positions, positions + k, sorted);
  std::sort(sorted, sorted + k);
  expand_masks e{};
  e.k = k;
  uint64_t prev = 0;  // mask of all bits below the current field
  for (int j = 0; j < k; ++j) {
    // field j holds bits [sorted[j-1]-(j-1), sorted[j]-j) of g
    uint64_t upto = (sorted[j] - j) >= 64 ? ~0ull : ((1ull << (sorted[j] - j)) - 1ull);
    e.m[j] = upto ^ prev;
    prev = upto;
  }
  e.m[k] = ~0ull ^ prev;
  return e;
}

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
// first reduction on a device and kept until the process ends.
struct reduce_scratch {
  double* device = nullptr;   // RED_MAX_BLOCKS * RED_MAX_VALUES doubles
  double* host = nullptr;     // pinned mirror
};
int get_reduce_scratch(reduce_scratch** out);

constexpr int RED_BLOCKS = 1024;      // fixed grid -> fixed summation order
constexpr int RED_THREADS = 256;
constexpr int RED_MAX_VALUES = 4;

}  // namespace brk
