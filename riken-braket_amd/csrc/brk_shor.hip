// SHORBOX: the modular-exponentiation state of Shor's algorithm.
//
// ket::shor_box (ket/include/ket/shor_box.hpp:60-98) overwrites the state with
//   2^{-m/2} sum_x |reverse_m(x)>_E |v_x>_L,
// E the top m qubits (exponent), L the n - m low qubits, v_0 = 1 and
// v_x = (v_{x-1} * base) mod divisor, only the low n - m bits of v_x kept
// (make_filtered_integer, shor_box.hpp:30-45).  Here the 2^m values v_x are
// computed in parallel by square-and-multiply into a table, then one
// write-only pass sets every amplitude: row r = i >> (n - m) belongs to
// exponent x = reverse_m(r), and its one non-zero entry is at low bits v_x.
#include "brk_internal.hpp"

#include <cmath>

namespace brk {

__device__ __forceinline__ uint64_t mulmod(uint64_t a, uint64_t b, uint64_t m) {
  return (uint64_t)(((unsigned __int128)a * b) % m);
}

// v_x for x < count.  Equal to the reference's recurrence whenever that
// recurrence does not wrap around 2^64, which the host checks.
__global__ __launch_bounds__(256) void k_shor_table(uint64_t* __restrict__ table, uint64_t count, uint64_t base,
                                                   uint64_t divisor) {
  const uint64_t x = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (x >= count) return;
  if (x == 0) {
    table[0] = 1;  // v_0 = 1, not reduced (shor_box.hpp:77)
    return;
  }
  uint64_t r = 1 % divisor, b = base % divisor;
  for (uint64_t e = x; e; e >>= 1) {
    if (e & 1) r = mulmod(r, b, divisor);
    b = mulmod(b, b, divisor);
  }
  table[x] = r;
}

template <typename R>
__global__ __launch_bounds__(256) void k_shor_fill(cpx<R>* __restrict__ s, uint64_t size,
                                                  const uint64_t* __restrict__ table, unsigned low_bits, unsigned m,
                                                  R coefficient) {
  const uint64_t low_mask = (1ull << low_bits) - 1ull;
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < size; i += stride) {
    const uint64_t r = i >> low_bits;
    const uint64_t x = m ? (__brevll(r) >> (64 - m)) : 0;
    const bool hit = (i & low_mask) == (table[x] & low_mask);
    store_amp<true>(s + i, cpx<R>{hit ? coefficient : R(0), R(0)});
  }
}

}  // namespace brk

using namespace brk;

extern "C" {

int brk_shor_box(void* first, uint64_t size, int dtype, unsigned num_exponent_qubits, uint64_t divisor,
                 uint64_t base, void* stream) {
  unsigned nq;
  if (int st = check_state(first, size, dtype, &nq)) return st;
  const unsigned m = num_exponent_qubits;
  if (m > nq) return fail(BRK_ERR_INVALID_ARGUMENT, "%u exponent qubits for a %u-qubit state", m, nq);
  if (divisor == 0) return fail(BRK_ERR_INVALID_ARGUMENT, "divisor is 0");
  // the reference multiplies v (< divisor after the first step) by base in uint64
  if (base != 0 && (divisor - 1) > UINT64_MAX / base)
    return fail(BRK_ERR_UNSUPPORTED, "(divisor - 1) * base overflows 64 bits");
  const uint64_t count = 1ull << m;
  hipStream_t s = as_stream(stream);
  uint64_t* table = nullptr;
  BRK_HIP_CHECK(hipMalloc((void**)&table, sizeof(uint64_t) * count));
  hipLaunchKernelGGL(k_shor_table, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, table, count, base,
                     divisor);
  int status = check_launch("shor table");
  if (!status) {
    const unsigned grid = (unsigned)std::min<uint64_t>((size + 255) / 256, 1u << 20);
    const unsigned low_bits = nq - m;
    if (dtype == BRK_C128)
      hipLaunchKernelGGL(k_shor_fill<double>, dim3(grid), dim3(256), 0, s, (cpx<double>*)first, size, table,
                         low_bits, m, std::pow(static_cast<double>(count), -0.5));
    else
      hipLaunchKernelGGL(k_shor_fill<float>, dim3(grid), dim3(256), 0, s, (cpx<float>*)first, size, table,
                         low_bits, m, std::pow(static_cast<float>(count), -0.5f));
    status = check_launch("shor fill");
  }
  (void)hipStreamSynchronize(s);
  (void)hipFree(table);
  return status;
}

}  // extern "C"
