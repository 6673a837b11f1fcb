// FP64 reductions over a state vector: inner product, norm, spin expectation
// values, Pauli-string expectation values and the open-chain Ising energy,
// plus H|psi> for that Hamiltonian.
//
// Every reduction runs on a fixed grid of RED_BLOCKS blocks with a
// grid-stride loop, so the assignment of amplitudes to threads, the in-block
// tree and the host-side sum over the block partials (in block order) are the
// same on every call: results are bit-reproducible.  Partial sums are kept in
// double whatever the element type (the reference accumulates per-thread
// partials and then std::accumulate's them, inner_product.hpp:43-55; spin
// values use long double, spin_expectation_value.hpp:31-64).
#include "brk_internal.hpp"

#include <complex>
#include <vector>

namespace brk {

template <int NV>
struct dvec {
  double v[NV];
};

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
  return x;
}

// Reduces `acc` over the block and writes it to partial[blockIdx.x * NV + v].
template <int NV>
__device__ __forceinline__ void block_store(dvec<NV> acc, double* __restrict__ partial) {
  __shared__ double lds[RED_THREADS / 64][NV];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    double x = wave_sum(acc.v[v]);
    if (lane == 0) lds[wave][v] = x;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < RED_THREADS / 64; ++w) t += lds[w][threadIdx.x];
    partial[(size_t)blockIdx.x * NV + threadIdx.x] = t;
  }
}

template <typename R1, typename R2>
__global__ __launch_bounds__(RED_THREADS) void k_inner(const cpx<R1>* __restrict__ a, const cpx<R2>* __restrict__ b,
                                                      uint64_t size, double* partial) {
  dvec<2> acc{{0.0, 0.0}};
  const uint64_t stride = (uint64_t)RED_BLOCKS * RED_THREADS;
  for (uint64_t i = (uint64_t)blockIdx.x * RED_THREADS + threadIdx.x; i < size; i += stride) {
    cpx<R1> x = a[i];
    cpx<R2> y = b[i];
    const double xr = x.re, xi = x.im, yr = y.re, yi = y.im;
    // conj(y) * x
    acc.v[0] += yr * xr + yi * xi;
    acc.v[1] += yr * xi - yi * xr;
  }
  block_store<2>(acc, partial);
}

template <typename R>
__global__ __launch_bounds__(RED_THREADS) void k_norm(const cpx<R>* __restrict__ a, uint64_t size, double* partial) {
  dvec<1> acc{{0.0}};
  const uint64_t stride = (uint64_t)RED_BLOCKS * RED_THREADS;
  for (uint64_t i = (uint64_t)blockIdx.x * RED_THREADS + threadIdx.x; i < size; i += stride) {
    cpx<R> x = a[i];
    const double xr = x.re, xi = x.im;
    acc.v[0] += xr * xr + xi * xi;
  }
  block_store<1>(acc, partial);
}

// Pairs of `qubit`: (Re conj(a0) a1, Im conj(a0) a1, |a0|^2 - |a1|^2).
template <typename R>
__global__ __launch_bounds__(RED_THREADS) void k_spin(const cpx<R>* __restrict__ a, uint64_t npairs, uint64_t lower,
                                                     uint64_t qmask, double* partial) {
  dvec<3> acc{{0.0, 0.0, 0.0}};
  const uint64_t stride = (uint64_t)RED_BLOCKS * RED_THREADS;
  for (uint64_t g = (uint64_t)blockIdx.x * RED_THREADS + threadIdx.x; g < npairs; g += stride) {
    const uint64_t i0 = ((g & ~lower) << 1) | (g & lower);
    cpx<R> x0 = a[i0];
    cpx<R> x1 = a[i0 | qmask];
    const double r0 = x0.re, m0 = x0.im, r1 = x1.re, m1 = x1.im;
    acc.v[0] += r0 * r1 + m0 * m1;
    acc.v[1] += r0 * m1 - m0 * r1;
    acc.v[2] += (r0 * r0 + m0 * m0) - (r1 * r1 + m1 * m1);
  }
  block_store<3>(acc, partial);
}

// sum_i conj(a_i) a_{i ^ flip} (-1)^{popcount(i & sign)}
template <typename R>
__global__ __launch_bounds__(RED_THREADS) void k_pauli(const cpx<R>* __restrict__ a, uint64_t size, uint64_t flip,
                                                      uint64_t sign, double* partial) {
  dvec<2> acc{{0.0, 0.0}};
  const uint64_t stride = (uint64_t)RED_BLOCKS * RED_THREADS;
  for (uint64_t i = (uint64_t)blockIdx.x * RED_THREADS + threadIdx.x; i < size; i += stride) {
    cpx<R> x = a[i];
    cpx<R> y = a[i ^ flip];
    const double xr = x.re, xi = x.im, yr = y.re, yi = y.im;
    double re = xr * yr + xi * yi;
    double im = xr * yi - xi * yr;
    if (__popcll(i & sign) & 1) {
      re = -re;
      im = -im;
    }
    acc.v[0] += re;
    acc.v[1] += im;
  }
  block_store<2>(acc, partial);
}

constexpr int MAX_CHAIN = 48;

struct ising_args {
  double J[MAX_CHAIN];
  double h[MAX_CHAIN];
  int n;
};

// Diagonal part of the Ising Hamiltonian at basis index k.
__device__ __forceinline__ double ising_diag(uint64_t k, const ising_args& a) {
  const uint64_t x = k ^ (k >> 1);  // bit i = b_i xor b_{i+1}
  double e = 0.0;
  for (int i = 0; i + 1 < a.n; ++i) e += ((x >> i) & 1ull) ? -a.J[i] : a.J[i];
  return e;
}

template <typename R>
__global__ __launch_bounds__(RED_THREADS) void k_ising_zz(const cpx<R>* __restrict__ a, uint64_t size,
                                                         ising_args args, double* partial) {
  dvec<1> acc{{0.0}};
  const uint64_t stride = (uint64_t)RED_BLOCKS * RED_THREADS;
  for (uint64_t i = (uint64_t)blockIdx.x * RED_THREADS + threadIdx.x; i < size; i += stride) {
    cpx<R> x = a[i];
    const double xr = x.re, xi = x.im;
    acc.v[0] += (xr * xr + xi * xi) * ising_diag(i, args);
  }
  block_store<1>(acc, partial);
}

template <typename R>
__global__ __launch_bounds__(256) void k_apply_ising(cpx<R>* __restrict__ out, const cpx<R>* __restrict__ psi,
                                                    uint64_t size, ising_args args) {
  const uint64_t k = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= size) return;
  cpx<R> x = psi[k];
  const double e = ising_diag(k, args);
  double re = e * (double)x.re, im = e * (double)x.im;
  for (int i = 0; i < args.n; ++i) {
    cpx<R> y = psi[k ^ (1ull << i)];
    re += args.h[i] * (double)y.re;
    im += args.h[i] * (double)y.im;
  }
  out[k] = cpx<R>{(R)re, (R)im};
}

// Launches `launch(partial)` on the fixed reduction grid and sums the NV
// partials per block on the host.
template <int NV, typename Launch>
static int run_reduction(void* stream, Launch launch, double* out) {
  reduce_scratch* sc;
  if (int st = get_reduce_scratch(&sc)) return st;
  hipStream_t s = as_stream(stream);
  launch(sc->device, s);
  if (int st = check_launch("reduction")) return st;
  BRK_HIP_CHECK(hipMemcpyAsync(sc->host, sc->device, sizeof(double) * RED_BLOCKS * NV, hipMemcpyDeviceToHost, s));
  BRK_HIP_CHECK(hipStreamSynchronize(s));
  for (int v = 0; v < NV; ++v) out[v] = 0.0;
  for (int b = 0; b < RED_BLOCKS; ++b)
    for (int v = 0; v < NV; ++v) out[v] += sc->host[b * NV + v];
  return BRK_OK;
}

static int spin_raw(const void* first, uint64_t size, int dtype, unsigned qubit, double out[3], void* stream) {
  const uint64_t qmask = 1ull << qubit, lower = qmask - 1ull, npairs = size >> 1;
  return run_reduction<3>(
      stream,
      [&](double* partial, hipStream_t s) {
        if (dtype == BRK_C128)
          hipLaunchKernelGGL(k_spin<double>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s,
                             (const cpx<double>*)first, npairs, lower, qmask, partial);
        else
          hipLaunchKernelGGL(k_spin<float>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, (const cpx<float>*)first,
                             npairs, lower, qmask, partial);
      },
      out);
}

static int make_ising_args(unsigned nq, const double* J, const double* h, ising_args* a) {
  if (nq > (unsigned)MAX_CHAIN) return fail(BRK_ERR_TOO_MANY_QUBITS, "Ising chain longer than %d", MAX_CHAIN);
  if (!J || !h) return fail(BRK_ERR_INVALID_ARGUMENT, "J or h is null");
  *a = ising_args{};
  a->n = (int)nq;
  for (unsigned i = 0; i + 1 < nq; ++i) a->J[i] = J[i];
  for (unsigned i = 0; i < nq; ++i) a->h[i] = h[i];
  return BRK_OK;
}

}  // namespace brk

using namespace brk;

extern "C" {

int brk_inner_product(const void* first1, int dtype1, const void* first2, int dtype2, uint64_t size,
                      double result[2], void* stream) {
  unsigned nq;
  if (int st = check_state(first1, size, dtype1, &nq)) return st;
  if (int st = check_state(first2, size, dtype2, &nq)) return st;
  if (!result) return fail(BRK_ERR_INVALID_ARGUMENT, "result is null");
  return run_reduction<2>(
      stream,
      [&](double* partial, hipStream_t s) {
        dim3 g(RED_BLOCKS), b(RED_THREADS);
        if (dtype1 == BRK_C128 && dtype2 == BRK_C128)
          hipLaunchKernelGGL((k_inner<double, double>), g, b, 0, s, (const cpx<double>*)first1,
                             (const cpx<double>*)first2, size, partial);
        else if (dtype1 == BRK_C128)
          hipLaunchKernelGGL((k_inner<double, float>), g, b, 0, s, (const cpx<double>*)first1,
                             (const cpx<float>*)first2, size, partial);
        else if (dtype2 == BRK_C128)
          hipLaunchKernelGGL((k_inner<float, double>), g, b, 0, s, (const cpx<float>*)first1,
                             (const cpx<double>*)first2, size, partial);
        else
          hipLaunchKernelGGL((k_inner<float, float>), g, b, 0, s, (const cpx<float>*)first1,
                             (const cpx<float>*)first2, size, partial);
      },
      result);
}

int brk_fidelity(const void* first1, int dtype1, const void* first2, int dtype2, uint64_t size, double* result,
                 void* stream) {
  if (!result) return fail(BRK_ERR_INVALID_ARGUMENT, "result is null");
  double ip[2];
  if (int st = brk_inner_product(first1, dtype1, first2, dtype2, size, ip, stream)) return st;
  *result = ip[0] * ip[0] + ip[1] * ip[1];
  return BRK_OK;
}

int brk_norm2(const void* first, uint64_t size, int dtype, double* result, void* stream) {
  unsigned nq;
  if (int st = check_state(first, size, dtype, &nq)) return st;
  if (!result) return fail(BRK_ERR_INVALID_ARGUMENT, "result is null");
  return run_reduction<1>(
      stream,
      [&](double* partial, hipStream_t s) {
        if (dtype == BRK_C128)
          hipLaunchKernelGGL(k_norm<double>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, (const cpx<double>*)first,
                             size, partial);
        else
          hipLaunchKernelGGL(k_norm<float>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, (const cpx<float>*)first,
                             size, partial);
      },
      result);
}

int brk_spin_expectation_value(const void* first, uint64_t size, int dtype, unsigned qubit, double result[3],
                               void* stream) {
  unsigned nq;
  if (int st = check_state(first, size, dtype, &nq)) return st;
  if (!result) return fail(BRK_ERR_INVALID_ARGUMENT, "result is null");
  if (qubit >= nq) return fail(BRK_ERR_QUBIT_OUT_OF_RANGE, "qubit %u out of range for %u qubits", qubit, nq);
  double raw[3];
  if (int st = spin_raw(first, size, dtype, qubit, raw, stream)) return st;
  result[0] = raw[0];
  result[1] = raw[1];
  result[2] = raw[2] * 0.5;  // spin_expectation_value.hpp:75-76
  return BRK_OK;
}

int brk_expectation_pauli_string(const void* first, uint64_t size, int dtype, const char* pauli_string,
                                 const unsigned* qubits, unsigned num_qubits, double result[2], void* stream) {
  unsigned nq;
  if (int st = check_state(first, size, dtype, &nq)) return st;
  if (!result || !pauli_string) return fail(BRK_ERR_INVALID_ARGUMENT, "null argument");
  unsigned ops[BRK_MAX_OPERATED_QUBITS];
  unsigned cnt = 0;
  if (int st = collect_qubits(qubits, num_qubits, nq, ops, &cnt, "operated")) return st;
  uint64_t flip = 0, sign = 0;
  unsigned ny = 0;
  for (unsigned k = 0; k < num_qubits; ++k) {
    const char c = pauli_string[k];
    const uint64_t bit = 1ull << qubits[k];
    if (c == 'X') flip |= bit;
    else if (c == 'Y') { flip |= bit; sign |= bit; ++ny; }
    else if (c == 'Z') sign |= bit;
    else if (c != 'I') return fail(BRK_ERR_INVALID_ARGUMENT, "bad Pauli character '%c'", c);
  }
  if (pauli_string[num_qubits] != '\0')
    return fail(BRK_ERR_INVALID_ARGUMENT, "Pauli string length differs from the number of qubits");
  double raw[2];
  int st = run_reduction<2>(
      stream,
      [&](double* partial, hipStream_t s) {
        if (dtype == BRK_C128)
          hipLaunchKernelGGL(k_pauli<double>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, (const cpx<double>*)first,
                             size, flip, sign, partial);
        else
          hipLaunchKernelGGL(k_pauli<float>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, (const cpx<float>*)first,
                             size, flip, sign, partial);
      },
      raw);
  if (st) return st;
  // coefficient (-i)^{#Y} (pauli_index_coeff.hpp:32-59: -i for Y on 0, +i for Y on 1)
  std::complex<double> r{raw[0], raw[1]};
  static const std::complex<double> phases[4] = {{1, 0}, {0, -1}, {-1, 0}, {0, 1}};
  r *= phases[ny & 3u];
  result[0] = r.real();
  result[1] = r.imag();
  return BRK_OK;
}

int brk_expectation_ising(const void* first, uint64_t size, int dtype, const double* J, const double* h,
                          double* result, void* stream) {
  unsigned nq;
  if (int st = check_state(first, size, dtype, &nq)) return st;
  if (!result) return fail(BRK_ERR_INVALID_ARGUMENT, "result is null");
  ising_args a;
  if (int st = make_ising_args(nq, J, h, &a)) return st;
  double zz = 0.0;
  int st = run_reduction<1>(
      stream,
      [&](double* partial, hipStream_t s) {
        if (dtype == BRK_C128)
          hipLaunchKernelGGL(k_ising_zz<double>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s,
                             (const cpx<double>*)first, size, a, partial);
        else
          hipLaunchKernelGGL(k_ising_zz<float>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s,
                             (const cpx<float>*)first, size, a, partial);
      },
      &zz);
  if (st) return st;
  double total = zz;
  for (unsigned q = 0; q < nq; ++q) {
    if (h[q] == 0.0) continue;
    double raw[3];
    if (int st2 = spin_raw(first, size, dtype, q, raw, stream)) return st2;
    total += h[q] * 2.0 * raw[0];  // <X_q> = 2 sum Re conj(a0) a1
  }
  *result = total;
  return BRK_OK;
}

int brk_apply_ising(void* out, const void* psi, uint64_t size, int dtype, const double* J, const double* h,
                    void* stream) {
  unsigned nq;
  if (int st = check_state(out, size, dtype, &nq)) return st;
  if (int st = check_state(psi, size, dtype, &nq)) return st;
  if (out == psi) return fail(BRK_ERR_INVALID_ARGUMENT, "apply_ising cannot work in place");
  ising_args a;
  if (int st = make_ising_args(nq, J, h, &a)) return st;
  uint64_t blocks = (size + 255) / 256;
  if (blocks > 0x7fffffffull) return fail(BRK_ERR_INVALID_ARGUMENT, "state too large");
  if (dtype == BRK_C128)
    hipLaunchKernelGGL(k_apply_ising<double>, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream),
                       (cpx<double>*)out, (const cpx<double>*)psi, size, a);
  else
    hipLaunchKernelGGL(k_apply_ising<float>, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream),
                       (cpx<float>*)out, (const cpx<float>*)psi, size, a);
  return check_launch("apply_ising");
}

}  // extern "C"
