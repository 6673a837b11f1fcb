// Differentiation form: the reverse-sweep step of the adjoint method.
//
// One launch handles a chain of up to ADJ_MAX_CHAIN single-qubit ops on the
// same qubit (e.g. the RY, RZ pair of an ansatz layer).  Per amplitude pair
// (psi0, psi1), (lam0, lam1) of that qubit it
//   * accumulates, for every parametric op k of the chain (in reverse order),
//     Re(conj(lam0) (D_k psi)_0 + conj(lam1) (D_k psi)_1) in FP64, where
//     D_k = (dU_k/dtheta) U_k^dagger maps the state after op k to the
//     derivative of the state, and then
//   * un-applies op k to both states (psi, lam := U_k^dagger psi, U_k^dagger lam),
// reading and writing each of the two states once for the whole chain.
// Partial sums are reduced per block and written to a fixed slot (fixed grid,
// fixed order: the gradient is bit-reproducible).
#include "brk_internal.hpp"
#include "brk_adjoint.hpp"

namespace brk {

constexpr int ADJ_THREADS = 256;

struct cd_dev {
  double re, im;
};

__device__ __forceinline__ cd_dev dmul(cd_dev a, cd_dev b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cd_dev dadd(cd_dev a, cd_dev b) { return {a.re + b.re, a.im + b.im}; }

template <typename R, int NOPS, bool NT>
__global__ __launch_bounds__(ADJ_THREADS) void k_adjoint_chain(cpx<R>* __restrict__ psi, cpx<R>* __restrict__ lam,
                                                              uint64_t npairs, uint64_t lower, uint64_t tbit,
                                                              adj_chain c, double* __restrict__ partials) {
  double acc[NOPS];
#pragma unroll
  for (int k = 0; k < NOPS; ++k) acc[k] = 0.0;
  const uint64_t stride = (uint64_t)ADJ_BLOCKS * ADJ_THREADS;
  for (uint64_t g = (uint64_t)blockIdx.x * ADJ_THREADS + threadIdx.x; g < npairs; g += stride) {
    const uint64_t i0 = ((g & ~lower) << 1) | (g & lower);
    const uint64_t i1 = i0 | tbit;
    const cpx<R> a0 = load_amp<NT>(psi + i0), a1 = load_amp<NT>(psi + i1);
    const cpx<R> b0 = load_amp<NT>(lam + i0), b1 = load_amp<NT>(lam + i1);
    cd_dev p0{a0.re, a0.im}, p1{a1.re, a1.im}, l0{b0.re, b0.im}, l1{b1.re, b1.im};
#pragma unroll
    for (int k = 0; k < NOPS; ++k) {
      const adj_op& o = c.ops[k];
      if (o.has_grad) {
        const cd_dev d00{o.d[0], o.d[1]}, d01{o.d[2], o.d[3]}, d10{o.d[4], o.d[5]}, d11{o.d[6], o.d[7]};
        const cd_dev q0 = dadd(dmul(d00, p0), dmul(d01, p1));
        const cd_dev q1 = dadd(dmul(d10, p0), dmul(d11, p1));
        // Re(conj(l) q)
        acc[k] += (l0.re * q0.re + l0.im * q0.im) + (l1.re * q1.re + l1.im * q1.im);
      }
      const cd_dev m00{o.m[0], o.m[1]}, m01{o.m[2], o.m[3]}, m10{o.m[4], o.m[5]}, m11{o.m[6], o.m[7]};
      const cd_dev np0 = dadd(dmul(m00, p0), dmul(m01, p1)), np1 = dadd(dmul(m10, p0), dmul(m11, p1));
      const cd_dev nl0 = dadd(dmul(m00, l0), dmul(m01, l1)), nl1 = dadd(dmul(m10, l0), dmul(m11, l1));
      p0 = np0;
      p1 = np1;
      l0 = nl0;
      l1 = nl1;
      if constexpr (sizeof(R) == 4) {  // complex64 states are rounded after every op, as a per-gate run would
        p0 = {(double)(R)p0.re, (double)(R)p0.im};
        p1 = {(double)(R)p1.re, (double)(R)p1.im};
        l0 = {(double)(R)l0.re, (double)(R)l0.im};
        l1 = {(double)(R)l1.re, (double)(R)l1.im};
      }
    }
    store_amp<NT>(psi + i0, cpx<R>{(R)p0.re, (R)p0.im});
    store_amp<NT>(psi + i1, cpx<R>{(R)p1.re, (R)p1.im});
    store_amp<NT>(lam + i0, cpx<R>{(R)l0.re, (R)l0.im});
    store_amp<NT>(lam + i1, cpx<R>{(R)l1.re, (R)l1.im});
  }
  __shared__ double red[ADJ_THREADS / 64][ADJ_MAX_CHAIN];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NOPS; ++k) {
    double x = acc[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    if (lane == 0) red[wave][k] = x;
  }
  __syncthreads();
  if (threadIdx.x < NOPS) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < ADJ_THREADS / 64; ++w) t += red[w][threadIdx.x];
    partials[(size_t)threadIdx.x * ADJ_BLOCKS + blockIdx.x] = t;
  }
}

template <typename R, int NOPS>
static void launch_n(const adj_chain& c, void* psi, void* lam, uint64_t size, bool nt, double* partials,
                     hipStream_t st) {
  const uint64_t tbit = 1ull << c.target, lower = tbit - 1ull, npairs = size >> 1;
  if (nt)
    hipLaunchKernelGGL((k_adjoint_chain<R, NOPS, true>), dim3(ADJ_BLOCKS), dim3(ADJ_THREADS), 0, st,
                       (cpx<R>*)psi, (cpx<R>*)lam, npairs, lower, tbit, c, partials);
  else
    hipLaunchKernelGGL((k_adjoint_chain<R, NOPS, false>), dim3(ADJ_BLOCKS), dim3(ADJ_THREADS), 0, st,
                       (cpx<R>*)psi, (cpx<R>*)lam, npairs, lower, tbit, c, partials);
}

template <typename R>
static void launch_r(const adj_chain& c, void* psi, void* lam, uint64_t size, bool nt, double* partials,
                     hipStream_t st) {
  switch (c.nops) {
    case 1: launch_n<R, 1>(c, psi, lam, size, nt, partials, st); break;
    case 2: launch_n<R, 2>(c, psi, lam, size, nt, partials, st); break;
    case 3: launch_n<R, 3>(c, psi, lam, size, nt, partials, st); break;
    default: launch_n<R, 4>(c, psi, lam, size, nt, partials, st); break;
  }
}

// Whole-state sign pass: every amplitude is read and written, two per thread
// 256 apart so that each wavefront instruction covers one contiguous run
// (whole lines), blocks in the XCD-contiguous order.
template <typename R, bool NT>
__global__ __launch_bounds__(256) void k_sign_terms(cpx<R>* __restrict__ s, uint64_t size, sign_terms t) {
  // size is a power of two: blocks of 512 amplitudes, or one block of `size` when smaller
  const uint64_t stride = size >= 512 ? 256 : size / 2;
  const uint64_t i = xcd_tile(blockIdx.x, gridDim.x, 1) * 512 + threadIdx.x;
  if (threadIdx.x >= stride) return;
  cpx<R> v[2];
#pragma unroll
  for (int u = 0; u < 2; ++u) v[u] = load_amp<NT>(s + i + u * stride);
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const uint64_t x = i + u * stride;
    uint32_t par = 0;
    for (int k = 0; k < t.n; ++k) par += (uint32_t)__popcll(x & (x >> t.d[k]) & t.a[k]);
    if (par & 1u) v[u] = cpx<R>{-v[u].re, -v[u].im};
    store_amp<NT>(s + x, v[u]);
  }
}

int launch_sign_terms(const sign_terms& t, void* first, uint64_t size, int dtype, void* stream) {
  if (t.n < 0 || t.n > SIGN_MAX_TERMS) return fail(BRK_ERR_INVALID_ARGUMENT, "bad sign term count");
  if (t.n == 0) return BRK_OK;
  const uint64_t blocks = (size + 511) / 512;  // size < 512: one block, threads [0, size/2) each take i, i + size/2
  if (blocks > 0xffffffffull) return fail(BRK_ERR_INVALID_ARGUMENT, "state too large for the sign pass");
  const bool nt = size * amp_bytes(dtype) >= get_tuning().nt_min_bytes;
  hipStream_t st = as_stream(stream);
  if (dtype == BRK_C128) {
    if (nt) hipLaunchKernelGGL((k_sign_terms<double, true>), dim3((unsigned)blocks), dim3(256), 0, st, (cpx<double>*)first, size, t);
    else hipLaunchKernelGGL((k_sign_terms<double, false>), dim3((unsigned)blocks), dim3(256), 0, st, (cpx<double>*)first, size, t);
  } else {
    if (nt) hipLaunchKernelGGL((k_sign_terms<float, true>), dim3((unsigned)blocks), dim3(256), 0, st, (cpx<float>*)first, size, t);
    else hipLaunchKernelGGL((k_sign_terms<float, false>), dim3((unsigned)blocks), dim3(256), 0, st, (cpx<float>*)first, size, t);
  }
  return check_launch("sign terms");
}

int launch_adjoint_chain(const adj_chain& c, void* psi, void* lam, uint64_t size, int dtype, double* partials,
                         void* stream) {
  if (c.nops < 1 || c.nops > ADJ_MAX_CHAIN) return fail(BRK_ERR_INVALID_ARGUMENT, "bad adjoint chain length");
  const bool nt = size * amp_bytes(dtype) >= get_tuning().nt_min_bytes;
  if (dtype == BRK_C128) launch_r<double>(c, psi, lam, size, nt, partials, as_stream(stream));
  else launch_r<float>(c, psi, lam, size, nt, partials, as_stream(stream));
  return check_launch("adjoint chain");
}

}  // namespace brk
