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
#include "brk_block.hpp"

#include <algorithm>
#include <complex>
#include <vector>

namespace brk {

// Grid-stride loops load RED_UNROLL strides ahead before adding them in stride order (the
// same sums, more loads in flight per thread).
constexpr int RED_UNROLL = 4;

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
  auto add = [&](cpx<R1> x, cpx<R2> y) {
    const double xr = x.re, xi = x.im, yr = y.re, yi = y.im;
    // conj(y) * x
    acc.v[0] += yr * xr + yi * xi;
    acc.v[1] += yr * xi - yi * xr;
  };
  uint64_t i = (uint64_t)blockIdx.x * RED_THREADS + threadIdx.x;
  for (; i + (RED_UNROLL - 1) * stride < size; i += RED_UNROLL * stride) {
    cpx<R1> x[RED_UNROLL];
    cpx<R2> y[RED_UNROLL];
#pragma unroll
    for (int u = 0; u < RED_UNROLL; ++u) {
      x[u] = a[i + u * stride];
      y[u] = b[i + u * stride];
    }
#pragma unroll
    for (int u = 0; u < RED_UNROLL; ++u) add(x[u], y[u]);
  }
  for (; i < size; i += stride) add(a[i], b[i]);
  block_store<2>(acc, partial);
}

template <typename R>
__global__ __launch_bounds__(RED_THREADS) void k_norm(const cpx<R>* __restrict__ a, uint64_t size, double* partial) {
  dvec<1> acc{{0.0}};
  const uint64_t stride = (uint64_t)RED_BLOCKS * RED_THREADS;
  auto add = [&](cpx<R> x) {
    const double xr = x.re, xi = x.im;
    acc.v[0] += xr * xr + xi * xi;
  };
  uint64_t i = (uint64_t)blockIdx.x * RED_THREADS + threadIdx.x;
  for (; i + (RED_UNROLL - 1) * stride < size; i += RED_UNROLL * stride) {
    cpx<R> x[RED_UNROLL];
#pragma unroll
    for (int u = 0; u < RED_UNROLL; ++u) x[u] = a[i + u * stride];
#pragma unroll
    for (int u = 0; u < RED_UNROLL; ++u) add(x[u]);
  }
  for (; i < size; i += stride) add(a[i]);
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

// P(qubit = 0) and P(qubit = 1): sums of |a|^2 over the two halves of the pairs
// (projective_measurement.hpp:36-72, clear.hpp zero_probability).
template <typename R>
__global__ __launch_bounds__(RED_THREADS) void k_zero_one(const cpx<R>* __restrict__ a, uint64_t npairs, uint64_t lower,
                                                         uint64_t qmask, double* partial) {
  dvec<2> acc{{0.0, 0.0}};
  const uint64_t stride = (uint64_t)RED_BLOCKS * RED_THREADS;
  auto add = [&](cpx<R> x0, cpx<R> x1) {
    acc.v[0] += (double)x0.re * x0.re + (double)x0.im * x0.im;
    acc.v[1] += (double)x1.re * x1.re + (double)x1.im * x1.im;
  };
  auto first = [&](uint64_t g) { return ((g & ~lower) << 1) | (g & lower); };
  uint64_t g = (uint64_t)blockIdx.x * RED_THREADS + threadIdx.x;
  for (; g + (RED_UNROLL - 1) * stride < npairs; g += RED_UNROLL * stride) {
    cpx<R> x0[RED_UNROLL], x1[RED_UNROLL];
#pragma unroll
    for (int u = 0; u < RED_UNROLL; ++u) {
      const uint64_t i0 = first(g + u * stride);
      x0[u] = a[i0];
      x1[u] = a[i0 | qmask];
    }
#pragma unroll
    for (int u = 0; u < RED_UNROLL; ++u) add(x0[u], x1[u]);
  }
  for (; g < npairs; g += stride) {
    const uint64_t i0 = first(g);
    add(a[i0], a[i0 | qmask]);
  }
  block_store<2>(acc, partial);
}

// Keeps the half of the pairs where the qubit equals `value`, scaled by m, and
// sets the other half to +0 (change_state_after_measuring_zero/one,
// projective_measurement.hpp:74-150; clear.hpp; set.hpp).
template <typename R, bool NT>
__global__ __launch_bounds__(256) void k_project(cpx<R>* __restrict__ a, uint64_t npairs, uint64_t lower,
                                                 uint64_t qmask, int value, R m) {
  const uint64_t g = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= npairs) return;
  const uint64_t i0 = ((g & ~lower) << 1) | (g & lower);
  const uint64_t keep = value ? (i0 | qmask) : i0, drop = value ? i0 : (i0 | qmask);
  const cpx<R> x = load_amp<NT>(a + keep);  // NT: large states stream through once
  store_amp<NT>(a + keep, cpx<R>{x.re * m, x.im * m});
  store_amp<NT>(a + drop, cpx<R>{R(0), R(0)});
}

// Sampling (generate_events.hpp:15-48).  The cumulative probability of basis
// index i is a nested left fold: the state is cut into chunks of SAMPLE_CHUNK
// amplitudes, a chunk into 16 waves, a wave into 64 threads and a thread into
// `per` consecutive amplitudes.  With S the running sum over whole chunks,
// W over the waves of the chunk, L over the threads of the wave and c over the
// thread's own amplitudes (each summed left to right from 0.0),
//   C(i) = S + (W + (L + c)).
// The last value of each level equals the next level's running sum exactly,
// so C is monotone and continuous across chunk, wave and thread boundaries:
// an index whose probability is zero repeats its predecessor's C and is
// never the first to exceed a target.  It matches a sequential scan up to
// rounding.
constexpr int SAMPLE_CHUNK = 1 << 13;
constexpr int SAMPLE_THREADS = 1024;
constexpr int SAMPLE_WAVES = SAMPLE_THREADS / 64;

__device__ inline double readlane_f64(double v, int lane) {
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, lane);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), lane);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Coalesced load of one chunk's |a|^2 into LDS.  A whole chunk issues all its
// loads before the first norm is stored (eight in flight per thread).
template <typename R>
__device__ inline void load_norms(const cpx<R>* __restrict__ a, uint64_t base, uint32_t chunk, double* nrm) {
  constexpr int PER = SAMPLE_CHUNK / SAMPLE_THREADS;
  if (chunk == (uint32_t)SAMPLE_CHUNK) {
    cpx<R> x[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) x[k] = a[base + threadIdx.x + (uint32_t)k * SAMPLE_THREADS];
#pragma unroll
    for (int k = 0; k < PER; ++k)
      nrm[threadIdx.x + (uint32_t)k * SAMPLE_THREADS] = (double)x[k].re * x[k].re + (double)x[k].im * x[k].im;
    return;
  }
  for (uint32_t j = threadIdx.x; j < chunk; j += SAMPLE_THREADS) {
    const cpx<R> x = a[base + j];
    nrm[j] = (double)x.re * x.re + (double)x.im * x.im;
  }
}

struct chunk_fold {
  double lane_before;  // L before this thread
  double wave_before;  // W before this wave
  double total;        // the chunk's sum
};

// Called by every thread after the norms are in LDS (and a barrier).
__device__ inline chunk_fold fold_chunk(const double* nrm, uint32_t chunk, uint32_t per, double* wave_sums) {
  const uint32_t first = threadIdx.x * per;
  double t = 0.0;
  for (uint32_t k = 0; k < per; ++k)
    if (first + k < chunk) t += nrm[first + k];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double run = 0.0, before = 0.0;
  for (int k = 0; k < 64; ++k) {
    if (k == lane) before = run;
    run += readlane_f64(t, k);
  }
  if (lane == 0) wave_sums[wave] = run;
  __syncthreads();
  chunk_fold f{before, 0.0, 0.0};
  for (int k = 0; k < SAMPLE_WAVES; ++k) {
    if (k == wave) f.wave_before = f.total;
    f.total += wave_sums[k];
  }
  return f;
}

template <typename R>
__global__ __launch_bounds__(SAMPLE_THREADS) void k_chunk_sums(const cpx<R>* __restrict__ a, uint32_t chunk,
                                                              double* __restrict__ sums) {
  __shared__ double nrm[SAMPLE_CHUNK];
  __shared__ double wave_sums[SAMPLE_WAVES];
  const uint32_t per = chunk >= SAMPLE_THREADS ? chunk / SAMPLE_THREADS : 1;
  load_norms(a, (uint64_t)blockIdx.x * chunk, chunk, nrm);
  __syncthreads();
  const chunk_fold f = fold_chunk(nrm, chunk, per, wave_sums);
  if (threadIdx.x == 0) sums[blockIdx.x] = f.total;
}

// One block per event: the first index of chunk chunk_of[e] whose C exceeds
// target[e] (chunk_start[e] is S before the chunk).  When none does (a target
// at the total, which the host only sends for the last chunk holding
// probability), the chunk's last index of non-zero probability.
template <typename R>
__global__ __launch_bounds__(SAMPLE_THREADS) void k_chunk_search(const cpx<R>* __restrict__ a, uint32_t chunk,
                                                                const uint64_t* __restrict__ chunk_of,
                                                                const double* __restrict__ chunk_start,
                                                                const double* __restrict__ target,
                                                                uint64_t* __restrict__ events) {
  __shared__ double nrm[SAMPLE_CHUNK];
  __shared__ double wave_sums[SAMPLE_WAVES];
  __shared__ unsigned found, last_nonzero;
  const uint64_t base = chunk_of[blockIdx.x] * chunk;
  const double S = chunk_start[blockIdx.x], r = target[blockIdx.x];
  const uint32_t per = chunk >= SAMPLE_THREADS ? chunk / SAMPLE_THREADS : 1;
  if (threadIdx.x == 0) {
    found = ~0u;
    last_nonzero = 0u;
  }
  load_norms(a, base, chunk, nrm);
  __syncthreads();
  const chunk_fold f = fold_chunk(nrm, chunk, per, wave_sums);
  const uint32_t first = threadIdx.x * per;
  double c = 0.0;
  for (uint32_t k = 0; k < per && first + k < chunk; ++k) {
    const double p = nrm[first + k];
    c += p;
    if (p > 0.0) atomicMax(&last_nonzero, first + k);
    if (S + (f.wave_before + (f.lane_before + c)) > r) {
      atomicMin(&found, first + k);
      break;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) events[blockIdx.x] = base + (found != ~0u ? found : last_nonzero);
}

// Launches `launch(partial)` on the fixed reduction grid and sums the NV
// partials per block on the host.
template <int NV, typename Launch>
static int run_reduction(void* stream, Launch launch, double* out) {
  reduce_scratch* sc;
  if (int st = get_reduce_scratch(&sc)) return st;
  // the device's scratch is ours from the launch until the partials are read back
  std::lock_guard<std::mutex> hold(sc->busy);
  hipStream_t s = as_stream(stream);
  launch(sc->device, s);
  if (int st = check_launch("reduction")) return st;
  const hipError_t copied =
      hipMemcpyAsync(sc->host, sc->device, sizeof(double) * RED_BLOCKS * NV, hipMemcpyDeviceToHost, s);
  const hipError_t synced = hipStreamSynchronize(s);
  if (copied != hipSuccess) return fail(BRK_ERR_HIP, "reduction read-back: %s", hipGetErrorString(copied));
  if (synced != hipSuccess) return fail(BRK_ERR_HIP, "reduction: %s", hipGetErrorString(synced));
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

// Per event: target r = u * total, its chunk c = the first with S[c] > r (the running
// sums S over whole chunks, computed on the host left to right), or the last chunk
// holding probability when none is (r at the total), and the running sum before c.
__global__ __launch_bounds__(256) void k_locate_events(const double* __restrict__ S, uint64_t nchunks,
                                                       double total, uint64_t last_chunk,
                                                       const double* __restrict__ canonical, unsigned num_events,
                                                       uint64_t* __restrict__ chunk_of,
                                                       double* __restrict__ chunk_start, double* __restrict__ target) {
  const unsigned e = blockIdx.x * 256 + threadIdx.x;
  if (e >= num_events) return;
  const double r = canonical[e] * total;
  uint64_t lo = 0, hi = nchunks;
  while (lo < hi) {  // upper bound: the first index with S > r
    const uint64_t mid = lo + (hi - lo) / 2;
    if (S[mid] > r) hi = mid;
    else lo = mid + 1;
  }
  const uint64_t c = lo >= nchunks ? last_chunk : lo;
  chunk_of[e] = c;
  chunk_start[e] = c ? S[c - 1] : 0.0;
  target[e] = r;
}

// Running sums over whole chunks for the no-allocation sampling entry point:
// sums[c] becomes S[c] = S[c - 1] + sums[c] (S[-1] = 0), added left to right
// exactly as brk_generate_events adds them on the host, so both give the same
// S bit for bit.  One wavefront: every lane adds the same values in the same
// order (run is uniform), lane l keeps the value after the l-th add of its
// 64-chunk window.  meta[0] = S[nchunks - 1] (the total), meta[1] = the last
// chunk whose sum is positive as an integer (-1: none, a zero-norm state).
__global__ __launch_bounds__(64) void k_running_sums(double* __restrict__ sums, uint64_t nchunks,
                                                     double* __restrict__ meta) {
  const int lane = threadIdx.x;
  double run = 0.0;
  int64_t last = -1;
  for (uint64_t base = 0; base < nchunks; base += 64) {
    const uint64_t c = base + (uint64_t)lane;
    const double v = c < nchunks ? sums[c] : 0.0;
    if (c < nchunks && v > 0.0) last = (int64_t)c;
    double mine = 0.0;
    const uint64_t left = nchunks - base;
    if (left >= 64) {
#pragma unroll
      for (int l = 0; l < 64; ++l) {
        run = run + readlane_f64(v, l);
        if (l == lane) mine = run;
      }
    } else {
      for (int l = 0; l < (int)left; ++l) {
        run = run + readlane_f64(v, l);
        if (l == lane) mine = run;
      }
    }
    if (c < nchunks) sums[c] = mine;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const int64_t o = __shfl_xor(last, off, 64);
    last = o > last ? o : last;
  }
  if (lane == 0) {
    meta[0] = run;
    meta[1] = (double)last;
  }
}

// Device scratch of brk_generate_events_noalloc.  Each device holds its own
// copy for the life of the process: the chunk sums (then their running sums)
// of a state of up to NOALLOC_MAX_CHUNKS chunks of SAMPLE_CHUNK amplitudes
// (2^32 amplitudes, 32 qubits in either dtype), per event its draw, chunk,
// chunk start, target and result for up to NOALLOC_MAX_EVENTS events, then
// the two values k_running_sums writes -- 6.5 MiB in all.  Every element a
// call reads is written earlier in the same call, so nothing is cleared.
// Host code never passes `scratch` itself to a kernel or a copy (on the host
// it names a stand-in, not device memory): it uses the per-device address
// noalloc_scratch() obtains from hipGetSymbolAddress.
constexpr uint64_t NOALLOC_MAX_CHUNKS = 1ull << 19;
constexpr unsigned NOALLOC_MAX_EVENTS = 1u << 16;
constexpr uint64_t NOALLOC_DOUBLES = NOALLOC_MAX_CHUNKS + 5ull * NOALLOC_MAX_EVENTS + 2;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

constexpr int NOALLOC_MAX_DEVICES = 64;
// One lock per device, held by a call from its first launch until its events
// are back on the host: calls on one device (from any host thread) take turns.
static std::mutex g_noalloc_busy[NOALLOC_MAX_DEVICES];
static double* g_noalloc_ptr[NOALLOC_MAX_DEVICES];  // written under g_noalloc_busy[device]

// The current device's address of `scratch`, looked up on the device's first
// call (its lock held by the caller).
static int noalloc_scratch(int device, double** out) {
  if (!g_noalloc_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(BRK_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    g_noalloc_ptr[device] = static_cast<double*>(p);
  }
  *out = g_noalloc_ptr[device];
  return BRK_OK;
}

constexpr int ROW_THREADS = 256;

__global__ __launch_bounds__(ROW_THREADS) void k_row_sums(const double* __restrict__ in, uint64_t cols,
                                                          double scale, double* __restrict__ out) {
  __shared__ double part[ROW_THREADS];
  const double* row = in + (uint64_t)blockIdx.x * cols;
  double s = 0.0;
  for (uint64_t c = threadIdx.x; c < cols; c += ROW_THREADS) s += row[c];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = ROW_THREADS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = scale * part[0];
}

int sum_gradient_partials(const double* partials, unsigned rows, uint64_t chunks, double scale, double* out,
                          void* stream) {
  if (!rows) return BRK_OK;
  if (!partials || !out || !chunks) return fail(BRK_ERR_INVALID_ARGUMENT, "gradient partials: null or empty");
  if (rows > 0x7fffffffu) return fail(BRK_ERR_INVALID_ARGUMENT, "too many gradient rows");
  hipLaunchKernelGGL(k_row_sums, dim3(rows), dim3(ROW_THREADS), 0, as_stream(stream), partials, chunks, scale, out);
  return check_launch("gradient partial sums");
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

int brk_zero_one_probabilities(const void* first, uint64_t size, int dtype, unsigned qubit, double result[2],
                               void* stream) {
  unsigned nq;
  if (int st = check_state(first, size, dtype, &nq)) return st;
  if (!result) return fail(BRK_ERR_INVALID_ARGUMENT, "result is null");
  if (qubit >= nq) return fail(BRK_ERR_QUBIT_OUT_OF_RANGE, "qubit %u out of range for %u qubits", qubit, nq);
  const uint64_t qmask = 1ull << qubit, lower = qmask - 1ull, npairs = size >> 1;
  return run_reduction<2>(
      stream,
      [&](double* partial, hipStream_t s) {
        if (dtype == BRK_C128)
          hipLaunchKernelGGL(k_zero_one<double>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s,
                             (const cpx<double>*)first, npairs, lower, qmask, partial);
        else
          hipLaunchKernelGGL(k_zero_one<float>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s,
                             (const cpx<float>*)first, npairs, lower, qmask, partial);
      },
      result);
}

int brk_project_qubit(void* first, uint64_t size, int dtype, unsigned qubit, int value, double multiplier,
                      void* stream) {
  unsigned nq;
  if (int st = check_state(first, size, dtype, &nq)) return st;
  if (qubit >= nq) return fail(BRK_ERR_QUBIT_OUT_OF_RANGE, "qubit %u out of range for %u qubits", qubit, nq);
  if (value != 0 && value != 1) return fail(BRK_ERR_INVALID_ARGUMENT, "value must be 0 or 1");
  const uint64_t qmask = 1ull << qubit, lower = qmask - 1ull, npairs = size >> 1;
  const uint64_t blocks = (npairs + 255) / 256;
  if (blocks > 0xffffffffull) return fail(BRK_ERR_INVALID_ARGUMENT, "state too large");
  const bool nt = size * amp_bytes(dtype) >= get_tuning().nt_min_bytes;
  hipStream_t s = as_stream(stream);
  if (dtype == BRK_C128) {
    if (nt)
      hipLaunchKernelGGL((k_project<double, true>), dim3((unsigned)blocks), dim3(256), 0, s, (cpx<double>*)first,
                         npairs, lower, qmask, value, multiplier);
    else
      hipLaunchKernelGGL((k_project<double, false>), dim3((unsigned)blocks), dim3(256), 0, s, (cpx<double>*)first,
                         npairs, lower, qmask, value, multiplier);
  } else {
    if (nt)
      hipLaunchKernelGGL((k_project<float, true>), dim3((unsigned)blocks), dim3(256), 0, s, (cpx<float>*)first,
                         npairs, lower, qmask, value, (float)multiplier);
    else
      hipLaunchKernelGGL((k_project<float, false>), dim3((unsigned)blocks), dim3(256), 0, s, (cpx<float>*)first,
                         npairs, lower, qmask, value, (float)multiplier);
  }
  return check_launch("project");
}

int brk_generate_events(const void* first, uint64_t size, int dtype, const double* canonical, unsigned num_events,
                        uint64_t* events, void* stream) {
  unsigned nq;
  if (int st = check_state(first, size, dtype, &nq)) return st;
  if (num_events == 0) return BRK_OK;
  if (!canonical || !events) return fail(BRK_ERR_INVALID_ARGUMENT, "null argument");
  for (unsigned e = 0; e < num_events; ++e)
    if (!(canonical[e] >= 0.0 && canonical[e] < 1.0))
      return fail(BRK_ERR_INVALID_ARGUMENT, "canonical[%u] = %g is not in [0, 1)", e, canonical[e]);
  const uint32_t chunk = (uint32_t)std::min<uint64_t>(size, SAMPLE_CHUNK);
  const uint64_t nchunks = size / chunk;
  if (nchunks > 0x7fffffffull) return fail(BRK_ERR_INVALID_ARGUMENT, "state too large");
  hipStream_t s = as_stream(stream);
  // device buffers: chunk sums, then their running sums | per event: draw, chunk, chunk start,
  // target, result
  const size_t bytes = sizeof(double) * nchunks + (sizeof(uint64_t) * 2 + sizeof(double) * 3) * num_events;
  void* buffers = nullptr;
  BRK_HIP_CHECK(hipMalloc(&buffers, bytes));
  double* d_sums = static_cast<double*>(buffers);
  double* d_canonical = d_sums + nchunks;
  uint64_t* d_chunk = reinterpret_cast<uint64_t*>(d_canonical + num_events);
  double* d_start = reinterpret_cast<double*>(d_chunk + num_events);
  double* d_target = d_start + num_events;
  uint64_t* d_events = reinterpret_cast<uint64_t*>(d_target + num_events);
  if (dtype == BRK_C128)
    hipLaunchKernelGGL(k_chunk_sums<double>, dim3((unsigned)nchunks), dim3(SAMPLE_THREADS), 0, s,
                       (const cpx<double>*)first, chunk, d_sums);
  else
    hipLaunchKernelGGL(k_chunk_sums<float>, dim3((unsigned)nchunks), dim3(SAMPLE_THREADS), 0, s,
                       (const cpx<float>*)first, chunk, d_sums);
  std::vector<double> sums(nchunks);
  int status = check_launch("chunk sums");
  if (!status && (hipMemcpyAsync(sums.data(), d_sums, sizeof(double) * nchunks, hipMemcpyDeviceToHost, s) != hipSuccess ||
                  hipStreamSynchronize(s) != hipSuccess))
    status = fail(BRK_ERR_HIP, "copying the chunk sums failed");
  std::vector<double> S(nchunks);  // running sum through each chunk, left to right
  double run = 0.0;
  for (uint64_t c = 0; c < nchunks; ++c) S[c] = run = run + sums[c];
  uint64_t last_chunk = nchunks;  // the last chunk holding probability
  for (uint64_t c = nchunks; c-- > 0;)
    if (sums[c] > 0.0) {
      last_chunk = c;
      break;
    }
  if (!status && last_chunk == nchunks) status = fail(BRK_ERR_INVALID_ARGUMENT, "the state has zero norm");
  if (!status) {
    // the running sums replace the chunk sums on the device; each event finds its chunk there
    if (hipMemcpyAsync(d_sums, S.data(), sizeof(double) * nchunks, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(d_canonical, canonical, sizeof(double) * num_events, hipMemcpyHostToDevice, s) !=
            hipSuccess)
      status = fail(BRK_ERR_HIP, "uploading the running sums failed");
    if (!status) {
      hipLaunchKernelGGL(k_locate_events, dim3((num_events + 255) / 256), dim3(256), 0, s, d_sums, nchunks,
                         S[nchunks - 1], last_chunk, d_canonical, num_events, d_chunk, d_start, d_target);
      status = check_launch("locate events");
    }
  }
  if (!status) {
    if (dtype == BRK_C128)
      hipLaunchKernelGGL(k_chunk_search<double>, dim3(num_events), dim3(SAMPLE_THREADS), 0, s,
                         (const cpx<double>*)first, chunk, d_chunk, d_start, d_target, d_events);
    else
      hipLaunchKernelGGL(k_chunk_search<float>, dim3(num_events), dim3(SAMPLE_THREADS), 0, s,
                         (const cpx<float>*)first, chunk, d_chunk, d_start, d_target, d_events);
    status = check_launch("chunk search");
  }
  if (!status && (hipMemcpyAsync(events, d_events, sizeof(uint64_t) * num_events, hipMemcpyDeviceToHost, s) !=
                      hipSuccess ||
                  hipStreamSynchronize(s) != hipSuccess))
    status = fail(BRK_ERR_HIP, "copying the events failed");
  (void)hipStreamSynchronize(s);
  (void)hipFree(buffers);
  return status;
}

int brk_generate_events_noalloc(const void* first, uint64_t size, int dtype, const double* canonical,
                                unsigned num_events, uint64_t* events, void* stream) {
  unsigned nq;
  if (int st = check_state(first, size, dtype, &nq)) return st;
  if (num_events == 0) return BRK_OK;
  if (!canonical || !events) return fail(BRK_ERR_INVALID_ARGUMENT, "null argument");
  for (unsigned e = 0; e < num_events; ++e)
    if (!(canonical[e] >= 0.0 && canonical[e] < 1.0))
      return fail(BRK_ERR_INVALID_ARGUMENT, "canonical[%u] = %g is not in [0, 1)", e, canonical[e]);
  const uint32_t chunk = (uint32_t)std::min<uint64_t>(size, SAMPLE_CHUNK);
  const uint64_t nchunks = size / chunk;
  if (nchunks > NOALLOC_MAX_CHUNKS)
    return fail(BRK_ERR_INVALID_ARGUMENT,
                "brk_generate_events_noalloc takes states of at most %llu amplitudes (%llu given)",
                (unsigned long long)(NOALLOC_MAX_CHUNKS * SAMPLE_CHUNK), (unsigned long long)size);
  if (num_events > NOALLOC_MAX_EVENTS)
    return fail(BRK_ERR_INVALID_ARGUMENT, "brk_generate_events_noalloc takes at most %u events per call (%u given)",
                NOALLOC_MAX_EVENTS, num_events);
  int dev = 0;
  BRK_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
    return fail(BRK_ERR_INVALID_ARGUMENT, "device %d: no-allocation scratch exists for devices 0..%d", dev,
                NOALLOC_MAX_DEVICES - 1);
  std::lock_guard<std::mutex> hold(g_noalloc_busy[dev]);
  double* d_sums = nullptr;
  if (int st = noalloc_scratch(dev, &d_sums)) return st;
  double* d_canonical = d_sums + NOALLOC_MAX_CHUNKS;
  uint64_t* d_chunk = reinterpret_cast<uint64_t*>(d_canonical + NOALLOC_MAX_EVENTS);
  double* d_start = reinterpret_cast<double*>(d_chunk + NOALLOC_MAX_EVENTS);
  double* d_target = d_start + NOALLOC_MAX_EVENTS;
  uint64_t* d_events = reinterpret_cast<uint64_t*>(d_target + NOALLOC_MAX_EVENTS);
  double* d_meta = reinterpret_cast<double*>(d_events + NOALLOC_MAX_EVENTS);
  hipStream_t s = as_stream(stream);
  // every exit after the first launch waits for the stream, so the next call
  // (which takes the lock) never overlaps this call's kernels on the scratch
  auto run = [&]() -> int {
    if (dtype == BRK_C128)
      hipLaunchKernelGGL(k_chunk_sums<double>, dim3((unsigned)nchunks), dim3(SAMPLE_THREADS), 0, s,
                         (const cpx<double>*)first, chunk, d_sums);
    else
      hipLaunchKernelGGL(k_chunk_sums<float>, dim3((unsigned)nchunks), dim3(SAMPLE_THREADS), 0, s,
                         (const cpx<float>*)first, chunk, d_sums);
    if (int st = check_launch("chunk sums")) return st;
    hipLaunchKernelGGL(k_running_sums, dim3(1), dim3(64), 0, s, d_sums, nchunks, d_meta);
    if (int st = check_launch("running sums")) return st;
    double meta[2] = {0.0, -1.0};  // total, last chunk holding probability
    if (hipMemcpyAsync(meta, d_meta, sizeof(meta), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return fail(BRK_ERR_HIP, "copying the running-sum total failed");
    if (meta[1] < 0.0) return fail(BRK_ERR_INVALID_ARGUMENT, "the state has zero norm");
    if (hipMemcpyAsync(d_canonical, canonical, sizeof(double) * num_events, hipMemcpyHostToDevice, s) !=
        hipSuccess)
      return fail(BRK_ERR_HIP, "uploading the draws failed");
    hipLaunchKernelGGL(k_locate_events, dim3((num_events + 255) / 256), dim3(256), 0, s, d_sums, nchunks, meta[0],
                       (uint64_t)meta[1], d_canonical, num_events, d_chunk, d_start, d_target);
    if (int st = check_launch("locate events")) return st;
    if (dtype == BRK_C128)
      hipLaunchKernelGGL(k_chunk_search<double>, dim3(num_events), dim3(SAMPLE_THREADS), 0, s,
                         (const cpx<double>*)first, chunk, d_chunk, d_start, d_target, d_events);
    else
      hipLaunchKernelGGL(k_chunk_search<float>, dim3(num_events), dim3(SAMPLE_THREADS), 0, s,
                         (const cpx<float>*)first, chunk, d_chunk, d_start, d_target, d_events);
    if (int st = check_launch("chunk search")) return st;
    if (hipMemcpyAsync(events, d_events, sizeof(uint64_t) * num_events, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return fail(BRK_ERR_HIP, "copying the events failed");
    return BRK_OK;
  };
  const int status = run();
  (void)hipStreamSynchronize(s);
  return status;
}

}  // extern "C"
