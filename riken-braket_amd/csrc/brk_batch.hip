// Batching form: one persistent launch runs a list of small circuits and
// writes the FP64 Ising energy of each (ancestor in the reference: bra's
// CIRCUITS mode, which advances several states one after another,
// bra/src/bra.cpp:315-336, bra/src/interpreter.cpp:411-419).
//
// One 1024-thread workgroup per CU takes circuits from a global queue (most
// expensive first).  The program layout is described in brk_batch.hpp.  Per
// chunk: load (or initialise) into LDS, apply the stage's instructions with a
// barrier between them, accumulate the energy terms the stage measures, store.
// Gate instructions hold up to 4 local positions: each thread keeps the 16
// amplitudes of one group in registers and applies up to 8 2x2 gates to them.
// LDS addresses are XOR-swizzled, phys(j) = j ^ ((j >> 5) & 31), so that
// threads working on low positions do not collide on banks; the swizzle is
// linear over GF(2), so phys(base | off) = phys(base) ^ phys(off) for disjoint
// bit sets and a group needs one per-thread address.
#include "brk_internal.hpp"
#include "brk_batch.hpp"

namespace brk {
namespace {

using f2 = cpx<float>;

__device__ __forceinline__ uint32_t lpad(uint32_t j) { return j ^ ((j >> 5) & 31u); }

// Scatters the bits of c into the set bits of mask (lowest first).
__device__ __forceinline__ uint32_t deposit(uint32_t c, uint32_t mask) {
  uint32_t r = 0;
  for (uint32_t bit = 1; mask; mask &= mask - 1, bit <<= 1)
    if (c & bit) r |= mask & (0u - mask);
  return r;
}

__device__ __forceinline__ f2 gmul_add(f2 m0, f2 a0, f2 m1, f2 a1) {
  f2 r;
  r.re = m0.re * a0.re - m0.im * a0.im + m1.re * a1.re - m1.im * a1.im;
  r.im = m0.re * a0.im + m0.im * a0.re + m1.re * a1.im + m1.im * a1.re;
  return r;
}

template <int NP, int S>
__device__ __forceinline__ void apply_slot(f2 (&a)[1 << NP], const uint32_t* op) {
  const f2 m00 = {__uint_as_float(op[1]), __uint_as_float(op[2])};
  const f2 m01 = {__uint_as_float(op[3]), __uint_as_float(op[4])};
  const f2 m10 = {__uint_as_float(op[5]), __uint_as_float(op[6])};
  const f2 m11 = {__uint_as_float(op[7]), __uint_as_float(op[8])};
#pragma unroll
  for (int e = 0; e < (1 << NP); ++e) {
    if (e & (1 << S)) continue;
    const f2 a0 = a[e], a1 = a[e | (1 << S)];
    a[e] = gmul_add(m00, a0, m01, a1);
    a[e | (1 << S)] = gmul_add(m10, a0, m11, a1);
  }
}

template <int NP>
__device__ void gate_pass(f2* chunk, uint32_t L, const uint32_t* w, uint32_t tid) {
  const uint32_t nops = (w[0] >> 16) & 0xffu;
  uint32_t p[NP];
#pragma unroll
  for (int k = 0; k < NP; ++k) p[k] = (w[1] >> (8 * k)) & 0xffu;
  const uint32_t ngroups = 1u << (L - NP);
  for (uint32_t g = tid; g < ngroups; g += BATCH_THREADS) {
    uint32_t base = g;
#pragma unroll
    for (int k = 0; k < NP; ++k) base = ((base >> p[k]) << (p[k] + 1)) | (base & ((1u << p[k]) - 1u));
    const uint32_t pb = lpad(base);
    f2 a[1 << NP];
#pragma unroll
    for (int e = 0; e < (1 << NP); ++e) {
      uint32_t off = 0;
#pragma unroll
      for (int k = 0; k < NP; ++k) off |= ((uint32_t)(e >> k) & 1u) << p[k];
      a[e] = chunk[pb ^ lpad(off)];
    }
    // recompute the store addresses instead of keeping 2^NP of them live across the gates
    uint32_t pbs = pb;
    asm volatile("" : "+v"(pbs));
    const uint32_t* op = w + 2;
    for (uint32_t o = 0; o < nops; ++o, op += BATCH_GATE_WORDS) {
      switch (op[0]) {
        case 0: apply_slot<NP, 0>(a, op); break;
        case 1: if constexpr (NP > 1) apply_slot<NP, 1>(a, op); break;
        case 2: if constexpr (NP > 2) apply_slot<NP, 2>(a, op); break;
        default: if constexpr (NP > 3) apply_slot<NP, 3>(a, op); break;
      }
    }
#pragma unroll
    for (int e = 0; e < (1 << NP); ++e) {
      uint32_t off = 0;
#pragma unroll
      for (int k = 0; k < NP; ++k) off |= ((uint32_t)(e >> k) & 1u) << p[k];
      chunk[pbs ^ lpad(off)] = a[e];
    }
  }
}

__device__ __forceinline__ uint32_t global_index(uint32_t cbase, const uint32_t* tlo, const uint32_t* thi,
                                                 uint32_t j) {
  return cbase | tlo[j & 127u] | thi[j >> 7];
}

__global__ __launch_bounds__(BATCH_THREADS, 1) void k_batch(batch_launch P) {
  extern __shared__ __align__(16) unsigned char smem[];
  f2* chunk = reinterpret_cast<f2*>(smem);
  uint32_t* tlo = reinterpret_cast<uint32_t*>(smem + sizeof(f2) * BATCH_CHUNK_PADDED);
  uint32_t* thi = tlo + 128;
  double* ham = reinterpret_cast<double*>(thi + 128);  // J[0..n-2], h[0..n-1]
  uint32_t* next = reinterpret_cast<uint32_t*>(ham + 2 * BATCH_MAX_QUBITS);
  double* red = reinterpret_cast<double*>(smem);       // reuses the chunk after the last stage
  const uint32_t tid = threadIdx.x;
  f2* slot = P.slots ? reinterpret_cast<f2*>(P.slots) + (size_t)blockIdx.x * P.slot_elems : nullptr;

  for (;;) {
    if (tid == 0) *next = atomicAdd(P.counter, 1u);
    __syncthreads();
    const uint32_t k = __builtin_amdgcn_readfirstlane(*next);  // uniform: scalar loads of the program
    if (k >= P.num_circuits) break;
    const uint32_t ci = P.order[k];
    const batch_circuit C = P.circuits[ci];
    if (tid == 0) P.t_start[ci] = (uint64_t)wall_clock64();
    if (tid < 2 * C.n - 1) ham[tid] = P.ham[C.ham_offset + tid];
    const double* J = ham;
    const double* h = ham + (C.n - 1);
    double acc = 0.0;

    for (uint32_t s = 0; s < C.num_stages; ++s) {
      const batch_stage& S = P.stages[C.first_stage + s];
      const uint32_t L = S.L;
      if (tid < 128) {
        uint32_t lo = 0, hi = 0;
        for (uint32_t q = 0; q < 7; ++q) {
          if (q < L && ((tid >> q) & 1u)) lo |= 1u << S.local[q];
          if (q + 7 < L && ((tid >> q) & 1u)) hi |= 1u << S.local[q + 7];
        }
        tlo[tid] = lo;
        thi[tid] = hi;
      }
      __syncthreads();
      const uint32_t csize = 1u << L;
      const uint32_t nchunks = 1u << (C.n - L);
      for (uint32_t c = 0; c < nchunks; ++c) {
        const uint32_t cbase = deposit(c, S.nonlocal_mask);
        if (S.flags & BS_INIT) {
          for (uint32_t j = tid; j < csize; j += BATCH_THREADS) {
            const uint32_t x = global_index(cbase, tlo, thi, j);
            chunk[lpad(j)] = f2{x == 0 ? 1.0f : 0.0f, 0.0f};
          }
        } else {
          for (uint32_t j = tid; j < csize; j += BATCH_THREADS)
            chunk[lpad(j)] = slot[global_index(cbase, tlo, thi, j)];
        }
        __syncthreads();
        const uint32_t* w = P.words + S.first_word;
        for (uint32_t i = 0; i < S.num_instr; ++i) {
          const uint32_t type = w[0] & 0xffu;
          if (type == BI_GATES) {
            const uint32_t npos = (w[0] >> 8) & 0xffu;
            switch (npos) {
              case 1: gate_pass<1>(chunk, L, w, tid); break;
              case 2: gate_pass<2>(chunk, L, w, tid); break;
              case 3: gate_pass<3>(chunk, L, w, tid); break;
              default: gate_pass<4>(chunk, L, w, tid); break;
            }
            w += 2 + BATCH_GATE_WORDS * ((w[0] >> 16) & 0xffu);
          } else {
            const uint32_t nm = (w[0] >> 16) & 0xffu;
            for (uint32_t j = tid; j < csize; j += BATCH_THREADS) {
              const uint32_t x = global_index(cbase, tlo, thi, j);
              uint32_t par = 0;
              for (uint32_t m = 0; m < nm; ++m) par ^= (x & w[1 + m]) == w[1 + m] ? 1u : 0u;
              if (par) {
                const f2 v = chunk[lpad(j)];
                chunk[lpad(j)] = f2{-v.re, -v.im};
              }
            }
            w += 1 + nm;
          }
          __syncthreads();
        }
        if (S.flags & BS_ZZ) {
          for (uint32_t j = tid; j < csize; j += BATCH_THREADS) {
            const uint32_t x = global_index(cbase, tlo, thi, j);
            const f2 v = chunk[lpad(j)];
            const double p = (double)v.re * (double)v.re + (double)v.im * (double)v.im;
            const uint32_t d = x ^ (x >> 1);  // bit i: Z_i Z_i+1 = -1
            double e = 0.0;
            for (uint32_t i = 0; i + 1 < C.n; ++i) e += ((d >> i) & 1u) ? -J[i] : J[i];
            acc += p * e;
          }
        }
        for (uint32_t mx = S.measure_x; mx; mx &= mx - 1) {
          const uint32_t pos = __builtin_ctz(mx);
          const double h2 = 2.0 * h[S.local[pos]];
          for (uint32_t g = tid; g < (csize >> 1); g += BATCH_THREADS) {
            const uint32_t j0 = ((g >> pos) << (pos + 1)) | (g & ((1u << pos) - 1u));
            const f2 a0 = chunk[lpad(j0)], a1 = chunk[lpad(j0 | (1u << pos))];
            acc += h2 * ((double)a0.re * (double)a1.re + (double)a0.im * (double)a1.im);
          }
        }
        if (S.flags & BS_STORE) {
          for (uint32_t j = tid; j < csize; j += BATCH_THREADS)
            slot[global_index(cbase, tlo, thi, j)] = chunk[lpad(j)];
        }
        __syncthreads();
      }
    }

    // fixed-shape tree over the workgroup: the energy does not depend on timing
    red[tid] = acc;
    __syncthreads();
    for (uint32_t half = BATCH_THREADS / 2; half > 0; half >>= 1) {
      if (tid < half) red[tid] += red[tid + half];
      __syncthreads();
    }
    if (tid == 0) {
      P.energies[ci] = red[0];
      P.t_end[ci] = (uint64_t)wall_clock64();
    }
    __syncthreads();
  }
}

__global__ void k_batch_reset(uint32_t* counter) {
  if (threadIdx.x == 0) *counter = 0u;
}

}  // namespace

size_t batch_lds_bytes() {
  return sizeof(f2) * BATCH_CHUNK_PADDED + 2 * 128 * sizeof(uint32_t) + 2 * BATCH_MAX_QUBITS * sizeof(double) +
         16;
}

int configure_batch() {
  BRK_HIP_CHECK(hipFuncSetAttribute((const void*)k_batch, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)batch_lds_bytes()));
  return BRK_OK;
}

int launch_batch(const batch_launch& L, unsigned grid, void* stream) {
  const size_t lds = batch_lds_bytes();
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, as_stream(stream), L.counter);
  if (int s = check_launch("batch queue reset")) return s;
  hipLaunchKernelGGL(k_batch, dim3(grid), dim3(BATCH_THREADS), lds, as_stream(stream), L);
  return check_launch("batch");
}

}  // namespace brk
