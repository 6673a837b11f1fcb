// Batching form: one persistent launch runs a list of small circuits and
// writes the FP64 Ising energy of each (ancestor in the reference: bra's
// CIRCUITS mode, which advances several states one after another,
// bra/src/bra.cpp:315-336, bra/src/interpreter.cpp:411-419).
//
// Two 512-thread workgroups per CU take circuits from a global queue (most
// expensive first); while one is in a barrier-separated LDS phase the other
// computes.  The program layout is described in brk_batch.hpp.  Per
// chunk: load (or initialise) into LDS, apply the stage's instructions with a
// barrier between them, accumulate the energy terms the stage measures, store.
// A gate group holds up to 4 local positions: each thread keeps the 16
// amplitudes of one group in registers, flips the signs of the CZ/Z terms
// attached before and after the gates, and applies one 2x2 matrix per slot.
// The program is read through the constant address space: it does not change
// during the launch, so the compiler may use scalar loads for it.
// LDS addresses are XOR-swizzled, phys(j) = j ^ ((j >> 5) & 31), so that
// threads working on low positions do not collide on banks; the swizzle is
// linear over GF(2), so phys(base | off) = phys(base) ^ phys(off) for disjoint
// bit sets and a group needs one per-thread address.
#include "brk_internal.hpp"
#include "brk_batch.hpp"

namespace brk {
namespace {

using f2 = cpx<float>;
typedef const __attribute__((address_space(4))) uint32_t* cu32;
typedef const __attribute__((address_space(4))) double* cf64;
typedef const __attribute__((address_space(4))) batch_stage* cstage;
typedef const __attribute__((address_space(4))) batch_circuit* ccircuit;

constexpr int ZZ_TABLES = 3;  // byte tables of the ZZ energy: n - 1 <= 23 bond bits

__device__ __forceinline__ uint32_t lpad(uint32_t j) { return j ^ ((j >> 5) & 31u); }

// Scatters the bits of c into the set bits of mask (lowest first).
__device__ __forceinline__ uint32_t deposit(uint32_t c, uint32_t mask) {
  uint32_t r = 0;
  for (uint32_t bit = 1; mask; mask &= mask - 1, bit <<= 1)
    if (c & bit) r |= mask & (0u - mask);
  return r;
}

__device__ __forceinline__ uint32_t term_bits(uint32_t x, uint32_t d, uint32_t A) {
  if (d < BATCH_TERM_ALL) return __builtin_popcount(x & (x >> d) & A);  // d = 0: x & x = x
  return (x & A) == A ? 1u : 0u;
}

__device__ __forceinline__ f2 flip(f2 a, uint32_t odd) {
  const uint32_t s = odd << 31;
  return f2{__uint_as_float(__float_as_uint(a.re) ^ s), __uint_as_float(__float_as_uint(a.im) ^ s)};
}

// Sign terms attached to a gate group are of the Z / CZ kinds only (d < 32).  Element e of a
// thread's group has the index x = xb | xoff[e]: the thread's bits xb and the element's bits
// xoff[e] (one qubit q_j per position j) are disjoint, so x & (x >> d) & A is the disjoint union
// of four parts and its bit count, mod 2, is
//   popcount(xb & (xb >> d) & A)                                  the thread's own sign
// + popcount(xoff[e] & M),  M = ((xb & A) << d) ^ ((xb >> d) & A)  linear in e: bit q_j of M per position
// + popcount(xoff[e] & (xoff[e] >> d) & A)                        the same for every thread.
// The planner sums the last part over the terms (bit e of `uniform`: its parity for element e);
// the others take a few instructions per thread and term, and the element signs follow from the
// positions' bits of M by one XOR each.

template <int NP>
__device__ __forceinline__ void apply_terms(f2 (&a)[1 << NP], uint32_t xb, const uint32_t (&q)[NP], cu32 t,
                                            uint32_t nterms, uint32_t uniform) {
  uint32_t own = 0, M = 0;
  for (uint32_t k = 0; k < nterms; ++k) {
    const uint32_t d = t[2 * k], A = t[2 * k + 1];
    own += __builtin_popcount(xb & (xb >> d) & A);
    M ^= ((xb & A) << d) ^ ((xb >> d) & A);
  }
  // sign bit (bit 31) of element e, without the uniform part: own + sum over the positions j set in e
  // of bit q_j of M
  uint32_t sg[1 << NP];
  sg[0] = own << 31;
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const uint32_t mj = (M >> q[j]) << 31;
#pragma unroll
    for (int e = 0; e < (1 << j); ++e) sg[e | (1 << j)] = sg[e] ^ mj;
  }
  if (uniform) {  // (the planner moves them into the real form's phase tables)
    // the element masks are formed here, not kept live across the pass's loop
    uint32_t u = uniform;
    asm volatile("" : "+s"(u));
#pragma unroll
    for (int e = 0; e < (1 << NP); ++e) sg[e] ^= ((u >> e) & 1u) << 31;
  }
#pragma unroll
  for (int e = 0; e < (1 << NP); ++e)
    a[e] = f2{__uint_as_float(__float_as_uint(a[e].re) ^ sg[e]), __uint_as_float(__float_as_uint(a[e].im) ^ sg[e])};
}

// m0 a0 + m1 a1 as one multiply and three fused multiply-adds per component
// (the library builds without FP contraction; the batching form fuses on purpose).
__device__ __forceinline__ f2 gmul_add(f2 m0, f2 a0, f2 m1, f2 a1) {
  f2 r;
  r.re = __builtin_fmaf(m0.re, a0.re, __builtin_fmaf(-m0.im, a0.im, __builtin_fmaf(m1.re, a1.re, -m1.im * a1.im)));
  r.im = __builtin_fmaf(m0.re, a0.im, __builtin_fmaf(m0.im, a0.re, __builtin_fmaf(m1.re, a1.im, m1.im * a1.re)));
  return r;
}

template <int NP, int S>
__device__ __forceinline__ void apply_slot(f2 (&a)[1 << NP], cu32 op) {
  const f2 m00 = {__uint_as_float(op[0]), __uint_as_float(op[1])};
  const f2 m01 = {__uint_as_float(op[2]), __uint_as_float(op[3])};
  const f2 m10 = {__uint_as_float(op[4]), __uint_as_float(op[5])};
  const f2 m11 = {__uint_as_float(op[6]), __uint_as_float(op[7])};
#pragma unroll
  for (int e = 0; e < (1 << NP); ++e) {
    if (e & (1 << S)) continue;
    const f2 a0 = a[e], a1 = a[e | (1 << S)];
    a[e] = gmul_add(m00, a0, m01, a1);
    a[e | (1 << S)] = gmul_add(m10, a0, m11, a1);
  }
}

// Rotation [[c, -s], [s, c]], |theta| <= pi/2, on the pairs of slot S as three shears
// (t = tan(theta/2), |t| <= 1): a0 -= t a1; a1 += s a0; a0 -= t a1, which leaves
// a0 = c a0 - s a1 and a1 = s a0 + c a1.  Three fused multiply-adds per component.
template <int NP, int S>
__device__ __forceinline__ void apply_real_slot(f2 (&a)[1 << NP], cu32 op) {
  const float t = __uint_as_float(op[0]), s = __uint_as_float(op[1]);
#pragma unroll
  for (int e = 0; e < (1 << NP); ++e) {
    if (e & (1 << S)) continue;
    f2 x = a[e], y = a[e | (1 << S)];
    x.re = __builtin_fmaf(-t, y.re, x.re);
    x.im = __builtin_fmaf(-t, y.im, x.im);
    y.re = __builtin_fmaf(s, x.re, y.re);
    y.im = __builtin_fmaf(s, x.im, y.im);
    x.re = __builtin_fmaf(-t, y.re, x.re);
    x.im = __builtin_fmaf(-t, y.im, x.im);
    a[e] = x;
    a[e | (1 << S)] = y;
  }
}

// a[e] *= table[e] (complex64 words, the same for every thread of the pass).
template <int NP>
__device__ __forceinline__ void apply_phases(f2 (&a)[1 << NP], cu32 table) {
#pragma unroll
  for (int e = 0; e < (1 << NP); ++e) {
    const float zr = __uint_as_float(table[2 * e]), zi = __uint_as_float(table[2 * e + 1]);
    const f2 v = a[e];
    a[e] = f2{__builtin_fmaf(zr, v.re, -zi * v.im), __builtin_fmaf(zr, v.im, zi * v.re)};
  }
}

struct chunk_ctx {
  f2* chunk;
  const uint32_t* tlo;
  const uint32_t* thi;
  uint32_t cbase;  // global index bits of the chunk
  uint32_t L;
  cstage S;
};

__device__ __forceinline__ uint32_t global_index(const chunk_ctx& c, uint32_t j) {
  return c.cbase | c.tlo[j & 127u] | c.thi[j >> 7];
}

// FORM: BI_GATES (complex 2x2 per position), BI_GATES_REAL, BI_GATES_REAL_PRE (brk_batch.hpp).
template <int NP, uint32_t FORM>
__device__ void gate_pass(const chunk_ctx& c, cu32 w, uint32_t tid) {
  const uint32_t w0 = w[0], pw = w[1], qw = w[2];
  const uint32_t npre = (w0 >> 16) & 0xffu, npost = (w0 >> 24) & 0xffu;
  const uint32_t pre_uniform = w[3] & 0xffffu, post_uniform = w[3] >> 16;
  const cu32 pre = w + 4;
  const cu32 post = pre + 2 * npre;
  const cu32 ops = post + 2 * npost;
  uint32_t p[NP], q[NP];
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    p[k] = (pw >> (8 * k)) & 0xffu;
    q[k] = (qw >> (8 * k)) & 0xffu;
  }
  // per group element (uniform): swizzled LDS byte offset, the XOR of its positions' offsets (the
  // swizzle is linear over GF(2))
  uint32_t boff[1 << NP];
  boff[0] = 0;
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    const uint32_t bk = lpad(1u << p[k]) * 8u;
#pragma unroll
    for (int e = 0; e < (1 << k); ++e) boff[e | (1 << k)] = boff[e] ^ bk;
  }
  const uint32_t ngroups = 1u << (c.L - NP);
  for (uint32_t g = tid; g < ngroups; g += BATCH_THREADS) {
    uint32_t base = g;
#pragma unroll
    for (int k = 0; k < NP; ++k) base = ((base >> p[k]) << (p[k] + 1)) | (base & ((1u << p[k]) - 1u));
    // byte addresses: (phys(base) ^ phys(off)) * 8 = phys(base) * 8 ^ phys(off) * 8
    const uint32_t pb = lpad(base) * 8u;
    char* const bytes = reinterpret_cast<char*>(c.chunk);
    f2 a[1 << NP];
#pragma unroll
    for (int e = 0; e < (1 << NP); ++e) a[e] = *reinterpret_cast<const f2*>(bytes + (pb ^ boff[e]));
    uint32_t xb = 0;
    if (npre | npost) xb = global_index(c, base);
    if (npre) apply_terms<NP>(a, xb, q, pre, npre, pre_uniform);
    if constexpr (FORM == BI_GATES) {
      apply_slot<NP, 0>(a, ops);
      if constexpr (NP > 1) apply_slot<NP, 1>(a, ops + BATCH_GATE_WORDS);
      if constexpr (NP > 2) apply_slot<NP, 2>(a, ops + 2 * BATCH_GATE_WORDS);
      if constexpr (NP > 3) apply_slot<NP, 3>(a, ops + 3 * BATCH_GATE_WORDS);
    } else {
      const cu32 post_table = ops + NP * BATCH_REAL_WORDS;
      if constexpr (FORM == BI_GATES_REAL_PRE) apply_phases<NP>(a, post_table + 2 * (1 << NP));
      apply_real_slot<NP, 0>(a, ops);
      if constexpr (NP > 1) apply_real_slot<NP, 1>(a, ops + BATCH_REAL_WORDS);
      if constexpr (NP > 2) apply_real_slot<NP, 2>(a, ops + 2 * BATCH_REAL_WORDS);
      if constexpr (NP > 3) apply_real_slot<NP, 3>(a, ops + 3 * BATCH_REAL_WORDS);
      apply_phases<NP>(a, post_table);
    }
    if (npost) apply_terms<NP>(a, xb, q, post, npost, post_uniform);
#pragma unroll
    for (int e = 0; e < (1 << NP); ++e) *reinterpret_cast<f2*>(bytes + (pb ^ boff[e])) = a[e];
  }
}

// A sign pass of its own: 8 amplitudes per thread per step, terms outermost.
__device__ void sign_pass(const chunk_ctx& c, cu32 w, uint32_t tid) {
  const uint32_t nterms = (w[0] >> 16) & 0xffffu;
  const cu32 t = w + 1;
  const uint32_t csize = 1u << c.L;
  for (uint32_t j0 = tid; j0 < csize; j0 += 8 * BATCH_THREADS) {
    uint32_t x[8], par[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const uint32_t j = j0 + r * BATCH_THREADS;
      x[r] = j < csize ? global_index(c, j) : 0u;
      par[r] = 0;
    }
    for (uint32_t k = 0; k < nterms; ++k) {
      const uint32_t d = t[2 * k], A = t[2 * k + 1];
#pragma unroll
      for (int r = 0; r < 8; ++r) par[r] += term_bits(x[r], d, A);
    }
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const uint32_t j = j0 + r * BATCH_THREADS;
      if (j < csize && (par[r] & 1u)) c.chunk[lpad(j)] = flip(c.chunk[lpad(j)], 1u);
    }
  }
}

constexpr int CHUNK_PER_THREAD = BATCH_CHUNK / BATCH_THREADS;
// (the second bound is waves per SIMD: BATCH_WG_PER_CU workgroups of BATCH_THREADS / 64 waves over
// 4 SIMDs, which caps the registers at 512 / 4 = 128 VGPRs)
constexpr int BATCH_WAVES_PER_SIMD = BATCH_WG_PER_CU * BATCH_THREADS / 64 / 4;

__global__ __launch_bounds__(BATCH_THREADS, BATCH_WAVES_PER_SIMD) void k_batch(batch_launch P) {
  // static LDS: addresses are compile-time constants, so group offsets fold into the instructions
  __shared__ __align__(16) f2 chunk[BATCH_CHUNK_PADDED];
  __shared__ uint32_t tlo[128], thi[128];
  __shared__ double zz[ZZ_TABLES * 256];  // ZZ energy byte tables
  __shared__ uint32_t next[1];
  double* red = reinterpret_cast<double*>(chunk);  // reuses the chunk after the last stage
  const uint32_t tid = threadIdx.x;
  f2* slot = P.slots ? reinterpret_cast<f2*>(P.slots) + (size_t)blockIdx.x * P.slot_elems : nullptr;
  const ccircuit circuits = (ccircuit)P.circuits;
  const cstage stages = (cstage)P.stages;
  const cu32 words = (cu32)P.words;
  const cf64 ham = (cf64)P.ham;
  const cu32 order = (cu32)P.order;

  for (;;) {
    if (tid == 0) *next = atomicAdd(P.counter, 1u);
    __syncthreads();
    const uint32_t k = __builtin_amdgcn_readfirstlane(*next);
    if (k >= P.num_circuits) break;
    const uint32_t ci = order[k];
    const uint32_t n = circuits[ci].n;
    const uint32_t num_stages = circuits[ci].num_stages;
    const uint32_t first_stage = circuits[ci].first_stage;
    const cf64 J = ham + circuits[ci].ham_offset;
    const cf64 h = J + (n - 1);
    if (tid == 0) P.t_start[ci] = (uint64_t)wall_clock64();
    // ZZ energy of a bond pattern d (bit i set: Z_i Z_i+1 = -1) = sum of its three byte-table entries
    for (uint32_t t = tid; t < ZZ_TABLES * 256; t += BATCH_THREADS) {
      const uint32_t byte = t >> 8, b = t & 255u;
      double e = 0.0;
      for (uint32_t i = 0; i < 8; ++i) {
        const uint32_t bond = 8 * byte + i;
        if (bond + 1 < n) e += ((b >> i) & 1u) ? -J[bond] : J[bond];
      }
      zz[t] = e;
    }
    double acc = 0.0;

    for (uint32_t s = 0; s < num_stages; ++s) {
      const cstage S = stages + first_stage + s;
      const uint32_t L = S->L, flags = S->flags, nonlocal = S->nonlocal_mask, measure_x = S->measure_x;
      const uint32_t num_instr = S->num_instr;
      const cu32 first_word = words + S->first_word;
      if (tid < 128) {
        uint32_t lo = 0, hi = 0;
        for (uint32_t q = 0; q < 7; ++q) {
          if (q < L && ((tid >> q) & 1u)) lo |= 1u << S->local[q];
          if (q + 7 < L && ((tid >> q) & 1u)) hi |= 1u << S->local[q + 7];
        }
        tlo[tid] = lo;
        thi[tid] = hi;
      }
      __syncthreads();
      const uint32_t csize = 1u << L;
      const uint32_t nchunks = 1u << (n - L);
      for (uint32_t cidx = 0; cidx < nchunks; ++cidx) {
        const chunk_ctx c{chunk, tlo, thi, deposit(cidx, nonlocal), L, S};
        if (flags & BS_INIT) {
          for (uint32_t j = tid; j < csize; j += BATCH_THREADS)
            chunk[lpad(j)] = f2{global_index(c, j) == 0 ? 1.0f : 0.0f, 0.0f};
        } else {
          // a circuit with stages after the first has n > BATCH_MAX_LOCAL, so its chunks are full
          // (csize = BATCH_CHUNK); all of the thread's loads go out at once (a loop would wait for
          // each in turn)
          f2 v[CHUNK_PER_THREAD];
#pragma unroll
          for (int r = 0; r < CHUNK_PER_THREAD; ++r) v[r] = slot[global_index(c, tid + BATCH_THREADS * r)];
#pragma unroll
          for (int r = 0; r < CHUNK_PER_THREAD; ++r) chunk[lpad(tid + BATCH_THREADS * r)] = v[r];
        }
        __syncthreads();
        cu32 w = first_word;
        for (uint32_t i = 0; i < num_instr; ++i) {
          const uint32_t w0 = w[0], op = w0 & 0xffu;
          if (op == BI_GATES) {
            const uint32_t npos = (w0 >> 8) & 0xffu;
            switch (npos) {
              case 1: gate_pass<1, BI_GATES>(c, w, tid); break;
              case 2: gate_pass<2, BI_GATES>(c, w, tid); break;
              case 3: gate_pass<3, BI_GATES>(c, w, tid); break;
              default: gate_pass<4, BI_GATES>(c, w, tid); break;
            }
            w += 4 + 2 * (((w0 >> 16) & 0xffu) + (w0 >> 24)) + BATCH_GATE_WORDS * npos;
          } else if (op == BI_GATES_REAL || op == BI_GATES_REAL_PRE) {
            // the planner emits BI_GATES_REAL for 1..4 positions, BI_GATES_REAL_PRE for 2..4 (for one
            // position with both tables the complex form is cheaper)
            const uint32_t npos = (w0 >> 8) & 0xffu;
            if (op == BI_GATES_REAL) {
              switch (npos) {
                case 1: gate_pass<1, BI_GATES_REAL>(c, w, tid); break;
                case 2: gate_pass<2, BI_GATES_REAL>(c, w, tid); break;
                case 3: gate_pass<3, BI_GATES_REAL>(c, w, tid); break;
                default: gate_pass<4, BI_GATES_REAL>(c, w, tid); break;
              }
            } else {
              switch (npos) {
                case 2: gate_pass<2, BI_GATES_REAL_PRE>(c, w, tid); break;
                case 3: gate_pass<3, BI_GATES_REAL_PRE>(c, w, tid); break;
                default: gate_pass<4, BI_GATES_REAL_PRE>(c, w, tid); break;
              }
            }
            w += 4 + 2 * (((w0 >> 16) & 0xffu) + (w0 >> 24)) + BATCH_REAL_WORDS * npos +
                 2 * (1u << npos) * (op == BI_GATES_REAL_PRE ? 2u : 1u);
          } else {
            sign_pass(c, w, tid);
            w += 1 + 2 * ((w0 >> 16) & 0xffffu);
          }
          __syncthreads();
        }
        if (flags & BS_ZZ) {
          const uint32_t bonds = n > 1 ? (1u << (n - 1)) - 1u : 0u;
          for (uint32_t j = tid; j < csize; j += BATCH_THREADS) {
            const uint32_t x = global_index(c, j);
            const f2 v = chunk[lpad(j)];
            const double p = (double)v.re * (double)v.re + (double)v.im * (double)v.im;
            const uint32_t d = (x ^ (x >> 1)) & bonds;
            acc += p * (zz[d & 255u] + zz[256 + ((d >> 8) & 255u)] + zz[512 + (d >> 16)]);
          }
        }
        for (uint32_t mx = measure_x; mx; mx &= mx - 1) {
          const uint32_t pos = __builtin_ctz(mx);
          const double h2 = 2.0 * h[S->local[pos]];
          for (uint32_t g = tid; g < (csize >> 1); g += BATCH_THREADS) {
            const uint32_t j0 = ((g >> pos) << (pos + 1)) | (g & ((1u << pos) - 1u));
            const f2 a0 = chunk[lpad(j0)], a1 = chunk[lpad(j0 | (1u << pos))];
            acc += h2 * ((double)a0.re * (double)a1.re + (double)a0.im * (double)a1.im);
          }
        }
        if (flags & BS_STORE) {
          f2 v[CHUNK_PER_THREAD];  // (full chunks, as for the loads)
#pragma unroll
          for (int r = 0; r < CHUNK_PER_THREAD; ++r) v[r] = chunk[lpad(tid + BATCH_THREADS * r)];
#pragma unroll
          for (int r = 0; r < CHUNK_PER_THREAD; ++r) slot[global_index(c, tid + BATCH_THREADS * r)] = v[r];
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
  return sizeof(f2) * BATCH_CHUNK_PADDED + 2 * 128 * sizeof(uint32_t) + ZZ_TABLES * 256 * sizeof(double) + 4;
}

int configure_batch() { return BRK_OK; }  // static LDS: nothing to opt in to

int launch_batch(const batch_launch& L, unsigned grid, void* stream) {
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, as_stream(stream), L.counter);
  if (int s = check_launch("batch queue reset")) return s;
  hipLaunchKernelGGL(k_batch, dim3(grid), dim3(BATCH_THREADS), 0, as_stream(stream), L);
  return check_launch("batch");
}

}  // namespace brk
