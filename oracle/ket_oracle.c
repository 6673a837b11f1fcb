/*
 * ket_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY; see ket_oracle.h).
 *
 * A plain-C restatement of the reference ket library's per-gate loops.
 * Each function cites the reference function it restates.  The reference
 * cannot be compiled in this environment (its gate headers include Boost,
 * which the image lacks), so this restatement is pinned by the tests in
 * tests/test_oracle_*.py against dense matrices, the reference tests'
 * fixture states, closed forms and invariants.
 *
 * Complex arithmetic is written out on (re, im) pairs with the operation
 * order of the reference's std::complex expressions.
 */
#include "ket_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

typedef struct { double re, im; } cd;

static inline cd cd_load(const double* s, uint64_t i) { cd v = {s[2 * i], s[2 * i + 1]}; return v; }
static inline void cd_store(double* s, uint64_t i, cd v) { s[2 * i] = v.re; s[2 * i + 1] = v.im; }
static inline cd cd_add(cd a, cd b) { cd r = {a.re + b.re, a.im + b.im}; return r; }
static inline cd cd_sub(cd a, cd b) { cd r = {a.re - b.re, a.im - b.im}; return r; }
static inline cd cd_mul(cd a, cd b) { cd r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; return r; }
static inline cd cd_scale(cd a, double s) { cd r = {a.re * s, a.im * s}; return r; }
static inline cd cd_conj(cd a) { cd r = {a.re, -a.im}; return r; }
static inline cd cd_make(double re, double im) { cd r = {re, im}; return r; }
static inline cd cd_expi(double phase) { return cd_make(cos(phase), sin(phase)); }

/* boost::math::constants::one_div_root_two<double>() (hadamard.hpp:66-70) */
static const double ONE_DIV_ROOT_TWO = 0.707106781186547524400844362104849039284835937688474036588;

void or_set_num_threads(int n) { if (n > 0) omp_set_num_threads(n); }
int or_get_num_threads(void) { return omp_get_max_threads(); }

/* ------------------------------------------------------------------------- */
/* Index masks -- gate_bitmasks.hpp:50-130 (runtime make_index_masks),       */
/* index_with_qubits.hpp:91-112 (bit-mask form)                               */
/* ------------------------------------------------------------------------- */

static int cmp_unsigned(const void* a, const void* b) {
  unsigned x = *(const unsigned*)a, y = *(const unsigned*)b;
  return (x > y) - (x < y);
}

void or_make_index_masks(const unsigned* qubits, unsigned k, uint64_t* index_masks) {
  if (k == 0) { index_masks[0] = ~(uint64_t)0; return; }
  unsigned sorted[64];
  memcpy(sorted, qubits, k * sizeof(unsigned));
  qsort(sorted, k, sizeof(unsigned), cmp_unsigned);
  uint64_t previous = ((uint64_t)1 << sorted[0]) - 1u;
  index_masks[0] = previous;
  for (unsigned j = 1; j < k; ++j) {
    uint64_t partial = ((uint64_t)1 << (sorted[j] - j)) - 1u;
    index_masks[j] = partial ^ previous;
    previous = partial;
  }
  index_masks[k] = (~(uint64_t)0) ^ previous;
}

static inline uint64_t base_index(uint64_t g, const uint64_t* im, unsigned k) {
  uint64_t r = 0;
  for (unsigned j = 0; j <= k; ++j) r |= (g & im[j]) << j;
  return r;
}

uint64_t or_index_with_qubits(uint64_t g, uint64_t v, const unsigned* qubits, unsigned k,
                              const uint64_t* index_masks) {
  uint64_t r = base_index(g, index_masks, k);
  for (unsigned j = 0; j < k; ++j)
    if ((v >> j) & 1u) r |= (uint64_t)1 << qubits[j];
  return r;
}

static int popcount64(uint64_t x) { return __builtin_popcountll(x); }

static uint64_t mask_of(const unsigned* q, unsigned n) {
  uint64_t m = 0;
  for (unsigned i = 0; i < n; ++i) m |= (uint64_t)1 << q[i];
  return m;
}

/* operated qubits = list1 ++ list2 */
static unsigned join_qubits(unsigned* out, const unsigned* a, unsigned na, const unsigned* b, unsigned nb) {
  for (unsigned i = 0; i < na; ++i) out[i] = a[i];
  for (unsigned i = 0; i < nb; ++i) out[na + i] = b[i];
  return na + nb;
}

/* ------------------------------------------------------------------------- */
/* Single-target pair loop: for every group, z = base | controls,            */
/* o = z | target.  The body gets (z, o).                                     */
/* ------------------------------------------------------------------------- */

#define PAIR_LOOP(SIZE, TARGET, CONTROLS, NC, ...)                                     \
  do {                                                                                 \
    unsigned ops_[64];                                                                 \
    unsigned k_ = join_qubits(ops_, &(TARGET), 1, (CONTROLS), (NC));                   \
    uint64_t im_[65];                                                                  \
    or_make_index_masks(ops_, k_, im_);                                                \
    const uint64_t cm_ = mask_of((CONTROLS), (NC));                                    \
    const uint64_t tm_ = (uint64_t)1 << (TARGET);                                      \
    const int64_t n_ = (int64_t)((SIZE) >> k_);                                        \
    _Pragma("omp parallel for schedule(static)")                                       \
    for (int64_t g_ = 0; g_ < n_; ++g_) {                                              \
      const uint64_t z = base_index((uint64_t)g_, im_, k_) | cm_;                      \
      const uint64_t o = z | tm_;                                                      \
      __VA_ARGS__                                                                       \
    }                                                                                  \
  } while (0)

/* H -- hadamard.hpp:54-71 (and the controlled forms :77-223) */
void or_hadamard(double* s, uint64_t size, unsigned target, const unsigned* controls, unsigned nc) {
  PAIR_LOOP(size, target, controls, nc, {
    const cd zv = cd_load(s, z);
    cd zn = cd_add(zv, cd_load(s, o));
    zn = cd_scale(zn, ONE_DIV_ROOT_TWO);
    cd on = cd_sub(zv, cd_load(s, o));
    on = cd_scale(on, ONE_DIV_ROOT_TWO);
    cd_store(s, z, zn);
    cd_store(s, o, on);
  });
}

/* U2 -- phase_shift.hpp:653-709: z -= e2*o; z *= 1/sqrt2; o *= e2; o += z_old; o *= e1/sqrt2 */
void or_phase_shift2(double* s, uint64_t size, double phase1, double phase2, unsigned target,
                     const unsigned* controls, unsigned nc) {
  const cd e1 = cd_expi(phase1), e2 = cd_expi(phase2);
  const cd me1 = cd_scale(e1, ONE_DIV_ROOT_TWO);
  PAIR_LOOP(size, target, controls, nc, {
    const cd zv = cd_load(s, z);
    cd ov = cd_load(s, o);
    cd zn = cd_sub(zv, cd_mul(e2, ov));
    zn = cd_scale(zn, ONE_DIV_ROOT_TWO);
    ov = cd_mul(ov, e2);
    ov = cd_add(ov, zv);
    ov = cd_mul(ov, me1);
    cd_store(s, z, zn);
    cd_store(s, o, ov);
  });
}

/* U3 -- phase_shift.hpp:1219-1252:
   z *= cos; z -= (sin e3) o; o *= (cos e3); o += sin z_old; o *= e2 */
void or_phase_shift3(double* s, uint64_t size, double phase1, double phase2, double phase3, unsigned target,
                     const unsigned* controls, unsigned nc) {
  const double sine = sin(0.5 * phase1), cosine = cos(0.5 * phase1);
  const cd e2 = cd_expi(phase2), e3 = cd_expi(phase3);
  const cd se3 = cd_scale(e3, sine), ce3 = cd_scale(e3, cosine);
  PAIR_LOOP(size, target, controls, nc, {
    const cd zv = cd_load(s, z);
    cd ov = cd_load(s, o);
    cd zn = cd_scale(zv, cosine);
    zn = cd_sub(zn, cd_mul(se3, ov));
    ov = cd_mul(ov, ce3);
    ov = cd_add(ov, cd_scale(zv, sine));
    ov = cd_mul(ov, e2);
    cd_store(s, z, zn);
    cd_store(s, o, ov);
  });
}

/* general 2x2: (z, o) := M (z, o) -- the functor form of gate_bitmasks.hpp:470-586 */
void or_gate_1q(double* s, uint64_t size, unsigned target, const double m[8], const unsigned* controls,
                unsigned nc) {
  const cd m00 = cd_make(m[0], m[1]), m01 = cd_make(m[2], m[3]);
  const cd m10 = cd_make(m[4], m[5]), m11 = cd_make(m[6], m[7]);
  PAIR_LOOP(size, target, controls, nc, {
    const cd zv = cd_load(s, z), ov = cd_load(s, o);
    cd_store(s, z, cd_add(cd_mul(m00, zv), cd_mul(m01, ov)));
    cd_store(s, o, cd_add(cd_mul(m10, zv), cd_mul(m11, ov)));
  });
}

/* ------------------------------------------------------------------------- */
/* Multi-target flips: pairs (i, i ^ T) with the first target's bit of i = 0  */
/* ------------------------------------------------------------------------- */

#define FLIP_LOOP(SIZE, TARGETS, NT, CONTROLS, NC, ...)                                \
  do {                                                                                 \
    unsigned ops_[64];                                                                 \
    unsigned k_ = join_qubits(ops_, (TARGETS), (NT), (CONTROLS), (NC));                \
    uint64_t im_[65];                                                                  \
    or_make_index_masks(ops_, k_, im_);                                                \
    const uint64_t cm_ = mask_of((CONTROLS), (NC));                                    \
    const uint64_t tm_ = mask_of((TARGETS), (NT));                                     \
    const unsigned nt_ = (NT);                                                         \
    const int64_t n_ = (int64_t)((SIZE) >> k_);                                        \
    _Pragma("omp parallel for schedule(static)")                                       \
    for (int64_t g_ = 0; g_ < n_; ++g_) {                                              \
      const uint64_t b_ = base_index((uint64_t)g_, im_, k_) | cm_;                     \
      for (uint64_t v_ = 0; v_ < ((uint64_t)1 << nt_); v_ += 2) {                      \
        uint64_t z = b_;                                                               \
        for (unsigned j_ = 1; j_ < nt_; ++j_)                                          \
          if ((v_ >> j_) & 1u) z |= (uint64_t)1 << (TARGETS)[j_];                      \
        const uint64_t o = z ^ tm_;                                                    \
        __VA_ARGS__                                                                   \
      }                                                                                \
    }                                                                                  \
  } while (0)

/* X...X under controls -- pauli_x.hpp:31-61 (iter_swap of the pair), :116-165 (CX),
   multi-target forms after :166 */
void or_pauli_x(double* s, uint64_t size, const unsigned* targets, unsigned nt, const unsigned* controls,
                unsigned nc) {
  FLIP_LOOP(size, targets, nt, controls, nc, {
    const cd zv = cd_load(s, z), ov = cd_load(s, o);
    cd_store(s, z, ov);
    cd_store(s, o, zv);
  });
}

/* Y...Y under controls (pauli_y.hpp; one target: a0' = -i a1, a1' = i a0,
   docs/bra.md "Y t"): (Y^m a)[i] = (-i)^m (-1)^{popcount(i & T)} a[i ^ T] */
void or_pauli_y(double* s, uint64_t size, const unsigned* targets, unsigned nt, const unsigned* controls,
                unsigned nc) {
  static const cd minus_i_pow[4] = {{1, 0}, {0, -1}, {-1, 0}, {0, 1}};
  const cd ph = minus_i_pow[nt & 3u];
  FLIP_LOOP(size, targets, nt, controls, nc, {
    const cd kz = (popcount64(z & tm_) & 1) ? cd_scale(ph, -1.0) : ph;
    const cd ko = (popcount64(o & tm_) & 1) ? cd_scale(ph, -1.0) : ph;
    const cd zv = cd_load(s, z), ov = cd_load(s, o);
    cd_store(s, z, cd_mul(kz, ov));
    cd_store(s, o, cd_mul(ko, zv));
  });
}

/* exp(i t X..X) -- exponential_pauli_x.hpp:36-84:
   z *= cos; z += o * (i sin); o *= cos; o += z_old * (i sin) */
void or_exponential_pauli_x(double* s, uint64_t size, double phase, const unsigned* targets, unsigned nt,
                            const unsigned* controls, unsigned nc) {
  const cd e = cd_expi(phase);
  const double c = e.re;
  const cd is = cd_make(0.0, e.im);
  FLIP_LOOP(size, targets, nt, controls, nc, {
    const cd zv = cd_load(s, z);
    cd ov = cd_load(s, o);
    cd zn = cd_scale(zv, c);
    zn = cd_add(zn, cd_mul(ov, is));
    ov = cd_scale(ov, c);
    ov = cd_add(ov, cd_mul(zv, is));
    cd_store(s, z, zn);
    cd_store(s, o, ov);
  });
}

/* exp(i t Y..Y) -- exponential_pauli_y.hpp:36-90 (one target:
   z *= cos; z += o sin; o *= cos; o -= z_old sin).  With m targets,
   (Y^m a)[i] = (-i)^m (-1)^{popcount(i & T)} a[i ^ T]. */
void or_exponential_pauli_y(double* s, uint64_t size, double phase, const unsigned* targets, unsigned nt,
                            const unsigned* controls, unsigned nc) {
  const cd e = cd_expi(phase);
  const double c = e.re;
  static const cd minus_i_pow[4] = {{1, 0}, {0, -1}, {-1, 0}, {0, 1}};
  const cd base_phase = cd_mul(cd_make(0.0, e.im), minus_i_pow[nt & 3u]); /* i sin (-i)^m */
  FLIP_LOOP(size, targets, nt, controls, nc, {
    const int pz = popcount64(z & tm_), po = popcount64(o & tm_);
    const cd kz = (pz & 1) ? cd_scale(base_phase, -1.0) : base_phase;
    const cd ko = (po & 1) ? cd_scale(base_phase, -1.0) : base_phase;
    const cd zv = cd_load(s, z);
    cd ov = cd_load(s, o);
    cd zn = cd_scale(zv, c);
    zn = cd_add(zn, cd_mul(ov, kz));
    ov = cd_scale(ov, c);
    ov = cd_add(ov, cd_mul(zv, ko));
    cd_store(s, z, zn);
    cd_store(s, o, ov);
  });
}

/* SWAP under controls -- swap.hpp:54-74: iter_swap(base|q1, base|q2) */
void or_swap(double* s, uint64_t size, unsigned q1, unsigned q2, const unsigned* controls, unsigned nc) {
  unsigned ops[64];
  ops[0] = q1;
  ops[1] = q2;
  unsigned k = 2 + join_qubits(ops + 2, controls, nc, NULL, 0);
  uint64_t im[65];
  or_make_index_masks(ops, k, im);
  const uint64_t cm = mask_of(controls, nc);
  const uint64_t m1 = (uint64_t)1 << q1, m2 = (uint64_t)1 << q2;
  const int64_t n = (int64_t)(size >> k);
#pragma omp parallel for schedule(static)
  for (int64_t g = 0; g < n; ++g) {
    const uint64_t b = base_index((uint64_t)g, im, k) | cm;
    const cd x = cd_load(s, b | m1), y = cd_load(s, b | m2);
    cd_store(s, b | m1, y);
    cd_store(s, b | m2, x);
  }
}

/* exp(i phase SWAP) under controls -- exponential_swap.hpp:31-90: |00>, |11> times
   e^{i phase}; then a_{10} <- cos a_{10} + i sin a_{01}, a_{01} <- cos a_{01} + i sin a_{10}(old),
   in the reference's order (qubit1's bit set = "qubit1_on") */
void or_exponential_swap(double* s, uint64_t size, double phase, unsigned q1, unsigned q2, const unsigned* controls,
                         unsigned nc) {
  unsigned ops[64];
  ops[0] = q1;
  ops[1] = q2;
  unsigned k = 2 + join_qubits(ops + 2, controls, nc, NULL, 0);
  uint64_t im[65];
  or_make_index_masks(ops, k, im);
  const uint64_t cm = mask_of(controls, nc);
  const uint64_t m1 = (uint64_t)1 << q1, m2 = (uint64_t)1 << q2;
  const cd e = cd_expi(phase), i_sin = cd_make(0.0, e.im);
  const int64_t n = (int64_t)(size >> k);
#pragma omp parallel for schedule(static)
  for (int64_t g = 0; g < n; ++g) {
    const uint64_t b = base_index((uint64_t)g, im, k) | cm;
    cd_store(s, b, cd_mul(cd_load(s, b), e));
    cd_store(s, b | m1 | m2, cd_mul(cd_load(s, b | m1 | m2), e));
    const cd one = cd_load(s, b | m1), two = cd_load(s, b | m2);
    cd_store(s, b | m1, cd_add(cd_scale(one, e.re), cd_mul(two, i_sin)));
    cd_store(s, b | m2, cd_add(cd_scale(two, e.re), cd_mul(one, i_sin)));
  }
}

/* ------------------------------------------------------------------------- */
/* Diagonal gates                                                            */
/* ------------------------------------------------------------------------- */

/* Z on control qubits: negate where every listed bit is 1 -- pauli_z.hpp:55-210 */
void or_pauli_z(double* s, uint64_t size, const unsigned* qubits, unsigned nq) {
  uint64_t im[65];
  or_make_index_masks(qubits, nq, im);
  const uint64_t qm = mask_of(qubits, nq);
  const int64_t n = (int64_t)(size >> nq);
#pragma omp parallel for schedule(static)
  for (int64_t g = 0; g < n; ++g) {
    const uint64_t i = base_index((uint64_t)g, im, nq) | qm;
    const cd v = cd_load(s, i);
    cd_store(s, i, cd_make(-v.re, -v.im));
  }
}

/* Z on ket::qubit targets under controls (case 2) -- pauli_z.hpp:265-303 (Z, ZZ), :392-470
   (C...CZ...Z): per group of the operated qubits (targets, then controls), every local value
   with all control bits set whose target part has an odd number of ones is multiplied by -1. */
void or_pauli_zn(double* s, uint64_t size, const unsigned* targets, unsigned nt, const unsigned* controls,
                 unsigned nc) {
  unsigned q[128];
  const unsigned k = join_qubits(q, targets, nt, controls, nc);
  uint64_t im[65];
  or_make_index_masks(q, k, im);
  const uint64_t base = (((uint64_t)1 << nc) - 1u) << nt;
  const int64_t n = (int64_t)(size >> k);
#pragma omp parallel for schedule(static)
  for (int64_t g = 0; g < n; ++g)
    for (uint64_t v = 0; v < ((uint64_t)1 << nt); ++v) {
      if ((popcount64(v) & 1) == 0) continue;
      const uint64_t i = or_index_with_qubits((uint64_t)g, base + v, q, k, im);
      const cd a = cd_load(s, i);
      cd_store(s, i, cd_make(-a.re, -a.im));
    }
}

/* sqrt(Z) (adjoint: sqrt(Z)^dagger) on ket::qubit targets under controls -- sqrt_pauli_z.hpp:
   565-700 (:846-1000): every local value with all control bits set is multiplied by
   1, i, -1, -i (1, -i, -1, i) as its target part has 0, 1, 2, 3 ones mod 4. */
void or_sqrt_pauli_zn(double* s, uint64_t size, const unsigned* targets, unsigned nt, const unsigned* controls,
                      unsigned nc, int adjoint) {
  unsigned q[128];
  const unsigned k = join_qubits(q, targets, nt, controls, nc);
  uint64_t im[65];
  or_make_index_masks(q, k, im);
  const uint64_t base = (((uint64_t)1 << nc) - 1u) << nt;
  const double sgn = adjoint ? -1.0 : 1.0;
  const cd coefficient[4] = {cd_make(1.0, 0.0), cd_make(0.0, sgn), cd_make(-1.0, 0.0), cd_make(0.0, -sgn)};
  const int64_t n = (int64_t)(size >> k);
#pragma omp parallel for schedule(static)
  for (int64_t g = 0; g < n; ++g)
    for (uint64_t v = 0; v < ((uint64_t)1 << nt); ++v) {
      const uint64_t i = or_index_with_qubits((uint64_t)g, base + v, q, k, im);
      cd_store(s, i, cd_mul(cd_load(s, i), coefficient[popcount64(v) & 3]));
    }
}

/* U1 via coefficient on control qubits -- phase_shift.hpp:37-93 (`*it *= coeff`) */
void or_phase_shift_coeff(double* s, uint64_t size, const double coeff[2], const unsigned* qubits, unsigned nq) {
  uint64_t im[65];
  or_make_index_masks(qubits, nq, im);
  const uint64_t qm = mask_of(qubits, nq);
  const cd c = cd_make(coeff[0], coeff[1]);
  const int64_t n = (int64_t)(size >> nq);
#pragma omp parallel for schedule(static)
  for (int64_t g = 0; g < n; ++g) {
    const uint64_t i = base_index((uint64_t)g, im, nq) | qm;
    cd_store(s, i, cd_mul(cd_load(s, i), c));
  }
}

/* exp(i t Z..Z) under controls -- exponential_pauli_z.hpp:36-78:
   even parity of the target bits *= e^{it}, odd parity *= e^{-it} */
void or_exponential_pauli_z(double* s, uint64_t size, double phase, const unsigned* targets, unsigned nt,
                            const unsigned* controls, unsigned nc) {
  const cd e = cd_expi(phase), ce = cd_conj(e);
  uint64_t im[65];
  or_make_index_masks(controls, nc, im);
  const uint64_t cm = mask_of(controls, nc), tm = mask_of(targets, nt);
  const int64_t n = (int64_t)(size >> nc);
#pragma omp parallel for schedule(static)
  for (int64_t g = 0; g < n; ++g) {
    const uint64_t i = base_index((uint64_t)g, im, nc) | cm;
    cd_store(s, i, cd_mul(cd_load(s, i), (popcount64(i & tm) & 1) ? ce : e));
  }
}

/* ------------------------------------------------------------------------- */
/* Local qubit swap and permutation                                          */
/* ------------------------------------------------------------------------- */

/* swap_permutated_local_qubits.hpp:38-77 with one data block: for every
   value of the bits outside (min, max) swap the run of 2^min amplitudes at
   ...1(max)...0(min)... with the one at ...0(max)...1(min)... */
void or_swap_permutated_local_qubits(double* s, uint64_t size, unsigned p1, unsigned p2) {
  const unsigned lo = p1 < p2 ? p1 : p2, hi = p1 < p2 ? p2 : p1;
  const uint64_t min_mask = (uint64_t)1 << lo, max_mask = (uint64_t)1 << hi;
  const uint64_t middle = ((uint64_t)1 << (hi - lo - 1u)) - 1u;
  const int64_t n = (int64_t)((size >> lo) >> 2u);
#pragma omp parallel for schedule(static)
  for (int64_t v = 0; v < n; ++v) {
    const uint64_t b = (((uint64_t)v & middle) << (lo + 1u)) | (((uint64_t)v & ~middle) << (lo + 2u));
    const uint64_t i1 = b | max_mask, i2 = b | min_mask;
    for (uint64_t r = 0; r < min_mask; ++r) {
      const cd x = cd_load(s, i1 + r), y = cd_load(s, i2 + r);
      cd_store(s, i1 + r, y);
      cd_store(s, i2 + r, x);
    }
  }
}

/* qubit_permutation::permutate -- qubit_permutation.hpp:242-252 */
void or_permutate(unsigned* perm, unsigned* inverse, unsigned n, unsigned q1, unsigned q2) {
  (void)n;
  unsigned t = inverse[perm[q1]];
  inverse[perm[q1]] = inverse[perm[q2]];
  inverse[perm[q2]] = t;
  t = perm[q1];
  perm[q1] = perm[q2];
  perm[q2] = t;
}

/* permutate_bits -- qubit_permutation.hpp:361-409: the bit at position b moves to perm[b] */
uint64_t or_permutate_bits(const unsigned* perm, unsigned n, uint64_t x) {
  unsigned char done[64] = {0};
  for (unsigned bit = 0; bit < n; ++bit) {
    if (done[bit]) continue;
    unsigned present = bit;
    uint64_t present_value = x & ((uint64_t)1 << bit);
    do {
      const unsigned previous = present;
      const uint64_t previous_value = present_value;
      present = perm[previous];
      present_value = x & ((uint64_t)1 << present);
      x = (x & ~((uint64_t)1 << present)) | ((previous_value >> previous) << present);
      done[present] = 1;
    } while (present != bit);
  }
  return x;
}

/* inverse_permutate_bits -- qubit_permutation.hpp:412-460 */
uint64_t or_inverse_permutate_bits(const unsigned* perm, const unsigned* inverse, unsigned n, uint64_t x) {
  (void)perm;
  unsigned char done[64] = {0};
  for (unsigned bit = 0; bit < n; ++bit) {
    if (done[bit]) continue;
    unsigned present = bit;
    uint64_t present_value = x & ((uint64_t)1 << bit);
    do {
      const unsigned previous = present;
      const uint64_t previous_value = present_value;
      present = inverse[previous];
      present_value = x & ((uint64_t)1 << present);
      x = (x & ~((uint64_t)1 << present)) | ((previous_value >> previous) << present);
      done[present] = 1;
    } while (present != bit);
  }
  return x;
}

/* ------------------------------------------------------------------------- */
/* Reductions: per-thread partial sums, then summed in thread order          */
/* ------------------------------------------------------------------------- */

#define MAX_THREADS 1024

/* <b|a> = sum conj(b_i) a_i -- inner_product.hpp:32-56 */
void or_inner_product(const double* a, const double* b, uint64_t size, double out[2]) {
  cd partial[MAX_THREADS];
  int nthreads = 1;
  for (int t = 0; t < MAX_THREADS; ++t) partial[t] = cd_make(0, 0);
#pragma omp parallel
  {
    const int t = omp_get_thread_num();
#pragma omp single
    nthreads = omp_get_num_threads();
    cd acc = cd_make(0, 0);
#pragma omp for schedule(static)
    for (int64_t i = 0; i < (int64_t)size; ++i)
      acc = cd_add(acc, cd_mul(cd_conj(cd_load(b, (uint64_t)i)), cd_load(a, (uint64_t)i)));
    partial[t] = acc;
  }
  cd r = cd_make(0, 0);
  for (int t = 0; t < nthreads; ++t) r = cd_add(r, partial[t]);
  out[0] = r.re;
  out[1] = r.im;
}

double or_norm2(const double* a, uint64_t size) {
  double r[2];
  or_inner_product(a, a, size, r);
  return r[0];
}

/* spin_expectation_value.hpp:20-79 (long double partials, z halved at the end) */
void or_spin_expectation_value(const double* s, uint64_t size, unsigned qubit, double out[3]) {
  long double partial[MAX_THREADS][3];
  int nthreads = 1;
  const uint64_t qm = (uint64_t)1 << qubit, lower = qm - 1u, upper = ~lower;
  memset(partial, 0, sizeof partial);
#pragma omp parallel
  {
    const int t = omp_get_thread_num();
#pragma omp single
    nthreads = omp_get_num_threads();
    long double a0 = 0, a1 = 0, a2 = 0;
#pragma omp for schedule(static)
    for (int64_t g = 0; g < (int64_t)(size >> 1); ++g) {
      const uint64_t z = (((uint64_t)g & upper) << 1u) | ((uint64_t)g & lower);
      const cd cz = cd_conj(cd_load(s, z)), ov = cd_load(s, z | qm);
      const cd p = cd_mul(cz, ov);
      a0 += (long double)p.re;
      a1 += (long double)p.im;
      a2 += (long double)(cz.re * cz.re + cz.im * cz.im) - (long double)(ov.re * ov.re + ov.im * ov.im);
    }
    partial[t][0] = a0;
    partial[t][1] = a1;
    partial[t][2] = a2;
  }
  long double r0 = 0, r1 = 0, r2 = 0;
  for (int t = 0; t < nthreads; ++t) {
    r0 += partial[t][0];
    r1 += partial[t][1];
    r2 += partial[t][2];
  }
  out[0] = (double)r0;
  out[1] = (double)r1;
  out[2] = (double)r2 * 0.5;
}

/* <psi| P |psi> for a Pauli string on operated qubits -- the functor of
   bra/src/nompi_state.cpp:1155-1202 over ket::expectation_value
   (expectation_value.hpp:53-131) with pauli_index_coeff.hpp:32-59 */
void or_expectation_pauli_string(const double* s, uint64_t size, const char* pauli, const unsigned* qubits,
                                 unsigned nq, double out[2]) {
  uint64_t im[65];
  or_make_index_masks(qubits, nq, im);
  cd partial[MAX_THREADS];
  int nthreads = 1;
  for (int t = 0; t < MAX_THREADS; ++t) partial[t] = cd_make(0, 0);
#pragma omp parallel
  {
    const int t = omp_get_thread_num();
#pragma omp single
    nthreads = omp_get_num_threads();
    cd acc = cd_make(0, 0);
#pragma omp for schedule(static)
    for (int64_t g = 0; g < (int64_t)(size >> nq); ++g) {
      cd group = cd_make(0, 0);
      for (uint64_t v = 0; v < ((uint64_t)1 << nq); ++v) {
        /* pauli_index_coeff(pauli, v) */
        uint64_t other = v;
        cd coeff = cd_make(1.0, 0.0);
        for (unsigned c = 0; c < nq; ++c) {
          if (pauli[c] == 'X') other ^= (uint64_t)1 << c;
          else if (pauli[c] == 'Y') {
            other ^= (uint64_t)1 << c;
            coeff = cd_mul(coeff, ((v >> c) & 1u) ? cd_make(0.0, 1.0) : cd_make(0.0, -1.0));
          } else if (pauli[c] == 'Z' && ((v >> c) & 1u))
            coeff = cd_scale(coeff, -1.0);
        }
        const uint64_t i = or_index_with_qubits((uint64_t)g, v, qubits, nq, im);
        const uint64_t j = or_index_with_qubits((uint64_t)g, other, qubits, nq, im);
        group = cd_add(group, cd_mul(cd_conj(cd_load(s, i)), cd_mul(coeff, cd_load(s, j))));
      }
      acc = cd_add(acc, group);
    }
    partial[t] = acc;
  }
  cd r = cd_make(0, 0);
  for (int t = 0; t < nthreads; ++t) r = cd_add(r, partial[t]);
  out[0] = r.re;
  out[1] = r.im;
}

static unsigned log2_size(uint64_t size) { return (unsigned)__builtin_ctzll(size); }

/* H = sum_{i<n-1} J_i Z_i Z_{i+1} + sum_i h_i X_i as a sum of Pauli-string terms */
double or_expectation_ising(const double* s, uint64_t size, const double* J, const double* h) {
  const unsigned n = log2_size(size);
  double total = 0.0, r[2];
  for (unsigned i = 0; i + 1 < n; ++i) {
    unsigned q[2] = {i, i + 1};
    or_expectation_pauli_string(s, size, "ZZ", q, 2, r);
    total += J[i] * r[0];
  }
  for (unsigned i = 0; i < n; ++i) {
    or_expectation_pauli_string(s, size, "X", &i, 1, r);
    total += h[i] * r[0];
  }
  return total;
}

/* out = H psi for the same Hamiltonian */
void or_apply_ising(double* out, const double* psi, uint64_t size, const double* J, const double* h) {
  const unsigned n = log2_size(size);
#pragma omp parallel for schedule(static)
  for (int64_t kk = 0; kk < (int64_t)size; ++kk) {
    const uint64_t k = (uint64_t)kk;
    double e = 0.0;
    for (unsigned i = 0; i + 1 < n; ++i) {
      const int zi = (k >> i) & 1u, zj = (k >> (i + 1)) & 1u;
      e += (zi == zj) ? J[i] : -J[i];
    }
    cd acc = cd_scale(cd_load(psi, k), e);
    for (unsigned i = 0; i < n; ++i) acc = cd_add(acc, cd_scale(cd_load(psi, k ^ ((uint64_t)1 << i)), h[i]));
    cd_store(out, k, acc);
  }
}

/* ---- measurement and sampling ------------------------------------------ */

/* MT19937-64 (Nishimura, ACM TOMACS 10(4), 2000); the parameters are those
   of std::mt19937_64 ([rand.predef]). */
void or_mt64_seed(or_mt64* g, uint64_t seed) {
  g->mt[0] = seed;
  for (unsigned i = 1; i < 312; ++i)
    g->mt[i] = 6364136223846793005ull * (g->mt[i - 1] ^ (g->mt[i - 1] >> 62)) + i;
  g->index = 312;
}

uint64_t or_mt64_next(or_mt64* g) {
  const uint64_t upper = ~0ull << 31, lower = ~upper;
  if (g->index >= 312) {
    for (unsigned i = 0; i < 312; ++i) {
      const uint64_t y = (g->mt[i] & upper) | (g->mt[(i + 1) % 312] & lower);
      g->mt[i] = g->mt[(i + 156) % 312] ^ (y >> 1) ^ ((y & 1u) ? 0xB5026F5AA96619E9ull : 0ull);
    }
    g->index = 0;
  }
  uint64_t z = g->mt[g->index++];
  z ^= (z >> 29) & 0x5555555555555555ull;
  z ^= (z << 17) & 0x71D67FFFEDA60000ull;
  z ^= (z << 37) & 0xFFF7EEE000000000ull;
  z ^= z >> 43;
  return z;
}

double or_mt64_canonical(or_mt64* g) {
  const double r = (double)or_mt64_next(g) / 18446744073709551616.0;
  return r >= 1.0 ? nextafter(1.0, 0.0) : r;
}

void or_zero_one_probabilities(const double* s, uint64_t size, unsigned qubit, double out[2]) {
  const uint64_t qmask = 1ull << qubit, lower = qmask - 1ull;
  long double p0 = 0.0L, p1 = 0.0L;
  for (uint64_t g = 0; g < size / 2; ++g) {
    const uint64_t i0 = ((g & ~lower) << 1) | (g & lower), i1 = i0 | qmask;
    p0 += (long double)(s[2 * i0] * s[2 * i0] + s[2 * i0 + 1] * s[2 * i0 + 1]);
    p1 += (long double)(s[2 * i1] * s[2 * i1] + s[2 * i1 + 1] * s[2 * i1 + 1]);
  }
  out[0] = (double)p0;
  out[1] = (double)p1;
}

void or_project(double* s, uint64_t size, unsigned qubit, int value, double multiplier) {
  const uint64_t qmask = 1ull << qubit, lower = qmask - 1ull;
#pragma omp parallel for schedule(static)
  for (uint64_t g = 0; g < size / 2; ++g) {
    const uint64_t i0 = ((g & ~lower) << 1) | (g & lower);
    const uint64_t keep = value ? (i0 | qmask) : i0, drop = value ? i0 : (i0 | qmask);
    s[2 * keep] *= multiplier;
    s[2 * keep + 1] *= multiplier;
    s[2 * drop] = 0.0;
    s[2 * drop + 1] = 0.0;
  }
}

int or_projective_measurement(double* s, uint64_t size, unsigned qubit, double canonical) {
  double p[2];
  or_zero_one_probabilities(s, size, qubit, p);
  const double total = p[0] + p[1];
  const int outcome = canonical * total < p[0] ? 0 : 1;
  or_project(s, size, qubit, outcome, pow(p[outcome], -0.5));
  return outcome;
}

void or_generate_events(const double* s, uint64_t size, const double* canonical, unsigned num_events,
                        uint64_t* events) {
  double* scan = (double*)malloc(sizeof(double) * size);
  double run = 0.0;
  for (uint64_t i = 0; i < size; ++i) {
    run += s[2 * i] * s[2 * i] + s[2 * i + 1] * s[2 * i + 1];
    scan[i] = run;
  }
  const double total = scan[size - 1];
  for (unsigned e = 0; e < num_events; ++e) {
    const double r = canonical[e] * total;
    uint64_t lo = 0, hi = size;  /* first i with scan[i] > r */
    while (lo < hi) {
      const uint64_t mid = lo + (hi - lo) / 2;
      if (r < scan[mid]) hi = mid;
      else lo = mid + 1;
    }
    events[e] = lo;
  }
  free(scan);
}

/* ---- Shor box ----------------------------------------------------------- */

/* ket::shor_box (shor_box.hpp:60-98): zero the state, then for x = 0 .. 2^m - 1
   set the amplitude at (reverse_m(x) << (n - m)) | (v_x & (2^(n-m) - 1)) to
   2^{-m/2}, with v_0 = 1 and v_{x+1} = (v_x * base) mod divisor in uint64. */
void or_shor_box(double* s, uint64_t size, unsigned m, uint64_t divisor, uint64_t base) {
  unsigned n = 0;
  while ((1ull << n) < size) ++n;
  memset(s, 0, sizeof(double) * 2 * size);
  const uint64_t count = 1ull << m, low_mask = (1ull << (n - m)) - 1ull;
  const double coefficient = pow((double)count, -0.5);
  uint64_t v = 1;
  for (uint64_t x = 0; x < count; ++x) {
    uint64_t r = 0;
    for (unsigned b = 0; b < m; ++b) r |= ((x >> (m - 1 - b)) & 1ull) << b;
    const uint64_t i = (r << (n - m)) | (v & low_mask);
    s[2 * i] = coefficient;
    s[2 * i + 1] = 0.0;
    v *= base;
    v %= divisor;
  }
}
