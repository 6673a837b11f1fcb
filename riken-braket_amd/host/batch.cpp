// Batching form, host side: compiles a list of small circuits into the stage
// programs of csrc/brk_batch.hpp, uploads them once, and runs the persistent
// kernel as a captured HIP graph (queue reset + one persistent launch).
//
// Per circuit:
//  1. ops -> planner ops: uncontrolled single-qubit ops become 2x2 gates
//     (consecutive gates on one qubit, with nothing else on that qubit in
//     between, are multiplied into one in FP64); Z-type ops (Z, CZ, CnZ on
//     qubits and controls alike, ket/include/ket/gate/pauli_z.hpp:90-137)
//     become sign masks; the Hamiltonian adds "measure h_q X_q" per qubit and
//     one "measure sum J_i Z_i Z_i+1" after all gates.
//  2. stages: the local set is qubits 0..3 plus the qubits of the first
//     pending gates, in order of appearance; the stage then takes every
//     pending op that can run now: a gate needs its qubit local and no earlier
//     pending op on that qubit; a sign mask (diagonal) needs no earlier pending
//     gate on its qubits, since diagonal ops commute with each other.  Ops on
//     disjoint qubits commute, so taking ops out of order never changes the
//     state.
//  3. instructions: gates are grouped (<= 4 local positions, <= 8 gates,
//     one product matrix per position);
//     a run of sign masks becomes popcount terms (a CZ ladder is one term)
//     folded into the neighbouring gate group, or a pass of its own.
#include "circuit.hpp"
#include "brk_batch.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

namespace brk {
int fail(int status, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
}

using cd = std::complex<double>;

namespace {

struct pop {
  int kind;       // BRK_BATCH_GATE, _SIGN, _MEASURE_X, _MEASURE_ZZ
  uint32_t mask;  // qubits
  unsigned q;     // gate / measure_x qubit
  cd m[4];        // gate matrix, row-major
  bool diagonal() const { return kind == BRK_BATCH_SIGN || kind == BRK_BATCH_MEASURE_ZZ; }
};

struct planned_stage {
  std::vector<unsigned> local;  // ascending
  std::vector<pop> ops;         // taken ops in program order
};

struct planned_circuit {
  unsigned n = 0;
  std::vector<planned_stage> stages;
  std::vector<double> ham;  // J (n-1), h (n)
};

uint32_t all_qubits(unsigned n) { return n >= 32 ? 0xffffffffu : ((1u << n) - 1u); }

int to_planner_ops(const brk_circuit& c, const double* J, const double* h, std::vector<pop>& out) {
  const unsigned n = c.num_qubits;
  std::vector<long> last_gate(n, -1);  // index in `out` of a gate on q with nothing on q after it
  for (size_t i = 0; i < c.ops.size(); ++i) {
    const auto& o = c.ops[i];
    if (o.code == BRK_OP_Z) {
      uint32_t mask = 0;
      for (unsigned q : o.targets) mask |= 1u << q;
      for (unsigned q : o.controls) mask |= 1u << q;
      out.push_back(pop{BRK_BATCH_SIGN, mask, 0, {}});
      for (unsigned q = 0; q < n; ++q)
        if (mask & (1u << q)) last_gate[q] = -1;
      continue;
    }
    const bool single = o.targets.size() == 1 && o.controls.empty() && o.code != BRK_OP_SWAP &&
                        o.code != BRK_OP_LOCAL_SWAP;
    if (!single)
      return brk::fail(BRK_ERR_UNSUPPORTED,
                       "batched circuits take uncontrolled single-qubit ops and Z/CZ/CnZ (op %zu is not one)", i);
    std::vector<unsigned> qs;
    std::vector<std::array<double, 2>> u;
    if (!brk::op_unitary(o, qs, u) || u.size() != 4)
      return brk::fail(BRK_ERR_UNSUPPORTED, "op %zu has no 2x2 form", i);
    const unsigned q = o.targets[0];
    cd m[4];
    for (int k = 0; k < 4; ++k) m[k] = cd(u[k][0], u[k][1]);
    if (last_gate[q] >= 0) {
      pop& g = out[(size_t)last_gate[q]];
      const cd a0 = g.m[0], a1 = g.m[1], a2 = g.m[2], a3 = g.m[3];
      g.m[0] = m[0] * a0 + m[1] * a2;  // new * old: the old gate acts first
      g.m[1] = m[0] * a1 + m[1] * a3;
      g.m[2] = m[2] * a0 + m[3] * a2;
      g.m[3] = m[2] * a1 + m[3] * a3;
      continue;
    }
    pop g{BRK_BATCH_GATE, 1u << q, q, {m[0], m[1], m[2], m[3]}};
    out.push_back(g);
    last_gate[q] = (long)out.size() - 1;
  }
  for (unsigned q = 0; q < n; ++q)
    if (h && h[q] != 0.0) out.push_back(pop{BRK_BATCH_MEASURE_X, 1u << q, q, {}});
  bool any_j = false;
  for (unsigned i = 0; J && i + 1 < n; ++i) any_j |= J[i] != 0.0;
  if (any_j) out.push_back(pop{BRK_BATCH_MEASURE_ZZ, all_qubits(n), 0, {}});
  return BRK_OK;
}

void plan_stages(unsigned n, std::vector<pop> pending, std::vector<planned_stage>& stages) {
  const unsigned L = std::min<unsigned>(n, brk::BATCH_MAX_LOCAL);
  const unsigned P = std::min<unsigned>(n, brk::BATCH_PASSENGERS);
  do {  // a circuit with no ops still gets one stage (it initialises and measures nothing)
    uint32_t local = all_qubits(P);
    unsigned count = P;
    for (const pop& o : pending) {
      if (count == L) break;
      if (o.diagonal()) continue;
      if (!(local & o.mask)) {
        local |= o.mask;
        ++count;
      }
    }
    for (unsigned q = 0; q < n && count < L; ++q)
      if (!(local & (1u << q))) {
        local |= 1u << q;
        ++count;
      }
    planned_stage st;
    for (unsigned q = 0; q < n; ++q)
      if (local & (1u << q)) st.local.push_back(q);
    uint32_t blocked_any = 0, blocked_gate = 0;
    std::vector<pop> rest;
    for (const pop& o : pending) {
      const bool ok = o.diagonal() ? !(o.mask & blocked_gate) : (!(o.mask & ~local) && !(o.mask & blocked_any));
      if (ok) {
        st.ops.push_back(o);
      } else {
        rest.push_back(o);
        blocked_any |= o.mask;
        if (!o.diagonal()) blocked_gate |= o.mask;
      }
    }
    stages.push_back(std::move(st));
    pending.swap(rest);
  } while (!pending.empty());
}

uint32_t float_bits(double v) {
  const float f = (float)v;
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}

// Sign terms (d, A) of a list of Z-type masks (encoding in brk_batch.hpp);
// contributions add mod 2, so masks are merged by XOR.
std::vector<std::pair<uint32_t, uint32_t>> sign_terms(const std::vector<uint32_t>& masks) {
  uint32_t by_distance[32] = {};
  std::vector<std::pair<uint32_t, uint32_t>> terms, generic;
  for (uint32_t m : masks) {
    const int bits = __builtin_popcount(m);
    if (bits == 1) {
      by_distance[0] ^= m;
    } else if (bits == 2) {
      const int a = __builtin_ctz(m), b = 31 - __builtin_clz(m);
      by_distance[b - a] ^= 1u << a;
    } else {
      generic.push_back({brk::BATCH_TERM_ALL, m});
    }
  }
  for (uint32_t d = 0; d < 32; ++d)
    if (by_distance[d]) terms.push_back({d, by_distance[d]});
  terms.insert(terms.end(), generic.begin(), generic.end());
  return terms;
}

// U = diag(p0, p1) R diag(1, q) with R a real rotation by |theta| <= pi/2, for a 2x2 unitary U
// (row-major).  Within a row of a unitary the two entries' phases differ by the same angle in both
// rows (mod pi), so q removes that difference and each row is then its own phase times a real row.
// q is taken in the right half-plane (a factor -1 moves into R), and exactly 1 when it is within
// rounding of 1, so the RZ RY products of an ansatz layer need no pre-phase.  The real R is
// orthogonal: a reflection becomes a rotation by negating its second row (and p1), and a rotation
// with cos theta < 0 by negating it (and both p), so |theta| <= pi/2 and the kernel can apply it as
// three shears with t = tan(theta / 2) = s / (1 + c), |t| <= 1:
//   x -= t y;  y += s x;  x -= t y      (x' = c x - s y, y' = s x + c y).
// ok is false when a row is not real after the phases come out (U not unitary to FP64 rounding).
struct real_form {
  double t, s;  // tan(theta / 2), sin(theta)
  cd p[2];
  cd q;
  bool ok;
};

real_form real_decomposition(const std::array<cd, 4>& u) {
  real_form f{};
  const double w0 = std::abs(u[0]) * std::abs(u[1]), w1 = std::abs(u[2]) * std::abs(u[3]);
  cd q(1.0, 0.0);
  if (std::max(w0, w1) > 0.0) {
    const int r = w0 >= w1 ? 0 : 2;
    q = (u[r + 1] / std::abs(u[r + 1])) / (u[r] / std::abs(u[r]));
    if (q.real() < 0.0) q = -q;
    if (std::abs(q - cd(1.0, 0.0)) <= 1e-14) q = cd(1.0, 0.0);
  }
  const cd v[4] = {u[0], u[1] * std::conj(q), u[2], u[3] * std::conj(q)};
  double scale = 0.0;
  for (const cd& x : u) scale = std::max(scale, std::abs(x));
  f.ok = scale > 0.0;
  double R[4];
  for (int r = 0; r < 2; ++r) {
    const cd big = std::abs(v[2 * r]) >= std::abs(v[2 * r + 1]) ? v[2 * r] : v[2 * r + 1];
    const double m = std::abs(big);
    f.p[r] = m > 0.0 ? big / m : cd(1.0, 0.0);
    for (int c = 0; c < 2; ++c) {
      const cd x = v[2 * r + c] * std::conj(f.p[r]);
      R[2 * r + c] = x.real();
      if (std::abs(x.imag()) > 1e-12 * scale) f.ok = false;
    }
  }
  if (R[0] * R[3] - R[1] * R[2] < 0.0) {  // reflection: diag(1, -1) times a rotation
    R[2] = -R[2];
    R[3] = -R[3];
    f.p[1] = -f.p[1];
  }
  // rotation [[c, -s], [s, c]] (the two estimates of c and of s agree to rounding)
  double c = 0.5 * (R[0] + R[3]), sn = 0.5 * (R[2] - R[1]);
  if (c < 0.0) {
    c = -c;
    sn = -sn;
    f.p[0] = -f.p[0];
    f.p[1] = -f.p[1];
  }
  const double h = std::hypot(c, sn);
  if (!(h > 0.0) || std::abs(h - 1.0) > 1e-12 || std::abs(R[0] - R[3]) > 1e-12 || std::abs(R[1] + R[2]) > 1e-12)
    f.ok = false;
  if (f.ok) {
    c /= h;
    sn /= h;
  }
  f.t = sn / (1.0 + c);
  f.s = sn;
  f.q = q;
  return f;
}

// Appends the instruction stream of one stage; returns the instruction count.
uint32_t emit_stage(const planned_stage& st, std::vector<uint32_t>& words, uint32_t& measure_x, uint32_t& flags) {
  int pos_of[32];
  for (int& p : pos_of) p = -1;
  for (size_t p = 0; p < st.local.size(); ++p) pos_of[st.local[p]] = (int)p;

  // segments: gate groups and runs of sign masks, in program order
  struct segment {
    bool gates;
    std::vector<const pop*> ops;
    std::vector<unsigned> pos;  // distinct local positions of the group
    std::vector<uint32_t> masks;
    std::vector<std::pair<uint32_t, uint32_t>> pre, post;
    bool attached = false;  // a sign run folded into a neighbouring group
  };
  std::vector<segment> segs;
  for (const pop& o : st.ops) {
    if (o.kind == BRK_BATCH_GATE) {
      const unsigned pos = (unsigned)pos_of[o.q];
      segment* g = (!segs.empty() && segs.back().gates) ? &segs.back() : nullptr;
      if (g) {
        const bool known = std::find(g->pos.begin(), g->pos.end(), pos) != g->pos.end();
        if (g->ops.size() == (size_t)brk::BATCH_MAX_GROUP_OPS ||
            (!known && g->pos.size() == (size_t)brk::BATCH_MAX_GROUP))
          g = nullptr;
      }
      if (!g) {
        segs.push_back(segment{true, {}, {}, {}, {}, {}});
        g = &segs.back();
      }
      g->ops.push_back(&o);
      if (std::find(g->pos.begin(), g->pos.end(), pos) == g->pos.end()) g->pos.push_back(pos);
    } else if (o.kind == BRK_BATCH_SIGN) {
      if (segs.empty() || segs.back().gates) segs.push_back(segment{false, {}, {}, {}, {}, {}});
      segs.back().masks.push_back(o.mask);
    } else if (o.kind == BRK_BATCH_MEASURE_X) {
      measure_x |= 1u << pos_of[o.q];
    } else {
      flags |= brk::BS_ZZ;
    }
  }
  // fold each sign run into the previous group (after its gates) or the next one (before its gates);
  // after the gates first: there the real form's post-phase table takes the run's thread-independent
  // signs
  for (size_t i = 0; i < segs.size(); ++i) {
    if (segs[i].gates) continue;
    auto terms = sign_terms(segs[i].masks);
    if (terms.empty()) {  // the masks cancelled
      segs[i].attached = true;
      continue;
    }
    // groups take Z / CZ terms only; CnZ (three or more qubits) keeps a pass of its own
    bool pairs_only = terms.size() <= (size_t)brk::BATCH_MAX_ATTACHED_TERMS;
    for (const auto& t : terms) pairs_only &= t.first < brk::BATCH_TERM_ALL;
    if (!pairs_only) continue;
    if (i > 0 && segs[i - 1].gates && segs[i - 1].post.empty()) {
      segs[i - 1].post = terms;
      segs[i].attached = true;
    } else if (i + 1 < segs.size() && segs[i + 1].gates && segs[i + 1].pre.empty()) {
      segs[i + 1].pre = terms;
      segs[i].attached = true;
    }
  }
  uint32_t count = 0;
  for (const segment& sg : segs) {
    if (sg.gates) {
      std::vector<unsigned> pos = sg.pos;
      std::sort(pos.begin(), pos.end());
      uint32_t pw = 0, qw = 0;
      for (size_t k = 0; k < pos.size(); ++k) {
        pw |= (uint32_t)pos[k] << (8 * k);
        qw |= (uint32_t)st.local[pos[k]] << (8 * k);
      }
      // one matrix per position: the product of that position's gates, latest on the left
      std::vector<std::array<cd, 4>> mats(pos.size(), std::array<cd, 4>{cd(1), cd(0), cd(0), cd(1)});
      for (const pop* g : sg.ops) {
        const size_t slot = (size_t)(std::find(pos.begin(), pos.end(), (unsigned)pos_of[g->q]) - pos.begin());
        const auto a = mats[slot];
        mats[slot] = {g->m[0] * a[0] + g->m[1] * a[2], g->m[0] * a[1] + g->m[1] * a[3],
                      g->m[2] * a[0] + g->m[3] * a[2], g->m[2] * a[1] + g->m[3] * a[3]};
      }
      // real form (a rotation per position between two uniform phase tables) when it takes fewer
      // lane operations than the complex products: per thread 3 np 2^np + 4 2^np per table
      // against 8 np 2^np
      const size_t np = pos.size(), ne = (size_t)1 << np;
      std::vector<real_form> forms;
      bool real_ok = true, pre_phase = false;
      for (const auto& m : mats) {
        forms.push_back(real_decomposition(m));
        real_ok &= forms.back().ok;
        pre_phase |= forms.back().q != cd(1.0, 0.0);
      }
      const bool use_real = real_ok && 3 * np * ne + 4 * ne * (pre_phase ? 2 : 1) < 8 * np * ne;
      const uint32_t op = use_real ? (pre_phase ? brk::BI_GATES_REAL_PRE : brk::BI_GATES_REAL) : brk::BI_GATES;
      words.push_back(op | (uint32_t)np << 8 | (uint32_t)sg.pre.size() << 16 | (uint32_t)sg.post.size() << 24);
      words.push_back(pw);
      words.push_back(qw);
      // the sign terms' parity over the group's own index bits, per element (the kernel adds the
      // parts that depend on the thread's bits); a phase table of the real form takes them instead
      // (the signs and the table are diagonal and commute), so the kernel then has none to apply
      auto uniform_parities = [&](const std::vector<std::pair<uint32_t, uint32_t>>& terms) {
        uint32_t bits = 0;
        for (uint32_t e = 0; e < (1u << pos.size()); ++e) {
          uint32_t xe = 0;
          for (size_t k = 0; k < pos.size(); ++k) xe |= ((e >> k) & 1u) << st.local[pos[k]];
          uint32_t par = 0;
          for (const auto& t : terms) par += (uint32_t)__builtin_popcount(xe & (xe >> t.first) & t.second);
          bits |= (par & 1u) << e;
        }
        return bits;
      };
      const uint32_t pre_uniform = uniform_parities(sg.pre), post_uniform = uniform_parities(sg.post);
      const bool post_in_table = use_real, pre_in_table = use_real && pre_phase;
      words.push_back((pre_in_table ? 0u : pre_uniform) | (post_in_table ? 0u : post_uniform) << 16);
      for (const auto& t : sg.pre) words.insert(words.end(), {t.first, t.second});
      for (const auto& t : sg.post) words.insert(words.end(), {t.first, t.second});
      if (!use_real) {
        for (const auto& m : mats)
          for (int k = 0; k < 4; ++k) {
            words.push_back(float_bits(m[k].real()));
            words.push_back(float_bits(m[k].imag()));
          }
      } else {
        for (const auto& f : forms) {
          words.push_back(float_bits(f.t));
          words.push_back(float_bits(f.s));
        }
        // tables over the group's elements e: bit j of e is the value at position j
        for (int table = 0; table < (pre_phase ? 2 : 1); ++table)
          for (size_t e = 0; e < ne; ++e) {
            cd z(1.0, 0.0);
            for (size_t j = 0; j < np; ++j) {
              const unsigned bit = (unsigned)((e >> j) & 1u);
              z *= table == 0 ? forms[j].p[bit] : (bit ? forms[j].q : cd(1.0, 0.0));
            }
            if ((((table == 0 ? post_uniform : pre_uniform) >> e) & 1u) != 0) z = -z;
            words.push_back(float_bits(z.real()));
            words.push_back(float_bits(z.imag()));
          }
      }
    } else {
      if (sg.attached) continue;
      const auto terms = sign_terms(sg.masks);
      words.push_back(brk::BI_SIGN | (uint32_t)terms.size() << 16);
      for (const auto& t : terms) words.insert(words.end(), {t.first, t.second});
    }
    ++count;
  }
  return count;
}

}  // namespace

struct brk_batch {
  std::vector<planned_circuit> plans;
  std::vector<brk::batch_circuit> circuits;
  std::vector<brk::batch_stage> stages;
  std::vector<uint32_t> words;
  std::vector<double> ham;
  std::vector<uint32_t> order;
  unsigned max_multi_n = 0;  // largest n with more than one stage (0: none)

  // device side, created on the first run
  int device = -1;
  unsigned grid = 0;
  void* d_program = nullptr;  // circuits, stages, words, ham, order, counter, energies, t_start, t_end
  void* d_slots = nullptr;
  brk::batch_launch launch{};
  hipStream_t stream = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev_in = nullptr, ev_begin = nullptr, ev_end = nullptr;
  double tick_ms = 0.0;  // wall-clock tick in ms

  ~brk_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (ev_in) (void)hipEventDestroy(ev_in);
    if (ev_begin) (void)hipEventDestroy(ev_begin);
    if (ev_end) (void)hipEventDestroy(ev_end);
    if (stream) (void)hipStreamDestroy(stream);
    if (d_slots) (void)hipFree(d_slots);
    if (d_program) (void)hipFree(d_program);
  }
};

namespace {

template <typename T>
size_t aligned_bytes(size_t count) {
  return (count * sizeof(T) + 255) & ~(size_t)255;
}

int ensure_device(brk_batch* b) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return brk::fail(BRK_ERR_HIP, "hipGetDevice failed");
  if (b->device >= 0) {
    if (dev != b->device) return brk::fail(BRK_ERR_INVALID_ARGUMENT, "batch was set up on device %d", b->device);
    return BRK_OK;
  }
  const size_t nc = b->circuits.size();
  const size_t sizes[] = {aligned_bytes<brk::batch_circuit>(nc), aligned_bytes<brk::batch_stage>(b->stages.size()),
                          aligned_bytes<uint32_t>(std::max<size_t>(1, b->words.size())),
                          aligned_bytes<double>(std::max<size_t>(1, b->ham.size())), aligned_bytes<uint32_t>(nc),
                          aligned_bytes<uint32_t>(1), aligned_bytes<double>(nc), aligned_bytes<uint64_t>(nc),
                          aligned_bytes<uint64_t>(nc)};
  size_t total = 0;
  for (size_t s : sizes) total += s;
  if (hipMalloc(&b->d_program, total) != hipSuccess) {
    b->d_program = nullptr;
    return brk::fail(BRK_ERR_OUT_OF_MEMORY, "cannot allocate %zu bytes for the batch program", total);
  }
  char* base = static_cast<char*>(b->d_program);
  char* ptr[9];
  for (int i = 0; i < 9; ++i) {
    ptr[i] = base;
    base += sizes[i];
  }
  const void* host[] = {b->circuits.data(), b->stages.data(), b->words.data(), b->ham.data(), b->order.data()};
  const size_t bytes[] = {nc * sizeof(brk::batch_circuit), b->stages.size() * sizeof(brk::batch_stage),
                          b->words.size() * sizeof(uint32_t), b->ham.size() * sizeof(double), nc * sizeof(uint32_t)};
  for (int i = 0; i < 5; ++i)
    if (bytes[i] && hipMemcpy(ptr[i], host[i], bytes[i], hipMemcpyHostToDevice) != hipSuccess)
      return brk::fail(BRK_ERR_HIP, "uploading the batch program failed");

  int cus = 0, rate_khz = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    return brk::fail(BRK_ERR_HIP, "cannot read the CU count");
  if (hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || rate_khz <= 0)
    return brk::fail(BRK_ERR_HIP, "cannot read the wall-clock rate");
  b->tick_ms = 1.0 / (double)rate_khz;
  b->grid = (unsigned)std::min<size_t>((size_t)cus * brk::BATCH_WG_PER_CU, nc);
  uint64_t slot_elems = 0;
  if (b->max_multi_n) {
    slot_elems = 1ull << b->max_multi_n;
    const size_t slot_bytes = (size_t)b->grid * slot_elems * 8;
    if (hipMalloc(&b->d_slots, slot_bytes) != hipSuccess) {
      b->d_slots = nullptr;
      return brk::fail(BRK_ERR_OUT_OF_MEMORY, "cannot allocate %zu bytes of batch state slots", slot_bytes);
    }
  }
  brk::batch_launch& L = b->launch;
  L.circuits = reinterpret_cast<const brk::batch_circuit*>(ptr[0]);
  L.stages = reinterpret_cast<const brk::batch_stage*>(ptr[1]);
  L.words = reinterpret_cast<const uint32_t*>(ptr[2]);
  L.ham = reinterpret_cast<const double*>(ptr[3]);
  L.order = reinterpret_cast<const uint32_t*>(ptr[4]);
  L.num_circuits = (uint32_t)nc;
  L.counter = reinterpret_cast<uint32_t*>(ptr[5]);
  L.slots = b->d_slots;
  L.slot_elems = slot_elems;
  L.energies = reinterpret_cast<double*>(ptr[6]);
  L.t_start = reinterpret_cast<uint64_t*>(ptr[7]);
  L.t_end = reinterpret_cast<uint64_t*>(ptr[8]);

  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
    return brk::fail(BRK_ERR_HIP, "cannot create the batch stream");
  if (hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming) != hipSuccess ||
      hipEventCreate(&b->ev_begin) != hipSuccess || hipEventCreate(&b->ev_end) != hipSuccess)
    return brk::fail(BRK_ERR_HIP, "cannot create the batch events");
  if (int s = brk::configure_batch()) return s;
  // graph: queue reset + persistent kernel, captured once and replayed
  if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return brk::fail(BRK_ERR_HIP, "hipStreamBeginCapture failed");
  const int launch_status = brk::launch_batch(L, b->grid, b->stream);
  hipGraph_t graph = nullptr;
  const hipError_t end = hipStreamEndCapture(b->stream, &graph);
  if (launch_status) {
    if (graph) (void)hipGraphDestroy(graph);
    return launch_status;
  }
  if (end != hipSuccess || !graph) return brk::fail(BRK_ERR_HIP, "hipStreamEndCapture failed");
  const hipError_t inst = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (inst != hipSuccess) {
    b->exec = nullptr;
    return brk::fail(BRK_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(inst));
  }
  b->device = dev;
  return BRK_OK;
}

int run(brk_batch* b, double* energies, double* elapsed_ms, double* start_ms, double* end_ms, void* stream) {
  if (int s = ensure_device(b)) return s;
  hipStream_t caller = reinterpret_cast<hipStream_t>(stream);
  if (hipEventRecord(b->ev_in, caller) != hipSuccess || hipStreamWaitEvent(b->stream, b->ev_in, 0) != hipSuccess)
    return brk::fail(BRK_ERR_HIP, "ordering the batch after the caller's stream failed");
  if (hipEventRecord(b->ev_begin, b->stream) != hipSuccess) return brk::fail(BRK_ERR_HIP, "hipEventRecord failed");
  const hipError_t e = hipGraphLaunch(b->exec, b->stream);
  if (e != hipSuccess) return brk::fail(BRK_ERR_HIP, "hipGraphLaunch failed: %s", hipGetErrorString(e));
  if (hipEventRecord(b->ev_end, b->stream) != hipSuccess) return brk::fail(BRK_ERR_HIP, "hipEventRecord failed");
  const size_t nc = b->circuits.size();
  if (energies &&
      hipMemcpyAsync(energies, b->launch.energies, nc * sizeof(double), hipMemcpyDeviceToHost, b->stream) != hipSuccess)
    return brk::fail(BRK_ERR_HIP, "copying the energies failed");
  std::vector<uint64_t> ts, te;
  if (start_ms || end_ms) {
    ts.resize(nc);
    te.resize(nc);
    if (hipMemcpyAsync(ts.data(), b->launch.t_start, nc * 8, hipMemcpyDeviceToHost, b->stream) != hipSuccess ||
        hipMemcpyAsync(te.data(), b->launch.t_end, nc * 8, hipMemcpyDeviceToHost, b->stream) != hipSuccess)
      return brk::fail(BRK_ERR_HIP, "copying the circuit timestamps failed");
  }
  if (hipStreamSynchronize(b->stream) != hipSuccess) return brk::fail(BRK_ERR_HIP, "batch run failed");
  if (elapsed_ms) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, b->ev_begin, b->ev_end) != hipSuccess)
      return brk::fail(BRK_ERR_HIP, "hipEventElapsedTime failed");
    *elapsed_ms = ms;
  }
  if (start_ms || end_ms) {
    const uint64_t t0 = *std::min_element(ts.begin(), ts.end());
    for (size_t i = 0; i < nc; ++i) {
      if (start_ms) start_ms[i] = (double)(ts[i] - t0) * b->tick_ms;
      if (end_ms) end_ms[i] = (double)(te[i] - t0) * b->tick_ms;
    }
  }
  return BRK_OK;
}

}  // namespace

using brk::fail;

extern "C" {

int brk_batch_create(const brk_circuit* const* circuits, unsigned num_circuits, const double* const* J,
                     const double* const* h, brk_batch** out) {
  if (!out) return fail(BRK_ERR_INVALID_ARGUMENT, "null output");
  *out = nullptr;
  if (!circuits || num_circuits == 0) return fail(BRK_ERR_INVALID_ARGUMENT, "empty batch");
  auto* b = new brk_batch();
  std::vector<double> cost(num_circuits);
  for (unsigned i = 0; i < num_circuits; ++i) {
    const brk_circuit* c = circuits[i];
    if (!c) {
      delete b;
      return fail(BRK_ERR_INVALID_ARGUMENT, "circuit %u is null", i);
    }
    const unsigned n = c->num_qubits;
    if (n < 1 || n > (unsigned)brk::BATCH_MAX_QUBITS) {
      delete b;
      return fail(BRK_ERR_INVALID_ARGUMENT, "circuit %u has %u qubits (batched circuits take 1..%d)", i, n,
                  brk::BATCH_MAX_QUBITS);
    }
    planned_circuit pc;
    pc.n = n;
    std::vector<pop> ops;
    if (int s = to_planner_ops(*c, J ? J[i] : nullptr, h ? h[i] : nullptr, ops)) {
      delete b;
      return s;
    }
    plan_stages(n, std::move(ops), pc.stages);
    for (unsigned k = 0; k + 1 < n; ++k) pc.ham.push_back(J && J[i] ? J[i][k] : 0.0);
    for (unsigned k = 0; k < n; ++k) pc.ham.push_back(h && h[i] ? h[i][k] : 0.0);

    brk::batch_circuit bc{};
    bc.n = n;
    bc.num_stages = (uint32_t)pc.stages.size();
    bc.first_stage = (uint32_t)b->stages.size();
    bc.ham_offset = (uint32_t)b->ham.size();
    b->ham.insert(b->ham.end(), pc.ham.begin(), pc.ham.end());
    double work = 0.0;
    for (size_t s = 0; s < pc.stages.size(); ++s) {
      const planned_stage& st = pc.stages[s];
      // the kernel moves the chunks of a multi-stage circuit as full chunks
      if (pc.stages.size() > 1 && st.local.size() != (size_t)brk::BATCH_MAX_LOCAL) {
        delete b;
        return fail(BRK_ERR_UNSUPPORTED, "circuit %u: stage %zu has %zu local qubits, not %d", i, s, st.local.size(),
                    brk::BATCH_MAX_LOCAL);
      }
      brk::batch_stage bs{};
      bs.L = (uint32_t)st.local.size();
      uint32_t local_mask = 0;
      for (size_t p = 0; p < st.local.size(); ++p) {
        bs.local[p] = (uint8_t)st.local[p];
        local_mask |= 1u << st.local[p];
      }
      bs.nonlocal_mask = all_qubits(n) & ~local_mask;
      bs.first_word = (uint32_t)b->words.size();
      bool changes = false;
      for (const pop& o : st.ops) changes |= o.kind == BRK_BATCH_GATE || o.kind == BRK_BATCH_SIGN;
      bs.num_instr = emit_stage(st, b->words, bs.measure_x, bs.flags);
      if (s == 0) bs.flags |= brk::BS_INIT;
      if (s + 1 < pc.stages.size() && (changes || s == 0)) bs.flags |= brk::BS_STORE;
      b->stages.push_back(bs);
      work += (double)(1ull << n) * (bs.num_instr + 3);
    }
    if (pc.stages.size() > 1) b->max_multi_n = std::max(b->max_multi_n, n);
    cost[i] = work;
    b->circuits.push_back(bc);
    b->plans.push_back(std::move(pc));
  }
  b->order.resize(num_circuits);
  for (unsigned i = 0; i < num_circuits; ++i) b->order[i] = i;
  std::stable_sort(b->order.begin(), b->order.end(), [&](uint32_t x, uint32_t y) { return cost[x] > cost[y]; });
  *out = b;
  return BRK_OK;
}

int brk_batch_destroy(brk_batch* batch) {
  delete batch;
  return BRK_OK;
}

int brk_batch_num_stages(const brk_batch* batch, unsigned circuit, unsigned* num_stages) {
  if (!batch || !num_stages) return fail(BRK_ERR_INVALID_ARGUMENT, "null argument");
  if (circuit >= batch->plans.size()) return fail(BRK_ERR_INVALID_ARGUMENT, "circuit %u out of range", circuit);
  *num_stages = (unsigned)batch->plans[circuit].stages.size();
  return BRK_OK;
}

int brk_batch_stage(const brk_batch* batch, unsigned circuit, unsigned stage, unsigned* local_qubits,
                    unsigned* num_local, unsigned* num_ops) {
  if (!batch || !num_local || !num_ops) return fail(BRK_ERR_INVALID_ARGUMENT, "null argument");
  if (circuit >= batch->plans.size()) return fail(BRK_ERR_INVALID_ARGUMENT, "circuit %u out of range", circuit);
  const auto& stages = batch->plans[circuit].stages;
  if (stage >= stages.size()) return fail(BRK_ERR_INVALID_ARGUMENT, "stage %u out of range", stage);
  const planned_stage& st = stages[stage];
  *num_local = (unsigned)st.local.size();
  *num_ops = (unsigned)st.ops.size();
  if (local_qubits)
    for (size_t p = 0; p < st.local.size(); ++p) local_qubits[p] = st.local[p];
  return BRK_OK;
}

int brk_batch_stage_op(const brk_batch* batch, unsigned circuit, unsigned stage, unsigned k, int* kind,
                       uint64_t* qubit_mask, double* matrix) {
  if (!batch || !kind || !qubit_mask) return fail(BRK_ERR_INVALID_ARGUMENT, "null argument");
  if (circuit >= batch->plans.size()) return fail(BRK_ERR_INVALID_ARGUMENT, "circuit %u out of range", circuit);
  const auto& stages = batch->plans[circuit].stages;
  if (stage >= stages.size() || k >= stages[stage].ops.size())
    return fail(BRK_ERR_INVALID_ARGUMENT, "stage %u / op %u out of range", stage, k);
  const pop& o = stages[stage].ops[k];
  *kind = o.kind;
  *qubit_mask = o.mask;
  if (matrix)
    for (int i = 0; i < 4; ++i) {
      matrix[2 * i] = o.kind == BRK_BATCH_GATE ? o.m[i].real() : 0.0;
      matrix[2 * i + 1] = o.kind == BRK_BATCH_GATE ? o.m[i].imag() : 0.0;
    }
  return BRK_OK;
}

int brk_batch_run(brk_batch* batch, double* energies, void* stream) {
  if (!batch || !energies) return fail(BRK_ERR_INVALID_ARGUMENT, "null argument");
  return run(batch, energies, nullptr, nullptr, nullptr, stream);
}

int brk_batch_run_timed(brk_batch* batch, double* energies, double* elapsed_ms, double* circuit_start_ms,
                        double* circuit_end_ms, void* stream) {
  if (!batch || !energies || !elapsed_ms) return fail(BRK_ERR_INVALID_ARGUMENT, "null argument");
  return run(batch, energies, elapsed_ms, circuit_start_ms, circuit_end_ms, stream);
}

}  // extern "C"
