// Gate-block planner and its C ABI (brk_block_plan_*, brk_circuit_run_blocked).
//
// The circuit's ops are walked in order with the logical -> physical qubit
// map (as brk_circuit_run does).  Consecutive ops are grouped greedily while
// the union of their qubits, together with the qubits that address one
// 128-byte line (so every chunk access is whole lines), fits in the local
// bits of a chunk; the local set is then filled up with the lowest remaining
// qubits.  A local swap joins a block as the SWAP gate of its two physical
// qubits (the map changes for the ops after it).  A group of two or more ops
// becomes one block launch; a single op runs on the plain kernels.  This is the role of bra's
// BEGIN/END FUSION (bra/src/state.cpp:2193-2217, nompi_state.cpp:1282-1357)
// and of ket's cache-aware gate functions (gate_bitmasks.hpp:587-2134): a
// run of gates applied per group of amplitudes held near the core, with the
// per-gate arithmetic unchanged.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "brk_block.hpp"
#include "braket_hip.h"
#include "circuit.hpp"

namespace brk {
int fail(int status, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
}

struct brk_block_plan {
  struct step {
    unsigned first_op = 0, num_ops = 0;  // ops of the circuit covered by this step
    int program = -1;                    // index into programs, or -1: plain op
    brk_circuit::op physical;            // the op at physical qubits (plain steps)
  };
  unsigned num_qubits = 0;
  int dtype = BRK_C128;
  std::vector<step> steps;
  std::vector<brk::block_program> programs;
  std::vector<unsigned> final_permutation;
  brk::block_program* device = nullptr;
  int device_id = -1;
  ~brk_block_plan() {
    if (device) (void)hipFree(device);
  }
};

namespace brk {

static unsigned popcount64(uint64_t x) { return (unsigned)__builtin_popcountll(x); }

struct block_builder {
  unsigned n, max_local, line_bits;
  uint64_t qubits = 0;  // union of the ops' qubits
  std::vector<gate_desc> descs;
  std::vector<brk_circuit::op> ops;
  unsigned first_op = 0;

  uint64_t line_mask() const { return (1ull << std::min(line_bits, n)) - 1ull; }
  // a chunk is 2^L amplitudes with L >= 9 (one per thread of a workgroup at least)
  bool enabled() const { return std::min(max_local, n) >= (unsigned)BLOCK_THREAD_BITS; }
  bool fits(const gate_desc& d) const {
    return enabled() && descs.size() < (size_t)BLOCK_MAX_OPS &&
           popcount64(qubits | d.qubits | line_mask()) <= max_local;
  }
};

// Groups the program's gates into passes over LDS.  A segment holds consecutive
// gates whose flipped bits (at most two per gate) all lie in the segment's
// L - 9 group bits; diagonal gates join any segment.  Other gates get a pass
// of their own.
static void make_passes(block_program& prog) {
  const unsigned gb = prog.num_local > (unsigned)BLOCK_THREAD_BITS ? prog.num_local - BLOCK_THREAD_BITS : 0;
  prog.num_passes = 0;
  uint32_t bits = 0;
  auto close = [&](uint32_t first, uint32_t count) {
    if (!count) return;
    block_pass& p = prog.passes[prog.num_passes++];
    p = block_pass{};
    if (count == 1) {  // one gate: the strided pass visits only what it changes
      p.kind = BLOCK_PASS_SINGLE;
      p.first_op = first;
      p.num_ops = 1;
      return;
    }
    p.kind = BLOCK_PASS_SEGMENT;
    p.first_op = first;
    p.num_ops = count;
    for (unsigned b = 0; b < prog.num_local && (unsigned)__builtin_popcount(bits) < gb; ++b) bits |= 1u << b;
    unsigned k = 0;
    for (unsigned b = 0; b < prog.num_local; ++b)
      if (bits >> b & 1u) p.group[k++] = b;
    for (uint32_t i = first; i < first + count; ++i) {  // flip / pair bit / zmask over the group index
      block_op& o = prog.ops[i];
      o.group_flip = o.group_pair = 0;
      uint32_t zg = 0, zrest = o.zmask;
      for (unsigned g = 0; g < k; ++g) {
        if (o.mode != BLOCK_DIAG && (o.flip >> p.group[g] & 1u)) o.group_flip |= 1u << g;
        if (o.mode != BLOCK_DIAG && p.group[g] == o.pair_bit) o.group_pair = 1u << g;
        if (o.zmask >> p.group[g] & 1u) {
          zg |= 1u << g;
          zrest &= ~(1u << p.group[g]);
        }
      }
      o.group_zmask = zrest == 0 ? zg : ~0u;
    }
  };
  uint32_t first = 0, count = 0;
  for (uint32_t i = 0; i < prog.num_ops; ++i) {
    const block_op& o = prog.ops[i];
    const uint32_t mixing = o.mode == BLOCK_DIAG ? 0u : o.flip;
    // a diagonal gate under conditions visits only its subspace in a pass of its own, unless
    // it is a negation (Z, CZ), which a segment applies with a select
    const bool alone = o.mode == BLOCK_DIAG ? o.cond != 0 && !o.negate : (unsigned)__builtin_popcount(mixing) > 2;
    if (alone || (unsigned)__builtin_popcount(mixing) > gb) {
      close(first, count);
      block_pass& p = prog.passes[prog.num_passes++];
      p = block_pass{};
      p.kind = BLOCK_PASS_SINGLE;
      p.first_op = i;
      p.num_ops = 1;
      first = i + 1;
      count = 0;
      bits = 0;
      continue;
    }
    if ((unsigned)__builtin_popcount(bits | mixing) > gb) {
      close(first, count);
      first = i;
      count = 0;
      bits = 0;
    }
    bits |= mixing;
    ++count;
  }
  close(first, count);
}

// Local positions: the block's qubits, the line qubits, then the lowest others up to L.
static block_program make_program(const block_builder& b) {
  block_program prog{};
  const unsigned L = std::min(b.max_local, b.n);
  uint64_t local = b.qubits | b.line_mask();
  for (unsigned q = 0; q < b.n && popcount64(local) < L; ++q) local |= 1ull << q;
  unsigned r = 0, rank[64] = {};
  for (unsigned q = 0; q < b.n; ++q)
    if (local >> q & 1u) {
      rank[q] = r;
      prog.pos[r++] = q;
    }
  prog.num_local = r;
  auto to_local = [&](uint64_t m) {
    uint32_t out = 0;
    for (unsigned q = 0; q < b.n; ++q)
      if (m >> q & 1u) out |= 1u << rank[q];
    return out;
  };
  // local index j = t + BLOCK_THREADS * k: the physical offset of the k part
  const uint32_t count = 1u << r, per = count > (uint32_t)BLOCK_THREADS ? count / BLOCK_THREADS : 1;
  for (uint32_t k = 0; k < per; ++k) {
    const uint32_t j = k << BLOCK_THREAD_BITS;
    uint64_t dep = 0;
    for (unsigned bit = 0; bit < r; ++bit)
      if (j >> bit & 1u) dep |= 1ull << prog.pos[bit];
    prog.deposit_high[k] = dep;
  }
  prog.num_ops = (uint32_t)b.descs.size();
  for (size_t i = 0; i < b.descs.size(); ++i) {
    const gate_desc& d = b.descs[i];
    block_op& o = prog.ops[i];
    o.mode = d.mode;
    o.pair_bit = d.mode == BLOCK_DIAG ? 0 : rank[d.pair_bit];
    o.flip = to_local(d.flip);
    o.cond = to_local(d.cond);
    o.sign = to_local(d.sign);
    // the pair's first index has the pair bit clear: that bit of the sign mask never counts
    if (d.mode == BLOCK_MATRIX) o.sign &= ~(1u << o.pair_bit);
    o.negate = d.mode == BLOCK_DIAG && d.u[0] == -1.0 && d.u[1] == 0.0 && d.u[2] == -1.0 && d.u[3] == 0.0;
    o.zmask = to_local(d.zmask);
    const uint32_t inserted = o.cond | (d.mode == BLOCK_DIAG ? 0u : 1u << o.pair_bit);
    o.nins = 0;
    for (unsigned bit = 0; bit < r; ++bit)
      if (inserted >> bit & 1u) o.ins[o.nins++] = bit;
    std::copy(d.u, d.u + 8, o.u);
  }
  make_passes(prog);
  return prog;
}

block_program make_block_program(const gate_desc* descs, unsigned count, unsigned n, unsigned num_local,
                                 unsigned line_bits) {
  block_builder b;
  b.n = n;
  b.max_local = num_local;
  b.line_bits = line_bits;
  for (unsigned i = 0; i < count; ++i) {
    b.qubits |= descs[i].qubits;
    b.descs.push_back(descs[i]);
  }
  return make_program(b);
}

op_dependences dependences_of(const std::vector<brk_circuit::op>& ops) {
  const size_t count = ops.size();
  op_dependences d;
  d.mask.assign(count, 0);
  d.succ.assign(count, {});
  d.indeg.assign(count, 0);
  std::vector<long> last(64, -1);
  std::vector<long> prev;
  for (size_t i = 0; i < count; ++i) {
    const auto& o = ops[i];
    for (unsigned q : o.targets) d.mask[i] |= 1ull << q;
    for (unsigned q : o.controls) d.mask[i] |= 1ull << q;
    if (o.code == BRK_OP_LOCAL_SWAP) d.mask[i] = ~0ull;  // a barrier: it moves qubits between slots
    prev.clear();
    for (unsigned q = 0; q < 64; ++q)
      if (d.mask[i] >> q & 1ull) prev.push_back(last[q]);
    std::sort(prev.begin(), prev.end());
    prev.erase(std::unique(prev.begin(), prev.end()), prev.end());
    for (long f : prev)
      if (f >= 0) {
        d.succ[(size_t)f].push_back(i);
        ++d.indeg[i];
      }
    for (unsigned q = 0; q < 64; ++q)
      if (d.mask[i] >> q & 1ull) last[q] = (long)i;
  }
  return d;
}

std::vector<size_t> block_order(const std::vector<brk_circuit::op>& ops, unsigned num_local, unsigned line_bits) {
  const size_t count = ops.size();
  op_dependences dep = dependences_of(ops);
  const std::vector<uint64_t>& mask = dep.mask;
  const std::vector<std::vector<size_t>>& succ = dep.succ;
  std::vector<unsigned>& indeg = dep.indeg;
  std::vector<size_t> ready, order;
  order.reserve(count);
  for (size_t i = 0; i < count; ++i)
    if (!indeg[i]) ready.push_back(i);
  const uint64_t line = line_bits >= 64 ? ~0ull : (1ull << line_bits) - 1ull;
  auto take = [&](size_t k) {  // schedule ready[k]
    const size_t i = ready[k];
    ready.erase(ready.begin() + (long)k);
    order.push_back(i);
    for (size_t j : succ[i])
      if (--indeg[j] == 0) ready.insert(std::lower_bound(ready.begin(), ready.end(), j), j);
  };
  // Lookahead: how many ops (at most `budget`) could run within the local qubits `loc` once
  // `first` runs, without adding qubits.
  std::vector<unsigned> tmp_deg(count, 0), stamp(count, 0);
  unsigned cur_stamp = 0;
  std::vector<size_t> pend;
  auto closure = [&](size_t first, uint64_t loc, unsigned budget) {
    ++cur_stamp;
    pend.assign(1, first);
    for (size_t r : ready)
      if (r != first && !(mask[r] & ~loc)) pend.push_back(r);
    unsigned got = 0;
    while (!pend.empty() && got < budget) {
      const size_t i = pend.back();
      pend.pop_back();
      ++got;
      for (size_t j : succ[i]) {
        if (stamp[j] != cur_stamp) {
          stamp[j] = cur_stamp;
          tmp_deg[j] = indeg[j];
        }
        if (--tmp_deg[j] == 0 && !(mask[j] & ~loc)) pend.push_back(j);
      }
    }
    return got;
  };
  while (!ready.empty()) {
    uint64_t local = line;
    unsigned taken = 0;
    while (taken < (unsigned)BLOCK_MAX_OPS && !ready.empty()) {
      // an op within the block's qubits first (the earliest); else the op whose qubits
      // let the most ops run per added qubit
      size_t best = ready.size();
      for (size_t k = 0; k < ready.size() && best == ready.size(); ++k)
        if (!(mask[ready[k]] & ~local)) best = k;
      if (best == ready.size()) {
        double best_score = -1.0;
        unsigned best_growth = 65;
        for (size_t k = 0; k < ready.size(); ++k) {
          const uint64_t m = mask[ready[k]];
          if (popcount64(local | m) > num_local) continue;
          const unsigned growth = popcount64(m & ~local);
          const double score = (double)closure(ready[k], local | m, (unsigned)BLOCK_MAX_OPS - taken) / growth;
          if (score > best_score || (score == best_score && growth < best_growth)) {
            best = k;
            best_score = score;
            best_growth = growth;
          }
        }
      }
      if (best == ready.size()) break;
      local |= mask[ready[best]];
      take(best);
      ++taken;
    }
    if (!taken) take(0);  // an op no block holds: a step of its own
  }
  return order;
}

static brk_circuit::op map_op(const brk_circuit::op& o, const permutation& perm) {
  brk_circuit::op m = o;
  for (auto& q : m.targets) q = perm.perm[q];
  for (auto& q : m.controls) q = perm.perm[q];
  return m;
}

static void flush(block_builder& b, brk_block_plan& plan) {
  if (b.ops.empty()) return;
  if (b.ops.size() == 1 || !b.enabled()) {
    brk_block_plan::step s;
    s.first_op = b.first_op;
    s.num_ops = 1;
    s.physical = b.ops[0];
    plan.steps.push_back(std::move(s));
  } else {
    brk_block_plan::step s;
    s.first_op = b.first_op;
    s.num_ops = (unsigned)b.ops.size();
    s.program = (int)plan.programs.size();
    plan.programs.push_back(make_program(b));
    plan.steps.push_back(std::move(s));
  }
  b.qubits = 0;
  b.descs.clear();
  b.ops.clear();
}

static int build_plan(const brk_circuit* circuit, int dtype, const unsigned* permutation_in, brk_block_plan* plan) {
  const unsigned n = circuit->num_qubits;
  plan->num_qubits = n;
  plan->dtype = dtype;
  permutation perm(n);
  if (permutation_in && !perm.load(permutation_in, n))
    return fail(BRK_ERR_INVALID_ARGUMENT, "permutation is not a permutation of 0..%u", n - 1);
  block_builder b;
  b.n = n;
  b.max_local = block_max_local(dtype);
  b.line_bits = block_run_qubits(dtype);  // default: one 128-byte line
  for (unsigned i = 0; i < circuit->ops.size(); ++i) {
    const auto& o = circuit->ops[i];
    gate_desc d;
    brk_circuit::op m;
    if (o.code == BRK_OP_LOCAL_SWAP) {
      // the data of a local swap is a SWAP gate on its two physical qubits
      // (swap_permutated_local_qubits.hpp:64-77); the map changes for the ops after it
      m = o;
      if (int st = describe_physical_op(BRK_OP_SWAP, o.targets.data(), 2, nullptr, 0, nullptr, &d)) return st;
      perm.permutate(perm.inverse[o.targets[0]], perm.inverse[o.targets[1]]);
    } else {
      m = map_op(o, perm);
      if (int st = describe_physical_op(m.code, m.targets.data(), (unsigned)m.targets.size(),
                                        m.controls.empty() ? nullptr : m.controls.data(),
                                        (unsigned)m.controls.size(), m.params.data(), &d))
        return st;
    }
    if (!b.fits(d)) flush(b, *plan);
    if (b.ops.empty()) b.first_op = i;
    b.qubits |= d.qubits;
    b.descs.push_back(d);
    b.ops.push_back(m);
  }
  flush(b, *plan);
  plan->final_permutation = perm.perm;
  return BRK_OK;
}

static int upload(brk_block_plan* plan) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(BRK_ERR_HIP, "hipGetDevice failed");
  if (plan->device && plan->device_id == dev) return BRK_OK;
  if (plan->device) (void)hipFree(plan->device);
  plan->device = nullptr;
  if (plan->programs.empty()) {
    plan->device_id = dev;
    return BRK_OK;
  }
  const size_t bytes = sizeof(block_program) * plan->programs.size();
  if (hipMalloc((void**)&plan->device, bytes) != hipSuccess) return fail(BRK_ERR_OUT_OF_MEMORY, "block programs");
  if (hipMemcpy(plan->device, plan->programs.data(), bytes, hipMemcpyHostToDevice) != hipSuccess)
    return fail(BRK_ERR_HIP, "uploading the block programs failed");
  plan->device_id = dev;
  return BRK_OK;
}

static int run_plan(brk_block_plan* plan, void* first, void* stream) {
  if (int st = upload(plan)) return st;
  const uint64_t size = 1ull << plan->num_qubits;
  for (const auto& s : plan->steps) {
    int st;
    if (s.program < 0) st = apply_physical_op(s.physical, first, size, plan->dtype, stream);
    else st = launch_block(plan->device + s.program, plan->programs[s.program], first, size, plan->dtype, stream);
    if (st) return st;
  }
  return BRK_OK;
}

}  // namespace brk

using namespace brk;

extern "C" {

int brk_block_plan_create(const brk_circuit* circuit, int dtype, const unsigned* permutation, brk_block_plan** out) {
  if (!circuit || !out) return fail(BRK_ERR_INVALID_ARGUMENT, "null argument");
  if (dtype != BRK_C128 && dtype != BRK_C64) return fail(BRK_ERR_UNSUPPORTED_DTYPE, "dtype %d", dtype);
  *out = nullptr;
  auto* plan = new (std::nothrow) brk_block_plan;
  if (!plan) return fail(BRK_ERR_OUT_OF_MEMORY, "cannot allocate the plan");
  if (int st = build_plan(circuit, dtype, permutation, plan)) {
    delete plan;
    return st;
  }
  *out = plan;
  return BRK_OK;
}

int brk_block_plan_destroy(brk_block_plan* plan) {
  delete plan;
  return BRK_OK;
}

int brk_block_plan_num_steps(const brk_block_plan* plan, unsigned* num_steps) {
  if (!plan || !num_steps) return fail(BRK_ERR_INVALID_ARGUMENT, "null argument");
  *num_steps = (unsigned)plan->steps.size();
  return BRK_OK;
}

int brk_block_plan_step(const brk_block_plan* plan, unsigned i, unsigned* first_op, unsigned* num_ops,
                        unsigned* num_local, unsigned* local_qubits) {
  if (!plan) return fail(BRK_ERR_INVALID_ARGUMENT, "plan is null");
  if (i >= plan->steps.size()) return fail(BRK_ERR_INVALID_ARGUMENT, "step %u out of range", i);
  const auto& s = plan->steps[i];
  if (first_op) *first_op = s.first_op;
  if (num_ops) *num_ops = s.num_ops;
  const unsigned L = s.program < 0 ? 0 : plan->programs[s.program].num_local;
  if (num_local) *num_local = L;
  if (local_qubits && L) std::copy(plan->programs[s.program].pos, plan->programs[s.program].pos + L, local_qubits);
  return BRK_OK;
}

int brk_block_plan_final_permutation(const brk_block_plan* plan, unsigned* permutation) {
  if (!plan || !permutation) return fail(BRK_ERR_INVALID_ARGUMENT, "null argument");
  std::copy(plan->final_permutation.begin(), plan->final_permutation.end(), permutation);
  return BRK_OK;
}

int brk_block_plan_run(brk_block_plan* plan, void* first, void* stream) {
  if (!plan) return fail(BRK_ERR_INVALID_ARGUMENT, "plan is null");
  return run_plan(plan, first, stream);
}

int brk_circuit_block_order(const brk_circuit* circuit, unsigned num_local, unsigned line_bits, unsigned* order) {
  if (!circuit || (!order && !circuit->ops.empty())) return fail(BRK_ERR_INVALID_ARGUMENT, "null argument");
  if (num_local < 1 || num_local > 64 || line_bits > num_local)
    return fail(BRK_ERR_INVALID_ARGUMENT, "need 1 <= num_local <= 64 and line_bits <= num_local");
  const std::vector<size_t> o = block_order(circuit->ops, num_local, line_bits);
  for (size_t k = 0; k < o.size(); ++k) order[k] = (unsigned)o[k];
  return BRK_OK;
}

int brk_circuit_run_blocked(const brk_circuit* circuit, void* first, int dtype, unsigned* permutation,
                            void* stream) {
  brk_block_plan* plan = nullptr;
  if (int st = brk_block_plan_create(circuit, dtype, permutation, &plan)) return st;
  int st = run_plan(plan, first, stream);
  // the programs must stay resident until the launches that read them are done
  if (hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream)) != hipSuccess && !st)
    st = fail(BRK_ERR_HIP, "hipStreamSynchronize failed");
  if (!st && permutation) std::copy(plan->final_permutation.begin(), plan->final_permutation.end(), permutation);
  delete plan;
  return st;
}

}  // extern "C"
