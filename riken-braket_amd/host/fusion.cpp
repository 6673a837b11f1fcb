// Precision form, host side: greedy gate fusion and the fused-plan objects of
// the C ABI.
//
// Planning (FP64 on the host):
//  1. The circuit's ops are put in block order (block_order: a dependence-
//     preserving order that a lookahead packs into few groups of `width`
//     qubits; ops on disjoint qubits commute) and mapped to physical qubits by
//     replaying the logical -> physical permutation; a local swap (which keeps
//     its place relative to every op) becomes a SWAP gate on its two physical
//     qubits (same physical result as the plain runner).  The block order of
//     up to FUSION_ORDER_TRIES - 1 random topological orders (fixed seeds) is
//     fused too, and the candidate with the fewest blocks is kept.
//  2. Greedy fusion: a block starts at the first unplaced gate and scans the
//     rest in order.  A gate joins the block when none of its qubits is
//     "blocked" and the union of qubits stays within `width`; otherwise its
//     qubits become blocked (a later gate on them must stay after it).  A gate
//     that joins therefore only moves past gates on disjoint qubits, with
//     which it commutes.
//  3. The block matrix is the in-order product of its gates (each gate
//     applied to every column), widened with identity qubits to exactly
//     `width` qubits, then rounded once to FP32 for the device.
//  4. Chunk passes (width 4 and 5, at least 13 qubits): the blocks are put in
//     block order themselves (a dependence-preserving order that packs blocks
//     into few groups of 13 local qubits; blocks on disjoint qubits commute),
//     and consecutive blocks whose qubits, with the 4 lowest (128-byte runs of
//     complex64), fit in 13 local qubits form one pass over LDS-resident
//     chunks of 2^13 amplitudes.
//
// The reference's fusion (bra BEGIN/END FUSION, nompi_state.cpp:1285-1357)
// applies the recorded gates one by one per 2^k-amplitude group; a dense
// block matrix gives the same result and runs as a GEMM.
#include "circuit.hpp"
#include "brk_fused.hpp"
#include "brk_bits.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <complex>
#include <new>
#include <random>
#include <set>

namespace brk {
int fail(int status, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
}

using cd = std::complex<double>;

struct brk_fused_plan {
  struct block {
    std::vector<unsigned> qubits;  // sorted physical positions
    std::vector<cd> u;             // row-major 2^w x 2^w
    unsigned num_gates = 0;
    brk::fused_launch launch{};
  };
  unsigned num_qubits = 0, width = 0;
  mutable std::vector<block> blocks;  // launch.matrix is filled in on the first apply
  std::vector<unsigned> final_perm;
  std::vector<float> host_matrices;   // FP32 embeddings, fused_matrix_floats(width) per block
  mutable float* device_matrices = nullptr;
  std::vector<float> host_split;      // chunk passes: FP32 three-product forms, fused_split_floats(width) per block
  mutable float* device_split = nullptr;
  std::vector<brk::fused_chunk_program> passes;  // chunk passes over the blocks in order (may be empty)
  mutable brk::fused_chunk_program* device_passes = nullptr;
  ~brk_fused_plan() {
    if (device_matrices) (void)hipFree(device_matrices);
    if (device_split) (void)hipFree(device_split);
    if (device_passes) (void)hipFree(device_passes);
  }
};

namespace brk {

struct physical_gate {
  std::vector<unsigned> qubits;  // gate-local bit j <-> qubits[j]
  std::vector<std::array<double, 2>> u;
};

// y := G x on the gate's local positions `pos` inside a 2^w vector (applied in place).
static void apply_small(std::vector<cd>& x, unsigned w, const std::vector<std::array<double, 2>>& g,
                        const std::vector<unsigned>& pos) {
  const unsigned m = (unsigned)pos.size(), d = 1u << m;
  unsigned gmask = 0;
  for (unsigned p : pos) gmask |= 1u << p;
  std::vector<cd> in(d), out(d);
  std::vector<unsigned> idx(d);
  for (unsigned base = 0; base < (1u << w); ++base) {
    if (base & gmask) continue;
    for (unsigned v = 0; v < d; ++v) {
      unsigned i = base;
      for (unsigned j = 0; j < m; ++j)
        if ((v >> j) & 1u) i |= 1u << pos[j];
      idx[v] = i;
      in[v] = x[i];
    }
    for (unsigned r = 0; r < d; ++r) {
      cd acc{0, 0};
      for (unsigned c = 0; c < d; ++c) acc += cd(g[(size_t)r * d + c][0], g[(size_t)r * d + c][1]) * in[c];
      out[r] = acc;
    }
    for (unsigned v = 0; v < d; ++v) x[idx[v]] = out[v];
  }
}

static uint64_t deposit(uint64_t v, const unsigned* bits, unsigned nbits) {
  uint64_t r = 0;
  for (unsigned b = 0; b < nbits; ++b)
    if ((v >> b) & 1ull) r |= 1ull << bits[b];
  return r;
}

// Tile geometry of a block: rows = block qubits, columns = the 5 lowest
// qubits outside the block, tiles = every other qubit.
static void build_launch(brk_fused_plan::block& blk, unsigned n, unsigned w) {
  fused_launch& L = blk.launch;
  L.width = (int)w;
  std::vector<unsigned> cols;
  for (unsigned q = 0; cols.size() < 5 && q < n; ++q)
    if (!std::binary_search(blk.qubits.begin(), blk.qubits.end(), q)) cols.push_back(q);
  std::vector<unsigned> tile(blk.qubits);
  tile.insert(tile.end(), cols.begin(), cols.end());
  expand_masks e = make_expand_masks(tile.data(), (int)tile.size());
  std::copy(e.m, e.m + BRK_MAX_OPERATED_QUBITS + 1, L.outer_masks);
  L.outer_k = (int)tile.size();
  L.ntiles = 1ull << (n - tile.size());
  for (unsigned r = 0; r < FUSED_MAX_ROWS; ++r) L.row_bits[r] = r < (1u << w) ? deposit(r, blk.qubits.data(), w) : 0;
  for (unsigned c = 0; c < FUSED_COLUMNS; ++c) L.col_bits[c] = deposit(c, cols.data(), 5);

  // staged tiles: the 7 lowest qubits outside the block are the columns
  L.sntiles = 0;
  std::vector<unsigned> scols;
  for (unsigned q = 0; scols.size() < 7 && q < n; ++q)
    if (!std::binary_search(blk.qubits.begin(), blk.qubits.end(), q)) scols.push_back(q);
  if (w <= 5 && scols.size() == 7) {
    std::vector<unsigned> stile(blk.qubits);
    stile.insert(stile.end(), scols.begin(), scols.end());
    std::sort(stile.begin(), stile.end());
    expand_masks se = make_expand_masks(stile.data(), (int)stile.size());
    std::copy(se.m, se.m + BRK_MAX_OPERATED_QUBITS + 1, L.souter_masks);
    L.souter_k = (int)stile.size();
    L.sntiles = 1ull << (n - stile.size());
    for (size_t i = 0; i < stile.size(); ++i) {
      const unsigned q = stile[i];
      L.stile_qubit[i] = (uint8_t)q;
      const auto row = std::lower_bound(blk.qubits.begin(), blk.qubits.end(), q);
      if (row != blk.qubits.end() && *row == q) {
        L.stile_lds_bit[i] = (uint8_t)(7 + (row - blk.qubits.begin()));
      } else {
        L.stile_lds_bit[i] = (uint8_t)(std::lower_bound(scols.begin(), scols.end(), q) - scols.begin());
      }
    }
  }
}

// Real embedding [[Ur, -Ui], [Ui, Ur]] of a 2^w block, stored transposed
// (element (r, c) at c*W + r) with odd columns XOR-swizzled by 32 when W >= 64.
static void embed(const std::vector<cd>& u, unsigned w, float* out) {
  const unsigned R = 1u << w, W = 2 * R;
  for (unsigned r = 0; r < W; ++r)
    for (unsigned c = 0; c < W; ++c) {
      const cd z = u[(size_t)(r % R) * R + (c % R)];
      double v;
      if ((r < R) == (c < R)) v = z.real();
      else if (r < R) v = -z.imag();  // top-right: -Ui
      else v = z.imag();              // bottom-left: Ui
      const unsigned swz = (W >= 64 && (c & 1u)) ? 32u : 0u;
      out[(size_t)c * W + (r ^ swz)] = (float)v;
    }
}

// Three-product form of a 2^w block for Gauss's complex product Y = U X:
//   T1 = Ur Xr, T2 = Ui Xi, T3 = (Ur + Ui)(Xr + Xi);  Yr = T1 - T2, Yi = T3 - T1 - T2.
// Ur, Ui and Ur + Ui (the sum formed in FP64), each stored transposed: element (i, c) at c*R + i.
static void split3(const std::vector<cd>& u, unsigned w, float* out) {
  const unsigned R = 1u << w;
  for (unsigned i = 0; i < R; ++i)
    for (unsigned c = 0; c < R; ++c) {
      const cd z = u[(size_t)i * R + c];
      const size_t at = (size_t)c * R + i;
      out[at] = (float)z.real();
      out[(size_t)R * R + at] = (float)z.imag();
      out[2 * (size_t)R * R + at] = (float)(z.real() + z.imag());
    }
}

}  // namespace brk

using namespace brk;

// The ops in a random topological order of their dependences (ops sharing a qubit keep their
// order; a local swap keeps its place relative to every op), drawn from `gen`.
static std::vector<size_t> random_topological_order(const std::vector<brk_circuit::op>& ops, std::mt19937_64& gen) {
  const size_t count = ops.size();
  brk::op_dependences dep = brk::dependences_of(ops);
  const std::vector<std::vector<size_t>>& succ = dep.succ;
  std::vector<unsigned>& indeg = dep.indeg;
  std::vector<size_t> ready, order;
  for (size_t i = 0; i < count; ++i)
    if (!indeg[i]) ready.push_back(i);
  while (!ready.empty()) {
    const size_t k = (size_t)(gen() % ready.size());
    const size_t i = ready[k];
    ready[k] = ready.back();
    ready.pop_back();
    order.push_back(i);
    for (size_t j : succ[i])
      if (--indeg[j] == 0) ready.push_back(j);
  }
  return order;
}

// The physical gate list of the ops in `order` (replaying the permutation from `perm`; a local
// swap becomes a SWAP gate on its two physical qubits).
static int physical_gates(const brk_circuit& circuit, const std::vector<size_t>& order, brk::permutation perm,
                          unsigned width, std::vector<physical_gate>& gates, std::vector<unsigned>& final_perm) {
  gates.clear();
  gates.reserve(order.size());
  for (size_t oi : order) {
    const auto& o = circuit.ops[oi];
    physical_gate g;
    if (o.code == BRK_OP_LOCAL_SWAP) {
      brk_circuit::op sw;
      sw.code = BRK_OP_SWAP;
      sw.targets = o.targets;
      std::vector<unsigned> q;
      op_unitary(sw, q, g.u);
      g.qubits = o.targets;
      perm.permutate(perm.inverse[o.targets[0]], perm.inverse[o.targets[1]]);
    } else {
      std::vector<unsigned> logical;
      if (!op_unitary(o, logical, g.u)) return fail(BRK_ERR_UNSUPPORTED, "op %d has no dense form", o.code);
      if (logical.size() > width)
        return fail(BRK_ERR_UNSUPPORTED, "op on %zu qubits does not fit fusion width %u", logical.size(), width);
      for (unsigned q : logical) g.qubits.push_back(perm.perm[q]);
    }
    gates.push_back(std::move(g));
  }
  final_perm = perm.perm;
  return BRK_OK;
}

struct fused_group {
  std::set<unsigned> qubits;
  std::vector<size_t> members;  // gate indices, in order
};

// Greedy fusion: a block starts at the first unplaced gate and takes every later gate whose
// qubits keep the union within `width` and none of which is blocked by a gate left out.
static std::vector<fused_group> group_gates(const std::vector<physical_gate>& gates, unsigned width) {
  std::vector<fused_group> groups;
  std::vector<char> placed(gates.size(), 0);
  size_t first = 0;
  while (first < gates.size()) {
    std::set<unsigned> S, blocked;
    std::vector<size_t> members;
    for (size_t gi = first; gi < gates.size(); ++gi) {
      if (placed[gi]) continue;
      const auto& q = gates[gi].qubits;
      bool free_of_blocked = true;
      std::set<unsigned> uni = S;
      for (unsigned x : q) {
        if (blocked.count(x)) free_of_blocked = false;
        uni.insert(x);
      }
      if (free_of_blocked && uni.size() <= width) {
        S = std::move(uni);
        members.push_back(gi);
        placed[gi] = 1;
      } else {
        blocked.insert(q.begin(), q.end());
        if (S.size() == width && std::includes(blocked.begin(), blocked.end(), S.begin(), S.end())) break;
      }
    }
    while (first < gates.size() && placed[first]) ++first;
    groups.push_back({std::move(S), std::move(members)});
  }
  return groups;
}

// The 4 lowest qubits (128-byte runs of complex64) are local in every chunk pass.
constexpr unsigned CHUNK_RUN_QUBITS = 4;

// One pass program: local qubits = the run qubits, the blocks' qubits, then the lowest free
// ones up to 13; each block's rows on the local bits of its qubits, its columns on the rest.
static brk::fused_chunk_program make_pass(const brk_fused_plan& plan, size_t first, size_t count, uint64_t local) {
  using namespace brk;
  fused_chunk_program p{};
  for (unsigned q = 0; __builtin_popcountll(local) < FUSED_CHUNK_LOCAL && q < plan.num_qubits; ++q) local |= 1ull << q;
  unsigned pos[FUSED_CHUNK_LOCAL], k = 0;
  for (unsigned q = 0; q < 64; ++q)
    if ((local >> q) & 1ull) pos[k++] = q;
  std::copy(pos, pos + FUSED_CHUNK_LOCAL, p.pos);
  const expand_masks e = make_expand_masks(pos, FUSED_CHUNK_LOCAL);
  std::copy(e.m, e.m + 17, p.outer_masks);
  for (unsigned h = 0; h < 16; ++h) {
    uint64_t d = 0;
    for (unsigned b = 0; b < 4; ++b)
      if ((h >> b) & 1u) d |= 1ull << pos[9 + b];
    p.deposit_high[h] = d;
  }
  p.num_ops = (uint32_t)count;
  for (size_t i = 0; i < count; ++i) {
    const auto& blk = plan.blocks[first + i];
    fused_chunk_op& op = p.ops[i];
    op.matrix = (uint32_t)(first + i);
    uint32_t rows = 0;
    for (size_t j = 0; j < blk.qubits.size(); ++j) {
      const unsigned r = (unsigned)(std::find(pos, pos + FUSED_CHUNK_LOCAL, blk.qubits[j]) - pos);
      op.row_bits[j] = r;
      rows |= 1u << r;
    }
    unsigned c = 0;
    for (unsigned r = 0; r < (unsigned)FUSED_CHUNK_LOCAL; ++r)
      if (!((rows >> r) & 1u)) op.col_bits[c++] = r;
  }
  return p;
}

// Consecutive blocks whose qubits, with the run qubits, fit in a chunk form one pass, up to
// the number of embeddings a pass keeps in LDS.
static void build_passes(brk_fused_plan& plan) {
  using namespace brk;
  const uint64_t run = (1ull << CHUNK_RUN_QUBITS) - 1ull;
  size_t first = 0;
  uint64_t local = run;
  for (size_t b = 0; b <= plan.blocks.size(); ++b) {
    uint64_t m = 0;
    if (b < plan.blocks.size())
      for (unsigned q : plan.blocks[b].qubits) m |= 1ull << q;
    const bool fits = b < plan.blocks.size() && __builtin_popcountll(local | m) <= FUSED_CHUNK_LOCAL &&
                      b - first < (size_t)fused_chunk_max_blocks((int)plan.width);
    if (!fits) {
      if (b > first) plan.passes.push_back(make_pass(plan, first, b - first, local));
      first = b;
      local = run | m;
    } else {
      local |= m;
    }
  }
}

// The block's qubits widened to `width` with the lowest others (the block acts as the
// identity on them), as a mask.
static uint64_t widened_mask(const std::set<unsigned>& S, unsigned width) {
  uint64_t m = 0;
  for (unsigned q : S) m |= 1ull << q;
  for (unsigned q = 0; (unsigned)__builtin_popcountll(m) < width; ++q) m |= 1ull << q;
  return m;
}

// The number of passes build_passes makes of these blocks, in this order.
static size_t count_passes(const std::vector<fused_group>& groups, unsigned width) {
  const uint64_t run = (1ull << CHUNK_RUN_QUBITS) - 1ull;
  const size_t cap = (size_t)brk::fused_chunk_max_blocks((int)width);
  size_t passes = 0, in_pass = 0;
  uint64_t local = run;
  for (const auto& g : groups) {
    const uint64_t m = widened_mask(g.qubits, width);
    if (in_pass == 0 || __builtin_popcountll(local | m) > brk::FUSED_CHUNK_LOCAL || in_pass >= cap) {
      ++passes;
      local = run | m;
      in_pass = 1;
    } else {
      local |= m;
      ++in_pass;
    }
  }
  return passes;
}

// Passes built one at a time: a pass starts with the run qubits and repeatedly takes, among the
// blocks whose earlier blocks on shared qubits are all placed, one that still fits in the chunk,
// preferring the block that adds the fewest qubits (ties and near-ties broken by the seeded noise:
// the score is added qubits + noise * u, u uniform in [0, 1)); the pass ends when no such block
// fits or it holds the blocks a pass can keep.  Returns the blocks in pass order; build_passes and
// count_passes cut this order at the same places, since every pass ends at a block that does not
// fit or at the block limit.
static std::vector<size_t> greedy_chunk_order(const std::vector<uint64_t>& masks, const std::vector<std::vector<size_t>>& succ,
                                              const std::vector<unsigned>& npred0, size_t cap, double noise,
                                              std::mt19937_64& gen) {
  const uint64_t run = (1ull << CHUNK_RUN_QUBITS) - 1ull;
  const size_t B = masks.size();
  std::vector<unsigned> npred = npred0;
  std::vector<char> placed(B, 0);
  std::vector<size_t> order;
  order.reserve(B);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  while (order.size() < B) {
    uint64_t local = run;
    size_t in_pass = 0;
    while (in_pass < cap) {
      size_t pick = B;
      double best = 0.0;
      for (size_t i = 0; i < B; ++i) {
        if (placed[i] || npred[i]) continue;
        const uint64_t m = local | masks[i];
        if (__builtin_popcountll(m) > brk::FUSED_CHUNK_LOCAL) continue;
        const double score = (double)__builtin_popcountll(m & ~local) + noise * u(gen);
        if (pick == B || score < best) {
          pick = i;
          best = score;
        }
      }
      if (pick == B) break;
      placed[pick] = 1;
      local |= masks[pick];
      order.push_back(pick);
      ++in_pass;
      for (size_t j : succ[pick]) --npred[j];
    }
  }
  return order;
}

// Orders the blocks for chunk passes: the given order, the block order for chunks of 13 local
// qubits (the 4 lowest always among them) of it and of random dependence-respecting orders of
// it (blocks sharing a qubit keep their order; fixed seeds, so plans are reproducible); the
// order giving the fewest passes wins (the given one on a tie).  Returns that number.
static size_t order_for_chunks(std::vector<fused_group>& groups, unsigned width, int tries, int greedy_tries) {
  std::vector<brk_circuit::op> pseudo(groups.size());
  for (size_t g = 0; g < groups.size(); ++g) {
    pseudo[g].code = BRK_OP_X;
    pseudo[g].targets.assign(groups[g].qubits.begin(), groups[g].qubits.end());
  }
  std::vector<size_t> best(groups.size());
  for (size_t g = 0; g < best.size(); ++g) best[g] = g;
  size_t best_passes = count_passes(groups, width);
  for (int t = 0; t < tries; ++t) {
    std::vector<size_t> seq(groups.size());
    if (t == 0) {
      for (size_t g = 0; g < seq.size(); ++g) seq[g] = g;
    } else {
      std::mt19937_64 gen((uint64_t)t);
      seq = random_topological_order(pseudo, gen);
    }
    std::vector<brk_circuit::op> ops;
    ops.reserve(seq.size());
    for (size_t g : seq) ops.push_back(pseudo[g]);
    std::vector<size_t> order;
    order.reserve(seq.size());
    for (size_t i : brk::block_order(ops, brk::FUSED_CHUNK_LOCAL, CHUNK_RUN_QUBITS)) order.push_back(seq[i]);
    std::vector<fused_group> trial;
    trial.reserve(order.size());
    for (size_t g : order) trial.push_back(groups[g]);
    const size_t p = count_passes(trial, width);
    if (p < best_passes) {
      best_passes = p;
      best = std::move(order);
    }
  }
  // randomized greedy passes (fixed seeds), kept when they need fewer passes
  if (greedy_tries > 0) {
    const size_t B = groups.size();
    std::vector<uint64_t> masks(B);
    for (size_t g = 0; g < B; ++g) masks[g] = widened_mask(groups[g].qubits, width);
    std::vector<std::vector<size_t>> succ(B);
    std::vector<unsigned> npred(B, 0);
    for (size_t j = 0; j < B; ++j)
      for (size_t i = 0; i < j; ++i)
        if (!groups[i].qubits.empty() && !groups[j].qubits.empty() &&
            std::any_of(groups[j].qubits.begin(), groups[j].qubits.end(),
                        [&](unsigned q) { return groups[i].qubits.count(q) != 0; })) {
          succ[i].push_back(j);
          ++npred[j];
        }
    const size_t cap = (size_t)brk::fused_chunk_max_blocks((int)width);
    for (int t = 0; t < greedy_tries; ++t) {
      std::mt19937_64 gen(0x9e3779b97f4a7c15ull + (uint64_t)t);
      const std::vector<size_t> order = greedy_chunk_order(masks, succ, npred, cap, 2.0, gen);
      std::vector<fused_group> trial;
      trial.reserve(order.size());
      for (size_t g : order) trial.push_back(groups[g]);
      const size_t p = count_passes(trial, width);
      if (p < best_passes) {
        best_passes = p;
        best = order;
      }
    }
  }
  std::vector<fused_group> ordered;
  ordered.reserve(best.size());
  for (size_t g : best) ordered.push_back(std::move(groups[g]));
  groups = std::move(ordered);
  return best_passes;
}

// Block orders tried per plan: the block order of the given sequence, then that of random
// topological orders (fixed seeds: the plan is reproducible); the fewest blocks win.  Long
// circuits try fewer (planning time grows with ops x tries).
constexpr int FUSION_ORDER_TRIES = 128;
constexpr int CHUNK_ORDER_TRIES = 8;  // chunk orders tried per candidate gate order
// Time of a chunk pass's HBM round trip in units of one block's work in a pass (measured at
// width 4 on the 32-qubit circuit: ~5.7 ms per pass, ~4.6 ms per block).
constexpr double CHUNK_PASS_COST = 1.25;
static int fusion_order_tries(size_t num_ops) {
  return num_ops <= 4096 ? FUSION_ORDER_TRIES : num_ops <= 65536 ? 4 : 1;
}
// Randomized greedy pass orders (greedy_chunk_order) tried per candidate and for the final plan;
// one try costs ~blocks^2 steps, so long plans try fewer.
constexpr int CHUNK_GREEDY_TRIES = 16;
constexpr int CHUNK_FINAL_GREEDY_TRIES = 1024;
static int chunk_greedy_tries(size_t blocks, int tries) {
  const size_t budget = (size_t)1 << 24;  // block-pair steps
  const size_t per = std::max<size_t>(1, blocks * blocks);
  return (int)std::min<size_t>((size_t)tries, budget / per);
}

extern "C" {

int brk_fused_plan_create(const brk_circuit* circuit, unsigned width, const unsigned* permutation,
                          brk_fused_plan** out) {
  if (!circuit || !out) return fail(BRK_ERR_INVALID_ARGUMENT, "null argument");
  if (width < 4 || width > 6) return fail(BRK_ERR_UNSUPPORTED, "fusion width %u not in 4..6", width);
  const unsigned n = circuit->num_qubits;
  if (n < width + 5) return fail(BRK_ERR_UNSUPPORTED, "fusion width %u needs at least %u qubits", width, width + 5);
  brk::permutation perm(n);
  if (permutation && !perm.load(permutation, n)) return fail(BRK_ERR_INVALID_ARGUMENT, "bad permutation");

  // 1-2. physical gate lists of candidate orders, greedy fusion of each; the fewest blocks win
  // (for chunk passes: the fewest passes, then the fewest blocks)
  const bool chunked = width <= 5 && n >= (unsigned)FUSED_CHUNK_LOCAL;
  double best_cost = 0.0;
  std::vector<physical_gate> gates, trial;
  std::vector<fused_group> groups;
  std::vector<unsigned> final_perm, trial_perm;
  const int tries = fusion_order_tries(circuit->ops.size());
  for (int t = 0; t < tries; ++t) {
    std::vector<size_t> order;
    if (t == 0) {
      order = brk::block_order(circuit->ops, width, 0);
    } else {
      std::mt19937_64 gen((uint64_t)t);
      const std::vector<size_t> shuffled = random_topological_order(circuit->ops, gen);
      std::vector<brk_circuit::op> ops;
      ops.reserve(shuffled.size());
      for (size_t i : shuffled) ops.push_back(circuit->ops[i]);
      for (size_t i : brk::block_order(ops, width, 0)) order.push_back(shuffled[i]);
    }
    if (int st = physical_gates(*circuit, order, perm, width, trial, trial_perm)) return st;
    std::vector<fused_group> g = group_gates(trial, width);
    // chunk passes: the fewest passes (each reads and writes the state once), then the fewest blocks
    const size_t passes = chunked ? order_for_chunks(g, width, CHUNK_ORDER_TRIES, chunk_greedy_tries(g.size(), CHUNK_GREEDY_TRIES)) : 0;
    const double cost = chunked ? CHUNK_PASS_COST * (double)passes + (double)g.size() : (double)g.size();
    if (t == 0 || cost < best_cost) {
      best_cost = cost;
      groups = std::move(g);
      gates.swap(trial);
      final_perm = trial_perm;
    }
  }

  // 4. blocks in an order for chunks of 13 local qubits that packs them into few passes
  if (chunked)
    (void)order_for_chunks(groups, width, fusion_order_tries(groups.size()),
                           chunk_greedy_tries(groups.size(), CHUNK_FINAL_GREEDY_TRIES));

  auto* plan = new (std::nothrow) brk_fused_plan;
  if (!plan) return fail(BRK_ERR_OUT_OF_MEMORY, "cannot allocate plan");
  plan->num_qubits = n;
  plan->width = width;
  plan->final_perm = final_perm;

  for (auto& grp : groups) {
    std::set<unsigned>& S = grp.qubits;
    const std::vector<size_t>& members = grp.members;
    // widen to exactly `width` qubits with the lowest unused ones
    for (unsigned q = 0; S.size() < width; ++q) S.insert(q);
    brk_fused_plan::block blk;
    blk.qubits.assign(S.begin(), S.end());
    blk.num_gates = (unsigned)members.size();
    // 3. block matrix: columns of the identity pushed through the gates
    const unsigned R = 1u << width;
    std::vector<cd> cols((size_t)R * R);  // column-major scratch
    std::vector<cd> col(R);
    for (unsigned c = 0; c < R; ++c) {
      std::fill(col.begin(), col.end(), cd{0, 0});
      col[c] = 1.0;
      for (size_t gi : members) {
        std::vector<unsigned> pos;
        for (unsigned x : gates[gi].qubits)
          pos.push_back((unsigned)(std::lower_bound(blk.qubits.begin(), blk.qubits.end(), x) - blk.qubits.begin()));
        apply_small(col, width, gates[gi].u, pos);
      }
      for (unsigned r = 0; r < R; ++r) cols[(size_t)c * R + r] = col[r];
    }
    blk.u.resize((size_t)R * R);
    for (unsigned r = 0; r < R; ++r)
      for (unsigned c = 0; c < R; ++c) blk.u[(size_t)r * R + c] = cols[(size_t)c * R + r];
    build_launch(blk, n, width);
    plan->blocks.push_back(std::move(blk));
  }

  // FP32 real embeddings for the device (uploaded on the first apply)
  const uint64_t per = fused_matrix_floats((int)width);
  plan->host_matrices.resize(per * plan->blocks.size());
  for (size_t b = 0; b < plan->blocks.size(); ++b)
    embed(plan->blocks[b].u, width, plan->host_matrices.data() + b * per);
  if (chunked) build_passes(*plan);
  if (!plan->passes.empty()) {
    const uint64_t per3 = fused_split_floats((int)width);
    plan->host_split.resize(per3 * plan->blocks.size());
    for (size_t b = 0; b < plan->blocks.size(); ++b)
      split3(plan->blocks[b].u, width, plan->host_split.data() + b * per3);
  }
  *out = plan;
  return BRK_OK;
}

// Copies `bytes` from host to a new device buffer in *out (nothing when it is already there).
static int upload_once(void** out, const void* host, size_t bytes, const char* what) {
  if (*out || bytes == 0) return BRK_OK;
  void* d = nullptr;
  if (hipMalloc(&d, bytes) != hipSuccess) return fail(BRK_ERR_OUT_OF_MEMORY, "cannot allocate %s", what);
  if (hipMemcpy(d, host, bytes, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(d);
    return fail(BRK_ERR_HIP, "cannot upload %s", what);
  }
  *out = d;
  return BRK_OK;
}

// Uploads the plan's device data to the current device once: the block embeddings and, for
// plans with chunk passes, the three-product forms and the pass programs.  Each piece is
// uploaded independently, so a call after a failed one retries what is missing.
static int ensure_device_matrices(const brk_fused_plan* plan) {
  void* d = plan->device_matrices;
  int st = upload_once(&d, plan->host_matrices.data(), plan->host_matrices.size() * sizeof(float), "fused matrices");
  plan->device_matrices = static_cast<float*>(d);
  if (st) return st;
  const uint64_t per = fused_matrix_floats((int)plan->width);
  for (size_t b = 0; b < plan->blocks.size(); ++b) plan->blocks[b].launch.matrix = plan->device_matrices + b * per;
  if (plan->passes.empty()) return BRK_OK;
  d = plan->device_split;
  st = upload_once(&d, plan->host_split.data(), plan->host_split.size() * sizeof(float), "fused three-product forms");
  plan->device_split = static_cast<float*>(d);
  if (st) return st;
  d = plan->device_passes;
  st = upload_once(&d, plan->passes.data(), plan->passes.size() * sizeof(fused_chunk_program), "fused chunk passes");
  plan->device_passes = static_cast<fused_chunk_program*>(d);
  return st;
}

int brk_fused_plan_destroy(brk_fused_plan* plan) {
  delete plan;
  return BRK_OK;
}

int brk_fused_plan_num_blocks(const brk_fused_plan* plan, unsigned* num_blocks) {
  if (!plan || !num_blocks) return fail(BRK_ERR_INVALID_ARGUMENT, "null argument");
  *num_blocks = (unsigned)plan->blocks.size();
  return BRK_OK;
}

int brk_fused_plan_block(const brk_fused_plan* plan, unsigned i, unsigned* qubits, unsigned* num_gates,
                         double* matrix) {
  if (!plan) return fail(BRK_ERR_INVALID_ARGUMENT, "plan is null");
  if (i >= plan->blocks.size()) return fail(BRK_ERR_INVALID_ARGUMENT, "block %u out of range", i);
  const auto& b = plan->blocks[i];
  if (qubits) std::copy(b.qubits.begin(), b.qubits.end(), qubits);
  if (num_gates) *num_gates = b.num_gates;
  if (matrix)
    for (size_t k = 0; k < b.u.size(); ++k) {
      matrix[2 * k] = b.u[k].real();
      matrix[2 * k + 1] = b.u[k].imag();
    }
  return BRK_OK;
}

int brk_fused_plan_num_passes(const brk_fused_plan* plan, unsigned* num_passes) {
  if (!plan || !num_passes) return fail(BRK_ERR_INVALID_ARGUMENT, "null argument");
  *num_passes = (unsigned)plan->passes.size();
  return BRK_OK;
}

int brk_fused_plan_final_permutation(const brk_fused_plan* plan, unsigned* permutation) {
  if (!plan || !permutation) return fail(BRK_ERR_INVALID_ARGUMENT, "null argument");
  std::copy(plan->final_perm.begin(), plan->final_perm.end(), permutation);
  return BRK_OK;
}

static int apply_impl(const brk_fused_plan* plan, void* first, int dtype, float* block_ms, void* stream) {
  if (!plan || !first) return fail(BRK_ERR_INVALID_ARGUMENT, "null argument");
  if (dtype != BRK_C64) return fail(BRK_ERR_UNSUPPORTED_DTYPE, "the fused form runs on complex64 states");
  if (int s = ensure_device_matrices(plan)) return s;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  std::vector<hipEvent_t> ev;
  if (block_ms) {
    ev.resize(plan->blocks.size() + 1);
    for (auto& e : ev)
      if (hipEventCreate(&e) != hipSuccess) return fail(BRK_ERR_HIP, "hipEventCreate failed");
    (void)hipEventRecord(ev[0], st);
  }
  int status = BRK_OK;
  const bool passes = !plan->passes.empty() && fused_chunks_enabled();
  const size_t launches = passes ? plan->passes.size() : plan->blocks.size();
  const uint64_t size = 1ull << plan->num_qubits;
  for (size_t b = 0; b < launches && status == BRK_OK; ++b) {
    status = passes ? launch_fused_chunk(plan->device_passes + b, plan->passes[b], (int)plan->width,
                                         plan->device_split, first, size, stream)
                    : launch_fused_block(plan->blocks[b].launch, first, stream);
    if (block_ms && status == BRK_OK) (void)hipEventRecord(ev[b + 1], st);
  }
  if (block_ms) {
    if (status == BRK_OK && hipEventSynchronize(ev[launches]) != hipSuccess)
      status = fail(BRK_ERR_HIP, "hipEventSynchronize failed");
    for (size_t b = 0; b < plan->blocks.size(); ++b) block_ms[b] = 0.0f;
    for (size_t b = 0; status == BRK_OK && b < launches; ++b)
      (void)hipEventElapsedTime(&block_ms[b], ev[b], ev[b + 1]);
    for (auto& e : ev) (void)hipEventDestroy(e);
  }
  return status;
}

int brk_fused_apply(const brk_fused_plan* plan, void* first, int dtype, void* stream) {
  return apply_impl(plan, first, dtype, nullptr, stream);
}

int brk_fused_apply_timed(const brk_fused_plan* plan, void* first, int dtype, float* block_ms, void* stream) {
  if (!block_ms) return fail(BRK_ERR_INVALID_ARGUMENT, "block_ms is null");
  return apply_impl(plan, first, dtype, block_ms, stream);
}

}  // extern "C"
