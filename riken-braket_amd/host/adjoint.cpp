// Differentiation form, host side: the adjoint method for
// E(theta) = <psi(theta)| H |psi(theta)> with the open-chain Ising H.
//
//   forward:  psi = U_N ... U_1 |psi0>        (plain per-gate kernels)
//             lam = H psi, E = <psi|lam>
//   reverse:  for i = N..1:  dE/dtheta_i = 2 Re <lam| D_i |psi>,
//             psi := U_i^dagger psi, lam := U_i^dagger lam,
// with D_i = (dU_i/dtheta) U_i^dagger (so D_i |psi_i> = dU_i/dtheta |psi_(i-1)>):
//   RX/RY/RZ(theta) = exp(-i theta G / 2):  D = -(i/2) G
//   EX/EY/EZ(phi)   = exp(i phi G):         D = i G
//   U1(lambda)      = diag(1, e^{i lambda}):  D = diag(0, i)
// Default (tuning adjoint_blocks = 1): the forward pass runs in gate blocks,
// and the reverse sweep in blocks of ops whose qubits fit in a chunk of 2^11
// (complex<double>) or 2^12 (complex<float>) amplitudes: psi and lambda chunks
// stay in LDS while the block's inverse ops are un-applied and its gradient
// terms summed (k_adjoint_block), one HBM round trip of both states per block.
// Within a block, register segments take runs of ops on at most L - 9 qubits
// (every thread holds a group of psi and lambda amplitudes) with one LDS round
// trip; RX / RY / RZ-type ops use the general formulas with their zero terms
// left out.  Per-chunk gradient terms are summed on the device in a fixed order.
// Otherwise: runs of consecutive single-qubit ops on the same qubit are swept
// by one fused kernel (csrc/brk_adjoint.hip); runs of Z / CZ ops (diagonal,
// self-inverse, mutually commuting) are one sign pass per state; other ops
// are un-applied to both states with the plain kernels.  That forward pass
// merges the same runs: a single-qubit run becomes one 2x2 product (FP64 on
// the host), a Z / CZ run one sign pass.
#include "circuit.hpp"
#include "brk_adjoint.hpp"
#include "brk_block.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <array>
#include <complex>
#include <memory>
#include <mutex>
#include <vector>

namespace brk {
int fail(int status, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
}

using cd = std::complex<double>;

namespace brk {

static bool is_parametric(int code) {
  return code == BRK_OP_RX || code == BRK_OP_RY || code == BRK_OP_RZ || code == BRK_OP_EX || code == BRK_OP_EY ||
         code == BRK_OP_EZ || code == BRK_OP_U1;
}

// Single-qubit op without controls: the adjoint kernel can sweep it.
static bool chainable(const brk_circuit::op& o) {
  if (!o.controls.empty() || o.targets.size() != 1) return false;
  return o.code != BRK_OP_SWAP && o.code != BRK_OP_LOCAL_SWAP;
}

static void put(double* d, int slot, cd v) {
  d[2 * slot] = v.real();
  d[2 * slot + 1] = v.imag();
}

// M = U^dagger and D for a chainable op.
static bool make_adj_op(const brk_circuit::op& o, adj_op& a) {
  std::vector<unsigned> q;
  std::vector<std::array<double, 2>> u;
  if (!op_unitary(o, q, u) || u.size() != 4) return false;
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 2; ++c) put(a.m, 2 * r + c, std::conj(cd(u[2 * c + r][0], u[2 * c + r][1])));
  a.has_grad = is_parametric(o.code);
  for (double& x : a.d) x = 0.0;
  const cd I(0, 1);
  cd g00 = 0, g01 = 0, g10 = 0, g11 = 0;  // generator G
  cd scale = 0;
  switch (o.code) {
    case BRK_OP_RX: case BRK_OP_EX: g01 = g10 = 1; break;
    case BRK_OP_RY: case BRK_OP_EY: g01 = -I; g10 = I; break;
    case BRK_OP_RZ: case BRK_OP_EZ: g00 = 1; g11 = -1; break;
    case BRK_OP_U1: g11 = 1; break;
    default: break;
  }
  if (o.code == BRK_OP_RX || o.code == BRK_OP_RY || o.code == BRK_OP_RZ) scale = -0.5 * I;
  else if (a.has_grad) scale = I;
  put(a.d, 0, scale * g00);
  put(a.d, 1, scale * g01);
  put(a.d, 2, scale * g10);
  put(a.d, 3, scale * g11);
  return true;
}

static brk_circuit::op inverse_op(const brk_circuit::op& o) {
  brk_circuit::op r = o;
  switch (o.code) {
    case BRK_OP_U1: case BRK_OP_EX: case BRK_OP_EY: case BRK_OP_EZ:
    case BRK_OP_RX: case BRK_OP_RY: case BRK_OP_RZ: case BRK_OP_ESWAP:
      r.params[0] = -o.params[0];
      break;
    case BRK_OP_U3:
      r.params[0] = -o.params[0];
      r.params[1] = -o.params[2];
      r.params[2] = -o.params[1];
      break;
    case BRK_OP_U2:  // U2(phi, lam) = U3(pi/2, phi, lam)
      r.code = BRK_OP_U3;
      r.num_params = 3;
      r.params[0] = -1.57079632679489661923132169163975144;
      r.params[1] = -o.params[1];
      r.params[2] = -o.params[0];
      break;
    case BRK_OP_PHASE_COEFF: r.params[1] = -o.params[1]; break;  // the conjugate coefficient
    case BRK_OP_MATRIX:  // conjugate transpose
      for (int rr = 0; rr < 2; ++rr)
        for (int c = 0; c < 2; ++c) {
          r.params[2 * (2 * rr + c)] = o.params[2 * (2 * c + rr)];
          r.params[2 * (2 * rr + c) + 1] = -o.params[2 * (2 * c + rr) + 1];
        }
      break;
    default: break;  // H, X, Y, Z, SWAP are their own inverses
  }
  return r;
}

// Z on one qubit or CZ on two (controls count as qubits): one sign term.
static bool sign_term_of(const brk_circuit::op& o, uint32_t* d, uint64_t* a) {
  if (o.code != BRK_OP_Z) return false;
  std::vector<unsigned> q(o.targets);
  q.insert(q.end(), o.controls.begin(), o.controls.end());
  if (q.size() == 1) {
    *d = 0;
    *a = 1ull << q[0];
    return true;
  }
  if (q.size() != 2) return false;
  const unsigned lo = std::min(q[0], q[1]), hi = std::max(q[0], q[1]);
  *d = hi - lo;
  *a = 1ull << lo;
  return true;
}

static bool is_sign(const brk_circuit::op& o) {
  uint32_t d;
  uint64_t a;
  return sign_term_of(o, &d, &a);
}

// Applies the sign ops ops[b, e) to one or two states, as few passes as the
// term limit allows (the ops commute, so any grouping is exact).
static int apply_sign_run(const std::vector<brk_circuit::op>& ops, size_t b, size_t e, void* s1, void* s2,
                          uint64_t size, int dtype, void* stream) {
  while (b < e) {
    sign_terms t;
    size_t k = b;
    for (; k < e; ++k) {
      uint32_t d;
      uint64_t a;
      sign_term_of(ops[k], &d, &a);
      int slot = 0;
      while (slot < t.n && t.d[slot] != d) ++slot;
      if (slot == t.n) {
        if (t.n == SIGN_MAX_TERMS) break;
        t.d[t.n] = d;
        t.a[t.n] = 0;
        ++t.n;
      }
      t.a[slot] ^= a;
    }
    if (int s = launch_sign_terms(t, s1, size, dtype, stream)) return s;
    if (s2)
      if (int s = launch_sign_terms(t, s2, size, dtype, stream)) return s;
    b = k;
  }
  return BRK_OK;
}

// psi := U psi with the runs merged (forward pass).
static int forward(const std::vector<brk_circuit::op>& ops, void* psi, uint64_t size, int dtype, void* stream) {
  size_t i = 0;
  while (i < ops.size()) {
    const auto& o = ops[i];
    if (chainable(o)) {
      std::vector<unsigned> q;
      std::vector<std::array<double, 2>> u;
      if (!op_unitary(o, q, u) || u.size() != 4) return fail(BRK_ERR_UNSUPPORTED, "op %zu has no 2x2 form", i);
      cd m[4];
      for (int k = 0; k < 4; ++k) m[k] = cd(u[k][0], u[k][1]);
      size_t j = i + 1;
      for (; j < ops.size() && chainable(ops[j]) && ops[j].targets[0] == o.targets[0]; ++j) {
        if (!op_unitary(ops[j], q, u) || u.size() != 4) return fail(BRK_ERR_UNSUPPORTED, "op %zu has no 2x2 form", j);
        const cd b0(u[0][0], u[0][1]), b1(u[1][0], u[1][1]), b2(u[2][0], u[2][1]), b3(u[3][0], u[3][1]);
        const cd a0 = m[0], a1 = m[1], a2 = m[2], a3 = m[3];
        m[0] = b0 * a0 + b1 * a2;  // later op on the left
        m[1] = b0 * a1 + b1 * a3;
        m[2] = b2 * a0 + b3 * a2;
        m[3] = b2 * a1 + b3 * a3;
      }
      double mm[8];
      for (int k = 0; k < 4; ++k) {
        mm[2 * k] = m[k].real();
        mm[2 * k + 1] = m[k].imag();
      }
      if (int s = brk_gate_1q(psi, size, dtype, o.targets[0], mm, nullptr, 0, stream)) return s;
      i = j;
    } else if (is_sign(o)) {
      size_t j = i + 1;
      while (j < ops.size() && is_sign(ops[j])) ++j;
      if (int s = apply_sign_run(ops, i, j, psi, nullptr, size, dtype, stream)) return s;
      i = j;
    } else {
      if (int s = apply_physical_op(o, psi, size, dtype, stream)) return s;
      ++i;
    }
  }
  return BRK_OK;
}

// Per-device FP64 partial-sum buffer, grown as needed and kept until exit.  A gradient
// holds the device's buffer (a lease) from its first launch until it has read the
// result back: calls from several host threads on one device take turns, and the
// buffer is never freed or regrown while another call's kernels still write it.
struct partials_slot {
  std::mutex busy;
  double* ptr = nullptr;
  size_t doubles = 0;
};
static std::mutex g_partials_mutex;
static std::vector<std::unique_ptr<partials_slot>> g_partials;

struct partials_lease {
  std::unique_lock<std::mutex> hold;
  double* ptr = nullptr;
  hipStream_t stream = nullptr;
  // an early return may leave kernels that write the buffer in flight: drain them first
  ~partials_lease() {
    if (hold.owns_lock()) (void)hipStreamSynchronize(stream);
  }
};

static int get_partials(size_t doubles, hipStream_t stream, partials_lease& lease) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(BRK_ERR_HIP, "hipGetDevice failed");
  partials_slot* slot = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_partials_mutex);
    if ((int)g_partials.size() <= dev) g_partials.resize(dev + 1);
    if (!g_partials[dev]) g_partials[dev] = std::make_unique<partials_slot>();
    slot = g_partials[dev].get();
  }
  std::unique_lock<std::mutex> hold(slot->busy);
  if (slot->doubles < doubles) {
    if (slot->ptr) (void)hipFree(slot->ptr);
    slot->ptr = nullptr;
    slot->doubles = 0;
    if (hipMalloc(&slot->ptr, doubles * sizeof(double)) != hipSuccess) {
      slot->ptr = nullptr;
      return fail(BRK_ERR_OUT_OF_MEMORY, "cannot allocate %zu gradient partials", doubles);
    }
    slot->doubles = doubles;
  }
  lease.hold = std::move(hold);
  lease.ptr = slot->ptr;
  lease.stream = stream;
  return BRK_OK;
}

// The arithmetic form of parametric op k (adjoint_grad_kind) from the exact zeros of D and of
// the inverse op's coefficients.
static int32_t grad_kind(const block_op& op, const double* d) {
  const bool d00 = d[0] == 0.0 && d[1] == 0.0, d01 = d[2] == 0.0 && d[3] == 0.0, d10 = d[4] == 0.0 && d[5] == 0.0,
             d11 = d[6] == 0.0 && d[7] == 0.0;
  if (op.mode == BLOCK_DIAG) {
    // RZ / EZ: parity of the target alone; U1: the target's 1 side alone
    const uint32_t target = 1u << op.pair_bit;
    const bool form = (op.cond == 0 && op.zmask == target) || (op.cond == target && op.zmask == 0);
    return form && d01 && d10 && d[0] == 0.0 && d[6] == 0.0 ? ADJ_GRAD_Z : ADJ_GRAD_GENERAL;
  }
  // a sign mask of the target bit alone never flips: the pair's first index has it clear
  const bool no_flip = (op.sign & ~(1u << op.pair_bit)) == 0;
  if (op.mode != BLOCK_MATRIX || !no_flip || op.cond != 0 || !d00 || !d11) return ADJ_GRAD_GENERAL;
  const double* u = op.u;
  if (u[1] == 0.0 && u[3] == 0.0 && u[5] == 0.0 && u[7] == 0.0 && d[3] == 0.0 && d[5] == 0.0) return ADJ_GRAD_Y;
  if (u[1] == 0.0 && u[2] == 0.0 && u[4] == 0.0 && u[7] == 0.0 && d[2] == 0.0 && d[4] == 0.0) return ADJ_GRAD_X;
  return ADJ_GRAD_GENERAL;
}

// Passes of an adjoint block: register segments (each thread holds one group of
// psi and lambda amplitudes, L - 9 group bits) take runs of ops whose flipped
// bits -- and, for a parametric op, its target -- are group bits, with at most
// ADJ_SEGMENT_GRADIENTS parametric ops; diagonal ops under conditions other than
// negations (Z, CZ) and everything else get a pass of their own, as does a segment
// of one op.
static void make_adjoint_passes(adjoint_block_program& ap) {
  block_program& prog = ap.base;
  const unsigned gb = prog.num_local - (unsigned)BLOCK_THREAD_BITS;
  prog.num_passes = 0;
  uint32_t bits = 0, first = 0, count = 0, grads = 0;
  auto single = [&](uint32_t i) {
    block_pass& p = prog.passes[prog.num_passes++];
    p = block_pass{};
    p.kind = BLOCK_PASS_SINGLE;
    p.first_op = i;
    p.num_ops = 1;
  };
  auto close = [&]() {
    if (count == 1) single(first);
    if (count > 1) {
      block_pass& p = prog.passes[prog.num_passes++];
      p = block_pass{};
      p.kind = BLOCK_PASS_SEGMENT;
      p.first_op = first;
      p.num_ops = count;
      for (unsigned b = 0; b < prog.num_local && (unsigned)__builtin_popcount(bits) < gb; ++b) bits |= 1u << b;
      unsigned k = 0;
      for (unsigned b = 0; b < prog.num_local; ++b)
        if (bits >> b & 1u) p.group[k++] = b;
      for (uint32_t i = first; i < first + count; ++i) {
        block_op& o = prog.ops[i];
        o.group_flip = o.group_pair = 0;
        o.group_zmask = ~0u;  // the reverse segments select diagonal coefficients per amplitude
        for (unsigned g = 0; g < k; ++g) {
          if (o.mode != BLOCK_DIAG && (o.flip >> p.group[g] & 1u)) o.group_flip |= 1u << g;
          if ((o.mode != BLOCK_DIAG || ap.grad_slot[i] >= 0) && p.group[g] == o.pair_bit) o.group_pair = 1u << g;
        }
      }
    }
    bits = 0;
    count = 0;
    grads = 0;
  };
  for (uint32_t i = 0; i < prog.num_ops; ++i) {
    const block_op& o = prog.ops[i];
    const bool grad = ap.grad_slot[i] >= 0;
    const uint32_t mixing = (o.mode == BLOCK_DIAG ? 0u : o.flip) | (grad ? 1u << o.pair_bit : 0u);
    const bool alone = (o.mode == BLOCK_DIAG && !grad && o.cond != 0 && !o.negate) ||
                       (unsigned)__builtin_popcount(mixing) > 2 || (unsigned)__builtin_popcount(mixing) > gb;
    if (alone) {
      close();
      single(i);
      continue;
    }
    if ((unsigned)__builtin_popcount(bits | mixing) > gb || (grad && grads == (uint32_t)ADJ_SEGMENT_GRADIENTS))
      close();
    if (count == 0) first = i;
    bits |= mixing;
    ++count;
    grads += grad ? 1 : 0;
  }
  close();
}

// The blocked gradient (see the top of this file).  pidx: parameter index of
// every op or -1.  Writes the energy and 2 Re <lam| D |psi> per parameter.
static int blocked_gradient(const brk_circuit* given, const std::vector<int>& given_pidx, unsigned nparams,
                            void* psi, void* lam, int dtype, const double* J, const double* h, double* energy,
                            double* gradient, void* stream) {
  const unsigned n = given->num_qubits;
  const uint64_t size = 1ull << n;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const unsigned L = dtype == BRK_C128 ? ADJ_BLOCK_LOCAL_C128 : ADJ_BLOCK_LOCAL_C64;
  const unsigned line_bits = dtype == BRK_C128 ? 3 : 4;  // one 128-byte line
  const uint64_t line = (1ull << line_bits) - 1ull;

  // The ops reordered into few blocks (ops on disjoint qubits commute): for the forward
  // blocks of block_max_local qubits, and for the reverse blocks of L, with their parameter
  // indices.  Any order of the same operator is a valid sequence for the adjoint method: the
  // reverse sweep un-applies its own order from the same final state.
  brk_circuit forward_order, ordered;
  forward_order.num_qubits = ordered.num_qubits = n;
  for (size_t i : block_order(given->ops, block_max_local(dtype), line_bits))
    forward_order.ops.push_back(given->ops[i]);

  // forward pass in gate blocks and lam = H psi on the device, while the host builds the reverse
  // sweep's programs; then E = <psi|lam> (which synchronises the stream)
  brk_block_plan* plan = nullptr;
  if (int s = brk_block_plan_create(&forward_order, dtype, nullptr, &plan)) return s;
  int status = brk_block_plan_run(plan, psi, stream);
  if (!status) status = brk_apply_ising(lam, psi, size, dtype, J, h, stream);

  // reverse sweep: blocks of inverse ops, last op first
  std::vector<adjoint_block_program> programs;
  std::vector<gate_desc> descs;
  std::vector<int> slots;
  std::vector<std::array<double, 8>> dmats;
  std::vector<unsigned> targets;
  uint64_t qubits = 0;
  auto flush = [&]() {
    if (descs.empty()) return;
    programs.emplace_back();
    adjoint_block_program& ap = programs.back();
    ap = adjoint_block_program{};
    ap.base = make_block_program(descs.data(), (unsigned)descs.size(), n, L, line_bits);
    for (size_t k = 0; k < descs.size(); ++k) {
      ap.grad_slot[k] = slots[k];
      std::copy(dmats[k].begin(), dmats[k].end(), ap.d[k]);
      if (slots[k] >= 0) {  // the target's local bit (diagonal ops carry no pair bit)
        for (unsigned r = 0; r < ap.base.num_local; ++r)
          if (ap.base.pos[r] == targets[k]) ap.base.ops[k].pair_bit = r;
        ap.grad_kind[k] = grad_kind(ap.base.ops[k], ap.d[k]);
      }
    }
    make_adjoint_passes(ap);
    descs.clear();
    slots.clear();
    dmats.clear();
    targets.clear();
    qubits = 0;
  };
  std::vector<int> pidx;
  auto build_programs = [&]() -> int {
    for (size_t i : block_order(given->ops, L, line_bits)) {
      ordered.ops.push_back(given->ops[i]);
      pidx.push_back(given_pidx[i]);
    }
    const auto& ops = ordered.ops;
    for (size_t i = ops.size(); i-- > 0;) {
      const brk_circuit::op inv = inverse_op(ops[i]);
      gate_desc d;
      if (int s = describe_physical_op(inv.code, inv.targets.data(), (unsigned)inv.targets.size(),
                                       inv.controls.empty() ? nullptr : inv.controls.data(),
                                       (unsigned)inv.controls.size(), inv.params.data(), &d))
        return s;
      if (__builtin_popcountll(qubits | d.qubits | line) > (int)L || descs.size() == (size_t)BLOCK_MAX_OPS) flush();
      std::array<double, 8> dm{};
      if (pidx[i] >= 0) {
        adj_op a;
        if (!make_adj_op(ops[i], a)) return fail(BRK_ERR_UNSUPPORTED, "op %zu has no 2x2 form", i);
        std::copy(a.d, a.d + 8, dm.begin());
      }
      descs.push_back(d);
      slots.push_back(pidx[i]);
      dmats.push_back(dm);
      targets.push_back(ops[i].targets[0]);
      qubits |= d.qubits;
    }
    flush();
    return BRK_OK;
  };

  if (!status) status = build_programs();
  double ip[2] = {0.0, 0.0};
  if (!status) status = brk_inner_product(lam, dtype, psi, dtype, size, ip, stream);
  else (void)hipStreamSynchronize(st);  // the plan's device programs are idle before they are freed
  (void)brk_block_plan_destroy(plan);   // (brk_inner_product synchronised the stream)
  if (status) return status;
  *energy = ip[0];

  const uint64_t chunks = size >> L;
  // one leased buffer: nparams x chunks partials, the nparams gradients, then the programs
  static_assert(alignof(adjoint_block_program) <= alignof(double), "programs follow doubles");
  const size_t nsums = std::max(1u, nparams);
  const size_t bytes = sizeof(adjoint_block_program) * programs.size();
  partials_lease lease;
  if (int s = get_partials((size_t)nparams * chunks + nsums + (bytes + sizeof(double) - 1) / sizeof(double), st, lease))
    return s;
  double* partials = lease.ptr;
  double* sums = partials + (size_t)nparams * chunks;
  auto* device = reinterpret_cast<adjoint_block_program*>(sums + nsums);
  if (!programs.empty() && hipMemcpyAsync(device, programs.data(), bytes, hipMemcpyHostToDevice, st) != hipSuccess)
    status = fail(BRK_ERR_HIP, "uploading the adjoint programs failed");
  for (size_t b = 0; !status && b < programs.size(); ++b)
    status = launch_adjoint_block(device + b, programs[b], psi, lam, size, dtype, partials, stream);
  if (!status) status = sum_gradient_partials(partials, nparams, chunks, 2.0, sums, stream);
  if (!status && nparams &&
      hipMemcpyAsync(gradient, sums, nparams * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess)
    status = fail(BRK_ERR_HIP, "copying the gradient failed");
  if (hipStreamSynchronize(st) != hipSuccess && !status) status = fail(BRK_ERR_HIP, "hipStreamSynchronize failed");
  return status;
}

}  // namespace brk

using namespace brk;

extern "C" {

int brk_circuit_num_parameters(const brk_circuit* circuit, unsigned* num_parameters) {
  if (!circuit || !num_parameters) return fail(BRK_ERR_INVALID_ARGUMENT, "null argument");
  unsigned n = 0;
  for (const auto& o : circuit->ops) n += is_parametric(o.code) ? 1u : 0u;
  *num_parameters = n;
  return BRK_OK;
}

int brk_adjoint_ising_gradient(const brk_circuit* circuit, void* psi, void* work, int dtype, const double* J,
                               const double* h, double* energy, double* gradient, unsigned num_gradient,
                               void* stream) {
  if (!circuit || !psi || !work || !energy || (num_gradient && !gradient))
    return fail(BRK_ERR_INVALID_ARGUMENT, "null argument");
  if (psi == work) return fail(BRK_ERR_INVALID_ARGUMENT, "psi and work must be different states");
  unsigned nparams = 0;
  (void)brk_circuit_num_parameters(circuit, &nparams);
  if (nparams != num_gradient)
    return fail(BRK_ERR_INVALID_ARGUMENT, "circuit has %u parameters, gradient has %u", nparams, num_gradient);
  const auto& ops = circuit->ops;
  // parameter index of every op (-1 if none); reject ops the sweep cannot handle
  std::vector<int> pidx(ops.size(), -1);
  int next = 0;
  for (size_t i = 0; i < ops.size(); ++i) {
    if (ops[i].code == BRK_OP_LOCAL_SWAP) return fail(BRK_ERR_UNSUPPORTED, "local swaps are not supported here");
    if (is_parametric(ops[i].code)) {
      if (!chainable(ops[i]))
        return fail(BRK_ERR_UNSUPPORTED, "parametric op %zu must act on one qubit without controls", i);
      pidx[i] = next++;
    }
  }
  const unsigned L = dtype == BRK_C128 ? ADJ_BLOCK_LOCAL_C128 : ADJ_BLOCK_LOCAL_C64;
  // blocks hold every op with the line qubits in L local qubits; an op on more qubits (a
  // many-controlled gate) sends the whole gradient to the per-run sweep
  const uint64_t line_mask = dtype == BRK_C128 ? 0x7ull : 0xfull;
  bool all_fit = true;
  for (const auto& o : ops) {
    uint64_t m = line_mask;
    for (unsigned q : o.targets) m |= 1ull << q;
    for (unsigned q : o.controls) m |= 1ull << q;
    if ((unsigned)__builtin_popcountll(m) > L) all_fit = false;
  }
  if (adjoint_blocks_enabled() && circuit->num_qubits >= L && all_fit)
    return blocked_gradient(circuit, pidx, nparams, psi, work, dtype, J, h, energy, gradient, stream);
  const uint64_t size = 1ull << circuit->num_qubits;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);

  // forward pass, lam = H psi, E = <psi|lam>
  if (int s = forward(ops, psi, size, dtype, stream)) return s;
  if (int s = brk_apply_ising(work, psi, size, dtype, J, h, stream)) return s;
  double ip[2];
  if (int s = brk_inner_product(work, dtype, psi, dtype, size, ip, stream)) return s;
  *energy = ip[0];

  // reverse sweep; one partial region of ADJ_MAX_CHAIN * ADJ_BLOCKS doubles per chain
  size_t nchains = 0;
  for (size_t i = 0; i < ops.size(); ++i) nchains += chainable(ops[i]) ? 1 : 0;  // upper bound
  partials_lease lease;
  if (int s = get_partials(std::max<size_t>(1, nchains) * ADJ_MAX_CHAIN * ADJ_BLOCKS, st, lease)) return s;
  double* partials = lease.ptr;
  struct slot {
    int param;
    size_t offset;
  };
  std::vector<slot> slots;
  size_t region = 0;
  long i = (long)ops.size() - 1;
  while (i >= 0) {
    const auto& o = ops[(size_t)i];
    if (chainable(o)) {
      long j = i;
      while (j - 1 >= 0 && (i - (j - 1)) < ADJ_MAX_CHAIN && chainable(ops[(size_t)(j - 1)]) &&
             ops[(size_t)(j - 1)].targets[0] == o.targets[0])
        --j;
      adj_chain c{};
      c.target = o.targets[0];
      c.nops = (int)(i - j + 1);
      for (long k = i; k >= j; --k) {
        const int pos = (int)(i - k);
        if (!make_adj_op(ops[(size_t)k], c.ops[pos])) return fail(BRK_ERR_UNSUPPORTED, "op %ld has no 2x2 form", k);
        if (pidx[(size_t)k] >= 0) slots.push_back({pidx[(size_t)k], region + (size_t)pos * ADJ_BLOCKS});
      }
      if (int s = launch_adjoint_chain(c, psi, work, size, dtype, partials + region, stream)) return s;
      region += (size_t)ADJ_MAX_CHAIN * ADJ_BLOCKS;
      i = j - 1;
    } else if (is_sign(o)) {
      // the run of sign ops ending at i: self-inverse and commuting, so one pass per state
      long j = i;
      while (j - 1 >= 0 && is_sign(ops[(size_t)(j - 1)])) --j;
      if (int s = apply_sign_run(ops, (size_t)j, (size_t)i + 1, psi, work, size, dtype, stream)) return s;
      i = j - 1;
    } else {
      const brk_circuit::op inv = inverse_op(o);
      if (int s = apply_physical_op(inv, psi, size, dtype, stream)) return s;
      if (int s = apply_physical_op(inv, work, size, dtype, stream)) return s;
      --i;
    }
  }

  // gradient: sum each parameter's block partials in block order
  std::vector<double> host(region);
  if (region) {
    if (hipMemcpyAsync(host.data(), partials, region * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess)
      return fail(BRK_ERR_HIP, "copying gradient partials failed");
  }
  if (hipStreamSynchronize(st) != hipSuccess) return fail(BRK_ERR_HIP, "hipStreamSynchronize failed");
  for (unsigned k = 0; k < num_gradient; ++k) gradient[k] = 0.0;
  for (const auto& s : slots) {
    double t = 0.0;
    for (int b = 0; b < ADJ_BLOCKS; ++b) t += host[s.offset + (size_t)b];
    gradient[s.param] = 2.0 * t;
  }
  return BRK_OK;
}

}  // extern "C"
