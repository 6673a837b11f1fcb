// Batching form: device program layout shared by the planner (host/batch.cpp)
// and the persistent kernel (csrc/brk_batch.hip).
//
// A batch is a list of small circuits.  Each circuit is compiled into stages;
// a stage names up to BATCH_MAX_LOCAL "local" qubits.  The kernel walks the
// 2^(n-L) chunks of the stage: a chunk is the 2^L amplitudes whose non-local
// index bits are fixed; it is staged in LDS, every instruction of the stage is
// applied to it there, and it is written back.  A circuit with n <= L has a
// single stage whose one chunk is the whole state: it never leaves LDS.
#pragma once

#include <cstdint>

namespace brk {

constexpr int BATCH_THREADS = 512;      // two workgroups per CU overlap each other's LDS and ALU phases
constexpr int BATCH_WG_PER_CU = 2;
constexpr int BATCH_MAX_LOCAL = 13;     // 2^13 complex64 = 64 KiB of LDS per chunk
constexpr int BATCH_PASSENGERS = 4;     // qubits 0..3 are always local: chunk rows are 128-byte runs
constexpr int BATCH_MAX_QUBITS = 24;    // per circuit (one 2^n slot per workgroup for multi-stage circuits)
constexpr int BATCH_MAX_GROUP = 4;      // local positions per gate instruction (16 amplitudes per thread)
constexpr int BATCH_MAX_GROUP_OPS = 8;  // 2x2 gates per gate instruction
constexpr int BATCH_CHUNK = 1 << BATCH_MAX_LOCAL;
constexpr int BATCH_CHUNK_PADDED = BATCH_CHUNK;  // LDS slots of a chunk (swizzled, not padded)

// Instruction stream (32-bit words):
//   gate group : w0 = BI_GATES | npos << 8 | npre << 16 | npost << 24
//                w1 = local positions (bytes, ascending), w2 = the qubits at those positions (bytes),
//                w3 = the thread-independent parts of the sign terms' parities: bit e (pre terms) and
//                bit 16 + e (post terms) is the parity of the sum over the terms (d, A) of
//                popcount(xe & (xe >> d) & A), xe the index bits of group element e (bit j of e at
//                the qubit of position j),
//                then npre sign terms, npost sign terms (2 words each), then one 2x2 matrix per position,
//                8 words each: m00.re, m00.im, m01.re, m01.im, m10.re, m10.im, m11.re, m11.im (FP32 bit
//                patterns; a0' = m00 a0 + m01 a1, a1' = m10 a0 + m11 a1 on the pairs of that position).
//                The gates of a group on one position are multiplied into that matrix (in FP64 on the
//                host): gates on different positions commute.  Pre-signs are applied to the group's
//                amplitudes before its gates, post-signs after them; attached terms have d < 32.
//   real group : w0 = BI_GATES_REAL or BI_GATES_REAL_PRE, then w1, w2, w3 and the sign terms as for a gate
//                group, then one real rotation per position (2 words: t = tan(theta/2), s = sin(theta),
//                FP32, |theta| <= pi/2), then the post-phase table (2^npos complex64, 2 words each), and
//                for BI_GATES_REAL_PRE the pre-phase table (2^npos complex64).  The product matrix U of
//                position j is written as U = diag(p0, p1) R diag(1, q) with R the rotation
//                [[c, -s], [s, c]] (every unitary has this form: within a row the phase difference of
//                the two entries is the same for both rows, mod pi; signs move between R and p), so the
//                group applies, to its element e (bit j of e: the value at position j):
//                a[e] *= prod_j q_j^(e_j) (pre table), R_j on position j's pairs (a0, a1) as three
//                shears a0 -= t a1, a1 += s a0, a0 -= t a1, then a[e] *= prod_j p_j[e_j] (post table).
//                The diagonal factors of different positions commute with each other and with the
//                sign terms, and the tables are the same for every thread of the pass.
//   signs      : w0 = BI_SIGN | nterms << 16, then nterms sign terms (a pass of its own).
// A sign term (d, A) negates an amplitude with index x when its bit count is odd:
//   d = 0      : x & A                  (Z on each qubit of A)
//   0 < d < 32 : x & (x >> d) & A       (CZ(a, a + d) for each bit a of A)
//   d = 32     : (x & A) == A           (Z on the all-ones subspace of A: CnZ)
// Terms are merged by XOR, so repeated masks cancel exactly as the gates do.
enum : uint32_t { BI_GATES = 1, BI_SIGN = 2, BI_GATES_REAL = 3, BI_GATES_REAL_PRE = 4 };
constexpr int BATCH_GATE_WORDS = 8;
constexpr int BATCH_REAL_WORDS = 2;
constexpr int BATCH_MAX_ATTACHED_TERMS = 15;  // per pre/post list of a gate group
constexpr uint32_t BATCH_TERM_ALL = 32;

// stage flags
enum : uint32_t {
  BS_INIT = 1,   // the chunk starts as |0...0> (first stage); otherwise it is loaded from the slot
  BS_STORE = 2,  // write the chunk back to the slot (a later stage reads it)
  BS_ZZ = 4,     // add sum_i J_i <Z_i Z_i+1> over the chunk (FP64)
};

struct batch_stage {
  uint8_t local[16];       // qubit at local position p (ascending), p < L
  uint32_t L;              // local qubits
  uint32_t flags;
  uint32_t nonlocal_mask;  // global index bits enumerated by the chunk counter
  uint32_t measure_x;      // local positions p whose h_(local[p]) <X> is added (FP64)
  uint32_t first_word;     // instruction stream offset
  uint32_t num_instr;
  uint32_t pad[2];
};

struct batch_circuit {
  uint32_t n;
  uint32_t num_stages;
  uint32_t first_stage;
  uint32_t ham_offset;     // J[0..n-2] then h[0..n-1] in the Hamiltonian array
};

struct batch_launch {
  const batch_circuit* circuits;
  const batch_stage* stages;
  const uint32_t* words;
  const double* ham;
  const uint32_t* order;   // circuit indices, most expensive first
  uint32_t num_circuits;
  uint32_t* counter;       // work queue head (zeroed by the reset kernel before each launch)
  void* slots;             // complex64 [grid][slot_elems]; null when no circuit needs more than one stage
  uint64_t slot_elems;
  double* energies;        // [num_circuits]
  uint64_t* t_start;       // wall-clock ticks when the circuit was taken from the queue
  uint64_t* t_end;         // ... and when its energy was written
};

// LDS bytes the kernel needs (chunk + index tables + ZZ energy tables + queue slot).
size_t batch_lds_bytes();
// Lets the persistent kernel use more than 64 KiB of dynamic LDS (call once per device, outside capture).
int configure_batch();
// Enqueues the queue reset and the persistent kernel on `stream` (capturable: no memset nodes).
int launch_batch(const batch_launch& L, unsigned grid, void* stream);

}  // namespace brk
