// bra front end (nompi mode): quantum-assembler text -> statements.
//
// Follows bra::interpreter::invoke (bra/src/interpreter.cpp:371-872):
// everything after '!' is a comment, lines are trimmed and split on white
// space, the first column is upper-cased, and controlled forms are a run of
// 'C's or 'C' plus a count in front of a base mnemonic
// (interpreter::interpret_controlled_gates, interpreter.cpp:3341-3369); the
// control qubits come first among the operands.  Qubits of Z, S, T, U1, R and
// U are "control" qubits in bra: the phase applies where all of them are 1.
// A Pauli string over several targets (XX, Y3, ZZ, EXX, ...) acts on all of
// them at once.  Gates between two directives are collected into one
// brk_circuit, in program order.
#include "circuit.hpp"

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

namespace brk {
int fail(int status, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
}

struct brk_qcx {
  // one gate of the program: the ops it became in its statement's circuit and the
  // qubits bra draws depolarizing noise for after it, in bra's order
  struct gate_record {
    unsigned first_op = 0, num_ops = 0;
    std::vector<unsigned> noise;
  };
  struct statement {
    int kind = BRK_QCX_GATES;
    brk_circuit* gates = nullptr;
    int precision = 3;
    std::vector<uint64_t> indices;
    std::vector<gate_record> records;  // kind BRK_QCX_GATES
  };
  unsigned num_qubits = 0;
  uint64_t initial_state = 0;
  std::vector<statement> statements;
  // DEPOLARIZING CHANNEL (bra/src/interpreter.cpp:431-445, :1890-1975)
  bool depolarizing = false;
  double px = 0.0, py = 0.0, pz = 0.0;
  long long depolarizing_seed = -1;
  ~brk_qcx() {
    for (auto& s : statements)
      if (s.gates) brk_circuit_destroy(s.gates);
  }
};

namespace {

constexpr double kPi = 3.14159265358979323846264338327950288;

struct parse_error {
  int status;
  std::string message;
};

std::string upper(std::string s) {
  for (char& c : s) c = (char)std::toupper((unsigned char)c);
  return s;
}

bool all_of(const std::string& s, size_t from, char c) {
  if (from >= s.size()) return false;
  for (size_t i = from; i < s.size(); ++i)
    if (s[i] != c) return false;
  return true;
}

bool digits_from(const std::string& s, size_t from, size_t to) {
  if (from >= to) return false;
  for (size_t i = from; i < to; ++i)
    if (!std::isdigit((unsigned char)s[i])) return false;
  return true;
}

class parser {
 public:
  explicit parser(brk_qcx& p) : prog_(p) {}

  void line(size_t number, const std::string& raw) {
    if (stopped_) return;  // bra stops reading after GENERATE EVENTS and EXIT (interpreter.cpp:579-590, 764-784)
    line_ = number;
    std::string text = raw.substr(0, raw.find('!'));
    std::istringstream in(text);
    cols_.clear();
    for (std::string w; in >> w;) cols_.push_back(w);
    if (cols_.empty()) return;
    const std::string m = upper(cols_[0]);
    if (m == "QUBITS") return qubits();
    if (m == "CIRCUITS") {
      if (cols_.size() != 2 || integer(1) != 1) unsupported("more than one circuit (CIRCUITS n > 1)");
      return;
    }
    if (m == "INITIAL") {
      if (cols_.size() != 3 || upper(cols_[1]) != "STATE") error("expected INITIAL STATE <value>");
      prog_.initial_state = (uint64_t)integer(2);
      return;
    }
    if (m == "BIT") return bit_assignment();
    if (m == "DEPOLARIZING") return depolarizing_statement();
    if (m == "DO") return do_statement();
    if (m == "BEGIN") return begin_statement();
    if (m == "END") return end_statement();
    need_qubits();
    if (m == "M" || m == "CLEAR" || m == "SET") return single_qubit_statement(m);
    if (m == "GENERATE") return generate_statement();
    if (m == "SHORBOX") return shor_box_statement();
    if (m == "EXIT") {
      expect_columns(1);
      not_in_fusion("EXIT");
      directive(BRK_QCX_EXIT);
      stopped_ = true;
      return;
    }
    gate(m);
  }

  void finish() { flush(); }

 private:
  brk_qcx& prog_;
  size_t line_ = 0;
  std::vector<std::string> cols_;
  brk_circuit* current_ = nullptr;
  std::vector<brk_qcx::gate_record> records_;  // of current_
  std::vector<unsigned> noise_;                // the gate being parsed: its noise qubits
  bool in_fusion_ = false, stopped_ = false;

  [[noreturn]] void error(const std::string& what) const {
    throw parse_error{BRK_ERR_INVALID_ARGUMENT, "line " + std::to_string(line_) + ": " + what};
  }
  [[noreturn]] void unsupported(const std::string& what) const {
    throw parse_error{BRK_ERR_UNSUPPORTED, "line " + std::to_string(line_) + ": " + what + " is not supported"};
  }
  void need_qubits() const {
    if (!prog_.num_qubits) error("QUBITS must come before the first gate");
  }

  long long integer(size_t i) const {
    if (i >= cols_.size()) error("missing operand");
    errno = 0;
    char* end = nullptr;
    const long long v = std::strtoll(cols_[i].c_str(), &end, 10);
    if (errno || !end || *end) error("'" + cols_[i] + "' is not an integer");
    return v;
  }
  double real(size_t i) const {
    if (i >= cols_.size()) error("missing operand");
    errno = 0;
    char* end = nullptr;
    const double v = std::strtod(cols_[i].c_str(), &end);
    if (errno == ERANGE || !end || *end) error("'" + cols_[i] + "' is not a number (classical variables are not supported)");
    return v;
  }
  unsigned qubit(size_t i) const {
    const long long q = integer(i);
    if (q < 0 || (unsigned long long)q >= prog_.num_qubits)
      error("qubit " + cols_[i] + " out of range for " + std::to_string(prog_.num_qubits) + " qubits");
    return (unsigned)q;
  }

  void qubits() {
    if (cols_.size() != 2) error("expected QUBITS <n>");
    const long long n = integer(1);
    if (n < 1 || n > 40) error("QUBITS must be 1..40");
    if (prog_.num_qubits) error("QUBITS given twice");
    prog_.num_qubits = (unsigned)n;
  }

  void bit_assignment() {
    // nompi mode ignores the initial permutation (only the MPI states use it);
    // it is still checked to be a permutation of the qubits
    if (cols_.size() < 2 || upper(cols_[1]) != "ASSIGNMENT") unsupported("BIT " + (cols_.size() > 1 ? cols_[1] : ""));
    need_qubits();
    if (cols_.size() != 2 + (size_t)prog_.num_qubits) error("BIT ASSIGNMENT needs one entry per qubit");
    std::vector<int> seen(prog_.num_qubits, 0);
    for (size_t i = 2; i < cols_.size(); ++i) {
      const unsigned q = qubit(i);
      if (seen[q]++) error("BIT ASSIGNMENT repeats qubit " + cols_[i]);
    }
  }

  void flush() {
    if (current_) {
      brk_qcx::statement s;
      s.kind = BRK_QCX_GATES;
      s.gates = current_;
      s.records = std::move(records_);
      prog_.statements.push_back(std::move(s));
      current_ = nullptr;
      records_.clear();
    }
  }

  // DEPOLARIZING CHANNEL P_X=px, P_Y=py, P_Z=pz, SEED=seed: keys in any order and case, each
  // at most once, '=' and ',' with or without spaces around them (interpreter.hpp:518-600);
  // probabilities in [0, 1] with px + py + pz <= 1; missing ones are 0, a missing seed -1.
  void depolarizing_statement() {
    if (cols_.size() <= 2 || upper(cols_[1]) != "CHANNEL") error("expected DEPOLARIZING CHANNEL P_X=..,P_Y=..,P_Z=..,SEED=..");
    std::string text;
    for (size_t i = 2; i < cols_.size(); ++i) text += cols_[i];
    double p[3] = {0.0, 0.0, 0.0};
    long long seed = -1;
    bool seen[4] = {false, false, false, false};
    size_t at = 0;
    while (true) {
      const size_t comma = text.find(',', at);
      const std::string item = text.substr(at, comma == std::string::npos ? std::string::npos : comma - at);
      const size_t eq = item.find('=');
      if (eq == std::string::npos || eq + 1 >= item.size()) error("DEPOLARIZING CHANNEL: expected KEY=value, got '" + item + "'");
      const std::string key = upper(item.substr(0, eq)), value = item.substr(eq + 1);
      const int slot = key == "P_X" ? 0 : key == "P_Y" ? 1 : key == "P_Z" ? 2 : key == "SEED" ? 3 : -1;
      if (slot < 0) error("DEPOLARIZING CHANNEL: unknown key '" + key + "'");
      if (seen[slot]) error("DEPOLARIZING CHANNEL: " + key + " given twice");
      seen[slot] = true;
      char* end = nullptr;
      errno = 0;
      if (slot == 3) {
        seed = std::strtoll(value.c_str(), &end, 10);
        if (errno || *end || seed < -0x80000000ll || seed > 0x7fffffffll) error("DEPOLARIZING CHANNEL: bad SEED '" + value + "'");
      } else {
        p[slot] = std::strtod(value.c_str(), &end);
        if (errno || *end || !(p[slot] >= 0.0 && p[slot] <= 1.0))
          error("DEPOLARIZING CHANNEL: " + key + " must be a probability, got '" + value + "'");
      }
      if (comma == std::string::npos) break;
      at = comma + 1;
    }
    if (p[0] + p[1] + p[2] > 1.0) error("DEPOLARIZING CHANNEL: P_X + P_Y + P_Z exceeds 1");
    prog_.depolarizing = true;
    prog_.px = p[0];
    prog_.py = p[1];
    prog_.pz = p[2];
    prog_.depolarizing_seed = seed;
  }
  void directive(int kind, int precision = 3, std::vector<uint64_t> indices = {}) {
    flush();
    brk_qcx::statement s;
    s.kind = kind;
    s.precision = precision;
    s.indices = std::move(indices);
    prog_.statements.push_back(std::move(s));
  }

  void do_statement() {
    if (cols_.size() < 2) error("DO needs a statement");
    const std::string what = upper(cols_[1]);
    need_qubits();
    if (what == "MEASUREMENT") {
      if (cols_.size() > 3) error("DO MEASUREMENT takes at most one precision");
      directive(BRK_QCX_MEASUREMENT, cols_.size() == 3 ? (int)integer(2) : 3);
    } else if (what == "AMPLITUDES") {
      std::vector<uint64_t> idx;
      for (size_t i = 2; i < cols_.size(); ++i) {
        const long long v = integer(i);
        if (v < 0 || (unsigned long long)v >= (1ull << prog_.num_qubits)) error("amplitude index out of range");
        idx.push_back((uint64_t)v);
      }
      directive(BRK_QCX_AMPLITUDES, 3, std::move(idx));
    } else {
      unsupported("DO " + what);
    }
  }
  // bra::state rejects these between BEGIN FUSION and END FUSION (bra/src/state.cpp:1973-1983, 2064-2067, 2219-2238)
  void not_in_fusion(const std::string& what) const {
    if (in_fusion_) unsupported(what + " inside BEGIN FUSION / END FUSION");
  }
  // M, CLEAR, SET <qubit> (interpreter.cpp:3199-3206, 3335-3339)
  void single_qubit_statement(const std::string& m) {
    expect_columns(2);
    not_in_fusion(m);
    const long long q = integer(1);
    if (q < 0 || q >= (long long)prog_.num_qubits) error("qubit " + cols_[1] + " out of range");
    directive(m == "M" ? BRK_QCX_PROJECTIVE_MEASUREMENT : m == "CLEAR" ? BRK_QCX_CLEAR : BRK_QCX_SET, 3,
              {(uint64_t)q});
  }
  // SHORBOX <num_exponent_qubits> <divisor> <base> (interpreter.cpp:3208-3216, 1861-1873)
  void shor_box_statement() {
    expect_columns(4);
    not_in_fusion("SHORBOX");
    const long long m = integer(1), divisor = integer(2), base = integer(3);
    if (m < 0 || m > (long long)prog_.num_qubits) error("number of exponent qubits out of range");
    if (divisor <= 0 || base < 0) error("SHORBOX needs a positive divisor and a non-negative base");
    directive(BRK_QCX_SHOR_BOX, 3, {(uint64_t)m, (uint64_t)divisor, (uint64_t)base});
  }
  // GENERATE EVENTS <num_events> <seed> (interpreter.cpp:764-784, 1875-1887)
  void generate_statement() {
    if (cols_.size() != 4 || upper(cols_[1]) != "EVENTS") error("expected GENERATE EVENTS <num_events> <seed>");
    not_in_fusion("GENERATE EVENTS");
    const long long n = integer(2), seed = integer(3);
    if (n < 0 || n > 0x7fffffffll) error("number of events out of range");
    if (seed < -0x80000000ll || seed > 0x7fffffffll) error("seed out of range");
    directive(BRK_QCX_GENERATE_EVENTS, (int)seed, {(uint64_t)n});
    stopped_ = true;
  }
  void begin_statement() {
    if (cols_.size() < 2) error("BEGIN needs a statement");
    const std::string what = upper(cols_[1]);
    need_qubits();
    if (what == "MEASUREMENT") directive(BRK_QCX_MEASUREMENT, 3);
    else if (what == "FUSION") {
      if (in_fusion_) error("BEGIN FUSION inside BEGIN FUSION");
      in_fusion_ = true;
      directive(BRK_QCX_BEGIN_FUSION);
    }
    else unsupported("BEGIN " + what);
  }
  void end_statement() {
    if (cols_.size() < 2) error("END needs a statement");
    const std::string what = upper(cols_[1]);
    if (what == "FUSION") {
      if (!in_fusion_) error("END FUSION without BEGIN FUSION");
      in_fusion_ = false;
      directive(BRK_QCX_END_FUSION);
    }
    else unsupported("END " + what);
  }

  void add(int op, const std::vector<unsigned>& targets, const std::vector<unsigned>& controls,
           const std::vector<double>& params = {}) {
    if (!current_ && brk_circuit_create(prog_.num_qubits, &current_) != BRK_OK) error("cannot create a circuit");
    const int s = brk_circuit_add(current_, op, targets.data(), (unsigned)targets.size(),
                                  controls.empty() ? nullptr : controls.data(), (unsigned)controls.size(),
                                  params.empty() ? nullptr : params.data(), (unsigned)params.size());
    if (s != BRK_OK) throw parse_error{s, "line " + std::to_string(line_) + ": " + brk_last_error()};
  }
  void add_matrix(unsigned t, const std::vector<unsigned>& controls, const double (&m)[8]) {
    add(BRK_OP_MATRIX, {t}, controls, std::vector<double>(m, m + 8));
  }

  void expect_columns(size_t n) const {
    if (cols_.size() != n)
      error("'" + cols_[0] + "' takes " + std::to_string(n - 1) + " operands, got " + std::to_string(cols_.size() - 1));
  }

  // Reads `nc` controls then `nt` targets starting at column 1.
  void operands(size_t nc, size_t nt, std::vector<unsigned>& controls, std::vector<unsigned>& targets) const {
    controls.clear();
    targets.clear();
    for (size_t i = 0; i < nc; ++i) controls.push_back(qubit(1 + i));
    for (size_t i = 0; i < nt; ++i) targets.push_back(qubit(1 + nc + i));
  }

  // Multiplicity of a Pauli-string mnemonic: "XXX" -> 3, "X4" -> 4, "X" -> 1; 0 if not of that form.
  static size_t multiplicity(const std::string& base, size_t from, char letter) {
    if (from >= base.size() || base[from] != letter) return 0;
    if (all_of(base, from, letter)) return base.size() - from;
    if (digits_from(base, from + 1, base.size())) return (size_t)std::atoi(base.c_str() + from + 1);
    return 0;
  }

  void gate(const std::string& m) {
    size_t nc = 0;
    std::string base = m;
    if (m.size() >= 2 && m[0] == 'C' && m != "CIRCUITS") {
      size_t cs = 0;
      while (cs < m.size() && m[cs] == 'C') ++cs;
      if (cs > 1) {
        nc = cs;
        base = m.substr(cs);
      } else {
        size_t d = 1;
        while (d < m.size() && std::isdigit((unsigned char)m[d])) ++d;
        nc = d > 1 ? (size_t)std::atoi(m.c_str() + 1) : 1;
        base = m.substr(d);
      }
      if (nc == 0 || base.empty()) unsupported("mnemonic " + m);
    }
    if (!current_ && brk_circuit_create(prog_.num_qubits, &current_) != BRK_OK) error("cannot create a circuit");
    unsigned before = 0;
    brk_circuit_num_ops(current_, &before);
    noise_.clear();
    apply_base(m, base, nc);
    unsigned after = 0;
    brk_circuit_num_ops(current_, &after);
    brk_qcx::gate_record r;
    r.first_op = before;
    r.num_ops = after - before;
    r.noise = noise_;
    records_.push_back(std::move(r));
  }

  // Noise qubits of a gate, in the order bra::state draws them (bra/src/state.cpp:1000-3330,
  // apply_noises(targets..., controls...)): targets then controls, or, for the gates bra
  // reads as "control" qubits only (Z, CZ, S, T, U1, R, U and their controlled forms), all
  // operands in program order.
  void noise_targets_then_controls(const std::vector<unsigned>& t, const std::vector<unsigned>& c) {
    noise_ = t;
    noise_.insert(noise_.end(), c.begin(), c.end());
  }
  void noise_operand_order(const std::vector<unsigned>& c, const std::vector<unsigned>& t) {
    noise_ = c;
    noise_.insert(noise_.end(), t.begin(), t.end());
  }

  // sqrt(Z) strings: S followed by k >= 2 Z's or by Z<k>, optionally '+' (the adjoint).
  static size_t sqrt_z_targets(const std::string& base, bool* adjoint) {
    std::string core = base;
    *adjoint = !core.empty() && core.back() == '+';
    if (*adjoint) core.pop_back();
    if (core.size() < 2 || core[0] != 'S' || core[1] != 'Z') return 0;
    return multiplicity(core, 1, 'Z');
  }

  void apply_base(const std::string& m, const std::string& base, size_t nc) {
    std::vector<unsigned> c, t;
    const double r2 = 0.707106781186547524400844362104849039;
    // identity forms: checked, no operation
    if (base == "I" || base == "IC") {
      expect_columns(2 + nc);
      operands(nc, 1, c, t);
      if (base == "I") noise_targets_then_controls(t, c);
      else noise_operand_order(c, t);
      return;
    }
    if (size_t k = multiplicity(base, 0, 'I'); k >= 2) {
      expect_columns(1 + nc + k);
      operands(nc, k, c, t);
      noise_targets_then_controls(t, c);
      return;
    }
    if (base == "H") {
      expect_columns(2 + nc);
      operands(nc, 1, c, t);
      noise_targets_then_controls(t, c);
      return add(BRK_OP_H, t, c);
    }
    if (base == "NOT" || base == "TOFFOLI") {
      const size_t extra = base == "TOFFOLI" ? 2 : 0;
      if (base == "TOFFOLI" && nc) unsupported("controlled TOFFOLI");
      expect_columns(2 + nc + extra);
      operands(nc + extra, 1, c, t);
      noise_targets_then_controls(t, c);
      return add(BRK_OP_X, t, c);
    }
    for (const char letter : {'X', 'Y'}) {
      if (size_t k = multiplicity(base, 0, letter); k >= 1) {
        expect_columns(1 + nc + k);
        operands(nc, k, c, t);
        noise_targets_then_controls(t, c);
        return add(letter == 'X' ? BRK_OP_X : BRK_OP_Y, t, c);
      }
    }
    if (base == "Z") {
      // Z on a control qubit; with n >= 1 control prefixes bra reads n + 1 control
      // qubits (CZ c1 c2, CCZ c1 c2 c3: interpreter::add_cz, interpreter.cpp:3762-3779)
      const size_t q = nc == 0 ? 1 : nc + 1;
      expect_columns(1 + q);
      operands(0, q, c, t);
      noise_operand_order({}, t);
      return add(BRK_OP_Z, t, {});
    }
    if (size_t k = multiplicity(base, 0, 'Z'); k >= 2) {
      expect_columns(1 + nc + k);
      operands(nc, k, c, t);
      noise_targets_then_controls(t, c);
      for (unsigned q : t) add(BRK_OP_Z, {q}, c);  // Z..Z = product of single-qubit Z's
      return;
    }
    // sqrt(Z) strings SZZ, SZZZ.., SZ<n> and their adjoints (interpreter.cpp:808-819,
    // 2553-2650; controlled: 3494-3505, 4754-4870; docs/bra.md:133): i (or -i) per target
    // that is 1, where every control is 1 -- a product of per-target phases, each exactly
    // the coefficient i (ket::utility::imaginary_unit, sqrt_pauli_z.hpp:462-700).  The
    // reference's SZ<n> readers check their digits from the 'Z' on and so reject every such
    // mnemonic; this parser takes them as docs/bra.md:133 describes them (C3SZ3 c1 c2 c3 t1 t2 t3).
    {
      bool adjoint = false;
      if (size_t k = sqrt_z_targets(base, &adjoint); k >= 2) {
        expect_columns(1 + nc + k);
        operands(nc, k, c, t);
        noise_targets_then_controls(t, c);
        for (unsigned q : t) add(BRK_OP_PHASE_COEFF, {q}, c, {0.0, adjoint ? -1.0 : 1.0});
        return;
      }
    }
    if (base == "SWAP") {
      expect_columns(3 + nc);
      operands(nc, 2, c, t);
      noise_targets_then_controls(t, c);
      return add(BRK_OP_SWAP, t, c);
    }
    // ESWAP t1 t2 theta = exp(i theta SWAP); ESWAP+ its adjoint (interpreter.cpp:3169-3187)
    if (base == "ESWAP" || base == "ESWAP+") {
      expect_columns(4 + nc);
      operands(nc, 2, c, t);
      noise_targets_then_controls(t, c);
      const double theta = real(3 + nc);
      return add(BRK_OP_ESWAP, t, c, {base == "ESWAP" ? theta : -theta});
    }
    // sqrt(Z) and its adjoint on control qubits: S, SZ, SZ1 (+) -- exactly i / -i
    // (ket::gate::sqrt_pauli_z, sqrt_pauli_z.hpp:35-420)
    {
      bool adjoint = false;
      if (base == "S" || base == "S+" || sqrt_z_targets(base, &adjoint) == 1) {
        adjoint = base.back() == '+';
        expect_columns(2 + nc);
        operands(nc, 1, c, t);
        noise_operand_order(c, t);
        return add(BRK_OP_PHASE_COEFF, t, c, {0.0, adjoint ? -1.0 : 1.0});
      }
    }
    // T = phase_shift(3): exp(2 pi i / 2^3) on control qubits (bra/src/t_gate.cpp, state.cpp:1405-1421)
    if (base == "T" || base == "T+") {
      expect_columns(2 + nc);
      operands(nc, 1, c, t);
      noise_operand_order(c, t);
      return add(BRK_OP_U1, t, c, {base == "T" ? kPi / 4 : -kPi / 4});
    }
    if (base == "U1" || base == "U1+") {
      expect_columns(3 + nc);
      operands(nc, 1, c, t);
      noise_operand_order(c, t);
      const double lambda = real(2 + nc);
      return add(BRK_OP_U1, t, c, {base == "U1" ? lambda : -lambda});
    }
    if (base == "R" || base == "R+" || base == "U" || base == "U+") {
      // phase 2 pi / 2^k (k < 0: the adjoint), bra/src/state.cpp:2650-2669
      const size_t q = (base[0] == 'U' ? 2 : 1) + nc;
      expect_columns(2 + q);
      operands(0, q, c, t);
      const long long k = integer(1 + q);
      if (k < -62 || k > 62) error("phase exponent out of range");
      noise_operand_order({}, t);
      const unsigned long long kk = (unsigned long long)(k < 0 ? -k : k);
      double phase = 2.0 * kPi / (double)(1ull << kk);
      if ((k < 0) != (base.back() == '+')) phase = -phase;
      return add(BRK_OP_U1, {t[0]}, std::vector<unsigned>(t.begin() + 1, t.end()), {phase});
    }
    if (base == "U2" || base == "U2+") {
      expect_columns(4 + nc);
      operands(nc, 1, c, t);
      noise_targets_then_controls(t, c);
      const double phi = real(2 + nc), lambda = real(3 + nc);
      if (base == "U2") return add(BRK_OP_U2, t, c, {phi, lambda});
      return add(BRK_OP_U3, t, c, {-kPi / 2, -lambda, -phi});  // U2(phi, lambda)^dagger
    }
    if (base == "U3" || base == "U3+") {
      expect_columns(5 + nc);
      operands(nc, 1, c, t);
      noise_targets_then_controls(t, c);
      const double theta = real(2 + nc), phi = real(3 + nc), lambda = real(4 + nc);
      if (base == "U3") return add(BRK_OP_U3, t, c, {theta, phi, lambda});
      return add(BRK_OP_U3, t, c, {-theta, -lambda, -phi});
    }
    // fixed one-qubit matrices (docs/bra.md:109-188), row-major (re, im)
    struct fixed_gate {
      const char* name;
      double m[8];
    };
    static const fixed_gate fixed[] = {
        {"+X", {r2, 0, 0, r2, 0, r2, r2, 0}},      {"-X", {r2, 0, 0, -r2, 0, -r2, r2, 0}},
        {"+Y", {r2, 0, r2, 0, -r2, 0, r2, 0}},     {"-Y", {r2, 0, -r2, 0, r2, 0, r2, 0}},
        {"SX", {0.5, 0.5, 0.5, -0.5, 0.5, -0.5, 0.5, 0.5}},
        {"SX+", {0.5, -0.5, 0.5, 0.5, 0.5, 0.5, 0.5, -0.5}},
        {"SY", {0.5, 0.5, -0.5, -0.5, 0.5, 0.5, 0.5, 0.5}},
        {"SY+", {0.5, -0.5, 0.5, -0.5, -0.5, 0.5, 0.5, -0.5}}};
    for (const auto& g : fixed)
      if (base == g.name) {
        expect_columns(2 + nc);
        operands(nc, 1, c, t);
        noise_targets_then_controls(t, c);
        return add_matrix(t[0], c, g.m);
      }
    // exponentials of Pauli strings: EX t theta, EXX t1 t2 theta, EX3 ..., and their adjoints
    if (base.size() >= 2 && base[0] == 'E' && base != "ESWAP" && base != "ESWAP+") {
      const bool adjoint = base.back() == '+';
      const std::string core = adjoint ? base.substr(0, base.size() - 1) : base;
      for (const char letter : {'X', 'Y', 'Z'}) {
        const size_t k = multiplicity(core, 1, letter);
        if (k >= 1) {
          expect_columns(2 + nc + k);
          operands(nc, k, c, t);
          noise_targets_then_controls(t, c);
          const double theta = real(1 + nc + k);
          const int op = letter == 'X' ? BRK_OP_EX : letter == 'Y' ? BRK_OP_EY : BRK_OP_EZ;
          return add(op, t, c, {adjoint ? -theta : theta});
        }
      }
    }
    unsupported("mnemonic " + m);
  }
};

}  // namespace

using brk::fail;

extern "C" {

int brk_qcx_parse(const char* text, brk_qcx** out) {
  if (!text || !out) return fail(BRK_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  auto prog = std::make_unique<brk_qcx>();
  parser p(*prog);
  try {
    std::istringstream in(text);
    std::string raw;
    size_t number = 0;
    while (std::getline(in, raw)) p.line(++number, raw);
    p.finish();
  } catch (const parse_error& e) {
    return fail(e.status, "%s", e.message.c_str());
  }
  if (!prog->num_qubits) return fail(BRK_ERR_INVALID_ARGUMENT, "the program has no QUBITS statement");
  if (prog->initial_state >> prog->num_qubits) return fail(BRK_ERR_INVALID_ARGUMENT, "INITIAL STATE out of range");
  *out = prog.release();
  return BRK_OK;
}

int brk_qcx_destroy(brk_qcx* program) {
  delete program;
  return BRK_OK;
}

int brk_qcx_num_qubits(const brk_qcx* program, unsigned* num_qubits) {
  if (!program || !num_qubits) return fail(BRK_ERR_INVALID_ARGUMENT, "null argument");
  *num_qubits = program->num_qubits;
  return BRK_OK;
}

int brk_qcx_initial_state(const brk_qcx* program, uint64_t* basis_index) {
  if (!program || !basis_index) return fail(BRK_ERR_INVALID_ARGUMENT, "null argument");
  *basis_index = program->initial_state;
  return BRK_OK;
}

int brk_qcx_num_statements(const brk_qcx* program, unsigned* num_statements) {
  if (!program || !num_statements) return fail(BRK_ERR_INVALID_ARGUMENT, "null argument");
  *num_statements = (unsigned)program->statements.size();
  return BRK_OK;
}

int brk_qcx_statement(const brk_qcx* program, unsigned i, int* kind, const brk_circuit** gates, int* precision,
                      unsigned* num_indices, const uint64_t** indices) {
  if (!program) return fail(BRK_ERR_INVALID_ARGUMENT, "program is null");
  if (i >= program->statements.size()) return fail(BRK_ERR_INVALID_ARGUMENT, "statement %u out of range", i);
  const auto& s = program->statements[i];
  if (kind) *kind = s.kind;
  if (gates) *gates = s.gates;
  if (precision) *precision = s.precision;
  if (num_indices) *num_indices = (unsigned)s.indices.size();
  if (indices) *indices = s.indices.empty() ? nullptr : s.indices.data();
  return BRK_OK;
}

int brk_qcx_depolarizing(const brk_qcx* program, int* enabled, double probabilities[3], int64_t* seed) {
  if (!program) return fail(BRK_ERR_INVALID_ARGUMENT, "program is null");
  if (enabled) *enabled = program->depolarizing ? 1 : 0;
  if (probabilities) {
    probabilities[0] = program->px;
    probabilities[1] = program->py;
    probabilities[2] = program->pz;
  }
  if (seed) *seed = program->depolarizing_seed;
  return BRK_OK;
}

int brk_qcx_num_gates(const brk_qcx* program, unsigned statement, unsigned* num_gates) {
  if (!program || !num_gates) return fail(BRK_ERR_INVALID_ARGUMENT, "null argument");
  if (statement >= program->statements.size())
    return fail(BRK_ERR_INVALID_ARGUMENT, "statement %u out of range", statement);
  *num_gates = (unsigned)program->statements[statement].records.size();
  return BRK_OK;
}

int brk_qcx_gate(const brk_qcx* program, unsigned statement, unsigned gate, unsigned* first_op, unsigned* num_ops,
                 unsigned* noise_qubits, unsigned* num_noise_qubits) {
  if (!program) return fail(BRK_ERR_INVALID_ARGUMENT, "program is null");
  if (statement >= program->statements.size())
    return fail(BRK_ERR_INVALID_ARGUMENT, "statement %u out of range", statement);
  const auto& records = program->statements[statement].records;
  if (gate >= records.size()) return fail(BRK_ERR_INVALID_ARGUMENT, "gate %u out of range", gate);
  const auto& r = records[gate];
  if (first_op) *first_op = r.first_op;
  if (num_ops) *num_ops = r.num_ops;
  if (num_noise_qubits) *num_noise_qubits = (unsigned)r.noise.size();
  if (noise_qubits) std::copy(r.noise.begin(), r.noise.end(), noise_qubits);
  return BRK_OK;
}

}  // extern "C"
