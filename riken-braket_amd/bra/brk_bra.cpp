// brk-bra: bra's nompi front end on the MI355X engine.
//
//   brk-bra [--file <path>] [--threads <n>] [--seed <n>] [--amplitude-precision <p>] [--fusion-c64 <width>]
//
// Reads a quantum-assembler program (from --file or standard input, as
// bra/src/bra.cpp does), runs it on one GPU in complex<double> and prints what
// bra prints for the supported statements:
//   DO MEASUREMENT [p] / BEGIN MEASUREMENT  bra::state::measurement (bra/src/state.cpp:1984-2038)
//   DO AMPLITUDES [i..]                     bra::nompi_state::do_amplitudes (bra/src/nompi_state.cpp:1108-1133)
//   GENERATE EVENTS n seed                  bra::state::generate_events (bra/src/state.cpp:2064-2092)
//   EXIT                                    bra::state::exit (bra/src/state.cpp:2094-2120)
// each framed by the "... finished: <seconds since start> (<seconds since last>)"
// lines; M, CLEAR and SET print nothing.  Runs of gates (inside BEGIN/END
// FUSION or not) are applied in LDS-resident blocks, which leaves every
// amplitude bit-identical to gate-by-gate application.  A DEPOLARIZING CHANNEL
// statement inserts X, Y or Z after every gate (and CLEAR / SET) with bra's draws
// (bra::state::apply_noise, bra/src/state.cpp:3331-3362).  --seed seeds the state's
// std::mt19937_64 (default 1, bra/src/bra.cpp:73); --threads is accepted for
// command-line compatibility.  Two options bra does not have: --amplitude-precision
// sets the significant digits DO AMPLITUDES prints (default: the stream's 6, as bra);
// --fusion-c64 <width> (4, 5 or 6) runs the gates between BEGIN FUSION and END FUSION
// through the precision form (dense width-qubit blocks on MFMA in complex64) instead of
// the exact gate blocks.
#include <chrono>
#include <cstdlib>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <iterator>
#include <sstream>
#include <string>
#include <vector>

#include "bra_hip_state.hpp"

namespace {

using clock_type = std::chrono::steady_clock;

std::string bits_string(std::uint64_t value, unsigned n) {  // bra/include/bra/state.hpp:101-117
  std::string s;
  for (unsigned b = n; b > 0; --b) s.push_back(((value >> (b - 1)) & 1u) ? '1' : '0');
  return s;
}

// fmt's "{:^W}" centring: the left pad is half the padding, rounded down
std::string centre(const std::string& s, int width) {
  const int pad = width - (int)s.size();
  if (pad <= 0) return s;
  return std::string(pad / 2, ' ') + s + std::string(pad - pad / 2, ' ');
}

std::string fixed_with_space_sign(double v, int precision) {  // fmt "{: .Pf}"
  char buf[64];
  std::snprintf(buf, sizeof buf, "% .*f", precision, v);
  return buf;
}

struct timer {
  clock_type::time_point start = clock_type::now(), last = start;
  std::string lap(const char* what) {
    const auto now = clock_type::now();
    std::ostringstream oss;
    oss << what << " finished: " << std::chrono::duration<double>(now - start).count() << " ("
        << std::chrono::duration<double>(now - last).count() << ")\n";
    last = now;
    return oss.str();
  }
};

void measurement(brk::bra::hip_state& st, int precision, timer& t) {
  std::cout << t.lap("Operations") << std::flush;
  const auto spins = st.do_expectation_values();
  std::ostringstream oss;
  if (precision > 0) {
    const int width = std::max(precision + 3, 8);
    oss << "Expectation values of spins:\n"
        << centre("<Qx>", width) << ' ' << centre("<Qy>", width) << ' ' << centre("<Qz>", width) << '\n';
    for (const auto& s : spins)
      oss << centre(fixed_with_space_sign(0.5 - s[0], precision), width) << ' '
          << centre(fixed_with_space_sign(0.5 - s[1], precision), width) << ' '
          << centre(fixed_with_space_sign(0.5 - s[2], precision), width) << '\n';
  } else {
    oss << "Expectation values of spins:\n<Qx> <Qy> <Qz>\n";
    if (precision < 0) oss.precision(-precision);
    for (const auto& s : spins) oss << (0.5 - s[0]) << ' ' << (0.5 - s[1]) << ' ' << (0.5 - s[2]) << '\n';
  }
  std::cout << oss.str() << t.lap("Expectation values") << std::flush;
}

void amplitudes(brk::bra::hip_state& st, const std::vector<std::uint64_t>& indices, int precision, timer& t) {
  std::cout << t.lap("Operations") << std::flush;
  const auto a = st.amplitudes(indices);
  std::ostringstream oss;
  if (precision > 0) oss.precision(precision);
  for (size_t i = 0; i < a.size(); ++i) {
    const std::uint64_t index = indices.empty() ? i : indices[i];
    oss << bits_string(index, st.num_qubits()) << " => " << a[i].real() << " + " << a[i].imag() << " i\n";
  }
  std::cout << oss.str() << t.lap("Amplitudes") << std::flush;
}

// "Events:" then "<index> <bits>" per event (bra/src/state.cpp:2078-2081)
void generate_events(brk::bra::hip_state& st, int num_events, int seed, timer& t) {
  std::cout << t.lap("Operations") << std::flush;
  const auto events = st.do_generate_events(num_events, seed);
  std::ostringstream oss;
  oss << "Events:\n";
  for (size_t i = 0; i < events.size(); ++i) oss << i << ' ' << bits_string(events[i], st.num_qubits()) << '\n';
  std::cout << oss.str() << t.lap("Events") << std::flush;
}

void exit_statement(brk::bra::hip_state& st, timer& t) {
  std::cout << t.lap("Operations") << std::flush;
  const std::uint64_t value = st.do_measure();
  std::cout << "Measurement result: " << value << '\n' << t.lap("Measurement") << std::flush;
}

int usage() {
  std::cerr << "usage: brk-bra [--file <path>] [--threads <n>] [--seed <n>] [--amplitude-precision <p>] "
               "[--fusion-c64 <width>]\n";
  return 2;
}

}  // namespace

int main(int argc, char** argv) {
  std::string file;
  brk::bra::hip_state::seed_type seed = 1;
  int amplitude_precision = 0;
  unsigned fusion_width = 0;  // 0: exact gate blocks inside BEGIN / END FUSION
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if ((a == "--file" || a == "-f") && i + 1 < argc) file = argv[++i];
    else if (a == "--threads" && i + 1 < argc) ++i;
    else if (a == "--fusion-c64" && i + 1 < argc) {
      fusion_width = (unsigned)std::atoi(argv[++i]);
      if (fusion_width < 4 || fusion_width > 6) return usage();
    } else if (a == "--amplitude-precision" && i + 1 < argc) {
      amplitude_precision = std::atoi(argv[++i]);
      if (amplitude_precision < 1 || amplitude_precision > 17) return usage();
    }
    else if (a == "--seed" && i + 1 < argc) {
      try {
        size_t used = 0;
        seed = std::stoull(argv[++i], &used);
        if (argv[i][used] != '\0') return usage();
      } catch (const std::exception&) {
        return usage();
      }
    } else return usage();
  }
  std::string text;
  if (file.empty()) {
    text.assign(std::istreambuf_iterator<char>(std::cin), std::istreambuf_iterator<char>());
  } else {
    std::ifstream in(file);
    if (!in) {
      std::cerr << "ERROR: cannot open an input file " << file << '\n';
      return 1;
    }
    text.assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
  }

  brk_qcx* prog = nullptr;
  if (brk_qcx_parse(text.c_str(), &prog) != BRK_OK) {
    std::cerr << "Error: " << brk_last_error() << '\n';
    return 1;
  }
  int rc = 0;
  try {
    unsigned n = 0, count = 0;
    std::uint64_t initial = 0;
    brk::bra::check(brk_qcx_num_qubits(prog, &n));
    brk::bra::check(brk_qcx_initial_state(prog, &initial));
    brk::bra::check(brk_qcx_num_statements(prog, &count));
    timer t;
    brk::bra::hip_state st(n, initial, seed);
    st.set_depolarizing_channel(prog);
    bool in_fusion = false;
    for (unsigned i = 0; i < count; ++i) {
      int kind = 0, precision = 3;
      const brk_circuit* gates = nullptr;
      unsigned nidx = 0;
      const std::uint64_t* idx = nullptr;
      brk::bra::check(brk_qcx_statement(prog, i, &kind, &gates, &precision, &nidx, &idx));
      switch (kind) {
        case BRK_QCX_GATES:
          if (in_fusion && fusion_width) st.apply_fused(prog, i, fusion_width);
          else st.apply_blocked(prog, i);
          break;
        case BRK_QCX_BEGIN_FUSION: in_fusion = true; break;
        case BRK_QCX_END_FUSION: in_fusion = false; break;
        case BRK_QCX_MEASUREMENT: measurement(st, precision, t); break;
        case BRK_QCX_AMPLITUDES:
          amplitudes(st, std::vector<std::uint64_t>(idx, idx + nidx), amplitude_precision, t);
          break;
        case BRK_QCX_PROJECTIVE_MEASUREMENT: st.do_projective_measurement((unsigned)idx[0]); break;
        case BRK_QCX_CLEAR:
          st.do_clear((unsigned)idx[0]);
          st.apply_noise((unsigned)idx[0]);
          break;
        case BRK_QCX_SET:
          st.do_set((unsigned)idx[0]);
          st.apply_noise((unsigned)idx[0]);
          break;
        case BRK_QCX_GENERATE_EVENTS: generate_events(st, (int)idx[0], precision, t); break;
        case BRK_QCX_EXIT: exit_statement(st, t); break;
        case BRK_QCX_SHOR_BOX: st.do_shor_box((unsigned)idx[0], idx[1], idx[2]); break;
        default: break;  // fusion brackets: every run of gates is applied in blocks already
      }
    }
    brk::bra::check(brk_stream_synchronize(nullptr));
  } catch (const std::exception& e) {
    std::cerr << "Error: " << e.what() << '\n';
    rc = 1;
  }
  brk_qcx_destroy(prog);
  return rc;
}
