"""GPU checks of brk-bra, bra's nompi front end on the engine: this repository's
test programs (tests/golden/qcx/*.qcx) and synthetic ones are run by the binary
and its printed expectation values / amplitudes (17 significant digits:
DO MEASUREMENT -17, --amplitude-precision 17) are compared with the CPU oracle
running the same parsed gates at 1e-12, with bra's random draws replayed for
measurements and depolarizing noise (and with closed forms where they exist)."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import brk
from oracle_circuits import run_ops

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "riken-braket_amd", "brk-bra")
GOLDEN = os.path.join(ROOT, "tests", "golden", "qcx")
TIMING = re.compile(r"^(Operations|Expectation values|Amplitudes|Events|Measurement) finished: [0-9.e+-]+ \([0-9.e+-]+\)$")
AMPLITUDE = re.compile(r"^([01]+) => (\S+) \+ (\S+) i$")


def run_bra(text, args=()):
    if not os.path.isfile(EXE):
        pytest.fail(f"{EXE} is not built (run __graft_entry__.build())")
    print(f"gpu bra: running brk-bra {' '.join(args)} on {text.splitlines()[0]!r} ...", flush=True)
    sys.stdout.flush()
    r = subprocess.run([EXE, *args], input=text, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout.splitlines()


def oracle_state(oracle, text):
    q = brk.Qcx(text)
    s = np.zeros(1 << q.num_qubits, dtype=np.complex128)
    s[q.initial_state] = 1
    for kind, v in q.statements:
        if kind == brk.QCX_GATES:
            run_ops(oracle, s, v)
        elif kind == brk.QCX_SHOR_BOX:
            oracle.shor_box(s, *v)
    return q, s


def spins(oracle, s, n):
    return [[0.5 - v for v in oracle.spin_expectation_value(s, k)] for k in range(n)]


def table(lines, n):
    i = lines.index("Expectation values of spins:")
    return lines[i + 1], [[float(x) for x in row.split()] for row in lines[i + 2:i + 2 + n]]


def full_table(lines, n, which=0):
    """The rows of the which-th "<Qx> <Qy> <Qz>" table (DO MEASUREMENT with a negative precision)."""
    starts = [k for k, l in enumerate(lines) if l == "<Qx> <Qy> <Qz>"]
    i = starts[which]
    return np.array([[float(x) for x in lines[i + 1 + k].split()] for k in range(n)])


def printed_amplitudes(lines, n):
    """Every "bits => re + im i" line, as {index: complex} per DO AMPLITUDES block, in order."""
    blocks, current = [], None
    for line in lines:
        m = AMPLITUDE.match(line)
        if m:
            if current is None:
                current = []
            assert len(m.group(1)) == n
            current.append((int(m.group(1), 2), complex(float(m.group(2)), float(m.group(3)))))
        elif current is not None:
            blocks.append(current)
            current = None
    if current is not None:
        blocks.append(current)
    return blocks


def test_h_wall_closed_form():
    lines = run_bra(open(os.path.join(GOLDEN, "h_wall_10.qcx")).read(), ["--amplitude-precision", "17"])
    assert TIMING.match(lines[0]) and TIMING.match(lines[-1])
    header, rows = table(lines, 10)
    # fmt "{:^8s}" columns for precision 3 (bra/src/state.cpp:1997-2006)
    assert header == " ".join(f"{h:^8}" for h in ("<Qx>", "<Qy>", "<Qz>"))
    # |+> on every qubit: <Qx> = 0, <Qy> = <Qz> = 1/2
    row = lines[lines.index(header) + 1]
    assert row.replace("-0.000", " 0.000") == " ".join(f"{v:^8}" for v in (" 0.000", " 0.500", " 0.500"))
    assert np.max(np.abs(full_table(lines, 10) - np.array([[0.0, 0.5, 0.5]] * 10))) < 1e-15
    blocks = printed_amplitudes(lines, 10)
    assert [i for i, _ in blocks[0]] == [0, 5, 1023] and len(blocks[1]) == 1024
    assert all(abs(a - 2.0 ** -5) < 1e-15 for b in blocks for _, a in b)


@pytest.mark.parametrize("name", ["toffoli_ladder_12.qcx", "draper_adder_8.qcx", "shor_box_21.qcx"])
def test_programs_vs_oracle(oracle, name):
    text = open(os.path.join(GOLDEN, name)).read()
    lines = run_bra(text, ["--amplitude-precision", "17"])
    q, s = oracle_state(oracle, text)
    n = q.num_qubits
    _, rows = table(lines, n)  # the 3-decimal table of DO MEASUREMENT's default precision
    ref = np.array(spins(oracle, s, n))
    assert np.max(np.abs(np.array(rows) - ref)) <= 0.5e-3 + 1e-12
    assert np.max(np.abs(full_table(lines, n) - ref)) < 1e-12
    (amps,) = printed_amplitudes(lines, n)
    assert [i for i, _ in amps] == list(range(1 << n))
    assert np.max(np.abs(np.array([a for _, a in amps]) - s)) < 1e-12


# --- depolarizing noise ------------------------------------------------------------------------

def replay_noisy(oracle, text, seed):
    """The oracle runs the program as bra::state does with a depolarizing channel: after every
    gate, one draw per noise qubit in bra's order (brk_qcx_gate), X / Y / Z by the thresholds
    (bra/src/state.cpp:3331-3362), from the program's own generator when SEED > 0, else from
    the state's generator (std::mt19937_64(--seed)), which M shares."""
    q = brk.Qcx(text)
    (px, py, pz), noise_seed = q.depolarizing
    s = np.zeros(1 << q.num_qubits, dtype=np.complex128)
    s[q.initial_state] = 1
    g = oracle.mt64(seed)
    gn = oracle.mt64(noise_seed) if noise_seed > 0 else g
    inserted = 0

    def noise(qubit):
        nonlocal inserted
        u = gn.canonical()
        if u < px:
            oracle.pauli_x(s, [qubit])
        elif u < px + py:
            oracle.pauli_y(s, [qubit])
        elif u < px + py + pz:
            oracle.pauli_z(s, [qubit])
        else:
            return
        inserted += 1

    amplitudes = []
    for i, (kind, v) in enumerate(q.statements):
        if kind == brk.QCX_GATES:
            for first, count, qubits in q.gate_records[i]:
                run_ops(oracle, s, v[first:first + count])
                for qubit in qubits:
                    noise(qubit)
        elif kind == brk.QCX_PROJECTIVE_MEASUREMENT:
            oracle.projective_measurement(s, v, g.canonical())
        elif kind in (brk.QCX_CLEAR, brk.QCX_SET):
            value = 0 if kind == brk.QCX_CLEAR else 1
            oracle.project(s, v, value, oracle.zero_one_probabilities(s, v)[value] ** -0.5)
            noise(v)
        elif kind == brk.QCX_AMPLITUDES:
            amplitudes.append(s.copy())
    return q, amplitudes, s, inserted


@pytest.mark.parametrize("name,args,seed", [("depolarizing_sz_9.qcx", [], 1),
                                            ("depolarizing_m_7.qcx", [], 1),
                                            ("depolarizing_m_7.qcx", ["--seed", "99"], 99)])
def test_depolarizing_programs_vs_oracle_replay(oracle, name, args, seed):
    text = open(os.path.join(GOLDEN, name)).read()
    lines = run_bra(text, ["--amplitude-precision", "17", *args])
    q, amplitudes, s, inserted = replay_noisy(oracle, text, seed)
    assert inserted > 0  # the channel did insert gates
    n = q.num_qubits
    (amps,) = printed_amplitudes(lines, n)
    assert np.max(np.abs(np.array([a for _, a in amps]) - amplitudes[0])) < 1e-12
    assert np.max(np.abs(full_table(lines, n) - np.array(spins(oracle, amplitudes[0], n)))) < 1e-12


def test_depolarizing_draws_differ_from_noise_free(oracle):
    # the same program without its DEPOLARIZING CHANNEL line gives a different state
    text = open(os.path.join(GOLDEN, "depolarizing_sz_9.qcx")).read()
    clean = "\n".join(l for l in text.splitlines() if not l.startswith("DEPOLARIZING")) + "\n"
    noisy = printed_amplitudes(run_bra(text, ["--amplitude-precision", "17"]), 9)[0]
    plain = printed_amplitudes(run_bra(clean, ["--amplitude-precision", "17"]), 9)[0]
    assert max(abs(a - b) for (_, a), (_, b) in zip(noisy, plain)) > 1e-3


def test_full_precision_measurement_and_amplitudes(oracle):
    text = ("QUBITS 5\nINITIAL STATE 3\nH 0\nCNOT 0 4\nU3 2 0.3 1.1 -0.4\nEXX 1 3 0.7\nCCZ 0 2 4\nSX 1\n"
            "+Y 3\nU 2 1 3\nT 4\nSWAP 0 3\nESWAP 1 4 0.3\nCESWAP 0 2 3 -0.6\nESWAP+ 3 4 1.2\n"
            "DO MEASUREMENT -17\nDO AMPLITUDES\nDO AMPLITUDES 3 17\n")
    lines = run_bra(text)
    q, s = oracle_state(oracle, text)
    i = lines.index("<Qx> <Qy> <Qz>")
    rows = np.array([[float(x) for x in lines[i + 1 + k].split()] for k in range(5)])
    assert np.max(np.abs(rows - np.array(spins(oracle, s, 5)))) < 1e-12
    amp = [l for l in lines if "=>" in l]
    assert len(amp) == 32 + 2
    pat = re.compile(r"^([01]{5}) => (\S+) \+ (\S+) i$")
    for k, line in enumerate(amp):
        m = pat.match(line)
        assert m, line
        index = int(m.group(1), 2)
        assert index == (k if k < 32 else (3, 17)[k - 32])
        # default stream formatting: 6 significant digits
        assert abs(complex(float(m.group(2)), float(m.group(3))) - s[index]) < 1e-5
    # the same run printing 17 digits
    blocks = printed_amplitudes(run_bra(text, ["--amplitude-precision", "17"]), 5)
    assert np.max(np.abs(np.array([a for _, a in blocks[0]]) - s)) < 1e-12


def test_initial_state_and_errors():
    lines = run_bra("QUBITS 3\nINITIAL STATE 5\nDO AMPLITUDES 5 0\n")
    assert [l for l in lines if "=>" in l] == ["101 => 1 + 0 i", "000 => 0 + 0 i"]
    r = subprocess.run([EXE], input="QUBITS 2\nVAR X INT\n", capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "VAR" in r.stderr


# --- measurement and sampling statements ----------------------------------------------------

def run_bra_args(text, args):
    print(f"gpu bra: running brk-bra {' '.join(args)} on {text.splitlines()[0]!r} ...", flush=True)
    r = subprocess.run([EXE] + args, input=text, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout.splitlines()


def replay(oracle, text, seed):
    """The oracle runs the program's statements with bra's generator (std::mt19937_64(seed)):
    M / CLEAR / SET / GENERATE EVENTS / EXIT as bra::nompi_state does."""
    q = brk.Qcx(text)
    s = np.zeros(1 << q.num_qubits, dtype=np.complex128)
    s[q.initial_state] = 1
    g = oracle.mt64(seed)
    out = {"events": None, "exit": None, "amplitudes": []}
    for kind, v in q.statements:
        if kind == brk.QCX_GATES:
            run_ops(oracle, s, v)
        elif kind == brk.QCX_PROJECTIVE_MEASUREMENT:
            oracle.projective_measurement(s, v, g.canonical())
        elif kind in (brk.QCX_CLEAR, brk.QCX_SET):
            value = 0 if kind == brk.QCX_CLEAR else 1
            oracle.project(s, v, value, oracle.zero_one_probabilities(s, v)[value] ** -0.5)
        elif kind == brk.QCX_AMPLITUDES:
            out["amplitudes"].append(s.copy())
        elif kind == brk.QCX_GENERATE_EVENTS:
            num, ev_seed = v
            src = g if ev_seed < 0 else oracle.mt64(ev_seed)
            out["events"] = oracle.generate_events(s, [src.canonical() for _ in range(num)])
        elif kind == brk.QCX_EXIT:
            out["exit"] = int(oracle.generate_events(s, [g.canonical()])[0])
    return q, out


def printed_events(lines, n):
    i = lines.index("Events:")
    assert TIMING.match(lines[i - 1]) and lines[i - 1].startswith("Operations finished")
    ev = []
    for k, line in enumerate(lines[i + 1:]):
        if line.startswith("Events finished"):
            break
        index, bits = line.split()
        assert int(index) == k and len(bits) == n
        ev.append(int(bits, 2))
    return ev


MEASURE_PROGRAM = ("QUBITS 7\nINITIAL STATE 9\nH 0\nH 1\nU3 2 0.3 1.1 -0.4\nCNOT 0 3\nEXX 4 5 0.7\nH 6\n"
                   "CCZ 0 2 4\nSX 5\nM 2\nCLEAR 4\nSET 1\nM 0\nH 4\nDO AMPLITUDES\n")


@pytest.mark.parametrize("seed_args,seed", [([], 1), (["--seed", "12345"], 12345)])
def test_measurement_statements_and_events(oracle, seed_args, seed):
    text = MEASURE_PROGRAM + "GENERATE EVENTS 300 -1\nH 0\n"
    lines = run_bra_args(text, seed_args)
    q, ref = replay(oracle, text, seed)
    amp = [l for l in lines if "=>" in l]
    assert len(amp) == 128
    for k, line in enumerate(amp):
        bits, rest = line.split(" => ")
        assert int(bits, 2) == k
        re_, im_ = rest[:-2].split(" + ")
        assert abs(complex(float(re_), float(im_)) - ref["amplitudes"][0][k]) < 1e-5
    ev = printed_events(lines, 7)
    assert len(ev) == 300
    assert ev == ref["events"].tolist()
    assert lines[-1].startswith("Events finished") and TIMING.match(lines[-1])


def test_generate_events_fresh_seed_and_exit(oracle):
    text = MEASURE_PROGRAM + "GENERATE EVENTS 50 77\n"
    lines = run_bra_args(text, [])
    _, ref = replay(oracle, text, 1)
    assert printed_events(lines, 7) == ref["events"].tolist()
    text = MEASURE_PROGRAM + "EXIT\nGARBAGE AFTER EXIT\n"
    lines = run_bra_args(text, ["--seed", "5"])
    _, ref = replay(oracle, text, 5)
    i = [k for k, l in enumerate(lines) if l.startswith("Measurement result: ")]
    assert len(i) == 1 and int(lines[i[0]].split(": ")[1]) == ref["exit"]
    assert lines[i[0] - 1].startswith("Operations finished") and lines[i[0] + 1].startswith("Measurement finished")


def test_fusion_bracket_on_the_precision_form(oracle):
    """--fusion-c64 5: the gates between BEGIN FUSION and END FUSION run as fused complex64
    blocks on MFMA; outside the bracket the exact gate blocks.  Against the oracle (FP64):
    infidelity <= 1e-6, amplitudes within 2e-6; without the option the same program is exact."""
    lines = ["QUBITS 12"] + [f"H {q}" for q in range(12)] + ["BEGIN FUSION"]
    rng = np.random.default_rng(8)
    for k in range(60):
        a, b = (int(x) for x in rng.choice(12, 2, replace=False))
        lines += [f"U3 {a} {rng.uniform(-3, 3)!r} {rng.uniform(-3, 3)!r} {rng.uniform(-3, 3)!r}", f"CNOT {a} {b}",
                  f"EZ {b} {rng.uniform(-3, 3)!r}", f"SZZ {a} {b}" if k % 3 == 0 else f"CZ {a} {b}"]
    lines += ["END FUSION", "T 3", "DO AMPLITUDES"]
    text = "\n".join(lines) + "\n"
    q, s = oracle_state(oracle, text)
    fused = np.array([a for _, a in printed_amplitudes(
        run_bra(text, ["--amplitude-precision", "17", "--fusion-c64", "5"]), 12)[0]])
    infidelity = 1 - abs(np.vdot(s, fused)) ** 2 / (np.vdot(fused, fused).real * np.vdot(s, s).real)
    assert infidelity < 1e-6 and np.max(np.abs(fused - s)) < 2e-6
    assert np.max(np.abs(fused - s)) > 1e-12  # complex64 rounding shows: the precision form ran
    exact = np.array([a for _, a in printed_amplitudes(run_bra(text, ["--amplitude-precision", "17"]), 12)[0]])
    assert np.max(np.abs(exact - s)) < 1e-12
