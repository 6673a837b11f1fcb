"""CPU checks of the bra front end (quantum-assembler parser, no device).

Pinned by: this repository's own test programs (tests/golden/qcx/*.qcx, written
by tests/golden/qcx/make_programs.py) run on the CPU oracle against closed
forms (an H wall, a QFT adder), the mnemonic semantics of docs/bra.md:109-188
checked through identities that hold independently of how the parser writes a
gate down (e.g. +X = exp(i pi/4 X), SX^2 = X, SZZ = sZ (x) sZ), bra's operand
order (controls first, interpreter.cpp:3341-3369) and the order in which
bra::state draws depolarizing noise (bra/src/state.cpp:1000-3330)."""
import math
import os

import numpy as np
import pytest

import brk
import dense as D
from oracle_circuits import run_ops

HERE = os.path.dirname(os.path.abspath(__file__))


def sample(name):
    with open(os.path.join(HERE, "golden", "qcx", name)) as f:
        return f.read()


def parse(text):
    return brk.Qcx(text)


def only_ops(body, n=8):
    return parse(f"QUBITS {n}\n{body}\n").ops()


def matrix_of(op):
    p = op[3]
    return np.array([[p[0] + 1j * p[1], p[2] + 1j * p[3]], [p[4] + 1j * p[5], p[6] + 1j * p[7]]])


X = np.array([[0, 1], [1, 0]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
I2 = np.eye(2)


def run_program(oracle, q):
    s = np.zeros(1 << q.num_qubits, dtype=np.complex128)
    s[q.initial_state] = 1
    for kind, v in q.statements:
        if kind == brk.QCX_GATES:
            run_ops(oracle, s, v)
        elif kind == brk.QCX_SHOR_BOX:
            oracle.shor_box(s, *v)
    return s


def test_h_wall_closed_form(oracle):
    q = parse(sample("h_wall_10.qcx"))
    assert q.num_qubits == 10 and q.initial_state == 0
    assert q.ops() == [("H", (i,), (), ()) for i in range(10)]
    assert q.statements[-1] == (brk.QCX_AMPLITUDES, []) and q.statements[-2] == (brk.QCX_MEASUREMENT, -17)
    s = run_program(oracle, q)
    assert np.max(np.abs(s - 2.0 ** -5)) < 1e-15


def test_draper_adder_closed_form(oracle):
    # |a>|5> -> |a>|a + 5 mod 16>, a uniform over 0..15 (qubits 0..3 hold a, 4..7 b)
    q = parse(sample("draper_adder_8.qcx"))
    kinds = [k for k, _ in q.statements]
    assert brk.QCX_BEGIN_FUSION in kinds and brk.QCX_END_FUSION in kinds
    s = run_program(oracle, q)
    expected = np.zeros(256, dtype=np.complex128)
    for a in range(16):
        expected[a | (((a + 5) % 16) << 4)] = 0.25
    assert np.max(np.abs(s - expected)) < 1e-14


def test_toffoli_ladder_ops():
    q = parse(sample("toffoli_ladder_12.qcx"))
    ops = q.ops()
    assert q.num_qubits == 12
    assert [o[0] for o in ops[:12]] == ["MATRIX"] * 12
    # TOFFOLI c1 c2 t: X on t under two controls; C3X: three controls
    assert ops[12] == ("X", (2,), (0, 1), ())
    assert ("X", (11,), (0, 4, 8), ()) in ops and ("Z", (3, 6, 9), (), ()) in ops
    assert ("SWAP", (5, 6), (0, 11), ()) in ops


def test_shor_box_program():
    q = parse(sample("shor_box_21.qcx"))
    assert q.num_qubits == 11
    assert q.statements[0] == (brk.QCX_SHOR_BOX, (6, 21, 2))
    ops = q.ops()
    # "U c1 c2 -k": phase -2 pi / 2^k where both qubits are 1
    assert ops[0] == ("H", (5,), (), ()) and ops[1] == ("U1", (6,), (5,), (-math.pi / 2,))


def test_unsupported_statement_names_line():
    with pytest.raises(brk.BrkError) as e:
        parse("QUBITS 4\nH 0\nVAR X INT\n")
    assert e.value.status == 8 and "line 3" in str(e.value) and "VAR" in str(e.value)


def test_fixed_gates_match_their_definitions():
    ops = only_ops("+X 0\n-X 0\n+Y 0\n-Y 0\nSX 0\nSX+ 0\nSY 0\nSY+ 0")
    m = {name: matrix_of(op) for name, op in zip(["+X", "-X", "+Y", "-Y", "SX", "SX+", "SY", "SY+"], ops)}
    c = s = 1 / math.sqrt(2)
    # +X rotates by -pi/2 about x: exp(i pi/4 X); -X by +pi/2: exp(-i pi/4 X); likewise for y
    assert np.allclose(m["+X"], c * I2 + 1j * s * X)
    assert np.allclose(m["-X"], c * I2 - 1j * s * X)
    assert np.allclose(m["+Y"], c * I2 + 1j * s * Y)
    assert np.allclose(m["-Y"], c * I2 - 1j * s * Y)
    # square roots and their adjoints
    assert np.allclose(m["SX"] @ m["SX"], X) and np.allclose(m["SY"] @ m["SY"], Y)
    assert np.allclose(m["SX+"], m["SX"].conj().T) and np.allclose(m["SY+"], m["SY"].conj().T)


def test_phase_gates():
    ops = only_ops("S 1\nS+ 1\nSZ 2\nT 3\nT+ 3\nU1 4 0.25\nU1+ 4 0.25\nR 5 3\nR+ 5 3\nR 5 -3\nCS 1 2\nCCR 0 1 2 1")
    # S = sqrt(Z): exactly i (ket::utility::imaginary_unit), S+ exactly -i
    assert ops[0] == ("PHASE", (1,), (), (0.0, 1.0))
    assert ops[1] == ("PHASE", (1,), (), (0.0, -1.0))
    assert ops[2] == ("PHASE", (2,), (), (0.0, 1.0))
    assert ops[3] == ("U1", (3,), (), (math.pi / 4,)) and ops[4] == ("U1", (3,), (), (-math.pi / 4,))
    assert ops[5] == ("U1", (4,), (), (0.25,)) and ops[6] == ("U1", (4,), (), (-0.25,))
    assert ops[7] == ("U1", (5,), (), (2 * math.pi / 8,))
    assert ops[8] == ops[9] == ("U1", (5,), (), (-2 * math.pi / 8,))
    # controlled S on control qubits 1 and 2; CCR: three control qubits
    assert ops[10] == ("PHASE", (2,), (1,), (0.0, 1.0))
    assert ops[11] == ("U1", (0,), (1, 2), (math.pi,))


def test_controlled_forms_and_pauli_strings():
    ops = only_ops("CNOT 1 0\nCCNOT 1 2 0\nC3X3 0 1 2 3 4 5\nCCXX 0 1 2 3\nCZ 2 5\nCCZ 1 2 3\nZZ 1 3\n"
                   "CZZ 0 1 2\nXX 4 5\nY3 0 1 2\nCH 0 1\nC2SWAP 0 1 2 3\nCU3 0 1 0.1 0.2 0.3\nCU2 2 3 0.4 0.5")
    assert ops[0] == ("X", (0,), (1,), ())
    assert ops[1] == ("X", (0,), (1, 2), ())
    assert ops[2] == ("X", (3, 4, 5), (0, 1, 2), ())
    assert ops[3] == ("X", (2, 3), (0, 1), ())
    assert ops[4] == ("Z", (2, 5), (), ())          # CZ: both are control qubits
    assert ops[5] == ("Z", (1, 2, 3), (), ())       # CnZ: n + 1 control qubits
    assert ops[6:8] == [("Z", (1,), (), ()), ("Z", (3,), (), ())]  # Z (x) Z
    assert ops[8:10] == [("Z", (1,), (0,), ()), ("Z", (2,), (0,), ())]
    assert ops[10] == ("X", (4, 5), (), ())
    assert ops[11] == ("Y", (0, 1, 2), (), ())
    assert ops[12] == ("H", (1,), (0,), ())
    assert ops[13] == ("SWAP", (2, 3), (0, 1), ())
    assert ops[14] == ("U3", (1,), (0,), (0.1, 0.2, 0.3))
    assert ops[15] == ("U2", (3,), (2,), (0.4, 0.5))


def test_exponentials_and_adjoints():
    ops = only_ops("EX 0 0.3\nEX+ 0 0.3\nEYY 1 2 0.2\nEZ3 0 1 2 0.1\nCEZ 3 4 0.5\nU2+ 1 0.3 0.4\nU3+ 1 0.1 0.2 0.3")
    assert ops[0] == ("EX", (0,), (), (0.3,)) and ops[1] == ("EX", (0,), (), (-0.3,))
    assert ops[2] == ("EY", (1, 2), (), (0.2,))
    assert ops[3] == ("EZ", (0, 1, 2), (), (0.1,))
    assert ops[4] == ("EZ", (4,), (3,), (0.5,))
    assert ops[5] == ("U3", (1,), (), (-math.pi / 2, -0.4, -0.3))
    assert ops[6] == ("U3", (1,), (), (-0.1, -0.3, -0.2))


def test_adjoint_pairs_cancel(oracle):
    # every gate followed by its adjoint form returns the state (checked on the oracle)
    body = ("H 0\nS 1\nS+ 1\nT 2\nT+ 2\nSX 3\nSX+ 3\nSY 0\nSY+ 0\n+X 1\n-X 1\n+Y 2\n-Y 2\nU1 3 0.7\nU1+ 3 0.7\n"
            "U2 0 0.3 0.9\nU2+ 0 0.3 0.9\nU3 1 0.1 0.2 0.3\nU3+ 1 0.1 0.2 0.3\nEXX 0 1 0.4\nEXX+ 0 1 0.4\n"
            "R 2 3\nR+ 2 3\nU 0 3 2\nU+ 0 3 2\nH 0")
    rng = np.random.default_rng(0)
    s = rng.normal(size=16) + 1j * rng.normal(size=16)
    s /= np.linalg.norm(s)
    t = s.copy()
    run_ops(oracle, t, only_ops(body, 4))
    assert np.max(np.abs(t - s)) < 1e-14


def test_directives_and_comments():
    q = parse("! header\ncircuits 1\nqubits 3\ninitial state 5\nbit assignment 2 0 1\nh 0 ! trailing comment\n"
              "BEGIN FUSION\nX 1\nEND FUSION\nDO AMPLITUDES 0 5 7\nDO MEASUREMENT -8\nDO AMPLITUDES\n")
    assert q.num_qubits == 3 and q.initial_state == 5
    kinds = [k for k, _ in q.statements]
    assert kinds == [brk.QCX_GATES, brk.QCX_BEGIN_FUSION, brk.QCX_GATES, brk.QCX_END_FUSION, brk.QCX_AMPLITUDES,
                     brk.QCX_MEASUREMENT, brk.QCX_AMPLITUDES]
    assert q.statements[4][1] == [0, 5, 7] and q.statements[5][1] == -8 and q.statements[6][1] == []


@pytest.mark.parametrize("text,status,fragment", [
    ("H 0\n", 1, "QUBITS must come before"),
    ("QUBITS 2\nH 2\n", 1, "line 2"),
    ("QUBITS 2\nCNOT 1 1\n", 3, "twice"),
    ("QUBITS 2\nCIRCUITS 2\n", 8, "more than one circuit"),
    ("QUBITS 2\nU1 0 pi\n", 1, "not a number"),
    ("QUBITS 2\nH 0 1\n", 1, "takes 1 operands"),
    ("QUBITS 2\nBIT ASSIGNMENT 0 0\n", 1, "repeats"),
    ("QUBITS 2\nDEPOLARIZING CHANNEL P = 0.1\n", 1, "unknown key"),
    ("QUBITS 2\nDEPOLARIZING CHANNEL P_X=0.6,P_Y=0.5\n", 1, "exceeds 1"),
    ("QUBITS 2\nDEPOLARIZING CHANNEL P_X=0.1,P_X=0.2\n", 1, "twice"),
    ("QUBITS 2\nDEPOLARIZING CHANNEL P_X=-0.1\n", 1, "probability"),
    ("QUBITS 2\nDEPOLARIZING CHANNEL P_X=0.1,\n", 1, "KEY=value"),
    ("QUBITS 2\nDEPOLARIZING NOISE P_X=0.1\n", 1, "CHANNEL"),
])
def test_errors(text, status, fragment):
    with pytest.raises(brk.BrkError) as e:
        parse(text)
    assert e.value.status == status and fragment in str(e.value)


def test_exponential_swap_forms(oracle):
    ops = only_ops("ESWAP 1 2 0.3\nESWAP+ 1 2 0.3\nCESWAP 0 3 4 -0.2\nC2ESWAP 5 6 0 1 0.7")
    assert ops[0] == ("ESWAP", (1, 2), (), (0.3,))
    assert ops[1] == ("ESWAP", (1, 2), (), (-0.3,))
    assert ops[2] == ("ESWAP", (3, 4), (0,), (-0.2,))
    assert ops[3] == ("ESWAP", (0, 1), (5, 6), (0.7,))
    # ESWAP(theta) ESWAP+(theta) = I, and ESWAP(pi/2) = i SWAP (checked on the oracle)
    rng = np.random.default_rng(1)
    s = rng.normal(size=8) + 1j * rng.normal(size=8)
    t = s.copy()
    run_ops(oracle, t, only_ops("ESWAP 0 2 0.9\nESWAP+ 0 2 0.9", 3))
    assert np.max(np.abs(t - s)) < 1e-15
    u = s.copy()
    run_ops(oracle, u, only_ops(f"ESWAP 0 2 {math.pi / 2!r}", 3))
    v = s.copy()
    oracle.swap(v, 0, 2)
    assert np.max(np.abs(u - 1j * v)) < 1e-15


def test_sqrt_z_strings():
    # SZZ / SZZZ.. / SZ<n> and adjoints: i (-i) per set target, under the controls; each a
    # per-target phase of exactly i.  SZ<n> follows docs/bra.md:133 (C3SZ3 c1 c2 c3 t1 t2 t3).
    ops = only_ops("SZZ 0 1\nSZZZ+ 2 3 4\nCSZZ 5 6 7\nSZ3 1 4 7\nC2SZ2 0 1 2 3\nSZ2+ 5 6\nCSZZ+ 0 1 2\nSZ1 3")
    i, mi = (0.0, 1.0), (0.0, -1.0)
    assert ops[0:2] == [("PHASE", (0,), (), i), ("PHASE", (1,), (), i)]
    assert ops[2:5] == [("PHASE", (q,), (), mi) for q in (2, 3, 4)]
    assert ops[5:7] == [("PHASE", (6,), (5,), i), ("PHASE", (7,), (5,), i)]
    assert ops[7:10] == [("PHASE", (q,), (), i) for q in (1, 4, 7)]
    assert ops[10:12] == [("PHASE", (2,), (0, 1), i), ("PHASE", (3,), (0, 1), i)]
    assert ops[12:14] == [("PHASE", (5,), (), mi), ("PHASE", (6,), (), mi)]
    assert ops[14:16] == [("PHASE", (1,), (0,), mi), ("PHASE", (2,), (0,), mi)]
    assert ops[16] == ("PHASE", (3,), (), i)


def test_sqrt_z_strings_match_their_definition(oracle):
    # C(sZ (x) sZ (x) sZ) on a random state equals the dense kron of diag(1, i) under the control
    rng = np.random.default_rng(3)
    s = D.random_state(5, rng)
    t = s.copy()
    run_ops(oracle, t, only_ops("CSZZZ 4 0 2 3\nSZZ+ 1 3", 5))
    sz = np.diag([1, 1j])
    u = D.apply(s, D.kron_all([sz, sz, sz]), [0, 2, 3], [4])
    u = D.apply(u, D.kron_all([sz.conj(), sz.conj()]), [1, 3])
    assert np.max(np.abs(t - u)) < 1e-15


def test_depolarizing_statement():
    q = parse("QUBITS 3\nDEPOLARIZING CHANNEL p_x = 0.125 , P_Y=0.25,P_z =0.0625, seed=77\nH 0\n")
    assert q.depolarizing == ([0.125, 0.25, 0.0625], 77)
    q = parse("DEPOLARIZING CHANNEL P_Z=0.5\nQUBITS 3\nH 0\n")
    assert q.depolarizing == ([0.0, 0.0, 0.5], -1)
    assert parse("QUBITS 3\nH 0\n").depolarizing is None


def test_noise_qubit_order():
    """bra::state draws noise per operated qubit after each gate: targets then controls
    (CNOT c t -> t, c; TOFFOLI c1 c2 t -> t, c1, c2), operand order for gates on "control"
    qubits only (CZ, S, U1, U), and for identities too."""
    q = parse("QUBITS 6\nCNOT 1 0\nTOFFOLI 3 4 2\nCZ 5 1\nCU1 2 3 0.5\nU 4 0 2\nCSZZ 5 0 1\nI 2\nIC 3\n"
              "CI 1 4\nII 0 5\nEXX 3 4 0.2\nCH 2 0\nC2SWAP 1 2 3 4\nZZ 0 5\n")
    (gi, _), = [(i, v) for i, (k, v) in enumerate(q.statements) if k == brk.QCX_GATES]
    records = q.gate_records[gi]
    noise = [r[2] for r in records]
    assert noise == [[0, 1], [2, 3, 4], [5, 1], [2, 3], [4, 0], [0, 1, 5], [2], [3], [4, 1], [0, 5], [3, 4],
                     [0, 2], [3, 4, 1, 2], [0, 5]]
    counts = [r[1] for r in records]
    assert counts == [1, 1, 1, 1, 1, 2, 0, 0, 0, 0, 1, 1, 1, 2]
    assert [r[0] for r in records] == list(np.cumsum([0] + counts[:-1]))
