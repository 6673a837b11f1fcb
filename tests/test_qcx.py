"""CPU checks of the bra front end (quantum-assembler parser, no device).

Pinned by: the reference's own sample programs (bra/qcx/*.qcx, copied under
tests/golden/qcx/), the mnemonic semantics of docs/bra.md:109-188 checked
through identities that hold independently of how the parser writes a gate
down (e.g. +X = exp(i pi/4 X), SX^2 = X), and bra's operand order
(controls first, interpreter.cpp:3341-3369)."""
import math
import os

import numpy as np
import pytest

import brk
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


def test_hadamards08():
    q = parse(sample("hadamards08.qcx"))
    assert q.num_qubits == 8 and q.initial_state == 0
    assert q.statements[-1] == (brk.QCX_MEASUREMENT, 3)
    assert q.ops() == [("H", (i,), (), ()) for i in range(8)]


def test_toffoli12():
    q = parse(sample("toffoli12.qcx"))
    ops = q.ops()
    assert q.num_qubits == 12 and len(ops) == 20
    assert [o[0] for o in ops[:12]] == ["MATRIX"] * 12
    # TOFFOLI c1 c2 t: X on t under two controls
    assert ops[12] == ("X", (0,), (1, 2), ())
    assert ops[-1] == ("X", (3,), (1, 10), ())
    assert q.statements[-1] == (brk.QCX_MEASUREMENT, 3)  # BEGIN MEASUREMENT


def test_adder_controlled_phase():
    q = parse(sample("adder6x2.qcx"))
    ops = q.ops()
    # "U 7 6 2": phase 2 pi / 2^2 where both qubits are 1; "U 7 6 -2": its adjoint
    assert ("U1", (7,), (6,), (math.pi / 2,)) in ops
    assert ops[-2] == ("U1", (7,), (6,), (-math.pi / 2,))


def test_shor12():
    q = parse(sample("shor12.qcx"))
    assert q.num_qubits == 12
    assert q.statements[0] == (brk.QCX_SHOR_BOX, (8, 15, 7))
    assert q.statements[-1] == (brk.QCX_MEASUREMENT, 3)
    ops = q.ops()
    assert ops[0] == ("H", (7,), (), ()) and ops[1] == ("U1", (7,), (6,), (-math.pi / 2,))


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
    assert ops[0] == ("U1", (1,), (), (math.pi / 2,))
    assert ops[1] == ("U1", (1,), (), (-math.pi / 2,))
    assert ops[2] == ("U1", (2,), (), (math.pi / 2,))
    assert ops[3] == ("U1", (3,), (), (math.pi / 4,)) and ops[4] == ("U1", (3,), (), (-math.pi / 4,))
    assert ops[5] == ("U1", (4,), (), (0.25,)) and ops[6] == ("U1", (4,), (), (-0.25,))
    assert ops[7] == ("U1", (5,), (), (2 * math.pi / 8,))
    assert ops[8] == ops[9] == ("U1", (5,), (), (-2 * math.pi / 8,))
    # controlled S on control qubits 1 and 2; CCR: three control qubits
    assert ops[10] == ("U1", (2,), (1,), (math.pi / 2,))
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
    ("QUBITS 2\nDEPOLARIZING CHANNEL P = 0.1\n", 8, "DEPOLARIZING"),
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
