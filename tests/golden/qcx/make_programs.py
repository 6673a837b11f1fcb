"""Writes the quantum-assembler test programs next to this script.

Each program is written for this repository's tests (none is taken from the
reference); together they exercise the statements of bra's front end
(docs/bra.md:109-188) that brk-bra implements:

  h_wall_10.qcx          H on all 10 qubits of |0..0>: every amplitude 2^-5
  toffoli_ladder_12.qcx  fixed rotations (+X -X +Y -Y SX SX+ SY SY+) then TOFFOLI,
                         CCNOT, C3X, CCZ, C2SWAP ladders
  draper_adder_8.qcx     |a>|b> -> |a>|a + b mod 16> with a QFT adder (H, U, U+,
                         BEGIN/END FUSION); a in uniform superposition, b = 5
  shor_box_21.qcx        SHORBOX 6 21 2 then the inverse QFT of the exponent
                         register (H, U with negative exponents, SWAP)
  depolarizing_sz_9.qcx  DEPOLARIZING CHANNEL with a seed, sqrt(Z) strings (SZZ,
                         SZZZ+, CSZZ, C2SZ2, SZ3, SZ2+), identities, CLEAR, SET
  depolarizing_m_7.qcx   DEPOLARIZING CHANNEL without a seed (noise draws share the
                         measurement generator), M between runs of gates

Every program ends with DO MEASUREMENT -17 and DO AMPLITUDES, so a test can
compare the printed values with the CPU oracle at full precision (run brk-bra
with --amplitude-precision 17).

  python tests/golden/qcx/make_programs.py
"""
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def write(name, lines):
    with open(os.path.join(HERE, name), "w") as f:
        f.write("\n".join(lines) + "\n")
    print("wrote", name)


def tail():
    return ["DO MEASUREMENT -17", "DO AMPLITUDES"]


def h_wall():
    out = ["! H on every qubit: uniform amplitudes 2^-5, <Qx> = 0, <Qy> = <Qz> = 1/2", "QUBITS 10"]
    out += [f"H {q}" for q in range(10)]
    return out + ["DO MEASUREMENT", "DO AMPLITUDES 0 5 1023"] + tail()


def toffoli_ladder():
    n = 12
    rot = ["+X", "-Y", "SX", "+Y", "SY+", "-X", "SY", "SX+"]
    out = ["! rotations on every qubit, then ladders of multi-controlled gates", f"QUBITS {n}"]
    out += [f"{rot[q % len(rot)]} {q}" for q in range(n)]
    out += [f"TOFFOLI {q} {q + 1} {q + 2}" for q in range(0, n - 2)]
    out += [f"CCNOT {q + 2} {q} {q + 1}" for q in range(0, n - 2, 3)]
    out += ["C3X 0 4 8 11", "CCCX 1 5 9 2", "CCZ 3 6 9", "C2SWAP 0 11 5 6", "CH 10 4", "CCH 2 3 7"]
    out += [f"TOFFOLI {q + 2} {q + 1} {q}" for q in range(n - 3, -1, -1)]
    return out + tail()


def qft_ops(qubits, inverse=False):
    """QFT without the final swaps on `qubits` (least significant first): qubit j ends up
    holding the phase e^{2 pi i b / 2^(j+1)}.  The inverse runs the gates backwards with
    U+ for each controlled phase."""
    ops = []
    for j in range(len(qubits) - 1, -1, -1):
        ops.append(("H", qubits[j]))
        for k in range(j - 1, -1, -1):
            ops.append(("U", qubits[j], qubits[k], j - k + 1))
    if not inverse:
        return [f"H {g[1]}" if g[0] == "H" else f"U {g[1]} {g[2]} {g[3]}" for g in ops]
    return [f"H {g[1]}" if g[0] == "H" else f"U+ {g[1]} {g[2]} {g[3]}" for g in reversed(ops)]


def draper_adder():
    a, b = [0, 1, 2, 3], [4, 5, 6, 7]
    out = ["! QFT adder: |a>|b> -> |a>|a + b mod 16>, a uniform over 0..15, b = 5", "QUBITS 8"]
    out += [f"H {q}" for q in a]
    out += ["X 4", "X 6"]  # b = 5
    out += ["BEGIN FUSION"] + qft_ops(b) + ["END FUSION"]
    # qubit b_j holds e^{2 pi i b / 2^(j+1)}: multiply by e^{2 pi i a / 2^(j+1)}
    for j in range(4):
        for i in range(j + 1):
            out.append(f"U {a[i]} {b[j]} {j + 1 - i}")
    out += qft_ops(b, inverse=True)
    return out + tail()


def shor_box():
    n, m = 11, 6
    e = list(range(n - m, n))  # exponent register: the top m qubits
    out = ["! modular exponentiation box, then the inverse QFT of the exponent register", f"QUBITS {n}",
           f"SHORBOX {m} 21 2"]
    for j in range(m):
        for k in range(j):
            out.append(f"U {e[j]} {e[k]} {-(j - k + 1)}")
        out.append(f"H {e[j]}")
    out += [f"SWAP {e[j]} {e[m - 1 - j]}" for j in range(m // 2)]
    return out + tail()


def depolarizing_sz():
    out = ["! depolarizing noise with its own seed; sqrt(Z) strings and identities",
           "DEPOLARIZING CHANNEL P_X = 0.05, P_Y=0.04 ,P_Z=0.07,SEED=2024", "QUBITS 9"]
    out += [f"H {q}" for q in range(9)]
    out += ["SZZ 0 1", "SZZZ+ 2 3 4", "SZZ+ 5 6", "SZZZ 6 7 8", "CSZZ 0 1 2", "CSZZZ 0 3 4 5", "SZZ 2 8",
            "CSZ 3 4", "CCSZ+ 0 1 2", "S 8", "S+ 0", "T 5", "I 3", "IC 4", "II 5 6", "CZ 1 7", "CNOT 2 6",
            "TOFFOLI 0 4 8", "U3 3 0.4 -1.2 2.2", "EZZ 0 5 0.6", "CCSZZZ 7 8 0 3 6"]
    out += ["CLEAR 2", "SET 5", "SZZZ 5 6 7", "SZ 8"]
    return out + tail()


def depolarizing_m():
    out = ["! depolarizing noise drawn from the measurement generator", "QUBITS 7",
           "DEPOLARIZING CHANNEL P_X=0.1,P_Y=0.1,P_Z=0.1"]
    out += [f"H {q}" for q in range(7)]
    out += ["CNOT 0 3", "SZZ 1 2", "U3 4 0.3 1.1 -0.4", "M 3", "CZ 2 5", "EXX 4 6 0.7", "M 0", "SET 6",
            "SX 5", "CCZ 1 4 6"]
    return out + tail()


if __name__ == "__main__":
    write("h_wall_10.qcx", h_wall())
    write("toffoli_ladder_12.qcx", toffoli_ladder())
    write("draper_adder_8.qcx", draper_adder())
    write("shor_box_21.qcx", shor_box())
    write("depolarizing_sz_9.qcx", depolarizing_sz())
    write("depolarizing_m_7.qcx", depolarizing_m())
