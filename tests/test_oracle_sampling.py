"""CPU checks of the oracle's measurement, sampling and Shor-box functions and
of the bra statements that use them (M, CLEAR, SET, GENERATE EVENTS, EXIT,
SHORBOX).

Pinned by:
- the C++ standard's known answer for std::mt19937_64 ([rand.predef]: the
  10000th draw of a default-constructed engine is 9981545732273789042);
- this image's libstdc++ itself: a g++-built probe prints generate_canonical and
  uniform_real_distribution draws, which the oracle must reproduce bit for bit
  (ket::utility::positive_random_value_upto draws through
  uniform_real_distribution, positive_random_value_upto.hpp:12-18);
- closed forms on product states (P(q = 1) = sin^2(theta_q / 2); projecting a
  product state replaces one factor by a basis vector);
- an independent numpy restatement of sampling (sequential cumsum +
  searchsorted(side="right"), which is std::upper_bound), plus the invariant
  that no event lands on an amplitude of probability zero;
- for the Shor box, Python's exact modular power pow(base, x, divisor) and an
  explicit bit reversal (v_0 = 1 unreduced, only the low n - m bits kept).
No reference test or fixture covers these paths; the test program
tests/golden/qcx/shor_box_21.qcx is run end to end by the GPU bra tests."""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

import brk

MT64_KAT = 9981545732273789042


def product_state(thetas, phis):
    """kron over qubits (qubit 0 is the least significant bit) of cos(t/2)|0> + e^{i phi} sin(t/2)|1>."""
    s = np.array([1.0 + 0j])
    for t, p in zip(thetas, phis):
        f = np.array([math.cos(t / 2), np.exp(1j * p) * math.sin(t / 2)])
        s = np.kron(f, s)
    return s


def random_state(n, seed):
    rng = np.random.default_rng(seed)
    s = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return s / np.linalg.norm(s)


def test_mt19937_64_known_answer(oracle):
    g = oracle.mt64(5489)
    for _ in range(9999):
        g.next()
    assert g.next() == MT64_KAT


PROBE = r"""
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <random>
int main(int argc, char** argv) {
  std::mt19937_64 g(std::strtoull(argv[1], nullptr, 10));
  for (int i = 0; i < 200; ++i)
    std::printf("%a\n", std::generate_canonical<double, std::numeric_limits<double>::digits>(g));
  for (int i = 0; i < 200; ++i) std::printf("%a\n", std::uniform_real_distribution<double>{0.0, 0.7}(g));
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ is not available")
@pytest.mark.parametrize("seed", [1, 0, 12345, 2**64 - 1])
def test_canonical_draws_match_libstdcxx(oracle, tmp_path, seed):
    src = tmp_path / "probe.cpp"
    src.write_text(PROBE)
    exe = tmp_path / "probe"
    subprocess.run(["g++", "-O2", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe), str(seed)], check=True, capture_output=True, text=True).stdout.split()
    values = [float.fromhex(v) for v in out]
    g = oracle.mt64(seed)
    assert [g.canonical() for _ in range(200)] == values[:200]
    assert [g.canonical() * 0.7 + 0.0 for _ in range(200)] == values[200:]


@pytest.mark.parametrize("n", [1, 3, 7])
def test_zero_one_probabilities_closed_form(oracle, n):
    rng = np.random.default_rng(n)
    thetas, phis = rng.uniform(0, math.pi, n), rng.uniform(-math.pi, math.pi, n)
    s = product_state(thetas, phis)
    for q in range(n):
        p0, p1 = oracle.zero_one_probabilities(s, q)
        assert abs(p1 - math.sin(thetas[q] / 2) ** 2) < 1e-14
        assert abs(p0 - math.cos(thetas[q] / 2) ** 2) < 1e-14


def test_zero_one_probabilities_random_state(oracle):
    n = 9
    s = random_state(n, 3)
    p = np.abs(s) ** 2
    for q in range(n):
        axes = p.reshape([2] * n)  # axis k is qubit n - 1 - k
        p1 = axes.take(1, axis=n - 1 - q).sum()
        assert abs(oracle.zero_one_probabilities(s, q)[1] - p1) < 1e-14


@pytest.mark.parametrize("q,value", [(0, 0), (0, 1), (2, 1), (4, 0)])
def test_project_replaces_one_factor(oracle, q, value):
    n = 5
    rng = np.random.default_rng(q + 10 * value)
    thetas, phis = rng.uniform(0.3, 2.8, n), rng.uniform(-math.pi, math.pi, n)
    s = product_state(thetas, phis)
    p = oracle.zero_one_probabilities(s, q)[value]
    oracle.project(s, q, value, p ** -0.5)
    t2 = thetas.copy()
    t2[q] = 0.0 if value == 0 else math.pi
    p2 = phis.copy()
    p2[q] = phis[q] if value else 0.0  # the kept factor keeps its phase
    assert np.max(np.abs(s - product_state(t2, p2))) < 1e-14
    # the dropped half is exactly +0
    idx = np.arange(s.size)
    dropped = s[((idx >> q) & 1) != value]
    assert np.all(dropped.real == 0) and np.all(np.signbit(dropped.real) == 0)
    assert np.all(np.signbit(dropped.imag) == 0)


def test_projective_measurement_decision(oracle):
    s0 = product_state([1.1, 0.4], [0.3, -0.2])
    p0 = math.cos(1.1 / 2) ** 2  # qubit 0
    below, above = (p0 - 1e-9), (p0 + 1e-9)
    s = s0.copy()
    assert oracle.projective_measurement(s, 0, below) == 0 and abs(np.vdot(s, s) - 1) < 1e-14
    s = s0.copy()
    assert oracle.projective_measurement(s, 0, above) == 1 and abs(np.vdot(s, s) - 1) < 1e-14


def numpy_events(s, canonical):
    scan = np.cumsum(np.abs(s) ** 2)  # sequential, left to right
    return np.searchsorted(scan, np.asarray(canonical) * scan[-1], side="right")


@pytest.mark.parametrize("n", [1, 4, 11])
def test_generate_events_vs_numpy(oracle, n):
    s = random_state(n, n)
    g = oracle.mt64(7)
    c = [g.canonical() for _ in range(3000)]
    assert np.array_equal(oracle.generate_events(s, c), numpy_events(s, c))


def test_generate_events_skips_zero_probability(oracle):
    s = np.zeros(64, dtype=np.complex128)
    s[[3, 17, 40, 41]] = [0.5, 0.5j, -0.5, 0.5]
    g = oracle.mt64(1)
    c = [g.canonical() for _ in range(4000)] + [0.0, np.nextafter(1.0, 0.0)]
    ev = oracle.generate_events(s, c)
    assert set(ev.tolist()) == {3, 17, 40, 41}
    counts = np.bincount(ev[:4000], minlength=64)[[3, 17, 40, 41]]
    assert np.all(np.abs(counts - 1000) < 5 * math.sqrt(4000 * 0.25 * 0.75))
    assert ev[-2] == 3 and ev[-1] == 41


# --- bra statements --------------------------------------------------------------------------

def test_parse_measurement_statements():
    q = brk.Qcx("QUBITS 4\nH 0\nM 0\nCLEAR 1\nset 2\nM 3\nGENERATE EVENTS 25 -1\nH 0\nNOT_A_GATE\n")
    kinds = [k for k, _ in q.statements]
    assert kinds == [brk.QCX_GATES, brk.QCX_PROJECTIVE_MEASUREMENT, brk.QCX_CLEAR, brk.QCX_SET,
                     brk.QCX_PROJECTIVE_MEASUREMENT, brk.QCX_GENERATE_EVENTS]
    assert [v for _, v in q.statements[1:]] == [0, 1, 2, 3, (25, -1)]


def test_exit_stops_reading():
    q = brk.Qcx("QUBITS 2\nH 1\nEXIT\nSHORBOX 1 2 3\n")
    assert [k for k, _ in q.statements] == [brk.QCX_GATES, brk.QCX_EXIT]


@pytest.mark.parametrize("text,status,fragment", [
    ("QUBITS 2\nBEGIN FUSION\nM 0\nEND FUSION\n", 8, "inside BEGIN FUSION"),
    ("QUBITS 2\nBEGIN FUSION\nGENERATE EVENTS 2 1\n", 8, "inside BEGIN FUSION"),
    ("QUBITS 2\nCLEAR 2\n", 1, "out of range"),
    ("QUBITS 2\nSET\n", 1, "takes 1 operands"),
    ("QUBITS 2\nGENERATE EVENTS 10\n", 1, "GENERATE EVENTS <num_events> <seed>"),
    ("QUBITS 2\nGENERATE SAMPLES 10 1\n", 1, "GENERATE EVENTS <num_events> <seed>"),
    ("QUBITS 2\nGENERATE EVENTS -1 1\n", 1, "out of range"),
    ("QUBITS 2\nEXIT now\n", 1, "takes 0 operands"),
    ("QUBITS 2\nEND FUSION\n", 1, "without BEGIN FUSION"),
])
def test_statement_errors(text, status, fragment):
    with pytest.raises(brk.BrkError) as e:
        brk.Qcx(text)
    assert e.value.status == status and fragment in str(e.value)


# --- Shor box --------------------------------------------------------------------------------

def shor_closed_form(n, m, divisor, base):
    """2^{-m/2} sum_x |reverse_m(x)> (top m qubits) |base^x mod divisor> (low bits kept), v_0 = 1."""
    s = np.zeros(1 << n, dtype=np.complex128)
    for x in range(1 << m):
        r = int(format(x, f"0{m}b")[::-1], 2) if m else 0
        v = 1 if x == 0 else pow(base, x, divisor)
        s[(r << (n - m)) | (v & ((1 << (n - m)) - 1))] = (2.0 ** m) ** -0.5
    return s


@pytest.mark.parametrize("n,m,divisor,base", [(12, 8, 15, 7), (10, 4, 21, 2), (6, 6, 5, 3), (5, 0, 7, 3),
                                              (7, 3, 1, 4), (9, 5, 9, 9), (8, 5, 200, 13)])
def test_shor_box_closed_form(oracle, n, m, divisor, base):
    s = np.full(1 << n, 0.3 + 0.1j)  # overwritten entirely
    oracle.shor_box(s, m, divisor, base)
    assert np.array_equal(s, shor_closed_form(n, m, divisor, base))
    assert abs(np.vdot(s, s).real - 1) < 1e-14


def test_parse_shor_box():
    q = brk.Qcx("QUBITS 6\nSHORBOX 3 5 2\nH 0\n")
    assert q.statements[0] == (brk.QCX_SHOR_BOX, (3, 5, 2))
    for text in ("QUBITS 6\nSHORBOX 7 5 2\n", "QUBITS 6\nSHORBOX 3 0 2\n", "QUBITS 6\nSHORBOX 3 5\n"):
        with pytest.raises(brk.BrkError) as e:
            brk.Qcx(text)
        assert e.value.status == 1
