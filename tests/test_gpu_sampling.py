"""GPU parity of measurement and sampling against the CPU oracle, through the
C ABI: zero/one probabilities, the projection used by M / CLEAR / SET, and
event generation (ket::generate_events).

Bars: probabilities within 1e-12 (FP64 sums in another order); projections
bit-exact (the same multiply per component); events equal to the oracle's
sequential scan, except where a target lies within rounding (1e-12 of the
total) of a cumulative boundary, which is checked for every differing event."""
import math
import sys

import numpy as np
import pytest

import brk
import dense as D
from devstate import DeviceState

pytestmark = pytest.mark.gpu

TOL = 1e-12


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def canonical_draws(oracle, seed, count):
    g = oracle.mt64(seed)
    return [g.canonical() for _ in range(count)]


def check_events(s, canonical, gpu, ora):
    gpu, ora = np.asarray(gpu, dtype=np.int64), np.asarray(ora, dtype=np.int64)
    p = np.abs(s.astype(np.complex128)) ** 2
    assert np.all(p[gpu] > 0), "an event landed on an amplitude of probability zero"
    diff = np.nonzero(gpu != ora)[0]
    scan = np.cumsum(p)
    for k in diff:
        r = canonical[k] * scan[-1]
        lo, hi = sorted((gpu[k], ora[k]))
        assert np.all(np.abs(scan[lo:hi] - r) <= 1e-12 * scan[-1]), (k, gpu[k], ora[k])
    assert diff.size <= max(2, len(canonical) // 1000)


@pytest.mark.parametrize("n,dtype", [(1, brk.C128), (10, brk.C128), (20, brk.C128), (16, brk.C64)])
def test_zero_one_probabilities(oracle, n, dtype):
    rng = np.random.default_rng(100 + n)
    s = D.random_state(n, rng)
    if dtype == brk.C64:
        s = s.astype(np.complex64).astype(np.complex128)
    st = DeviceState.from_host(s, dtype)
    qubits = range(n) if n <= 10 else sorted({0, 1, 7, n // 2, n - 1})
    for q in qubits:
        note(f"gpu zero/one n={n} dtype={dtype} q={q}")
        g = brk.zero_one_probabilities(st.ptr, st.size, dtype, q)
        o = oracle.zero_one_probabilities(s, q)
        assert abs(g[0] - o[0]) < TOL and abs(g[1] - o[1]) < TOL
    st.free()


@pytest.mark.parametrize("q,value", [(0, 0), (0, 1), (5, 1), (11, 0)])
def test_project_bit_exact(oracle, q, value):
    n = 12
    note(f"gpu project n={n} q={q} value={value}")
    rng = np.random.default_rng(7 * q + value)
    s = D.random_state(n, rng)
    m = oracle.zero_one_probabilities(s, q)[value] ** -0.5
    st = DeviceState.from_host(s)
    brk.project_qubit(st.ptr, st.size, brk.C128, q, value, m)
    ref = s.copy()
    oracle.project(ref, q, value, m)
    out = st.download()
    assert np.array_equal(out.view(np.uint64), ref.view(np.uint64))
    st.free()
    # complex64: one single-precision multiply per component by (float)m
    s32 = s.astype(np.complex64)
    st = DeviceState.from_host(s32, brk.C64)
    brk.project_qubit(st.ptr, st.size, brk.C64, q, value, m)
    keep = ((np.arange(s32.size) >> q) & 1) == value
    ref32 = np.zeros_like(s32)
    ref32.real[keep] = s32.real[keep] * np.float32(m)
    ref32.imag[keep] = s32.imag[keep] * np.float32(m)
    assert np.array_equal(st.download().view(np.uint32), ref32.view(np.uint32))
    st.free()


@pytest.mark.parametrize("n,dtype", [(1, brk.C128), (5, brk.C128), (10, brk.C128), (13, brk.C128),
                                     (14, brk.C128), (20, brk.C128), (18, brk.C64)])
def test_generate_events_vs_oracle(oracle, n, dtype):
    note(f"gpu events n={n} dtype={dtype}")
    rng = np.random.default_rng(200 + n)
    s = D.random_state(n, rng)
    if dtype == brk.C64:
        s = s.astype(np.complex64).astype(np.complex128)
    c = canonical_draws(oracle, n, 4096)
    st = DeviceState.from_host(s, dtype)
    gpu = brk.generate_events(st.ptr, st.size, dtype, c)
    check_events(s, c, gpu, oracle.generate_events(s, c))
    # deterministic, and the state is left unchanged
    assert brk.generate_events(st.ptr, st.size, dtype, c) == gpu
    assert np.array_equal(st.download().astype(np.complex128), s)
    st.free()


def test_generate_events_sparse_states(oracle):
    n = 16
    s = np.zeros(1 << n, dtype=np.complex128)
    nz = [0, 8191, 8192, 40000, 65535]
    s[nz] = [0.3, 0.4j, -0.5, 0.2 + 0.1j, 0.6]
    s /= np.linalg.norm(s)
    c = canonical_draws(oracle, 3, 5000) + [0.0, float(np.nextafter(1.0, 0.0))]
    st = DeviceState.from_host(s)
    note("gpu events sparse: nonzeros across chunk boundaries")
    gpu = brk.generate_events(st.ptr, st.size, brk.C128, c)
    check_events(s, c, gpu, oracle.generate_events(s, c))
    assert set(gpu) == set(nz) and gpu[-2] == 0 and gpu[-1] == 65535
    st.free()
    # all the probability in the first chunk: a target at the very top stays inside it
    s = np.zeros(1 << n, dtype=np.complex128)
    s[[1, 5, 4000]] = [0.6, 0.0, 0.8]
    st = DeviceState.from_host(s)
    note("gpu events sparse: one populated chunk")
    gpu = brk.generate_events(st.ptr, st.size, brk.C128, c)
    assert set(gpu) == {1, 4000} and gpu[-2] == 1 and gpu[-1] == 4000
    check_events(s, c, gpu, oracle.generate_events(s, c))
    st.free()


def test_generate_events_large(oracle):
    n = 26
    note(f"gpu events n={n} (1 GiB state, 20000 events)")
    rng = np.random.default_rng(26)
    s = D.random_state(n, rng)
    c = canonical_draws(oracle, 26, 20000)
    st = DeviceState.from_host(s)
    gpu = brk.generate_events(st.ptr, st.size, brk.C128, c)
    check_events(s, c, gpu, oracle.generate_events(s, c))
    st.free()


def test_invalid_requests():
    st = DeviceState.from_host(np.array([1, 0, 0, 0], dtype=np.complex128))
    for bad in (1.0, -0.5, float("nan")):
        with pytest.raises(brk.BrkError) as e:
            brk.generate_events(st.ptr, st.size, brk.C128, [0.5, bad])
        assert e.value.status == 1  # BRK_ERR_INVALID_ARGUMENT
    with pytest.raises(brk.BrkError):
        brk.project_qubit(st.ptr, st.size, brk.C128, 2, 0, 1.0)
    with pytest.raises(brk.BrkError):
        brk.project_qubit(st.ptr, st.size, brk.C128, 0, 2, 1.0)
    st.upload(np.zeros(4, dtype=np.complex128))
    note("gpu events: zero-norm state")
    with pytest.raises(brk.BrkError) as e:
        brk.generate_events(st.ptr, st.size, brk.C128, [0.5])
    assert "zero norm" in str(e.value)
    st.free()


def test_projective_measurement_sequence(oracle):
    """M on every qubit in turn with bra's draws: outcomes and the collapsed state match."""
    n = 14
    rng = np.random.default_rng(9)
    s = D.random_state(n, rng)
    st = DeviceState.from_host(s)
    g = oracle.mt64(1)
    for q in (3, 0, 13, 7):
        u = g.canonical()
        p = brk.zero_one_probabilities(st.ptr, st.size, brk.C128, q)
        outcome = 0 if u * (p[0] + p[1]) < p[0] else 1
        brk.project_qubit(st.ptr, st.size, brk.C128, q, outcome, p[outcome] ** -0.5)
        assert oracle.projective_measurement(s, q, u) == outcome
    note("gpu projective measurement sequence: compare states")
    assert np.max(np.abs(st.download() - s)) < 1e-12
    assert abs(np.vdot(s, s).real - 1) < 1e-12
    st.free()


@pytest.mark.parametrize("n,m,divisor,base,dtype", [(12, 8, 15, 7, brk.C128), (20, 10, 1021, 3, brk.C128),
                                                    (24, 12, 4093, 5, brk.C128), (10, 10, 5, 3, brk.C128),
                                                    (9, 0, 7, 3, brk.C128), (18, 9, 511, 2, brk.C64)])
def test_shor_box_exact(oracle, n, m, divisor, base, dtype):
    note(f"gpu shor box n={n} m={m} divisor={divisor} base={base} dtype={dtype}")
    st = DeviceState(n, dtype)
    brk.fill_modular(st.ptr, st.size, dtype, (1e-3, 2e-3), (3, 1, 97, 5, 2, 89))  # overwritten entirely
    brk.shor_box(st.ptr, st.size, dtype, m, divisor, base)
    ref = np.zeros(1 << n, dtype=np.complex128)
    oracle.shor_box(ref, m, divisor, base)
    out = st.download()
    if dtype == brk.C128:
        assert np.array_equal(out.view(np.uint64), ref.view(np.uint64))
    else:  # the coefficient is std::pow in float: equal to the rounded double up to an ulp
        assert np.array_equal(out != 0, ref != 0)
        assert np.max(np.abs(out - ref)) < 1e-8
    st.free()


def test_shor_box_errors():
    st = DeviceState(4)
    for args, status in (((5, 15, 7), 1), ((2, 0, 7), 1), ((2, 2**40, 2**30), 8)):
        with pytest.raises(brk.BrkError) as e:
            brk.shor_box(st.ptr, st.size, brk.C128, *args)
        assert e.value.status == status
    st.free()
