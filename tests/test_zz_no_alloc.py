"""brk_generate_events_noalloc (sampling without device allocation: its
buffers are a static device array) against brk_generate_events, through the
C ABI.  Events are indices, so the bar is equality: the same events, bit for
bit, from both entry points on the same state and draws.  Two calls in a row
share the device array, the second with a smaller state and fewer draws, so a
value left over from the first call cannot stand in for one the second call
must compute.  The oracle's sequential scan is checked beside it, as for
brk_generate_events."""
import sys

import numpy as np
import pytest

import brk
import dense as D
from devstate import DeviceState

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
        assert np.all(np.abs(scan[lo:hi] - r) <= TOL * scan[-1]), (k, gpu[k], ora[k])
    assert diff.size <= max(2, len(canonical) // 1000)


def _both(s, dtype, draws):
    st = DeviceState.from_host(s, dtype)
    try:
        ref = brk.generate_events(st.ptr, st.size, dtype, draws)
        got = brk.generate_events_noalloc(st.ptr, st.size, dtype, draws)
        # the state is left unchanged
        assert np.array_equal(st.download().astype(np.complex128), s)
    finally:
        st.free()
    return ref, got


@pytest.mark.gpu
def test_noalloc_events_two_calls_second_smaller(oracle):
    rng = np.random.default_rng(71)
    # first call: 2^22 amplitudes (512 chunks: eight full 64-chunk windows of
    # the running sums), 8192 draws
    s1 = D.random_state(22, rng)
    c1 = canonical_draws(oracle, 71, 8192)
    note("noalloc events: call 1, n=22 c128, 8192 draws")
    ref1, got1 = _both(s1, brk.C128, c1)
    assert got1 == ref1
    check_events(s1, c1, got1, oracle.generate_events(s1, c1))
    # second call, smaller: 2^16 complex64 amplitudes (8 chunks: a partial
    # window), 1000 draws
    s2 = D.random_state(16, rng).astype(np.complex64).astype(np.complex128)
    c2 = canonical_draws(oracle, 72, 1000)
    note("noalloc events: call 2, n=16 c64, 1000 draws")
    ref2, got2 = _both(s2, brk.C64, c2)
    assert got2 == ref2
    check_events(s2, c2, got2, oracle.generate_events(s2, c2))


@pytest.mark.gpu
@pytest.mark.parametrize("n,dtype,count", [(1, brk.C128, 50), (12, brk.C128, 500), (19, brk.C128, 3000),
                                           (20, brk.C64, 3000), (26, brk.C128, 20000)])
def test_noalloc_events_sizes(oracle, n, dtype, count):
    note(f"noalloc events: n={n} dtype={dtype} draws={count}")
    rng = np.random.default_rng(300 + n)
    s = D.random_state(n, rng)
    if dtype == brk.C64:
        s = s.astype(np.complex64).astype(np.complex128)
    c = canonical_draws(oracle, n, count)
    ref, got = _both(s, dtype, c)
    assert got == ref


@pytest.mark.gpu
def test_noalloc_events_sparse_and_edges(oracle):
    # probability only in a few chunks, zero chunks at the end; the draws 0 and
    # 1 - 2^-53 select the first and the last index holding probability
    n = 16
    s = np.zeros(1 << n, dtype=np.complex128)
    s[[0, 8191, 8192, 40000]] = [0.3, 0.4j, -0.5, 0.2 + 0.1j]
    s /= np.linalg.norm(s)
    c = canonical_draws(oracle, 3, 5000) + [0.0, float(np.nextafter(1.0, 0.0))]
    note("noalloc events: sparse state, trailing zero chunks")
    ref, got = _both(s, brk.C128, c)
    assert got == ref
    assert got[-2] == 0 and got[-1] == 40000


@pytest.mark.gpu
def test_noalloc_events_errors():
    st = DeviceState.from_host(np.zeros(1 << 14, dtype=np.complex128))
    try:
        note("noalloc events: zero-norm state")
        for entry in ("brk_generate_events", "brk_generate_events_noalloc"):
            with pytest.raises(brk.BrkError) as e:
                brk.generate_events(st.ptr, st.size, brk.C128, [0.5], entry=entry)
            assert "zero norm" in str(e.value)
        st.upload(np.full(1 << 14, 2.0 ** -7, dtype=np.complex128))
        # the scratch holds the failed call's sums; the next call is still exact
        ref = brk.generate_events(st.ptr, st.size, brk.C128, [0.25, 0.75])
        assert brk.generate_events_noalloc(st.ptr, st.size, brk.C128, [0.25, 0.75]) == ref == [4096, 12288]
        with pytest.raises(brk.BrkError) as e:
            brk.generate_events_noalloc(st.ptr, st.size, brk.C128, [0.5] * (65536 + 1))
        assert e.value.status == 1 and "at most 65536 events" in str(e.value)
    finally:
        st.free()


def test_noalloc_limits_checked_on_host():
    # rejected before any device call: a state above 2^32 amplitudes (the
    # pointer is never dereferenced)
    with pytest.raises(brk.BrkError) as e:
        brk.generate_events_noalloc(0x1000, 1 << 33, brk.C64, [0.5])
    assert e.value.status == 1 and "at most 4294967296 amplitudes" in str(e.value)
    with pytest.raises(brk.BrkError) as e:
        brk.generate_events_noalloc(0x1000, 1 << 10, brk.C128, [0.5] * 65537)
    assert e.value.status == 1 and "at most 65536 events" in str(e.value)
