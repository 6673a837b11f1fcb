"""GPU checks of the precision form (fused complex64 blocks on MFMA).

Against the FP64 oracle run of the same circuit: infidelity 1 - |<psi_ref|psi>|^2
<= 1e-6 (the target the workload states) and max-abs amplitude error <= 2e-6 on
normalised states.  Against the plan's own FP64 block matrices (applied by the
numpy model) the error isolates the FP32 GEMM: <= 2e-6.
"""
import sys

import numpy as np
import pytest

import brk
import circuits as W
import dense as D
from devstate import DeviceState
from oracle_circuits import run_ops

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _run_fused(plan, host):
    st = DeviceState.from_host(host.astype(np.complex64), brk.C64)
    try:
        plan.apply(st.ptr)
        return st.download().astype(np.complex128)
    finally:
        st.free()


@pytest.mark.parametrize("n,width", [(11, 4), (11, 5), (11, 6), (14, 5), (16, 4), (16, 6), (18, 5)])
def test_fused_vs_oracle(oracle, n, width):
    note(f"gpu fused n={n} width={width}")
    ops = W.random_circuit(n, 120, 6, seed=n * 10 + width)
    plan = brk.FusedPlan(brk.Circuit(n, ops), width)
    s = D.random_state(n, np.random.default_rng(n + width))
    expected = s.copy()
    run_ops(oracle, expected, ops)
    got = _run_fused(plan, s)
    infidelity = 1 - abs(np.vdot(expected, got)) ** 2 / np.vdot(got, got).real
    assert infidelity < 1e-6, infidelity
    assert np.max(np.abs(got - expected)) < 2e-6


@pytest.mark.parametrize("width", [4, 5, 6])
def test_fused_single_blocks_vs_fp64_matrices(width):
    # one block at a time: every block qubit position pattern the plan produced
    n = 13
    ops = W.random_circuit(n, 40, 2, seed=77 + width)
    plan = brk.FusedPlan(brk.Circuit(n, ops), width)
    s = D.random_state(n, np.random.default_rng(width))
    expected = s.copy()
    for i in range(plan.num_blocks):
        q, _, m = plan.block(i)
        expected = D.apply(expected, m, q)
    note(f"gpu fused blocks width={width} ({plan.num_blocks} blocks)")
    got = _run_fused(plan, s)
    assert np.max(np.abs(got - expected)) < 2e-6


def test_fused_rejects_c128():
    n = 11
    plan = brk.FusedPlan(brk.Circuit(n, [("H", (0,), (), ())]), 4)
    st = DeviceState(n)
    with pytest.raises(brk.BrkError):
        plan.apply(st.ptr, brk.C128)
    st.free()


@pytest.mark.parametrize("n,width", [(12, 5), (15, 4), (17, 5)])
def test_staged_and_register_direct_kernels_agree(n, width):
    # the two kernel forms run the same MFMA chain per column: identical results
    ops = W.random_circuit(n, 80, 4, seed=5 * n + width)
    plan = brk.FusedPlan(brk.Circuit(n, ops), width)
    s = D.random_state(n, np.random.default_rng(n))
    note(f"gpu fused staged vs register-direct n={n} width={width}")
    old = brk.get_tune("fused_chunks")
    brk.tune("fused_chunks", 0)  # one launch per block
    try:
        brk.tune("fused_staged", 0)
        try:
            direct = _run_fused(plan, s)
        finally:
            brk.tune("fused_staged", 1)
        staged = _run_fused(plan, s)
    finally:
        brk.tune("fused_chunks", old)
    assert np.array_equal(direct, staged)


@pytest.mark.parametrize("n,width,seed", [(13, 4, 1), (13, 5, 2), (16, 4, 3), (17, 5, 4), (20, 4, 5), (20, 5, 6)])
def test_chunk_passes_match_block_launches(oracle, n, width, seed):
    """Chunk passes (several blocks per LDS-resident chunk, Gauss's three-product complex GEMM)
    against one launch per block (the real-embedding GEMM): the same blocks in FP32 with another
    product, equal within FP32 rounding (2e-6 max-abs on the unit-norm state); both against the
    oracle: infidelity <= 1e-6 and max-abs <= 2e-6."""
    ops = W.random_circuit(n, 160, 6, seed=seed)
    plan = brk.FusedPlan(brk.Circuit(n, ops), width)
    assert plan.num_passes > 0 and plan.num_passes < plan.num_blocks
    s = D.random_state(n, np.random.default_rng(seed))
    note(f"gpu fused chunk passes n={n} width={width}: {plan.num_blocks} blocks in {plan.num_passes} passes")
    old = brk.get_tune("fused_chunks")
    try:
        brk.tune("fused_chunks", 1)
        chunked = _run_fused(plan, s)
        brk.tune("fused_chunks", 0)
        per_block = _run_fused(plan, s)
    finally:
        brk.tune("fused_chunks", old)
    assert np.max(np.abs(chunked - per_block)) < 2e-6
    expected = s.copy()
    run_ops(oracle, expected, ops)
    for got in (chunked, per_block):
        infidelity = 1 - abs(np.vdot(expected, got)) ** 2 / np.vdot(got, got).real
        assert infidelity < 1e-6, infidelity
        assert np.max(np.abs(got - expected)) < 2e-6
