"""GPU checks of the differentiation form (adjoint gradient through the C ABI).

The adjoint method has no reference counterpart; it is pinned by
  * a closed form: E(theta) = h <X> = h sin(theta) for RY(theta)|0> on one qubit,
  * the parameter-shift rule on the FP64 plain per-gate path (exact for
    rotations exp(-i theta G / 2): dE = (E(theta + pi/2) - E(theta - pi/2)) / 2),
  * central finite differences with step 1e-5 on the FP64 plain path,
  * the CPU oracle's energy,
and wrapped as a torch.autograd.Function (test harness only) whose backward
must return the adjoint gradient.
"""
import math
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


def plain_energy(n, ops, J, h, psi0=None):
    st = DeviceState(n)
    if psi0 is None:
        brk.set_basis_state(st.ptr, st.size, brk.C128, 0)
    else:
        st.upload(psi0)
    brk.Circuit(n, ops).run(st.ptr)
    e = brk.expectation_ising(st.ptr, st.size, brk.C128, J, h)
    st.free()
    return e


def adjoint(n, ops, J, h, dtype=brk.C128, psi0=None):
    psi, work = DeviceState(n, dtype), DeviceState(n, dtype)
    if psi0 is None:
        brk.set_basis_state(psi.ptr, psi.size, dtype, 0)
    else:
        psi.upload(psi0)
    e, g = brk.adjoint_ising_gradient(brk.Circuit(n, ops), psi.ptr, work.ptr, dtype, J, h)
    back = psi.download()
    psi.free()
    work.free()
    return e, g, back


def test_closed_form_one_qubit():
    note("gpu adjoint: closed form")
    th, hx = 0.7, 0.9
    ops = [("RY", (0,), (), (th,))]
    e, g, _ = adjoint(1, ops, [], [hx])
    assert abs(e - hx * math.sin(th)) < 1e-14
    assert abs(g[0] - hx * math.cos(th)) < 1e-14


@pytest.mark.parametrize("dtype", [brk.C128, brk.C64])
def test_small_ansatz_vs_shift_fd_and_oracle(oracle, dtype):
    n, layers = 10, 3
    rng = np.random.default_rng(3)
    ops = W.hea_layers(n, layers, rng) + [("RX", (2,), (), (0.4,)), ("H", (5,), (), ()), ("X", (7,), (3,), ())]
    J, h = rng.uniform(-1, 1, n - 1), rng.uniform(-1, 1, n)
    theta = np.array(W.parameters(ops))
    note(f"gpu adjoint n={n} params={theta.size} dtype={dtype}")
    e, g, back = adjoint(n, ops, J, h, dtype)
    # energy vs the oracle
    s = np.zeros(1 << n, dtype=np.complex128)
    s[0] = 1
    run_ops(oracle, s, ops)
    e_ref = oracle.expectation_ising(s, J, h)
    tol = 1e-12 if dtype == brk.C128 else 2e-5
    assert abs(e - e_ref) < tol
    # the sweep un-computes psi back to |0..0>
    assert abs(back[0] - 1) < (1e-12 if dtype == brk.C128 else 1e-5)
    # parameter shift on the FP64 plain path (all parameters are exp(-i t G/2) rotations)
    shift = np.empty_like(theta)
    fd = np.empty_like(theta)
    for k in range(theta.size):
        tp, tm = theta.copy(), theta.copy()
        tp[k] += math.pi / 2
        tm[k] -= math.pi / 2
        shift[k] = 0.5 * (plain_energy(n, W.with_parameters(ops, tp), J, h) -
                          plain_energy(n, W.with_parameters(ops, tm), J, h))
        tp, tm = theta.copy(), theta.copy()
        tp[k] += 1e-5
        tm[k] -= 1e-5
        fd[k] = (plain_energy(n, W.with_parameters(ops, tp), J, h) -
                 plain_energy(n, W.with_parameters(ops, tm), J, h)) / 2e-5
    if dtype == brk.C128:
        assert np.max(np.abs(g - shift)) < 1e-12
    else:
        assert np.max(np.abs(g - shift)) < 5e-5
    assert np.max(np.abs(g - fd)) < 5e-5  # FD truncation + rounding at step 1e-5


def mixed_ansatz(n, rng, layers=3):
    """every parametric kind (RX RY RZ EX EY EZ U1) beside H, CNOT, CZ and SWAP"""
    ops = []
    for layer in range(layers):
        for q in range(n):
            kind = ("RX", "RY", "RZ", "EX", "EY", "EZ", "U1")[(q + layer) % 7]
            ops.append((kind, (q,), (), (float(rng.uniform(-math.pi, math.pi)),)))
        for q in range(layer % 2, n - 1, 2):
            ops.append(("Z", (q, q + 1), (), ()))
        ops.append(("H", (int(rng.integers(n)),), (), ()))
        a, b = (int(x) for x in rng.choice(n, 2, replace=False))
        ops.append(("X", (b,), (a,), ()))
        a, b = (int(x) for x in rng.choice(n, 2, replace=False))
        ops.append(("SWAP", (a, b), (), ()))
    return ops


@pytest.fixture(params=["temporal", "nontemporal"])
def access(request):
    """Default access policy, then the non-temporal instantiations large states take."""
    if request.param == "temporal":
        yield request.param
        return
    old = brk.get_tune("nt_min_bytes")
    brk.tune("nt_min_bytes", 0)
    try:
        yield request.param
    finally:
        brk.tune("nt_min_bytes", old)


@pytest.mark.parametrize("n,dtype", [(14, brk.C128), (13, brk.C64)])
def test_blocked_sweep_matches_chain_sweep(n, dtype, access):
    """The blocked reverse sweep (default) against the per-run passes (tuning adjoint_blocks = 0)."""
    rng = np.random.default_rng(20 + n)
    ops = mixed_ansatz(n, rng)
    J, h = rng.uniform(-1, 1, n - 1), rng.uniform(-1, 1, n)
    note(f"gpu adjoint blocked vs chain n={n} dtype={dtype} ({access})")
    e1, g1, back1 = adjoint(n, ops, J, h, dtype)
    brk.tune("adjoint_blocks", 0)
    try:
        e0, g0, back0 = adjoint(n, ops, J, h, dtype)
    finally:
        brk.tune("adjoint_blocks", 1)
    tol = 1e-12 if dtype == brk.C128 else 2e-5
    assert abs(e1 - e0) < tol
    assert np.max(np.abs(g1 - g0)) < tol
    assert abs(back1[0] - 1) < (1e-12 if dtype == brk.C128 else 1e-5)
    if dtype == brk.C128:
        theta = np.array(W.parameters(ops))
        for k in (0, 5, theta.size // 2, theta.size - 1):
            kind = [op for op in ops if op[0] in W.PARAMETRIC][k][0]
            tp, tm = theta.copy(), theta.copy()
            # exp(-i t G/2) rotations: shift pi/2; exp(i p G): shift pi/4 with factor 1; U1: central FD
            if kind in ("RX", "RY", "RZ"):
                tp[k] += math.pi / 2
                tm[k] -= math.pi / 2
                ref = 0.5 * (plain_energy(n, W.with_parameters(ops, tp), J, h) -
                             plain_energy(n, W.with_parameters(ops, tm), J, h))
                assert abs(g1[k] - ref) < 1e-11
            else:
                tp[k] += 1e-5
                tm[k] -= 1e-5
                ref = (plain_energy(n, W.with_parameters(ops, tp), J, h) -
                       plain_energy(n, W.with_parameters(ops, tm), J, h)) / 2e-5
                assert abs(g1[k] - ref) < 1e-7


@pytest.mark.slow
def test_24q_12_layers_tfim():
    # the differentiation workload: 24 qubits, 12 layers (576 parameters), J = 1.0, h = 0.7
    n, layers = 24, 12
    ops, theta = W.adjoint_ansatz(n, layers, seed=5)
    theta = np.array(theta)
    assert theta.size == 576
    J, h = W.tfim(n)
    note("gpu adjoint 24q: adjoint sweep")
    e, g, _ = adjoint(n, ops, J, h)
    assert abs(e - plain_energy(n, ops, J, h)) < 1e-10
    rng = np.random.default_rng(1)
    for k in sorted(rng.choice(theta.size, 4, replace=False)):
        note(f"gpu adjoint 24q: parameter {k}")
        tp, tm = theta.copy(), theta.copy()
        tp[k] += math.pi / 2
        tm[k] -= math.pi / 2
        shift = 0.5 * (plain_energy(n, W.with_parameters(ops, tp), J, h) -
                       plain_energy(n, W.with_parameters(ops, tm), J, h))
        assert abs(g[k] - shift) < 1e-10
        tp, tm = theta.copy(), theta.copy()
        tp[k] += 1e-5
        tm[k] -= 1e-5
        fd = (plain_energy(n, W.with_parameters(ops, tp), J, h) -
              plain_energy(n, W.with_parameters(ops, tm), J, h)) / 2e-5
        assert abs(g[k] - fd) < 1e-7


@pytest.mark.slow
@pytest.mark.timeout(600)
def test_24q_12_layers_full_gradient_vs_parameter_shift():
    """All 576 parameters of the differentiation workload against the parameter-shift rule
    on the FP64 path: each shifted energy runs the circuit in gate blocks, which give the
    per-gate path's amplitudes bit for bit."""
    n, layers = 24, 12
    ops, theta = W.adjoint_ansatz(n, layers, seed=5)
    theta = np.array(theta)
    J, h = W.tfim(n)
    note("gpu adjoint 24q full gradient: adjoint sweep")
    e, g, _ = adjoint(n, ops, J, h)
    st = DeviceState(n)

    def energy(values):
        brk.set_basis_state(st.ptr, st.size, brk.C128, 0)
        brk.Circuit(n, W.with_parameters(ops, values)).run_blocked(st.ptr)
        return brk.expectation_ising(st.ptr, st.size, brk.C128, J, h)

    try:
        assert abs(e - energy(theta)) < 1e-10
        shift = np.empty_like(theta)
        for k in range(theta.size):
            if k % 64 == 0:
                note(f"gpu adjoint 24q full gradient: parameters {k}..{min(k + 63, theta.size - 1)}")
            tp, tm = theta.copy(), theta.copy()
            tp[k] += math.pi / 2
            tm[k] -= math.pi / 2
            shift[k] = 0.5 * (energy(tp) - energy(tm))
    finally:
        st.free()
    err = float(np.max(np.abs(g - shift)))
    note(f"gpu adjoint 24q full gradient: max |adjoint - parameter shift| over 576 = {err:.3e}")
    assert err <= 1e-10


def test_torch_autograd_wrapper():
    torch = pytest.importorskip("torch")
    n = 8
    rng = np.random.default_rng(11)
    ops = W.hea_layers(n, 2, rng)
    J, h = rng.uniform(-1, 1, n - 1), rng.uniform(-1, 1, n)

    class IsingEnergy(torch.autograd.Function):
        """E(theta) through the C ABI; backward uses the adjoint gradient."""

        @staticmethod
        def forward(ctx, theta):
            e, g, _ = adjoint(n, W.with_parameters(ops, theta.detach().cpu().numpy()), J, h)
            ctx.save_for_backward(torch.from_numpy(g))
            return torch.tensor(e, dtype=torch.float64)

        @staticmethod
        def backward(ctx, grad_out):
            (g,) = ctx.saved_tensors
            return grad_out * g

    theta = torch.tensor(W.parameters(ops), dtype=torch.float64, requires_grad=True)
    note("gpu adjoint: torch.autograd wrapper")
    energy = IsingEnergy.apply(theta)
    energy.backward()
    # central differences of the same Function
    k = 5
    with torch.no_grad():
        tp, tm = theta.clone(), theta.clone()
        tp[k] += 1e-5
        tm[k] -= 1e-5
        fd = (IsingEnergy.apply(tp) - IsingEnergy.apply(tm)).item() / 2e-5
    assert abs(theta.grad[k].item() - fd) < 1e-7
    assert abs(energy.item() - plain_energy(n, ops, J, h)) < 1e-12


def test_rejects_local_swaps_and_controlled_parameters():
    with pytest.raises(brk.BrkError):
        adjoint(4, [("LOCAL_SWAP", (3, 0), (), ()), ("RY", (0,), (), (0.1,))], [0.1] * 3, [0.2] * 4)
    with pytest.raises(brk.BrkError):
        adjoint(4, [("RY", (0,), (1,), (0.1,))], [0.1] * 3, [0.2] * 4)


def test_op_wider_than_a_block_uses_the_run_sweep():
    """A gate on more qubits than a reverse block holds (X with 10 controls at 14 qubits:
    11 qubits besides the 3 line qubits) sends the gradient to the per-run sweep: the result
    equals the one with blocks switched off, and parameter shift."""
    n = 14
    rng = np.random.default_rng(31)
    ops = mixed_ansatz(n, rng, layers=2)
    ops.insert(len(ops) // 2, ("X", (3,), tuple(range(4, 14)), ()))
    J, h = rng.uniform(-1, 1, n - 1), rng.uniform(-1, 1, n)
    note("gpu adjoint: op wider than a block")
    e1, g1, _ = adjoint(n, ops, J, h)
    brk.tune("adjoint_blocks", 0)
    try:
        e0, g0, _ = adjoint(n, ops, J, h)
    finally:
        brk.tune("adjoint_blocks", 1)
    assert e1 == e0 and np.array_equal(g1, g0)
    theta = np.array(W.parameters(ops))
    k = 1  # an RY (mixed_ansatz cycles RX, RY, RZ, ... over the qubits)
    assert [op for op in ops if op[0] in W.PARAMETRIC][k][0] == "RY"
    tp, tm = theta.copy(), theta.copy()
    tp[k] += math.pi / 2
    tm[k] -= math.pi / 2
    ref = 0.5 * (plain_energy(n, W.with_parameters(ops, tp), J, h) - plain_energy(n, W.with_parameters(ops, tm), J, h))
    assert abs(g1[k] - ref) < 1e-11
