#!/usr/bin/env python3
"""Benchmark: 32-qubit plain per-gate complex128 circuit on one MI355X.

Workload: the seeded random circuit (200 gates, 60 % single-qubit from
{H, X, RZ, U3}, 40 % CNOT/CZ, plus 8 local qubit swaps) on 2^32 complex128
amplitudes (64 GiB) resident in HBM.  A step is one pass of the circuit.

  value      = sum over the circuit's ops of the algorithmic bytes
               2 * 16 B * (amplitudes the op can change), divided by the step time
  roofline   = the kernel instantiation with the most device time, its
               algorithmic bytes per launch over its mean launch time (HIP events
               recorded on the stream the engine launches on)
  cpu_baseline = the reference CPU case -- 20 qubits complex128, H on every
               qubit then 10 layers of RZ on each qubit and a CNOT brick (315
               gates) -- run by the CPU oracle (ket's algorithm, -Ofast -fopenmp
               like bra's nompi-release build, all host cores): gate loop only,
               1 warm-up, median of 5; the GPU's per-gate and gate-block times
               for the same circuit beside it.  cpu_baseline_30q: the oracle on
               the benchmark circuit's generator at 30 qubits (bounded sample)
  fused_c64  = the precision form on the same circuit: greedy fusion into dense
               4/5/6-qubit blocks (width autotuned by timing each once), complex64
               state, FP32 MFMA GEMMs; circuit time, speedup over the complex128
               per-gate path and infidelity 1 - |<psi_c128|psi_c64>|^2 (FP64)
  batch_c64  = the batching form: 4096 ansatz circuits of 10..20 qubits in one
               graph-captured persistent launch; circuits/s, p50/p99 latency
  adjoint_c128 = the differentiation form: adjoint gradient of a 24-qubit,
               576-parameter ansatz; gradients/s, max error vs parameter shift over
               all 576 parameters and vs central differences on a sample

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--qubits 32] [--no-cpu] [--no-fused]
                       [--no-batch] [--no-adjoint]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "riken-braket_amd", "py"))

METRIC = "32q per-gate HBM GB/s (% peak); fused-c64 vs c128 circuit time at 1-F≤1e-6"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


LINE_BITS = 2  # 64-byte lines of complex128 amplitudes (the engine's default line_bytes)
LANE_BITS = 6  # wavefront of 64 lanes


def kernel_name(op, perm, n):
    """The kernel template instantiation the engine launches for `op` (c128 state of n qubits,
    non-temporal access), named as rocprofv3 demangles it without the argument list.  Mirrors
    the dispatch in riken-braket_amd/csrc/brk_gates.hip."""
    name, t, c, _ = op
    if name == "LOCAL_SWAP":
        lo, hi = sorted(t)
        if hi < LANE_BITS and lo < LINE_BITS:
            return "k_lane<double, 1, 0, true, false, true>"
        if lo < LINE_BITS and hi >= LANE_BITS:
            return "k_lane_swap<double, 1, true, false>"
        return "k_pair<double, 1, 2, true, false>"
    pt = [perm[q] for q in t]
    pc = [perm[q] for q in c]
    K = lambda k: str(k if k <= 3 else -1)
    if name in ("Z", "U1", "RZ", "EZ"):
        ctrl = pt + pc if name in ("Z", "U1") else pc
        high = [q for q in ctrl if q >= LINE_BITS]
        select = "1" if name in ("RZ", "EZ") else "0"  # coefficient by target parity / one coefficient
        cond = "true" if len(high) < len(ctrl) else "false"
        return f"k_diag<double, {K(len(high))}, {select}, true, {cond}>"
    mode = {"H": 0, "X": 1}.get(name, 2)
    if len(pt) == 1 and pt[0] < LINE_BITS and n >= LANE_BITS:
        high = [q for q in pc if q >= LANE_BITS]
        cond = "true" if len(high) < len(pc) else "false"
        return f"k_lane<double, {mode}, {K(len(high))}, true, {cond}, false>"
    high = [q for q in pc if q >= LINE_BITS]
    cond = "true" if len(high) < len(pc) else "false"
    return f"k_pair<double, {mode}, {K(1 + len(high))}, true, {cond}>"


def physical_kernel_names(ops, n):
    """kernel_name for every op, replaying the logical->physical permutation."""
    perm, inv = list(range(n)), list(range(n))
    names = []
    for op in ops:
        names.append(kernel_name(op, perm, n))
        if op[0] == "LOCAL_SWAP":
            a, b = inv[op[1][0]], inv[op[1][1]]
            inv[perm[a]], inv[perm[b]] = inv[perm[b]], inv[perm[a]]
            perm[a], perm[b] = perm[b], perm[a]
    return names


def kind_and_band(ops, n, line_qubits=3):
    """(gate kind, target band) for every op, replaying the logical->physical permutation.
    Kinds: the op name, CNOT for one-control X, CZ for Z on two qubits; band "low" when a physical qubit the
    op touches is one of the line_qubits lowest (they share 128-byte lines), else "high"."""
    perm, inv = list(range(n)), list(range(n))
    out = []
    for name, t, c, _ in ops:
        if name == "LOCAL_SWAP":
            phys = list(t)
            a, b = inv[t[0]], inv[t[1]]
            inv[perm[a]], inv[perm[b]] = inv[perm[b]], inv[perm[a]]
            perm[a], perm[b] = perm[b], perm[a]
        else:
            phys = [perm[q] for q in list(t) + list(c)]
        kind = name
        if name == "X" and len(t) == 1 and len(c) == 1:
            kind = "CNOT"
        elif name == "Z" and len(t) + len(c) == 2:  # the two-qubit phase flip, CZ
            kind = "CZ"
        out.append((kind, "low" if min(phys) < line_qubits else "high"))
    return out


def _host_description():
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return cores, model


def cpu_baseline_20q(stream, runs=5):
    """The reference's CPU case on the host: 20 qubits, H wall + 10 RZ / CNOT-brick layers (315
    gates), the CPU oracle built like bra's nompi-release (-Ofast -fopenmp), all host cores.
    Gate loop only, 1 warm-up, median of `runs`.  The same circuit on the GPU (per-gate and
    gate blocks, median of `runs` after a warm-up) goes beside it."""
    import numpy as np

    import brk
    import circuits as W
    import hip
    from devstate import DeviceState

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle import Oracle
    from oracle_circuits import run_ops

    n = 20
    ops = W.layered_rz_cnot(n, 10, seed=1)
    nbytes = sum(W.algorithmic_bytes(op, n) for op in ops)
    o = Oracle(fast=True)
    cores, model = _host_description()
    o.set_num_threads(cores)
    s = np.zeros(1 << n, dtype=np.complex128)
    times = []
    for k in range(runs + 1):
        s[:] = 0
        s[0] = 1
        t0 = time.perf_counter()
        run_ops(o, s, ops)
        if k:  # the first run is the warm-up
            times.append(time.perf_counter() - t0)
    cpu_s = float(np.median(times))
    circuit = brk.Circuit(n, ops)
    plan = brk.BlockPlan(circuit, brk.C128)
    st = DeviceState(n)
    gpu = {"per_gate": [], "blocks": []}
    for k in range(runs + 1):
        for key, run in (("per_gate", lambda: circuit.run(st.ptr, stream=stream)),
                         ("blocks", lambda: plan.run(st.ptr, stream))):
            brk.set_basis_state(st.ptr, st.size, brk.C128, 0, stream)
            hip.synchronize(stream)
            e0, e1 = hip.Event(), hip.Event()
            e0.record(stream)
            run()
            e1.record(stream)
            e1.synchronize()
            if k:
                gpu[key].append(e0.elapsed_ms(e1))
    got = st.download()
    st.free()
    err = float(np.max(np.abs(got - s)))
    gpu_ms = {key: float(np.median(v)) for key, v in gpu.items()}
    log(f"cpu baseline 20q: {cpu_s * 1e3:.1f} ms on {cores} threads; GPU per-gate {gpu_ms['per_gate']:.3f} ms, "
        f"blocks {gpu_ms['blocks']:.3f} ms")
    return {
        "value": round(nbytes / cpu_s / 1e9, 2),
        "unit": "GB/s",
        "cores": cores,
        "kind": "port",
        "sample": f"the reference CPU case, whole: 20 qubits complex128, H on every qubit then 10 layers of "
                  f"RZ on each qubit + CNOT brick ({len(ops)} gates, seed 1), oracle -Ofast -fopenmp, "
                  f"{cores} threads, gate loop only, median of {runs} after 1 warm-up; host {model}",
        "seconds": round(cpu_s, 5),
        "algorithmic_GB": round(nbytes / 1e9, 3),
        "nproc": os.cpu_count(),
        "cpu_model": model,
        "gpu_same_circuit_ms": {k: round(v, 4) for k, v in gpu_ms.items()},
        "gpu_speedup": {k: round(cpu_s * 1e3 / v, 1) for k, v in gpu_ms.items()},
        "gpu_max_abs_diff_vs_cpu": err,
    }


def cpu_baseline_30q(n_cpu=30, budget_s=15.0, seed=2):
    """The benchmark circuit's generator at 30 qubits (16 GiB; 32 qubits need more host
    memory than a bounded sample should take) on the CPU oracle, first ops within budget_s."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import circuits as W
    from oracle import Oracle
    from oracle_circuits import run_ops

    o = Oracle(fast=True)
    cores, model = _host_description()
    o.set_num_threads(cores)
    ops = W.random_circuit(n_cpu, 200, 8, seed=seed)
    s = np.zeros(1 << n_cpu, dtype=np.complex128)
    s[0] = 1
    # warm-up: first op (page faults of the fresh array)
    run_ops(o, s, ops[:1])
    s[:] = 0
    s[0] = 1
    t0 = time.perf_counter()
    _, done = run_ops(o, s, ops, time_budget=budget_s)
    dt = time.perf_counter() - t0
    nbytes = sum(W.algorithmic_bytes(op, n_cpu) for op in ops[:done])
    return {
        "value": round(nbytes / dt / 1e9, 2),
        "unit": "GB/s",
        "cores": cores,
        "kind": "port",
        "sample": f"first {done} of {len(ops)} ops of the benchmark circuit's generator at {n_cpu} qubits "
                  f"(2^{n_cpu} c128 = {2 ** n_cpu * 16 / 2 ** 30:.0f} GiB), {dt:.1f} s, "
                  f"oracle -Ofast -fopenmp, {cores} threads, {model}",
    }


def load_traffic(kernel, n):
    """HBM bytes per launch of `kernel` measured by the rocprofv3 FETCH_SIZE / WRITE_SIZE
    runs recorded in profiles/traffic.json (null when that file does not cover it)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        return None
    entry = data.get("kernels", {}).get(kernel)
    if not entry or data.get("qubits") != n:
        return None
    return {"bytes_per_launch": entry["bytes_per_launch"], "source": "profiles/traffic.json"}


def fused_leg(circuit, plain_state, args, stream):
    """Precision form on the same circuit: plans of width 4, 5, 6 are each timed once on a
    complex64 copy of |0..0> (the autotuner), the fastest is timed over `steps` runs, and
    its final state is compared with the complex128 plain-path state `plain_state`
    (same physical layout) through an FP64 inner product."""
    import time as _time

    import brk
    import hip
    from devstate import DeviceState

    n = circuit.num_qubits
    st = DeviceState(n, brk.C64)
    widths = {}
    plans = {}
    per_launch = {}
    for w in (4, 5, 6):
        t0 = _time.perf_counter()
        plans[w] = brk.FusedPlan(circuit, w)
        plan_ms = (_time.perf_counter() - t0) * 1e3
        brk.set_basis_state(st.ptr, st.size, brk.C64, 0, stream)
        ms = plans[w].apply_timed(st.ptr, stream=stream)
        per_launch[w] = [round(x, 2) for x in ms]
        widths[w] = {"blocks": plans[w].num_blocks, "chunk_passes": plans[w].num_passes, "ms": round(sum(ms), 2),
                     "plan_ms": round(plan_ms, 1)}
        log(f"fused width {w}: {plans[w].num_blocks} blocks in {plans[w].num_passes or plans[w].num_blocks} "
            f"launches, {sum(ms):.1f} ms")
    best = min(widths, key=lambda w: widths[w]["ms"])
    plan = plans[best]
    run_ms = []
    for _ in range(max(1, args.steps)):
        brk.set_basis_state(st.ptr, st.size, brk.C64, 0, stream)
        hip.synchronize(stream)
        e0, e1 = hip.Event(), hip.Event()
        e0.record(stream)
        plan.apply(st.ptr, stream=stream)
        e1.record(stream)
        e1.synchronize()
        run_ms.append(e0.elapsed_ms(e1))
    ip = brk.inner_product(st.ptr, brk.C64, plain_state.ptr, brk.C128, st.size, stream)
    n64 = brk.norm2(st.ptr, st.size, brk.C64, stream)
    n128 = brk.norm2(plain_state.ptr, st.size, brk.C128, stream)
    st.free()
    ms = sum(run_ms) / len(run_ms)
    blocks = plan.num_blocks
    passes = plan.num_passes if brk.get_tune("fused_chunks") else 0
    flops = 8 * (1 << best) * (1 << n) * blocks  # complex GEMM count: 8 real FLOPs per complex MAC
    hbm = 2 * 8 * (1 << n) * (passes or blocks)  # one read and one write of the c64 state per launch
    # per launch: the larger of its HBM time at 8 TB/s and its executed MFMA FLOPs at 157.3 TFLOP/s
    # (chunk passes run Gauss's three-product form, 6 real FLOPs per complex MAC; one launch per
    # block runs the real embedding, 8)
    launches = passes or blocks
    executed = (6 if passes else 8) * (1 << best) * (1 << n) * blocks
    bound_ms = max(hbm / launches / (HBM_PEAK_GBS * 1e9), executed / launches / 157.3e12) * 1e3
    launch_ms = ms / launches
    return {
        "width": best,
        "blocks": blocks,
        "chunk_passes": passes,
        "ms_per_circuit": round(ms, 2),
        "infidelity": 1.0 - abs(ip) ** 2 / (n64 * n128),
        "widths": widths,
        "mfma_tflops": round(flops / (ms * 1e-3) / 1e12, 1),
        "mfma_fraction_fp32": round(flops / (ms * 1e-3) / 157.3e12, 4),
        "flop_count": "8 * 2^width * 2^n per block (complex GEMM); chunk passes execute 6 * 2^width * 2^n "
                      "on MFMA (Gauss's three-product form)",
        "hbm_GBps": round(hbm / (ms * 1e-3) / 1e9, 1),
        "roofline_per_launch": {
            "kernel": "k_fused_chunk" if passes else "k_fused_staged / k_fused",
            "launch_ms": round(launch_ms, 3),
            "bound_ms": round(bound_ms, 3),
            "bound": "mfma" if executed / 157.3e12 > hbm / (HBM_PEAK_GBS * 1e9) else "hbm",
            "frac": round(bound_ms / launch_ms, 4),
            "each_ms_autotune_run": per_launch[best],
        },
        "dtype": "c64 state, FP32 MFMA (v_mfma_f32_16x16x4_f32 in chunk passes, v_mfma_f32_32x32x2_f32 in per-block launches), FP32 accumulate; FP64 fidelity",
    }


def blocked_leg(circuit, plain_state, args, stream):
    """The same circuit through gate blocks (brk_block_plan_run) on a second complex128 state.

    Amplitudes must equal the per-gate run bit for bit: 4096 seeded indices are compared
    exactly and the inner product with the per-gate state must be 1 within 1e-12."""
    import numpy as np

    import brk
    import hip
    from devstate import DeviceState
    n = circuit.num_qubits
    t0 = time.perf_counter()
    plan = brk.BlockPlan(circuit, brk.C128)
    plan_s = time.perf_counter() - t0
    blocks = [k for _, k, local in plan.steps if local]
    st = DeviceState(n)
    times = []
    for k in range(args.warmup + args.steps):
        brk.set_basis_state(st.ptr, st.size, brk.C128, 0, stream)
        hip.synchronize(stream)
        e0, e1 = hip.Event(), hip.Event()
        e0.record(stream)
        plan.run(st.ptr, stream)
        e1.record(stream)
        e1.synchronize()
        if k >= args.warmup:
            times.append(e0.elapsed_ms(e1))
            log(f"blocked step {k - args.warmup + 1}/{args.steps}: {times[-1]:.1f} ms")
    rng = np.random.default_rng(3)
    idx = sorted(set(int(i) for i in rng.integers(0, 1 << n, 4096)))
    same = bool(np.array_equal(st.download_indices(idx).view(np.uint64),
                               plain_state.download_indices(idx).view(np.uint64)))
    ip = brk.inner_product(st.ptr, brk.C128, plain_state.ptr, brk.C128, st.size, stream)
    # the same circuit in block order (brk_circuit_block_order): fewer blocks, same operator,
    # amplitudes equal to the per-gate run within rounding
    ordered = circuit.reordered(circuit.block_order(12, 3))
    plan2 = brk.BlockPlan(ordered, brk.C128)
    blocks2 = [k for _, k, local in plan2.steps if local]
    times2 = []
    for k in range(args.warmup + args.steps):
        brk.set_basis_state(st.ptr, st.size, brk.C128, 0, stream)
        hip.synchronize(stream)
        e0, e1 = hip.Event(), hip.Event()
        e0.record(stream)
        plan2.run(st.ptr, stream)
        e1.record(stream)
        e1.synchronize()
        if k >= args.warmup:
            times2.append(e0.elapsed_ms(e1))
            log(f"blocked (block order) step {k - args.warmup + 1}/{args.steps}: {times2[-1]:.1f} ms")
    diff2 = float(np.max(np.abs(st.download_indices(idx) - plain_state.download_indices(idx))))
    ip2 = brk.inner_product(st.ptr, brk.C128, plain_state.ptr, brk.C128, st.size, stream)
    st.free()
    ms = sum(times) / len(times)
    return {
        "ms_per_circuit": round(ms, 2),
        "blocks": len(blocks),
        "ops_in_blocks": int(sum(blocks)),
        "plain_steps": len(plan.steps) - len(blocks),
        "block_hbm_GB_per_launch": round(2 * 16 * (1 << n) / 1e9, 2),
        "plan_s": round(plan_s, 4),
        "sampled_amplitudes_bit_identical": same,
        "inner_product_error": abs(ip - 1.0),
        "block_order": {
            "ms_per_circuit": round(sum(times2) / len(times2), 2),
            "blocks": len(blocks2),
            "ops_in_blocks": int(sum(blocks2)),
            "max_abs_diff_sampled": diff2,
            "inner_product_error": abs(ip2 - 1.0),
        },
    }


def sampling_leg(state, args, stream):
    """Measurement and sampling on the circuit's final complex128 state (overwritten at the end):
    GENERATE EVENTS with 1 and 100 000 seeded canonical draws, one projective measurement
    (P(0), P(1) then the projection) and a Shor box.  Every event must land on an amplitude of
    non-zero probability."""
    import numpy as np

    import brk
    n = state.num_qubits
    rng = np.random.default_rng(4)
    out = {}
    for count in (1, 100000):
        c = rng.random(count)  # float64 array: passed to the C ABI without conversion
        brk.generate_events(state.ptr, state.size, brk.C128, c, stream)  # warm-up
        t0 = time.perf_counter()
        ev = brk.generate_events(state.ptr, state.size, brk.C128, c, stream)
        out[f"events_{count}_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
    sample = sorted(set(int(e) for e in ev[:256]))
    amps = state.download_indices(sample)
    out["events_on_nonzero_amplitudes"] = bool(np.all(np.abs(amps) > 0))
    out["events_per_s"] = round(100000 / (out["events_100000_ms"] * 1e-3), 1)
    t0 = time.perf_counter()
    p = brk.zero_one_probabilities(state.ptr, state.size, brk.C128, n // 2, stream)
    brk.project_qubit(state.ptr, state.size, brk.C128, n // 2, 0, p[0] ** -0.5, stream)
    brk.call("brk_stream_synchronize", stream)
    out["projective_measurement_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
    out["norm_after_projection_error"] = abs(brk.norm2(state.ptr, state.size, brk.C128, stream) - 1.0)
    t0 = time.perf_counter()
    brk.shor_box(state.ptr, state.size, brk.C128, n // 2, 65521, 7, stream)
    out["shor_box_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
    log(f"sampling: 100000 events in {out['events_100000_ms']} ms, projective measurement "
        f"{out['projective_measurement_ms']} ms, Shor box {out['shor_box_ms']} ms")
    return out


def plain_energy(n, ops, J, h, stream):
    import brk
    from devstate import DeviceState

    st = DeviceState(n)
    brk.set_basis_state(st.ptr, st.size, brk.C128, 0, stream)
    brk.Circuit(n, ops).run(st.ptr, stream=stream)
    e = brk.expectation_ising(st.ptr, st.size, brk.C128, J, h, stream)
    st.free()
    return e


def batch_leg(args, stream):
    """Batching form: 4096 hardware-efficient-ansatz circuits of 10..20 qubits (8 layers),
    complex64 states, one graph-captured persistent launch per batch; FP64 energies."""
    import numpy as np

    import brk
    import circuits as W

    items = W.batch_ansatz()
    t0 = time.perf_counter()
    batch = brk.Batch(items)
    plan_s = time.perf_counter() - t0
    energies = batch.run(stream)  # first run uploads the program and captures the graph
    runs = []
    for _ in range(max(1, args.steps)):
        e, ms, start, end = batch.run_timed(stream)
        runs.append((ms, end - start, end))
        assert np.array_equal(e, energies)
    ms = sum(r[0] for r in runs) / len(runs)
    service = np.concatenate([r[1] for r in runs])
    completion = np.concatenate([r[2] for r in runs])
    rng = np.random.default_rng(0)
    sample = sorted(int(i) for i in rng.choice(len(items), 8, replace=False))
    err = max(abs(energies[i] - plain_energy(*items[i], stream)) for i in sample)
    log(f"batch: {len(items)} circuits in {ms:.2f} ms ({len(items) / ms * 1e3:.0f} circuits/s)")
    return {
        "workload": "4096 circuits, qubits uniform in [10, 20], 8 layers of RY, RZ on each qubit + CZ ladder; "
                    "open-chain Ising energy, J, h ~ U[-1, 1], seed 4",
        "circuits_per_s": round(len(items) / ms * 1e3, 1),
        "ms_per_batch": round(ms, 3),
        "latency_ms": {
            "service_p50": round(float(np.percentile(service, 50)), 4),
            "service_p99": round(float(np.percentile(service, 99)), 4),
            "completion_p50": round(float(np.percentile(completion, 50)), 4),
            "completion_p99": round(float(np.percentile(completion, 99)), 4),
        },
        "latency_definition": "service: from a workgroup taking the circuit off the queue to its energy "
                              "being written; completion: from the first take in the launch to that write",
        "plan_s": round(plan_s, 2),
        "max_abs_energy_err_vs_c128_plain": err,
        "energy_check_sample": len(sample),
        "dtype": "c64 state, FP32 gate arithmetic, FP64 energy accumulation",
    }


def adjoint_leg(args, stream):
    """Differentiation form: 24 qubits, 12 layers (576 parameters), TFIM J = 1.0, h = 0.7,
    complex128; adjoint gradient timed, checked against the parameter-shift rule on every
    parameter and central differences (step 1e-5) on a sample, both on the FP64 plain path."""
    import math

    import numpy as np

    import brk
    import circuits as W
    from devstate import DeviceState

    n, layers = 24, 12
    ops, theta = W.adjoint_ansatz(n, layers, seed=5)
    theta = np.array(theta)
    J, h = W.tfim(n)
    circuit = brk.Circuit(n, ops)
    psi, work = DeviceState(n), DeviceState(n)
    times = []
    for k in range(max(1, args.steps) + 1):
        brk.set_basis_state(psi.ptr, psi.size, brk.C128, 0, stream)
        t0 = time.perf_counter()
        e, g = brk.adjoint_ising_gradient(circuit, psi.ptr, work.ptr, brk.C128, J, h, stream)
        if k:  # the first call is the warm-up
            times.append(time.perf_counter() - t0)
    work.free()
    g = np.asarray(g)

    def energy(values):
        # FP64 plain path: gate blocks give the per-gate amplitudes bit for bit
        brk.set_basis_state(psi.ptr, psi.size, brk.C128, 0, stream)
        brk.Circuit(n, W.with_parameters(ops, values)).run_blocked(psi.ptr, stream=stream)
        return brk.expectation_ising(psi.ptr, psi.size, brk.C128, J, h, stream)

    # every parameter against the parameter-shift rule
    t0 = time.perf_counter()
    shift = np.empty_like(theta)
    for k in range(theta.size):
        tp, tm = theta.copy(), theta.copy()
        tp[k] += math.pi / 2
        tm[k] -= math.pi / 2
        shift[k] = 0.5 * (energy(tp) - energy(tm))
    err_shift = float(np.max(np.abs(g - shift)))
    log(f"adjoint: parameter shift over all {theta.size} parameters in {time.perf_counter() - t0:.1f} s, "
        f"max |adjoint - shift| = {err_shift:.3e}")
    # central differences (step 1e-5) on a sample
    rng = np.random.default_rng(1)
    sample = sorted(int(i) for i in rng.choice(theta.size, 6, replace=False))
    err_fd = 0.0
    for k in sample:
        tp, tm = theta.copy(), theta.copy()
        tp[k] += 1e-5
        tm[k] -= 1e-5
        fd = (plain_energy(n, W.with_parameters(ops, tp), J, h, stream) -
              plain_energy(n, W.with_parameters(ops, tm), J, h, stream)) / 2e-5
        err_fd = max(err_fd, abs(g[k] - fd))
    psi.free()
    ms = sum(times) / len(times) * 1e3
    log(f"adjoint: {theta.size} parameters in {ms:.1f} ms")
    return {
        "workload": "24 qubits complex128, 12 layers of RY, RZ on each qubit + CZ ladder (576 parameters, "
                    "U[-pi, pi), seed 5), TFIM open chain J = 1.0, h = 0.7",
        "gradients_per_s": round(1e3 / ms, 2),
        "ms_per_gradient": round(ms, 2),
        "energy": e,
        "max_abs_err_vs_parameter_shift": err_shift,
        "parameter_shift_checked": f"all {theta.size} parameters (FP64 plain path, gate blocks)",
        "max_abs_err_vs_central_fd_1e-5": err_fd,
        "fd_checked_parameters": sample,
        "dtype": "c128 state, FP64 accumulation",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=32)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fused", action="store_true", help="skip the fused complex64 (precision form) leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batching-form leg")
    ap.add_argument("--no-blocked", action="store_true", help="skip the gate-block (complex128, LDS) leg")
    ap.add_argument("--no-sampling", action="store_true", help="skip the measurement and sampling leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiation-form leg")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run the streaming copy kernel once on 2^30 c128 amplitudes (16 GiB read + "
                         "16 GiB written) so a counter run can scale FETCH_SIZE/WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              f"(the state-vector path runs on one device)", file=sys.stderr)
        return 2

    import brk
    import circuits as W
    import hip
    from devstate import DeviceState

    n = args.qubits
    ops = W.random_circuit(n, 200, 8, seed=args.seed)
    circuit = brk.Circuit(n, ops)
    op_bytes = [W.algorithmic_bytes(op, n) for op in ops]
    total_bytes = sum(op_bytes)
    log(f"device {brk.device_name()}; {n} qubits c128 ({2 ** n * 16 / 2 ** 30:.0f} GiB), {len(ops)} ops, "
        f"{total_bytes / 1e12:.3f} TB algorithmic per step")

    stream = hip.create_stream()
    if args.copy_calibration:
        a, b = DeviceState(30), DeviceState(30)
        brk.copy(b.ptr, a.ptr, a.size, brk.C128, stream)
        hip.synchronize(stream)
        a.free()
        b.free()
        log("copy calibration: 17179869184 B read, 17179869184 B written")
    st = DeviceState(n)
    brk.set_basis_state(st.ptr, st.size, brk.C128, 0, stream)
    hip.synchronize(stream)

    for w in range(args.warmup):
        brk.set_basis_state(st.ptr, st.size, brk.C128, 0, stream)
        circuit.run(st.ptr, stream=stream)
        hip.synchronize(stream)
        log(f"warmup {w + 1}/{args.warmup} done")

    step_ms, per_op = [], [[] for _ in ops]
    for k in range(args.steps):
        brk.set_basis_state(st.ptr, st.size, brk.C128, 0, stream)
        hip.synchronize(stream)
        e0, e1 = hip.Event(), hip.Event()
        e0.record(stream)
        _, ms = circuit.run_timed(st.ptr, stream=stream)
        e1.record(stream)
        e1.synchronize()
        step_ms.append(e0.elapsed_ms(e1))
        for i, m in enumerate(ms):
            per_op[i].append(m)
        log(f"step {k + 1}/{args.steps}: {step_ms[-1]:.1f} ms, {total_bytes / (step_ms[-1] * 1e-3) / 1e9:.0f} GB/s")
    hip.synchronize(stream)
    norm = brk.norm2(st.ptr, st.size, brk.C128, stream)

    fused = None
    if not args.no_fused:
        fused = fused_leg(circuit, st, args, stream)
    blocked = None
    if not args.no_blocked:
        blocked = blocked_leg(circuit, st, args, stream)
    sampling = None
    if not args.no_sampling:
        sampling = sampling_leg(st, args, stream)
    st.free()

    ms_per_step = sum(step_ms) / len(step_ms)
    value = total_bytes / (ms_per_step * 1e-3) / 1e9

    # per kernel instantiation: launches, bytes per launch, mean launch time
    kernels = {}
    for kn, b, ms in zip(physical_kernel_names(ops, n), op_bytes, per_op):
        d = kernels.setdefault(kn, {"launches": 0, "bytes": 0.0, "ms": 0.0, "bytes_per_launch": set()})
        d["launches"] += len(ms)
        d["bytes"] += b * len(ms)
        d["ms"] += sum(ms)
        d["bytes_per_launch"].add(b)
    table = {}
    for kn, d in kernels.items():
        table[kn] = {
            "launches_per_step": d["launches"] // args.steps,
            "mean_ms": round(d["ms"] / d["launches"], 4),
            "GB/s": round(d["bytes"] / (d["ms"] * 1e-3) / 1e9, 1),
            "share_of_time": round(d["ms"] / (ms_per_step * args.steps), 4),
        }
    # per gate kind and target band (qubits 0..2 share 128-byte lines)
    kinds = {}
    for (kind, band), b, ms in zip(kind_and_band(ops, n), op_bytes, per_op):
        d = kinds.setdefault(f"{kind}/{band}", {"gates": 0, "bytes": 0.0, "ms": 0.0})
        d["gates"] += 1
        d["bytes"] += b * len(ms)
        d["ms"] += sum(ms)
    by_kind = {k: {"gates_per_step": d["gates"], "GB/s": round(d["bytes"] / (d["ms"] * 1e-3) / 1e9, 1),
                   "hbm_fraction": round(d["bytes"] / (d["ms"] * 1e-3) / 8e12, 4)}
               for k, d in sorted(kinds.items())}
    uniform = {kn: d for kn, d in kernels.items() if len(d["bytes_per_launch"]) == 1}
    dom_name = max(uniform, key=lambda kn: uniform[kn]["ms"])
    dom = uniform[dom_name]
    per_launch = next(iter(dom["bytes_per_launch"]))
    mean_ms = dom["ms"] / dom["launches"]
    achieved = per_launch / (mean_ms * 1e-3) / 1e9

    result = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "ms_per_step_median": round(sorted(step_ms)[len(step_ms) // 2], 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic",
        "config": {
            "workload": f"random circuit, {n} qubits complex128 (2^{n} amplitudes), 200 gates "
                        f"(120 from H/X/RZ/U3, 80 CNOT/CZ) + 8 local qubit swaps, seed {args.seed}, from |0..0>",
            "qubits": n,
            "ops_per_step": len(ops),
            "algorithmic_TB_per_step": round(total_bytes / 1e12, 4),
            "parallelism": "single GPU",
        },
        "hbm_fraction": round(value / HBM_PEAK_GBS, 4),
        "roofline": {
            "bound": "hbm",
            "kernel": dom_name,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "bytes_per_launch": per_launch,
            "mean_launch_ms": round(mean_ms, 4),
            "traffic": None,
        },
        "device": brk.device_name(),
        "kernels": table,
        "gate_kinds": by_kind,
        "norm_error": abs(norm - 1.0),
    }
    if fused is not None:
        fused["speedup_vs_c128"] = round(ms_per_step / fused["ms_per_circuit"], 2)
        result["fused_c64"] = fused
    if blocked is not None:
        blocked["speedup_vs_per_gate"] = round(ms_per_step / blocked["ms_per_circuit"], 2)
        result["blocked_c128"] = blocked
    if sampling is not None:
        result["sampling_c128"] = sampling
    if not args.no_batch:
        result["batch_c64"] = batch_leg(args, stream)
    if not args.no_adjoint:
        result["adjoint_c128"] = adjoint_leg(args, stream)
    traffic = load_traffic(dom_name, n)
    if traffic is not None:
        result["roofline"]["traffic"] = traffic["bytes_per_launch"]
        result["roofline"]["traffic_source"] = traffic["source"]
    if not args.no_cpu:
        log("cpu baseline (oracle: the 20-qubit reference case, then 30 qubits for <= 15 s)")
        result["cpu_baseline"] = cpu_baseline_20q(stream)
        result["cpu_baseline_30q"] = cpu_baseline_30q(seed=args.seed)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
