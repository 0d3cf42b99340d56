#!/usr/bin/env python3
"""Gate-sweep benchmark: algorithmic HBM GB/s of a 30-qubit complex128 sweep on one MI355X.

One step = one pass of the 200-gate sweep (1q/2q Haar dense, RX/RY/RZ, CZ/CNOT,
fused 5q dense; seeded) over a 30-qubit state (16 GiB) that stays resident in
HBM.  value = sum over gates of 2 x 16 B x (amplitudes the gate can change),
divided by the time of the timed steps.  Gates are issued through the C ABI
(libqulacs_mi355) on one HIP stream in stream-ordered mode.

Besides the headline the JSON line carries:
  roofline      -- the dominant kernel's achieved GB/s (algorithmic bytes per
                   launch / mean launch time from HIP events on the launch
                   stream) against the 8 TB/s HBM peak;
  per_kind      -- the same per gate kind and target position;
  dm_sweep      -- the 14-qubit density-matrix variant (depolarizing p = 0.01
                   after every gate on each touched qubit);
  cpu_baseline  -- the reference csim compiled from its sources
                   (oracle/_ref/libcsim_ref.so; the in-repo restatement when
                   that is absent) on the host cores, on a time-boxed prefix of
                   the same gate list at the same size.

Usage: python bench.py [--gpus 1] [--steps 3] [--warmup 1]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBPS = 8000.0


class HipEvents:
    """hipEvent timing on an explicit stream (libamdhip64 through ctypes)."""

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")
        self.h.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.h.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.h.hipEventSynchronize.argtypes = [C.c_void_p]
        self.h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.h.hipEventDestroy.argtypes = [C.c_void_p]
        self.h.hipStreamSynchronize.argtypes = [C.c_void_p]
        self.h.hipDeviceSynchronize.argtypes = []
        self.h.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]

    def create(self, n: int) -> list:
        evs = []
        for _ in range(n):
            e = C.c_void_p()
            assert self.h.hipEventCreate(C.byref(e)) == 0
            evs.append(e)
        return evs

    def record(self, e, stream) -> None:
        assert self.h.hipEventRecord(e, stream) == 0

    def elapsed_ms(self, a, b) -> float:
        self.h.hipEventSynchronize(b)
        f = C.c_float()
        assert self.h.hipEventElapsedTime(C.byref(f), a, b) == 0
        return float(f.value)

    def destroy(self, evs) -> None:
        for e in evs:
            self.h.hipEventDestroy(e)


TRAFFIC_PROFILE = ROOT / "profiles" / "hbm_traffic_gate_sweep_30q.md"


def measured_traffic(kernel: str, qubits: int):
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE / WRITE_SIZE
    summary (calibrated against a copy of known size), or None when that summary
    does not cover this kernel at this size."""
    if qubits != 30 or not TRAFFIC_PROFILE.exists():
        return None
    want = kernel.replace(" ", "").replace(">>", ">")
    for line in TRAFFIC_PROFILE.read_text().splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 7 and cells[0].startswith("`"):
            name = cells[0].strip("`").replace(" ", "").replace(">>", ">")
            if name == want:
                return round(float(cells[4]) * 1e9)
    return None


def kernel_of(g) -> str:
    """Name of the kernel instantiation that executes gate g (qulacs_amd/csrc/qm_gates.hip): the
    library streams with non-temporal accesses when no acted-on qubit is below bit 3; CZ with a
    qubit below 3 and CNOT with a control below 3 run the line kernel (k_lane, LM = 1 when the
    target is below 3); CZ on higher qubits takes two tuples per thread."""
    nt = "true" if min(g.targets) >= 3 else "false"
    if g.kind == "CZ":
        if min(g.targets) < 3:
            return f"k_lane<1, {1 if g.targets[1] < 3 else 0}u, 2, OpNeg1>"
        return "k_gather<1, 2, true, OpNeg1>"
    if g.kind == "CNOT" and g.targets[0] < 3:
        return f"k_lane<1, {1 if g.targets[1] < 3 else 0}u, 2, OpSwap01>"
    return {
        "dense1": f"k_gather<1, 1, {nt}, OpDenseArg<1>>", "RX": f"k_gather<1, 1, {nt}, OpDenseArg<1>>",
        "RY": f"k_gather<1, 1, {nt}, OpDenseArg<1>>", "RZ": f"k_gather<1, 1, {nt}, OpDiag<1>>",
        "dense2": f"k_gather<2, 1, {nt}, OpDenseArg<2>>", "dense5": f"k_dense_mfma<5, {nt}>",
        "CNOT": f"k_gather<1, 1, {nt}, OpSwap01>",
    }[g.kind]


def host_cpu_info() -> dict:
    """lscpu model, sockets and cores of the host, the CPUs this process may use
    and the OpenMP placement the CPU baselines run with."""
    info = {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core", "NUMA node(s)": "numa_nodes"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                v = v.strip()
                info[keys[k.strip()]] = int(v) if v.isdigit() else v
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def cpu_threads() -> int:
    """OpenMP threads for the CPU baselines: OMP_NUM_THREADS when set (the CPU
    share of this machine), else the physical cores this process may use."""
    v = os.environ.get("OMP_NUM_THREADS", "")
    if v.isdigit() and int(v) > 0:
        return int(v)
    info = host_cpu_info()
    tpc = info.get("threads_per_core", 1) if isinstance(info.get("threads_per_core"), int) else 1
    return max(1, info["affinity_cpus"] // max(1, tpc))


def run_cpu_worker(leg: str, args, threads: int) -> dict:
    """Runs a CPU-baseline leg in a child process with OMP_PROC_BIND=close,
    OMP_PLACES=cores and `threads` OpenMP threads (so the binding never touches
    this process's threads), and returns its JSON."""
    env = dict(os.environ)
    env["OMP_NUM_THREADS"] = str(threads)
    env.setdefault("OMP_PROC_BIND", "close")
    env.setdefault("OMP_PLACES", "cores")
    cmd = [sys.executable, str(ROOT / "bench.py"), "--cpu-worker", leg, "--qubits", str(args.qubits), "--gates",
           str(args.gates), "--seed", str(args.seed), "--cpu-seconds", str(args.cpu_seconds), "--small-count",
           str(args.small_count), "--small-plain", str(args.small_plain)]
    r = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=900)
    if r.returncode != 0 or not r.stdout.strip():
        return {"error": f"cpu worker {leg} failed ({r.returncode}): {r.stderr[-400:]}"}
    out = json.loads(r.stdout.strip().splitlines()[-1])
    out["omp"] = {k: env.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")}
    return out


def _cpu_kernels():
    from oracle import oracle as O

    if O.reference_available():
        return O.Reference(), "reference"
    return O.Oracle(), "port"


def cpu_worker(args) -> dict:
    """The CPU legs (run by run_cpu_worker in a child process)."""
    from oracle import oracle as O

    k, kind = _cpu_kernels()
    threads = int(os.environ.get("OMP_NUM_THREADS", "1"))
    k.set_num_threads(threads)
    toucher = O.Oracle()
    toucher.set_num_threads(threads)
    if args.cpu_worker == "sweep":
        from qulacs_amd.workloads import gate_sweep

        return cpu_baseline(args, gate_sweep(args.qubits, args.gates, seed=args.seed), args.qubits, k, kind,
                            threads, toucher)
    if args.cpu_worker == "brickwork":
        from qulacs_amd.workloads import brickwork_circuit

        n = 20
        gates, obs, u = brickwork_circuit(n, 10, 16, 1024, seed=args.seed + 2)
        uu = np.ascontiguousarray(u)

        def once():
            psi = np.empty(1 << n, dtype=np.complex128)
            toucher.parallel_fill(psi, 0.0)
            psi[0] = 1.0
            for kind_g, q, m in gates:
                if kind_g == "CNOT":
                    k.two("CNOT", q[0], q[1], psi)
                else:
                    k.dense1(q[0], m, psi)
            e = 0.0
            for x in obs:
                tl, pl = _mask_lists(x[0], x[1], n)
                e += x[3] * k.expectation(tl, pl, psi)
            toucher.sampling(psi, uu)
            return e

        once()
        best = float("inf")
        for _ in range(3):
            t0 = time.perf_counter()
            once()
            best = min(best, time.perf_counter() - t0)
        return {"ms": round(best * 1e3, 2), "cores": threads, "kind": kind,
                "sample": "whole workload (circuit, 16-term expectation, 1024 samples by the in-repo port), best "
                          "of 3"}
    if args.cpu_worker == "small":
        from qulacs_amd.workloads import many_small_circuits

        cs = many_small_circuits(args.small_count, seed=args.seed + 1001)
        sub = cs[::max(1, len(cs) // args.small_plain)]
        total_gates = sum(len(c.gates) for c in cs)
        spent, done_circ, done_gates = 0.0, 0, 0
        for c in sub:
            t0 = time.perf_counter()
            psi = np.empty(1 << c.n, dtype=np.complex128)
            toucher.parallel_fill(psi, 0.0)
            psi[0] = 1.0
            for kind_g, q, m in c.gates:
                if kind_g == "CZ":
                    k.two("CZ", q[0], q[1], psi)
                else:
                    k.dense1(q[0], m, psi)
            for bf, pf, rot, coef in c.terms:
                tl, pl = _mask_lists(bf, pf, c.n)
                k.expectation(tl, pl, psi)
            spent += time.perf_counter() - t0
            done_circ += 1
            done_gates += len(c.gates)
            if spent >= args.cpu_seconds / 2:
                break
        cpu_s = spent * total_gates / done_gates
        return {"circuits_per_s": round(len(cs) / cpu_s, 1), "cores": threads, "kind": kind,
                "sample": f"first {done_circ} of {len(sub)} evenly spaced circuits ({done_gates} gates, "
                          f"{spent:.1f} s, {threads} OpenMP threads), scaled by gate count to all {len(cs)}"}
    raise SystemExit(f"unknown cpu worker leg {args.cpu_worker}")


def sv_sweep(args, lib_mod, hip):
    from qulacs_amd.workloads import GateLauncher, alg_bytes, gate_sweep

    L = lib_mod.lib()
    n = args.qubits
    dim = 1 << n
    gates = gate_sweep(n, args.gates, seed=args.seed)
    launcher = GateLauncher(lib_mod, gates)
    stream = lib_mod.call("allocate_cuda_stream_host", 1, 0)
    hstream = C.c_void_p.from_address(stream).value
    state = lib_mod.call("allocate_quantum_state_host", dim, 0)
    lib_mod.call("initialize_Haar_random_state_with_seed_host", state, dim, 1, stream, 0)
    lib_mod.call("qm_set_synchronous", 0)

    def sweep():
        for i in range(len(gates)):
            launcher.launch(i, state, dim, stream)

    for _ in range(args.warmup):
        sweep()
    lib_mod.call("qm_synchronize", stream, 0)
    lib_mod.check()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sweep()
    lib_mod.call("qm_synchronize", stream, 0)
    t1 = time.perf_counter()
    lib_mod.check()
    total_bytes = sum(alg_bytes(g, n) for g in gates)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = total_bytes * args.steps / (t1 - t0) / 1e9

    # per-gate device time: events between consecutive launches on the launch stream
    evs = hip.create(len(gates) + 1)
    hip.record(evs[0], hstream)
    for i in range(len(gates)):
        launcher.launch(i, state, dim, stream)
        hip.record(evs[i + 1], hstream)
    lib_mod.call("qm_synchronize", stream, 0)
    durs = [hip.elapsed_ms(evs[i], evs[i + 1]) for i in range(len(gates))]
    hip.destroy(evs)
    norm = lib_mod.call("state_norm_squared_host", state, dim, stream, 0)

    per_kind = {}
    per_kernel = {}
    for g, d in zip(gates, durs):
        key = f"{g.kind}/{g.position}"
        e = per_kind.setdefault(key, {"count": 0, "ms": 0.0, "bytes": 0})
        e["count"] += 1
        e["ms"] += d
        e["bytes"] += alg_bytes(g, n)
        k = per_kernel.setdefault(kernel_of(g), {"count": 0, "ms": 0.0, "bytes": 0})
        k["count"] += 1
        k["ms"] += d
        k["bytes"] += alg_bytes(g, n)
    for e in list(per_kind.values()) + list(per_kernel.values()):
        e["GBps"] = round(e["bytes"] / (e["ms"] * 1e-3) / 1e9, 1)
        e["frac_of_peak"] = round(e["GBps"] / HBM_PEAK_GBPS, 4)
        e["mean_ms"] = round(e["ms"] / e["count"], 4)
        e["ms"] = round(e["ms"], 3)
    dom_name, dom = max(per_kernel.items(), key=lambda kv: kv[1]["ms"])
    per_launch = dom["bytes"] / dom["count"]
    achieved = per_launch / (dom["mean_ms"] * 1e-3) / 1e9

    # measured HBM ceilings on this device, from the library's own kernels:
    # a streaming copy (copy_quantum_state_from_device_to_device: half the state
    # onto the other half, read + write bytes) and an in-place read-modify-write
    # pass over the whole state (state_multiply_host by 1, the access pattern of
    # a dense gate with every qubit >= 3)
    half_dim = dim // 2
    one = lib_mod.QmComplex(1.0, 0.0)
    ev2 = hip.create(2)
    copy_ms, rmw_ms = [], []
    for _ in range(5):
        hip.record(ev2[0], hstream)
        L.copy_quantum_state_from_device_to_device(C.c_void_p(state + half_dim * 16), C.c_void_p(state), half_dim,
                                                   stream, 0)
        hip.record(ev2[1], hstream)
        copy_ms.append(hip.elapsed_ms(ev2[0], ev2[1]))
        hip.record(ev2[0], hstream)
        L.state_multiply_host(one, state, dim, stream, 0)
        hip.record(ev2[1], hstream)
        rmw_ms.append(hip.elapsed_ms(ev2[0], ev2[1]))
    hip.destroy(ev2)
    lib_mod.check()
    copy_gbps = 2 * half_dim * 16 / (min(copy_ms) * 1e-3) / 1e9
    rmw_gbps = 2 * dim * 16 / (min(rmw_ms) * 1e-3) / 1e9
    hbm_measured = {"copy_kernel_GBps": round(copy_gbps, 1), "inplace_rmw_GBps": round(rmw_gbps, 1),
                    "copy_bytes": 2 * half_dim * 16, "rmw_bytes": 2 * dim * 16, "method": "best of 5, HIP events",
                    "dominant_kernel_frac_of_rmw": round(achieved / rmw_gbps, 4)}

    lib_mod.call("qm_set_synchronous", 1)
    lib_mod.call("release_quantum_state_host", state, 0)
    lib_mod.call("release_cuda_stream_host", stream, 1, 0)
    return {
        "value": value, "ms_per_step": ms_per_step, "gates": gates, "per_kind": per_kind, "per_kernel": per_kernel,
        "norm_after": norm,
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
                     "traffic": measured_traffic(dom_name, n),
                     "bytes_per_launch": per_launch, "mean_launch_ms": dom["mean_ms"]},
        "hbm_measured": hbm_measured,
    }


def dm_sweep(args, lib_mod, hip):
    """14-qubit density matrix: every gate followed by depolarizing(0.01) on its qubits."""
    from qulacs_amd.workloads import gate_sweep, rotation_matrix

    n = args.dm_qubits
    dim = 1 << n
    gates = gate_sweep(n, args.gates, seed=args.seed + 1)
    stream = lib_mod.call("allocate_cuda_stream_host", 1, 0)
    hstream = C.c_void_p.from_address(stream).value
    rho = lib_mod.call("allocate_quantum_state_host", dim * dim, 0)
    lib_mod.call("dm_initialize_quantum_state_host", rho, dim, stream, 0)
    lib_mod.call("qm_set_synchronous", 0)
    L = lib_mod.lib()
    from qulacs_amd.workloads import _cmat
    X = np.array([[0, 1], [1, 0]])
    Z = np.diag([1, -1])
    prepared = []
    for g in gates:
        if g.kind in ("CNOT", "CZ"):
            m = _cmat(lib_mod, X if g.kind == "CNOT" else Z)
            prepared.append(("c", (C.c_uint * 1)(g.targets[0]), (C.c_uint * 1)(1), (C.c_uint * 1)(g.targets[1]), m))
        else:
            mat = g.matrix if g.matrix is not None else rotation_matrix(g.kind, g.angle)
            prepared.append(("u", (C.c_uint * len(g.targets))(*g.targets), len(g.targets), _cmat(lib_mod, mat)))

    def run():
        for g, p in zip(gates, prepared):
            if p[0] == "c":
                L.dm_multi_qubit_control_multi_qubit_dense_matrix_gate_host(p[1], p[2], 1, p[3], 1, p[4], rho, dim,
                                                                             stream, 0)
            elif p[2] == 1:
                L.dm_single_qubit_dense_matrix_gate_host(g.targets[0], p[3], rho, dim, stream, 0)
            else:
                L.dm_multi_qubit_dense_matrix_gate_host(p[1], p[2], p[3], rho, dim, stream, 0)
            for q in g.targets:
                L.dm_depolarizing_host(q, 0.01, rho, dim, stream, 0)

    run()
    lib_mod.call("qm_synchronize", stream, 0)
    lib_mod.check()
    ev = hip.create(2)
    hip.record(ev[0], hstream)
    run()
    hip.record(ev[1], hstream)
    ms = hip.elapsed_ms(ev[0], ev[1])
    hip.destroy(ev)
    lib_mod.check()
    tr = lib_mod.call("dm_state_norm_squared_host", rho, dim, stream, 0)
    ops = len(gates) + sum(len(g.targets) for g in gates)
    alg = ops * 32 * dim * dim  # each gate / channel may change every entry once
    lib_mod.call("qm_set_synchronous", 1)
    lib_mod.call("release_quantum_state_host", rho, 0)
    lib_mod.call("release_cuda_stream_host", stream, 1, 0)

    # The same noisy circuit through the library's circuit loop
    # (QuantumCircuit.update_quantum_state on a DensityMatrix): consecutive gates
    # and channels whose qubits fit in two are composed on the host into one
    # superoperator and applied in one pass.
    import qulacs_amd as qa

    circ = qa.QuantumCircuit(n)
    for g in gates:
        if g.kind == "CNOT":
            circ.add_gate(qa.gate.CNOT(g.targets[0], g.targets[1]))
        elif g.kind == "CZ":
            circ.add_gate(qa.gate.CZ(g.targets[0], g.targets[1]))
        else:
            mat = g.matrix if g.matrix is not None else rotation_matrix(g.kind, g.angle)
            circ.add_gate(qa.gate.DenseMatrix(list(g.targets), mat))
        for q in g.targets:
            circ.add_gate(qa.gate.DepolarizingNoise(q, 0.01))
    dmst = qa.DensityMatrix(n)
    qa.set_synchronous(False)
    circ.update_quantum_state(dmst)  # warm-up (kernel loading)
    qa.synchronize(dmst)
    dmst.set_zero_state()
    qa.synchronize(dmst)
    t0 = time.perf_counter()
    circ.update_quantum_state(dmst)
    qa.synchronize(dmst)
    loop_ms = (time.perf_counter() - t0) * 1e3
    qa.set_synchronous(True)
    tr_loop = dmst.get_squared_norm()
    # HBM passes the loop issues, by its grouping rule: three to five
    # single-qubit channels in a row on distinct qubits are one pass; otherwise a
    # run of gates/channels on at most two qubits together is one pass (a lone
    # gate on one or two qubits too); a gate on more qubits is two passes (rows,
    # then columns)
    seq = []
    for g in gates:
        seq.append(("gate", set(g.targets)))
        seq += [("chan", {q}) for q in g.targets]
    passes, cur, k = 0, set(), 0
    while k < len(seq):
        run = 0
        while k + run < len(seq) and run < 5 and seq[k + run][0] == "chan" and \
                not any(seq[k + run][1] <= seq[k + j][1] for j in range(run)):
            run += 1
        if run >= 3:
            passes += (1 if cur else 0) + 1
            cur = set()
            k += run
            continue
        part = seq[k][1]
        k += 1
        if len(part) > 2:
            passes += (1 if cur else 0) + 2
            cur = set()
            continue
        if len(cur | part) > 2:
            passes += 1
            cur = set()
        cur |= part
    passes += 1 if cur else 0
    moved = passes * 32 * dim * dim
    return {"qubits": n, "gates": len(gates), "channels": ops - len(gates), "ms": round(ms, 2),
            "alg_GBps": round(alg / (ms * 1e-3) / 1e9, 1), "trace_after": tr,
            "circuit_loop": {"ms": round(loop_ms, 2), "passes": passes,
                             "alg_GBps": round(alg / (loop_ms * 1e-3) / 1e9, 1),
                             "pass_GBps": round(moved / (loop_ms * 1e-3) / 1e9, 1),
                             "pass_frac_of_peak": round(moved / (loop_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4),
                             "trace_after": tr_loop,
                             "method": "QuantumCircuit.update_quantum_state on a DensityMatrix (gates and "
                                       "depolarizing channels fused into superoperators), one run after a warm-up; "
                                       "pass_GBps = passes x 32 B x 4^n / time"}}


FP64_PEAK_TFLOPS = 78.6


def many_small(args, lib_mod, hip):
    """Many small circuits (n ~ U{8..16}, 4..12 layers of RY RZ + CZ chain, TFIM
    energy each) through the batched executor: one launch per step."""
    from qulacs_amd.batch import Batch
    from qulacs_amd.workloads import many_small_circuits, small_circuit_alg_flops

    cs = many_small_circuits(args.small_count, seed=args.seed + 1001)
    b = Batch()
    t0 = time.perf_counter()
    for c in cs:
        b.add(c.n, c.gates, c.terms)
    t1 = time.perf_counter()
    tune = b.autotune(repeats=2)
    t2 = time.perf_counter()
    stream = lib_mod.call("allocate_cuda_stream_host", 1, 0)
    hstream = C.c_void_p.from_address(stream).value
    b.run(stream)
    ev = hip.create(2)
    hip.record(ev[0], hstream)
    for _ in range(args.small_steps):
        b.run_async(stream)
    hip.record(ev[1], hstream)
    ms = hip.elapsed_ms(ev[0], ev[1]) / args.small_steps
    hip.destroy(ev)
    energies = b.run(stream)
    # the same launch captured in a HIP graph and replayed
    b.capture(stream)
    g_e = b.run_graph(stream)
    ev = hip.create(2)
    hip.record(ev[0], hstream)
    for _ in range(args.small_steps):
        lib_mod.lib().qm_batch_run_graph(b.h, None, stream, 0)
    hip.record(ev[1], hstream)
    graph_ms = hip.elapsed_ms(ev[0], ev[1]) / args.small_steps
    hip.destroy(ev)
    graph_same = bool(np.array_equal(g_e, energies))
    flops = sum(small_circuit_alg_flops(c) for c in cs)
    stats = b.stats()
    b.close()

    # the per-gate path on an evenly spaced subset: one C ABI call per gate and
    # one observable reduction per circuit, stream-ordered, as a plain loop would
    L = lib_mod.lib()
    stride = max(1, len(cs) // args.small_plain)
    sub = cs[::stride]
    lib_mod.call("qm_set_synchronous", 0)
    t3 = time.perf_counter()
    plain_gates = 0
    for c in sub:
        dim = 1 << c.n
        st = lib_mod.call("allocate_quantum_state_host", dim, 0)
        lib_mod.call("initialize_quantum_state_host", st, dim, stream, 0)
        for kind, q, m in c.gates:
            if kind == "CZ":
                L.CZ_gate_host(q[0], q[1], st, dim, stream, 0)
            else:
                L.single_qubit_dense_matrix_gate_host(q[0], _cm(lib_mod, m), st, dim, stream, 0)
        plain_gates += len(c.gates)
        t = len(c.terms)
        bf = (C.c_ulonglong * t)(*[x[0] for x in c.terms])
        pf = (C.c_ulonglong * t)(*[x[1] for x in c.terms])
        rot = (C.c_uint * t)(*[x[2] for x in c.terms])
        co = (lib_mod.QmComplex * t)(*[lib_mod.QmComplex(x[3], 0.0) for x in c.terms])
        L.qm_observable_expectation_host(bf, pf, rot, co, t, st, dim, stream, 0)
        lib_mod.call("release_quantum_state_host", st, 0)
    lib_mod.check()
    t4 = time.perf_counter()
    lib_mod.call("qm_set_synchronous", 1)
    lib_mod.call("release_cuda_stream_host", stream, 1, 0)
    total_gates = sum(len(c.gates) for c in cs)
    plain_s = (t4 - t3) * total_gates / plain_gates

    out = {
        "circuits": len(cs), "qubits": "U{8..16}", "layers": "U{4..12}", "gates": total_gates,
        "ms_per_step": round(ms, 3), "circuits_per_s": round(len(cs) / (ms * 1e-3), 1),
        "fp64": {"bound": "fp64-valu", "achieved": round(flops / (ms * 1e-3) / 1e12, 2), "peak": FP64_PEAK_TFLOPS,
                 "unit": "TFLOP/s", "frac": round(flops / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 4),
                 "flops_per_step": flops},
        "graph_ms_per_step": round(graph_ms, 3), "graph_matches_eager": graph_same,
        "autotune_s": {f"K{k[0]}/window{k[1]}": round(v, 5) for k, v in tune.items()}, "schedule": stats,
        "host_add_s": round(t1 - t0, 2), "host_compile_and_tune_s": round(t2 - t1, 2),
        "plain_gpu": {"circuits_per_s": round(len(cs) / plain_s, 1),
                      "sample": f"{len(sub)} evenly spaced circuits ({plain_gates} gates), one C ABI call per gate, "
                                f"scaled by gate count to all {len(cs)}"},
        "energy_checksum": float(np.sum(energies)),
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = run_cpu_worker("small", args, cpu_threads())
    return out


def fused_large(args, lib_mod, hip):
    """The fused stage executor alone on one large state: an n-qubit brickwork circuit
    (default 26 qubits, 40 layers: a Haar 2x2 on every qubit, then a CNOT brick) from |0>,
    at register-block widths 3 and 4 with the 13-qubit window, timed with HIP events on the
    launch stream (best of three runs after a warm-up).  Reports the stage count, the state
    traffic (2 x 16 B x 2^n per stage) and the FP64 rate of the fused arithmetic (one 2x2
    complex matrix per qubit per layer: 16 flops per amplitude) against the 78.6 TFLOP/s
    vector peak; the two widths must agree (|<psi_3|psi_4>| within 1e-10 of 1)."""
    from qulacs_amd.batch import FusedCircuit
    from qulacs_amd.workloads import brickwork_circuit

    n, layers = args.fused_qubits, args.fused_layers
    dim = 1 << n
    gates, _, _ = brickwork_circuit(n, layers, 1, 1, seed=args.seed + 5)
    L = lib_mod.lib()
    stream = lib_mod.call("allocate_cuda_stream_host", 1, 0)
    hstream = C.c_void_p.from_address(stream).value  # the hipStream_t behind the handle
    flops = 16.0 * layers * n * dim
    out = {"workload": f"{n}-qubit brickwork, {layers} layers ({n * layers} Haar 1q + CNOT bricks), from |0>, "
                       "13-qubit window", "qubits": n, "layers": layers, "gates": len(gates)}
    states = {}
    lib_mod.call("qm_set_synchronous", 0)
    for K in (3, 4):
        fc = FusedCircuit(n, gates, block_qubits=K)
        st = lib_mod.call("allocate_quantum_state_host", dim, 0)
        e0, e1 = hip.create(2)
        L.initialize_quantum_state_host(st, dim, stream, 0)
        L.qm_circuit_apply(fc.h, st, dim, stream, 0)
        best = float("inf")
        for _ in range(3):
            L.initialize_quantum_state_host(st, dim, stream, 0)
            hip.record(e0, hstream)
            L.qm_circuit_apply(fc.h, st, dim, stream, 0)
            hip.record(e1, hstream)
            best = min(best, hip.elapsed_ms(e0, e1))
        hip.destroy([e0, e1])
        lib_mod.check()
        stats = fc.stats()
        traffic = stats["stages"] * 2.0 * 16 * dim
        out[f"K{K}"] = {"ms": round(best, 2), **stats, "state_traffic_GBps": round(traffic / best / 1e6, 1),
                        "fp64_TFLOPs": round(flops / best / 1e9, 2), "fp64_frac_of_peak": round(flops / best / 1e9 / 78.6, 4)}
        states[K] = st
        fc.close()
    lib_mod.call("qm_set_synchronous", 1)
    z = lib_mod.call("inner_product_host", states[3], states[4], dim, stream, 0)
    out["widths_agree"] = bool(abs(abs(complex(z.re, z.im)) - 1.0) < 1e-10)
    for st in states.values():
        lib_mod.call("release_quantum_state_host", st, 0)
    lib_mod.call("release_cuda_stream_host", stream, 1, 0)
    return out


def brickwork(args, lib_mod, hip):
    """20-qubit brickwork (10 layers of Haar 1q on every qubit + a CNOT brick),
    16-term observable and 1024 samples: the per-gate path end to end, the fused
    executor, the drop-in QuantumCircuit API, the batched executor on the same
    circuit, and the reference csim on the host."""
    from oracle import oracle as O
    from qulacs_amd.batch import Batch
    from qulacs_amd.workloads import brickwork_circuit

    n = 20
    dim = 1 << n
    gates, obs, u = brickwork_circuit(n, 10, 16, 1024, seed=args.seed + 2)
    L = lib_mod.lib()
    stream = lib_mod.call("allocate_cuda_stream_host", 1, 0)
    st = lib_mod.call("allocate_quantum_state_host", dim, 0)
    calls = []
    for kind, q, m in gates:
        if kind == "CNOT":
            calls.append((L.CNOT_gate_host, (q[0], q[1])))
        else:
            calls.append((L.single_qubit_dense_matrix_gate_host, (q[0], _cm(lib_mod, m))))
    t = len(obs)
    bf = (C.c_ulonglong * t)(*[x[0] for x in obs])
    pf = (C.c_ulonglong * t)(*[x[1] for x in obs])
    rot = (C.c_uint * t)(*[x[2] for x in obs])
    co = (lib_mod.QmComplex * t)(*[lib_mod.QmComplex(x[3], 0.0) for x in obs])
    uu = np.ascontiguousarray(u)
    samples = np.zeros(u.size, dtype=np.uint64)

    def run():
        L.initialize_quantum_state_host(st, dim, stream, 0)
        for fn, a in calls:
            fn(*a, st, dim, stream, 0)
        e = L.qm_observable_expectation_host(bf, pf, rot, co, t, st, dim, stream, 0)
        L.qm_sampling_host(uu.ctypes.data_as(lib_mod.PD), uu.size, samples.ctypes.data_as(lib_mod.PI), st, dim,
                           stream, 0)
        return e.re

    from qulacs_amd.batch import FusedCircuit

    fused = FusedCircuit(n, gates)
    fused_samples = np.zeros(u.size, dtype=np.uint64)

    def run_fused():
        L.initialize_quantum_state_host(st, dim, stream, 0)
        L.qm_circuit_apply(fused.h, st, dim, stream, 0)
        e = L.qm_observable_expectation_host(bf, pf, rot, co, t, st, dim, stream, 0)
        L.qm_sampling_host(uu.ctypes.data_as(lib_mod.PD), uu.size, fused_samples.ctypes.data_as(lib_mod.PI), st,
                           dim, stream, 0)
        return e.re

    lib_mod.call("qm_set_synchronous", 0)
    run()
    run_fused()
    lib_mod.check()
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        e_fused = run_fused()
    t1 = time.perf_counter()
    fused_ms = (t1 - t0) * 1e3 / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        e_gpu = run()
    t1 = time.perf_counter()
    lib_mod.check()
    lib_mod.call("qm_set_synchronous", 1)
    gpu_ms = (t1 - t0) * 1e3 / reps
    psi_gpu = np.zeros(dim, dtype=np.complex128)  # the per-gate path's final state (the variants reuse st)
    lib_mod.call("copy_quantum_state_from_device_to_host", psi_gpu.ctypes.data, st, dim, stream, 0)
    fused_stats = fused.stats()
    fused.close()
    # the fused executor's register-block width and LDS window
    variants = {}
    lib_mod.call("qm_set_synchronous", 0)
    for kq, lq in ((3, 12), (4, 12), (3, 13), (4, 13)):
        fv = FusedCircuit(n, gates, block_qubits=kq, local_qubits=lq)
        L.initialize_quantum_state_host(st, dim, stream, 0)
        L.qm_circuit_apply(fv.h, st, dim, stream, 0)
        lib_mod.call("qm_synchronize", stream, 0)
        t0 = time.perf_counter()
        for _ in range(reps):
            L.qm_circuit_apply(fv.h, st, dim, stream, 0)
        lib_mod.call("qm_synchronize", stream, 0)
        variants[f"K{kq}/window{lq}"] = round((time.perf_counter() - t0) * 1e3 / reps, 3)
        fv.close()
    lib_mod.call("qm_set_synchronous", 1)
    lib_mod.check()
    lib_mod.call("release_quantum_state_host", st, 0)
    lib_mod.call("release_cuda_stream_host", stream, 1, 0)

    # the drop-in API: QuantumCircuit.update_quantum_state (runs of fusable gates go
    # through the stage executor), Observable.get_expectation_value, sampling
    import qulacs_amd as qa

    api_c = qa.QuantumCircuit(n)
    for kind_g, q, m in gates:
        if kind_g == "CNOT":
            api_c.add_CNOT_gate(q[0], q[1])
        else:
            api_c.add_gate(qa.gate.DenseMatrix(q[0], m))
    api_o = qa.Observable(n)
    for x in obs:
        tl, pl = _mask_lists(x[0], x[1], n)
        api_o.add_operator(qa.PauliOperator(list(tl), list(pl), float(x[3])))
    api_s = qa.QuantumState(n)
    uu_list = uu.tolist()

    def run_api():
        api_s.set_zero_state()
        api_c.update_quantum_state(api_s)
        e = api_o.get_expectation_value(api_s)
        return e, api_s.sampling_with_uniforms(uu_list)

    run_api()
    t0 = time.perf_counter()
    for _ in range(reps):
        e_api, s_api = run_api()
    api_ms = (time.perf_counter() - t0) * 1e3 / reps

    b = Batch()
    b.add(n, gates, obs)
    b.compile(4)
    e_batch = b.run()[0]
    t2 = time.perf_counter()
    for _ in range(reps):
        b.run()
    batch_ms = (time.perf_counter() - t2) * 1e3 / reps
    b.close()

    scale = sum(abs(x[3]) for x in obs)
    out = {"workload": "20-qubit brickwork, 10 layers (200 Haar 1q + 95 CNOT), 16-term observable, 1024 samples",
           "gpu_ms": round(gpu_ms, 3), "fused_ms": round(fused_ms, 3), "fused_schedule": fused_stats,
           "fused_circuit_only_ms": variants,
           "batched_executor_ms": round(batch_ms, 3), "api_circuit_ms": round(api_ms, 3), "energy": e_gpu,
           "fused_matches": bool(abs(e_fused - e_gpu) <= 1e-10 * scale and np.array_equal(fused_samples, samples))}
    out["batched_energy_matches"] = bool(abs(e_batch - e_gpu) <= 1e-10 * scale)
    out["api_matches"] = bool(abs(e_api - e_gpu) <= 1e-10 * scale and np.array_equal(np.asarray(s_api, dtype=np.uint64), samples))
    if not args.no_cpu_baseline:
        # parity against the CPU (untimed, in this process) ...
        k, _ = _cpu_kernels()
        o = O.Oracle()
        psi = np.zeros(dim, dtype=np.complex128)
        psi[0] = 1.0
        for kind_g, q, m in gates:
            if kind_g == "CNOT":
                k.two("CNOT", q[0], q[1], psi)
            else:
                k.dense1(q[0], m, psi)
        e_cpu = 0.0
        for x in obs:
            tl, pl = _mask_lists(x[0], x[1], n)
            e_cpu += x[3] * k.expectation(tl, pl, psi)
        s_cpu = o.sampling(psi, uu)
        out["energy_matches"] = bool(abs(e_cpu - e_gpu) <= 1e-10 * scale)
        out["samples_identical"] = bool(np.array_equal(s_cpu, samples))
        out["state_rel_l2"] = float(np.linalg.norm(psi_gpu - psi) / np.linalg.norm(psi))
        # ... and its timing in child processes with bound OpenMP threads, at the
        # machine's CPU share and at one thread (the reference csim fixes its
        # thread count when first used, so each count needs its own process)
        out["cpu_baseline"] = run_cpu_worker("brickwork", args, cpu_threads())
        out["cpu_baseline_one_thread"] = run_cpu_worker("brickwork", args, 1)
    return out


def serving(args, lib_mod, hip):
    """Poisson stream of mixed-size brickwork circuits through the request
    server, against two sequential baselines over the same requests (FIFO from
    the same arrival times): the per-gate path (one C ABI call per gate, the
    reference's circuit loop) and the fused executor one request at a time (the
    same per-request kernels as the server, without batching or overlap)."""
    import threading

    from qulacs_amd.batch import _marshal
    from qulacs_amd.serving import Server
    from qulacs_amd.workloads import serving_requests

    reqs = serving_requests(args.serve_count, rate=args.serve_rate, seed=args.seed + 5,
                            n_range=(10, args.serve_max_qubits))
    L = lib_mod.lib()
    # pre-marshal so submission costs only the call
    prepared = []
    for r in reqs:
        arr, keep = _marshal(r.gates)
        t = len(r.terms)
        prepared.append((arr, keep, (C.c_ulonglong * t)(*[x[0] for x in r.terms]),
                         (C.c_ulonglong * t)(*[x[1] for x in r.terms]), (C.c_uint * t)(*[x[2] for x in r.terms]),
                         (C.c_double * t)(*[x[3] for x in r.terms]), np.ascontiguousarray(r.uniforms)))
    budget = int(args.serve_budget_gb * 1e9)
    srv = Server(hbm_budget_bytes=budget, large_slots=4, small_max_qubits=16)
    # warm-up: one small and one large request (kernel loading, allocator pools)
    for r in (min(reqs, key=lambda x: x.n), max(reqs, key=lambda x: x.n)):
        srv.wait(srv.submit(r.n, r.gates[:8], r.terms, r.uniforms[:4]))
    ids = [None] * len(reqs)
    submit_at = [0.0] * len(reqs)

    def submitter():
        t0 = time.perf_counter()
        for i, (r, p) in enumerate(zip(reqs, prepared)):
            delay = r.arrival - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            submit_at[i] = time.perf_counter() - t0
            ids[i] = L.qm_server_submit(srv.h, r.n, p[0], len(r.gates), p[2], p[3], p[4], p[5], len(r.terms),
                                        p[6].ctypes.data_as(lib_mod.PD), p[6].size)

    th = threading.Thread(target=submitter)
    th.start()
    th.join()
    lib_mod.check()
    served = []
    for i, r in enumerate(reqs):
        e = C.c_double()
        lat = C.c_double()
        nflag = C.c_uint()
        out = np.zeros(r.uniforms.size, dtype=np.uint64)
        flags = np.zeros(r.uniforms.size, dtype=np.uint8)
        L.qm_server_wait_checked(srv.h, ids[i], C.byref(e), out.ctypes.data_as(lib_mod.PI),
                                 flags.ctypes.data_as(lib_mod.PB), C.byref(nflag), C.byref(lat))
        served.append((e.value, out, flags, lat.value))
    lib_mod.check()
    stats = srv.stats()
    srv.close()
    lat_served = np.array([x[3] for x in served])
    makespan = max(submit_at[i] + served[i][3] for i in range(len(reqs)))

    stream = lib_mod.call("allocate_cuda_stream_host", 1, 0)
    lib_mod.call("qm_set_synchronous", 0)

    # (1) the per-gate path, one request at a time
    service = []
    match_e, match_s, match_unflagged, flagged_total, diff_unflagged = 0, 0, 0, 0, 0
    for r, p, sv in zip(reqs, prepared, served):
        t0 = time.perf_counter()
        dim = 1 << r.n
        st = lib_mod.call("allocate_quantum_state_host", dim, 0)
        L.initialize_quantum_state_host(st, dim, stream, 0)
        for kind, q, m in r.gates:
            if kind == "CNOT":
                L.CNOT_gate_host(q[0], q[1], st, dim, stream, 0)
            else:
                L.single_qubit_dense_matrix_gate_host(q[0], _cm(lib_mod, m), st, dim, stream, 0)
        co = (lib_mod.QmComplex * len(r.terms))(*[lib_mod.QmComplex(x[3], 0.0) for x in r.terms])
        e = L.qm_observable_expectation_host(p[2], p[3], p[4], co, len(r.terms), st, dim, stream, 0).re
        out = np.zeros(r.uniforms.size, dtype=np.uint64)
        L.qm_sampling_host(p[6].ctypes.data_as(lib_mod.PD), p[6].size, out.ctypes.data_as(lib_mod.PI), st, dim,
                           stream, 0)
        lib_mod.call("release_quantum_state_host", st, 0)
        service.append(time.perf_counter() - t0)
        scale = sum(abs(x[3]) for x in r.terms)
        ok = sv[2] == 0
        match_e += int(abs(e - sv[0]) <= 1e-10 * scale)
        match_s += int(np.array_equal(out, sv[1]))
        match_unflagged += int(np.array_equal(out[ok], sv[1][ok]))
        diff_unflagged += int(np.sum(out[ok] != sv[1][ok]))
        flagged_total += int((~ok).sum())
    lib_mod.check()

    # (2) the fused executor, one request at a time (host compile included, as in the server)
    from qulacs_amd.batch import FusedCircuit

    fused_service = []
    for r, p in zip(reqs, prepared):
        t0 = time.perf_counter()
        dim = 1 << r.n
        st = lib_mod.call("allocate_quantum_state_host", dim, 0)
        L.initialize_quantum_state_host(st, dim, stream, 0)
        if r.n >= 4:
            fc = FusedCircuit(r.n, r.gates, block_qubits=4)
            L.qm_circuit_apply(fc.h, st, dim, stream, 0)
        else:
            fc = None
            for kind, q, m in r.gates:
                if kind == "CNOT":
                    L.CNOT_gate_host(q[0], q[1], st, dim, stream, 0)
                else:
                    L.single_qubit_dense_matrix_gate_host(q[0], _cm(lib_mod, m), st, dim, stream, 0)
        co = (lib_mod.QmComplex * len(r.terms))(*[lib_mod.QmComplex(x[3], 0.0) for x in r.terms])
        L.qm_observable_expectation_host(p[2], p[3], p[4], co, len(r.terms), st, dim, stream, 0)
        out = np.zeros(r.uniforms.size, dtype=np.uint64)
        L.qm_sampling_host(p[6].ctypes.data_as(lib_mod.PD), p[6].size, out.ctypes.data_as(lib_mod.PI), st, dim,
                           stream, 0)
        lib_mod.call("release_quantum_state_host", st, 0)
        if fc is not None:
            fc.close()
        fused_service.append(time.perf_counter() - t0)
    lib_mod.check()
    lib_mod.call("qm_set_synchronous", 1)
    lib_mod.call("release_cuda_stream_host", stream, 1, 0)

    def fifo(svc):
        finish, lat = 0.0, []
        for r, sv in zip(reqs, svc):
            start = max(r.arrival, finish)
            finish = start + sv
            lat.append(finish - r.arrival)
        return np.array(lat), finish - reqs[0].arrival

    lat_plain, plain_makespan = fifo(service)
    lat_fused, fused_makespan = fifo(fused_service)

    def pct(a, q):
        return round(float(np.percentile(a, q)) * 1e3, 2)

    n_samples = int(sum(r.uniforms.size for r in reqs))
    return {
        "workload": f"{len(reqs)} Poisson requests at {args.serve_rate}/s, brickwork circuits, n = round(exp(U[ln 10, "
                    f"ln {args.serve_max_qubits}])), depth U{{20..200}}, 8-term observable, 1000 samples each; "
                    f"admission budget {args.serve_budget_gb} GB",
        "qubits": [r.n for r in reqs], "gates": int(sum(len(r.gates) for r in reqs)),
        "served": {"circuits_per_s": round(len(reqs) / makespan, 2), "p50_ms": pct(lat_served, 50),
                   "p99_ms": pct(lat_served, 99), "makespan_s": round(makespan, 3), "stats": stats},
        "plain_sequential": {"circuits_per_s": round(len(reqs) / plain_makespan, 2), "p50_ms": pct(lat_plain, 50),
                             "p99_ms": pct(lat_plain, 99), "service_s_total": round(float(np.sum(service)), 3),
                             "method": "each request alone through the per-gate C ABI path; latency from the same "
                                       "arrival times with first-in first-out queueing"},
        "fused_sequential": {"circuits_per_s": round(len(reqs) / fused_makespan, 2), "p50_ms": pct(lat_fused, 50),
                             "p99_ms": pct(lat_fused, 99), "service_s_total": round(float(np.sum(fused_service)), 3),
                             "method": "each request alone through the fused executor (compile, |0>, circuit, "
                                       "observable, sampling), one at a time, FIFO from the same arrival times"},
        "speedup_circuits_per_s": round(plain_makespan / makespan, 2),
        "speedup_over_fused_sequential": round(fused_makespan / makespan, 2),
        "energies_match": f"{match_e}/{len(reqs)}", "samples_identical": f"{match_s}/{len(reqs)}",
        "unflagged_samples_identical": f"{match_unflagged}/{len(reqs)}",
        "unflagged_samples_differing": diff_unflagged,
        "flagged_samples": f"{flagged_total}/{n_samples}",
    }


def adjoint(args, lib_mod, hip):
    """Adjoint gradients: 20-qubit hardware-efficient ansatz (12 layers of RY, RZ
    on every qubit + CZ chain = 480 parameters), Heisenberg XXX open chain (57
    terms), `adj_batch` parameter sets in one forward/backward."""
    import torch

    from oracle import oracle as O
    from qulacs_amd.adjoint import AdjointCircuit, hardware_efficient_ansatz, heisenberg_terms

    n, layers, B = 20, 12, args.adj_batch
    gates, P = hardware_efficient_ansatz(n, layers)
    terms = heisenberg_terms(n)
    rng = np.random.default_rng(args.seed + 7)
    theta_np = rng.uniform(0, 2 * np.pi, (B, P))
    theta = torch.tensor(theta_np, device="cuda:0")
    circ = AdjointCircuit(n, gates, P)
    circ.run(theta[:2], terms)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    energy, grad = circ.run(theta, terms)
    torch.cuda.synchronize()
    gpu_s = time.perf_counter() - t0
    energy, grad = energy.cpu().numpy(), grad.cpu().numpy()
    # central differences on two parameters of set 0 (step 1e-6), through the same forward
    fd = []
    for p in (0, P - 1):
        tp = theta[:1].clone()
        tm = theta[:1].clone()
        tp[0, p] += 1e-6
        tm[0, p] -= 1e-6
        ep, _ = circ.run(tp, terms)
        em, _ = circ.run(tm, terms)
        fd.append(abs((ep.item() - em.item()) / 2e-6 - grad[0, p]))
    scale = sum(abs(t[3]) for t in terms)
    out = {"workload": f"{n}-qubit ansatz, {layers} layers, {P} parameters, {len(terms)}-term Heisenberg chain, "
                       f"{B} parameter sets", "gates": len(gates),
           "gpu_s": round(gpu_s, 3), "gradients_per_s": round(B / gpu_s, 2),
           "fd_max_abs_diff": max(fd), "fd_within_1e-8": bool(max(fd) <= 1e-8 * scale)}
    if not args.no_cpu_baseline and O.reference_available():
        # the reference's backprop loop (src/vqcsim/parametric_circuit.cpp:205-335) driven
        # through the compiled csim kernels, one parameter set
        k = O.Reference()
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
        threads = min(threads, os.cpu_count() or threads)
        k.set_num_threads(threads)
        th = theta_np[0]
        t0 = time.perf_counter()

        def apply(g, psi, inverse=False):
            if g[0] == "rot":
                _, qs, ps, p, c = g
                k.pauli_rotation(qs, ps, -c * th[p] if inverse else c * th[p], psi)
            else:
                k.two("CZ", g[1][0], g[1][1], psi)

        psi = np.zeros(1 << n, dtype=np.complex128)
        psi[0] = 1
        for g in gates:
            apply(g, psi)
        lam = np.zeros_like(psi)
        work = np.empty_like(psi)
        for bf, pf, rot, c in terms:
            work[:] = psi
            tl, pl = _mask_lists(bf, pf, n)
            k.pauli(tl, pl, work)
            k.add_with_coef(c, work, lam)
        k.multiply(2.0, lam)
        g_cpu = np.zeros(P)
        for g in reversed(gates):
            if g[0] == "rot":
                work[:] = psi
                k.pauli_rotation(g[1], g[2], np.pi, work)
                g_cpu[g[3]] += g[4] * k.inner_product(lam, work).real / 2
            apply(g, lam, inverse=True)
            apply(g, psi, inverse=True)
        cpu_s = time.perf_counter() - t0
        out["cpu_baseline"] = {"s_per_gradient": round(cpu_s, 3), "cores": threads, "kind": "reference",
                               "sample": "one parameter set, the reference backprop loop over the compiled csim "
                                         "kernels"}
        out["gpu_speedup_per_gradient"] = round(cpu_s / (gpu_s / B), 1)
        out["cpu_gpu_max_abs_diff"] = float(np.abs(g_cpu - grad[0]).max())
        out["cpu_gpu_within_1e-10"] = bool(np.abs(g_cpu - grad[0]).max() <= 1e-10 * scale)
    circ.close()
    return out


def _mask_lists(bf, pf, n):
    t, p = [], []
    for q in range(n):
        x, z = (bf >> q) & 1, (pf >> q) & 1
        if x or z:
            t.append(q)
            p.append(2 if (x and z) else (1 if x else 3))
    return t, p


def _cm(lib_mod, m):
    a = np.ascontiguousarray(m, dtype=np.complex128).ravel()
    arr = (lib_mod.QmComplex * a.size)()
    C.memmove(arr, a.ctypes.data, a.nbytes)
    return arr


def cpu_baseline(args, gates, n, k, kind, threads, toucher):
    """Reference csim on the host cores, on a time-boxed prefix of the same sweep
    (the state first touched by the OpenMP threads)."""
    from qulacs_amd.workloads import alg_bytes, rotation_matrix

    dim = 1 << n
    psi = np.empty(dim, dtype=np.complex128)
    toucher.parallel_fill(psi, 1.0 / np.sqrt(dim))
    done_bytes = 0
    spent = 0.0
    count = 0
    for g in gates:
        t0 = time.perf_counter()
        if g.kind == "dense1":
            k.dense1(g.targets[0], g.matrix, psi)
        elif g.kind == "dense2":
            k.dense2(g.targets[0], g.targets[1], g.matrix, psi)
        elif g.kind == "dense5":
            k.dense(g.targets, g.matrix, psi)
        elif g.kind in ("RX", "RY"):
            k.dense1(g.targets[0], rotation_matrix(g.kind, g.angle), psi)
        elif g.kind == "RZ":
            m = rotation_matrix("RZ", g.angle)
            k.diag1(g.targets[0], [m[0, 0], m[1, 1]], psi)
        else:
            k.two(g.kind, g.targets[0], g.targets[1], psi)
        spent += time.perf_counter() - t0
        done_bytes += alg_bytes(g, n)
        count += 1
        if spent >= args.cpu_seconds:
            break
    return {"value": round(done_bytes / spent / 1e9, 2), "unit": "GB/s", "cores": threads, "kind": kind,
            "sample": f"first {count} of {len(gates)} gates of the same {n}-qubit sweep, {spent:.1f} s, "
                      f"{threads} OpenMP threads"}


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--gates", type=int, default=200)
    ap.add_argument("--dm-qubits", type=int, default=14)
    ap.add_argument("--seed", type=int, default=2002)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-dm", action="store_true")
    ap.add_argument("--small-count", type=int, default=4096)
    ap.add_argument("--small-steps", type=int, default=5)
    ap.add_argument("--small-plain", type=int, default=64)
    ap.add_argument("--no-small", action="store_true")
    ap.add_argument("--serve-count", type=int, default=200)
    ap.add_argument("--serve-rate", type=float, default=100.0)
    ap.add_argument("--serve-max-qubits", type=int, default=28)
    ap.add_argument("--serve-budget-gb", type=float, default=288.0)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--fused-qubits", type=int, default=26)
    ap.add_argument("--fused-layers", type=int, default=40)
    ap.add_argument("--no-fused", action="store_true")
    ap.add_argument("--adj-batch", type=int, default=256)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--cpu-worker", choices=["sweep", "brickwork", "small"], help=argparse.SUPPRESS)
    ap.add_argument("--only", choices=["sweep", "dm", "brickwork", "fused", "small", "adjoint", "serve"],
                    help="run one leg alone (for profiling): 'sweep' prints the full line with the other legs "
                         "left out, the others print their leg's JSON")
    args = ap.parse_args()
    if args.cpu_worker:
        print(json.dumps(cpu_worker(args)), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the sweep is a single 30-qubit "
              "state on one device)", file=sys.stderr)
        sys.exit(2)

    from qulacs_amd import _lib

    hip = HipEvents()
    if args.only == "sweep":
        args.no_dm = args.no_small = args.no_serve = args.no_adjoint = args.no_fused = True
    elif args.only:
        leg = {"dm": dm_sweep, "brickwork": brickwork, "fused": fused_large, "small": many_small,
               "adjoint": adjoint, "serve": serving}[args.only]
        print(json.dumps({args.only: leg(args, _lib, hip)}, default=str), flush=True)
        return
    print(f"[bench] {args.qubits}-qubit sweep, {args.gates} gates, warmup {args.warmup}, steps {args.steps}",
          file=sys.stderr, flush=True)
    sv = sv_sweep(args, _lib, hip)
    print(f"[bench] sweep {sv['value']:.1f} GB/s, {sv['ms_per_step']:.1f} ms/step", file=sys.stderr, flush=True)
    dm = None
    if not args.no_dm:
        dm = dm_sweep(args, _lib, hip)
        print(f"[bench] dm sweep {dm['ms']} ms", file=sys.stderr, flush=True)
    c20 = None
    if args.only != "sweep":
        c20 = brickwork(args, _lib, hip)
        print(f"[bench] 20q brickwork {c20['gpu_ms']} ms", file=sys.stderr, flush=True)
    fused = None
    if not args.no_fused:
        fused = fused_large(args, _lib, hip)
        print(f"[bench] fused {args.fused_qubits}q brickwork K3 {fused['K3']['ms']} ms, K4 {fused['K4']['ms']} ms",
              file=sys.stderr, flush=True)
    small = None
    if not args.no_small:
        small = many_small(args, _lib, hip)
        print(f"[bench] many small circuits {small['circuits_per_s']} circuits/s ({small['ms_per_step']} ms)",
              file=sys.stderr, flush=True)
    adj = None
    if not args.no_adjoint:
        adj = adjoint(args, _lib, hip)
        print(f"[bench] adjoint {adj}", file=sys.stderr, flush=True)
    serve = None
    if not args.no_serve:
        serve = serving(args, _lib, hip)
        print(f"[bench] serving {serve['served']} vs plain {serve['plain_sequential']}", file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = run_cpu_worker("sweep", args, cpu_threads())
        cpu["host"] = host_cpu_info()
        print(f"[bench] cpu baseline {cpu}", file=sys.stderr, flush=True)
    line = {
        "metric": "30q gate-sweep HBM GB/s",
        "value": round(sv["value"], 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sv["ms_per_step"], 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Haar gates, Haar-random initial state)",
        "config": {"workload": f"single large state: {args.qubits}-qubit complex128 state vector, {args.gates}-gate "
                               "sweep over 1q/2q Haar dense, RX/RY/RZ, CZ/CNOT and fused 5q dense",
                   "qubits": args.qubits, "gates": args.gates, "seed": args.seed},
        "roofline": sv["roofline"],
        "cpu_baseline": cpu,
        "hbm_measured": sv["hbm_measured"],
        "per_kind": sv["per_kind"],
        "per_kernel": sv["per_kernel"],
        "dm_sweep": dm,
        "brickwork_20q": c20,
        "fused_executor": fused,
        "many_small_circuits": small,
        "serving": serve,
        "adjoint": adj,
        "norm_after": sv["norm_after"],
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
