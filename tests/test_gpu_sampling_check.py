"""Sampling with the boundary check on the GPU (qm_sampling_checked_host,
dm_sampling_checked_host, qm_server_wait_checked).

The reference's contract (src/cppsim/state.hpp:909-930) returns
lower_bound(cdf, u) - 1 over the sequential CDF of its own state.  A state that
a fused circuit produced, or the CPU's, differs from the gate-by-gate device
state by rounding, so its CDF differs by a few units in the last place; a
uniform that close to a step may select the neighbouring index.  These tests
check that (1) every such uniform is flagged, (2) every unflagged sample is the
one the oracle's own state gives, at sizes up to the serving workload's, and
(3) the measured CDF differences stay well inside the margin.
"""
from __future__ import annotations

import threading

import numpy as np
import pytest

from qulacs_amd import _lib
from qulacs_amd.batch import _marshal
from qulacs_amd.serving import Server
from qulacs_amd.workloads import brickwork_circuit, serving_requests
from tests.batch_util import oracle_run
from tests.gpu_util import DevState, cmat

pytestmark = pytest.mark.gpu


def _pergate(n, gates) -> DevState:
    """The gate-by-gate path (one C ABI call per gate), from |0>."""
    dim = 1 << n
    s = DevState(dim)
    _lib.call("initialize_quantum_state_host", s.ptr, dim, None, 0)
    for kind, q, m in gates:
        if kind == "CNOT":
            _lib.call("CNOT_gate_host", q[0], q[1], s.ptr, dim, None, 0)
        else:
            _lib.call("single_qubit_dense_matrix_gate_host", q[0], cmat(m), s.ptr, dim, None, 0)
    return s


def _fused(n, gates, block_qubits=3) -> DevState:
    """The fused stage executor (qm_circuit_*), from |0>."""
    dim = 1 << n
    s = DevState(dim)
    _lib.call("initialize_quantum_state_host", s.ptr, dim, None, 0)
    arr, keep = _marshal(gates)
    c = _lib.call("qm_circuit_create", n, arr, len(gates), block_qubits, 0)
    assert c
    try:
        assert _lib.call("qm_circuit_apply", c, s.ptr, dim, None, 0) == 0
    finally:
        _lib.call("qm_circuit_destroy", c)
    del keep
    return s


def _cdf(s: DevState) -> np.ndarray:
    out = np.zeros(s.dim + 1)
    _lib.call("qm_sampling_cdf_host", out.ctypes.data_as(_lib.PD), s.ptr, s.dim, None, 0)
    return out


def _checked(s: DevState, us, margin, dm_dim=None):
    us = np.ascontiguousarray(us, dtype=np.float64)
    out = np.zeros(us.size, dtype=np.uint64)
    flags = np.zeros(us.size, dtype=np.uint8)
    k = _lib.C.c_uint()
    name, dim = ("dm_sampling_checked_host", dm_dim) if dm_dim else ("qm_sampling_checked_host", s.dim)
    _lib.call(name, us.ctypes.data_as(_lib.PD), us.size, float(margin), out.ctypes.data_as(_lib.PI),
              flags.ctypes.data_as(_lib.PB), _lib.C.byref(k), s.ptr, dim, None, 0)
    assert k.value == int(flags.sum())
    return out, flags


def _unchecked(s: DevState, us) -> np.ndarray:
    us = np.ascontiguousarray(us, dtype=np.float64)
    out = np.zeros(us.size, dtype=np.uint64)
    _lib.call("qm_sampling_host", us.ctypes.data_as(_lib.PD), us.size, out.ctypes.data_as(_lib.PI), s.ptr, s.dim,
              None, 0)
    return out


def test_dm_cdf_bit_identical_with_complex_diagonal(oracle):
    """|rho_ii| on the device rounds as the reference's std::abs (libm hypot):
    the device CDF equals the oracle's bit for bit on diagonals whose imaginary
    parts are far from zero, across magnitudes."""
    rng = np.random.default_rng(41)
    for n in (3, 7, 10):
        dim = 1 << n
        print("dm cdf", n, flush=True)
        rho = np.zeros((dim, dim), dtype=np.complex128)
        rho += (rng.standard_normal((dim, dim)) + 1j * rng.standard_normal((dim, dim))) * 1e-3
        mag = rng.random(dim) * np.exp2(-rng.integers(0, 40, dim).astype(np.float64))
        ph = rng.uniform(0, 2 * np.pi, dim)
        diag = mag * np.exp(1j * ph)
        diag[: dim // 8] = diag[: dim // 8].real  # some purely real entries
        rho[np.arange(dim), np.arange(dim)] = diag
        flat = rho.ravel().copy()
        want = oracle.dm_sampling_cdf(flat, dim)
        with DevState.from_host(flat) as s:
            got = np.zeros(dim + 1)
            _lib.call("dm_sampling_cdf_host", got.ctypes.data_as(_lib.PD), s.ptr, dim, None, 0)
            assert np.array_equal(got, want), (n, int(np.sum(got != want)))
            us = np.concatenate([rng.random(400) * want[-1], want[1:], np.nextafter(want[1:], 2.0)])
            out, flags = _checked(s, us, 0.0, dm_dim=dim)
            assert not flags.any()
            assert np.array_equal(out, oracle.dm_sampling(flat, dim, us))


def test_checked_sampling_flags_every_uniform_on_a_step(oracle):
    """A uniform equal to a CDF value, or one unit in the last place from it, is
    flagged; samples are those of qm_sampling_host whatever the margin, and a
    zero margin flags nothing."""
    rng = np.random.default_rng(42)
    for n in (12, 17):
        print("flags", n, flush=True)
        gates, _, _ = brickwork_circuit(n, layers=6, terms=1, samples=1, seed=100 + n)
        with _pergate(n, gates) as s:
            cdf = _cdf(s)
            idx = rng.integers(1, s.dim + 1, 300)
            on = cdf[idx]
            near = np.concatenate([np.nextafter(on, 2.0), np.nextafter(on, -1.0)])
            mid = (cdf[idx - 1] + cdf[idx]) / 2
            wide = (cdf[1:] - cdf[:-1]) > 1e-9 * cdf[-1]  # steps far wider than the margin
            mid = mid[wide[idx - 1]]
            us = np.concatenate([on, near, mid, rng.random(2000)])
            margin = _lib.call("qm_sampling_margin", len(gates))
            out, flags = _checked(s, us, margin)
            assert np.array_equal(out, _unchecked(s, us))
            k = on.size
            assert flags[:3 * k].all()
            assert not flags[3 * k:3 * k + mid.size].any()
            out0, flags0 = _checked(s, us, 0.0)
            assert np.array_equal(out0, out) and not flags0.any()
            psi = s.download()
        # the checked samples of the device state are the oracle's on that state
        assert np.array_equal(out, oracle.sampling(psi, us))


@pytest.mark.parametrize("n,layers", [(16, 40), (20, 30), (22, 20)])
def test_fused_and_pergate_unflagged_samples_match_oracle(oracle, n, layers):
    """The fused executor's state, the gate-by-gate device state and the oracle's
    CPU state differ by rounding only; their CDFs differ by far less than the
    margin, and every sample left unflagged by the check is the oracle's."""
    gates, terms, _ = brickwork_circuit(n, layers=layers, terms=4, samples=1, seed=200 + n)
    rng = np.random.default_rng(43 + n)
    print("fused vs oracle", n, layers, len(gates), flush=True)
    psi_ref, _ = oracle_run(oracle, n, gates, [])
    cdf_ref = oracle.sampling_cdf(psi_ref)
    margin = _lib.call("qm_sampling_margin", len(gates))
    # uniforms: random, plus exactly on the oracle's CDF steps and one ulp away
    idx = rng.integers(1, (1 << n) + 1, 200)
    us = np.concatenate([rng.random(20000), cdf_ref[idx], np.nextafter(cdf_ref[idx], 2.0)])
    want = oracle.sampling(psi_ref, us)
    eps = 2.0 ** -53
    for name, make in (("fused", _fused), ("pergate", _pergate)):
        with make(n, gates) as s:
            cdf = _cdf(s)
            dev = float(np.max(np.abs(cdf - cdf_ref)))
            print(f"  {name}: max|dcdf| = {dev:.3e} = {dev / (np.sqrt(len(gates)) * eps):.1f} sqrt(G) eps, "
                  f"margin*total = {margin * cdf[-1]:.3e}, ratio {dev / (margin * cdf[-1]):.4f}", flush=True)
            assert dev <= margin * cdf[-1] / 4
            out, flags = _checked(s, us, margin)
            ok = flags == 0
            assert np.array_equal(out[ok], want[ok]), (name, int(np.sum(out[ok] != want[ok])))
            # every uniform exactly on one of this state's own steps is flagged
            on_own = cdf[idx]
            _, f_own = _checked(s, on_own, margin)
            assert f_own.all()
            # random uniforms are rarely flagged (expected count 2 * margin * 2^n * 20000)
            assert int(flags[:20000].sum()) <= 20 + 4 * 2 * margin * (1 << n) * 20000, int(flags[:20000].sum())


def test_served_requests_at_serving_scale_against_oracle(oracle):
    """Requests shaped like the serving workload (brickwork, depth 100, 1000
    samples) through the request server: a 24-qubit request against the oracle's
    own CPU evolution, a 26-qubit request against the gate-by-gate device path,
    and small batched requests against the oracle.  Energies agree to 1e-10 and
    every unflagged sample is identical."""
    big = serving_requests(1, rate=1000.0, seed=77, n_range=(24, 24), depth_range=(100, 100), samples=1000)
    big += serving_requests(1, rate=1000.0, seed=78, n_range=(26, 26), depth_range=(100, 100), samples=1000)
    small = serving_requests(6, rate=1000.0, seed=79, n_range=(10, 16), depth_range=(100, 200), samples=1000)
    reqs = big + small
    print("sizes", [(r.n, r.depth, len(r.gates)) for r in reqs], flush=True)
    # the oracle's CPU evolution of the first 24-qubit request runs beside the GPU work
    ref = {}
    th = threading.Thread(target=lambda: ref.setdefault(0, oracle_run(oracle, reqs[0].n, reqs[0].gates,
                                                                      reqs[0].terms)))
    th.start()
    with Server(hbm_budget_bytes=48 << 30, large_slots=3, small_max_qubits=16) as srv:
        ids = [srv.submit(r.n, r.gates, r.terms, r.uniforms) for r in reqs]
        res = [srv.wait_checked(i) for i in ids]
        st = srv.stats()
    print("stats", st, "flagged", [int(f.sum()) for _, _, f, _ in res], flush=True)
    assert st["failed"] == 0
    total_flagged = 0
    for k, (r, (e, smp, flags, _)) in enumerate(zip(reqs, res)):
        scale = max(1.0, sum(abs(t[3]) for t in r.terms))
        ok = flags == 0
        total_flagged += int((~ok).sum())
        if r.n == 26:
            with _pergate(r.n, r.gates) as s:
                want = _unchecked(s, r.uniforms)
                tn = [(t[0], t[1], t[2], t[3]) for t in r.terms]
                bf = (_lib.C.c_ulonglong * len(tn))(*[t[0] for t in tn])
                pf = (_lib.C.c_ulonglong * len(tn))(*[t[1] for t in tn])
                rot = (_lib.C.c_uint * len(tn))(*[t[2] for t in tn])
                co = cmat([t[3] for t in tn])
                ew = _lib.call("qm_observable_expectation_host", bf, pf, rot, co, len(tn), s.ptr, s.dim, None, 0).re
            assert abs(e - ew) <= 1e-10 * scale, (r.n, e, ew)
            assert np.array_equal(smp[ok], want[ok]), (r.n, int(np.sum(smp[ok] != want[ok])))
            continue
        if k == 0:
            th.join()
            psi, ew = ref[0]
        else:
            psi, ew = oracle_run(oracle, r.n, r.gates, r.terms)
        assert abs(e - ew) <= 1e-10 * scale, (r.n, e, ew)
        want = oracle.sampling(psi, r.uniforms)
        assert np.array_equal(smp[ok], want[ok]), (r.n, int(np.sum(smp[ok] != want[ok])))
    th.join()
    assert total_flagged <= 10, total_flagged
