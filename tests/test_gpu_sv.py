"""State-vector parity on the GPU, through the C ABI.

Bar (complex128 throughout): states within 1e-10 relative L2 of the oracle (the
golden vectors of the compiled reference are met at 1e-12); expectations within
1e-10 * max(|E|, sum |c_k|); sampled indices identical for identical uniforms.
"""
from __future__ import annotations

import numpy as np
import pytest

from qulacs_amd import _lib
from tests.golden.gen import haar_unitary, pauli_masks, random_pauli_terms, random_state
from tests.gpu_util import DevState, apply_golden_case, cmat, cval, rel, to_complex, u32
from tests.test_oracle import golden_sv_cases

pytestmark = pytest.mark.gpu

STATE_TOL = 1e-10


def test_golden_sv_gates_through_c_abi(golden):
    g = golden("sv_gates.npz")
    dim = 1 << int(g["n"])
    with DevState(dim) as s:
        for i, case in golden_sv_cases(g):
            print("case", i, case[0], case[2], flush=True)
            s.upload(g["psi_in"][i])
            apply_golden_case(case, s, dim)
            assert rel(s.download(), g["psi_out"][i]) < 1e-12, (i, case[0], case[2])


def _random_gate(rng, n):
    kind = int(rng.integers(0, 11))
    if kind == 0:
        t = int(rng.integers(n))
        return ("dense", [t], [], [], haar_unitary(rng, 1))
    if kind == 1:
        return ("dense", rng.choice(n, 2, replace=False).tolist(), [], [], haar_unitary(rng, 2))
    if kind == 2:
        k = int(rng.integers(3, 6))
        return ("dense", rng.choice(n, k, replace=False).tolist(), [], [], haar_unitary(rng, k))
    if kind == 3:
        c, t = rng.choice(n, 2, replace=False).tolist()
        return ("CNOT", [t], [c], [1], None)
    if kind == 4:
        c, t = rng.choice(n, 2, replace=False).tolist()
        return ("CZ", [t], [c], [1], None)
    if kind == 5:
        a, b = rng.choice(n, 2, replace=False).tolist()
        return ("SWAP", [a, b], [], [], None)
    if kind == 6:
        k = int(rng.integers(1, 6))
        return ("pauli_rot", rng.choice(n, k, replace=False).tolist(), [], rng.integers(0, 4, k).tolist(),
                float(rng.uniform(0, 2 * np.pi)))
    if kind == 7:
        t = int(rng.integers(n))
        th = float(rng.uniform(0, 2 * np.pi))
        return ("diag1", [t], [], [], np.array([np.exp(1j * th / 2), np.exp(-1j * th / 2)]))
    if kind == 8:
        q = rng.choice(n, 4, replace=False).tolist()
        return ("cdense", q[2:], q[:2], rng.integers(0, 2, 2).tolist(), haar_unitary(rng, 2))
    if kind == 9:
        q = rng.choice(n, 3, replace=False).tolist()
        return ("cdense", q[1:2], q[:1], [int(rng.integers(0, 2))], haar_unitary(rng, 1))
    q = rng.choice(n, 3, replace=False).tolist()
    return ("cdiag", q[:2], q[2:], [1], np.exp(1j * rng.uniform(0, 6.3, 4)))


def _apply_oracle(o, gate, psi):
    kind, tg, cl, cv, m = gate
    if kind == "dense":
        o.dense(tg, m, psi)
    elif kind in ("CNOT", "CZ"):
        o.two(kind, cl[0], tg[0], psi)
    elif kind == "SWAP":
        o.two("SWAP", tg[0], tg[1], psi)
    elif kind == "pauli_rot":
        o.pauli_rotation(tg, cv, m, psi)
    elif kind == "diag1":
        o.diag1(tg[0], m, psi)
    elif kind == "cdense":
        o.controlled_dense(cl, cv, tg, m, psi)
    elif kind == "cdiag":
        o.controlled_diag(cl, cv, tg, m, psi)


def _apply_gpu(gate, s, dim):
    kind, tg, cl, cv, m = gate
    c = _lib.call
    if kind == "dense":
        c("multi_qubit_dense_matrix_gate_host", u32(tg), len(tg), cmat(m), s.ptr, dim, None, 0)
    elif kind in ("CNOT", "CZ"):
        c(f"{kind}_gate_host", cl[0], tg[0], s.ptr, dim, None, 0)
    elif kind == "SWAP":
        c("SWAP_gate_host", tg[0], tg[1], s.ptr, dim, None, 0)
    elif kind == "pauli_rot":
        c("multi_qubit_Pauli_rotation_gate_partial_list_host", u32(tg), u32(cv), len(tg), m, s.ptr, dim, None, 0)
    elif kind == "diag1":
        c("single_qubit_diagonal_matrix_gate_host", tg[0], cmat(m), s.ptr, dim, None, 0)
    elif kind == "cdense":
        c("multi_qubit_control_multi_qubit_dense_matrix_gate_host", u32(cl), u32(cv), len(cl), u32(tg), len(tg),
          cmat(m), s.ptr, dim, None, 0)
    elif kind == "cdiag":
        c("qm_multi_qubit_control_multi_qubit_diagonal_matrix_gate_host", u32(cl), u32(cv), len(cl), u32(tg),
          len(tg), cmat(m), s.ptr, dim, None, 0)


@pytest.mark.parametrize("n,seed", [(10, 1), (16, 2), (20, 3)])
def test_random_circuits_match_oracle(oracle, n, seed):
    rng = np.random.default_rng(seed)
    dim = 1 << n
    psi = random_state(rng, n)
    gates = [_random_gate(rng, n) for _ in range(80)]
    ref = psi.copy()
    for gte in gates:
        _apply_oracle(oracle, gte, ref)
    with DevState.from_host(psi) as s:
        for gte in gates:
            _apply_gpu(gte, s, dim)
        assert rel(s.download(), ref) < STATE_TOL


@pytest.mark.parametrize("k", [4, 5])
def test_dense_on_matrix_cores_with_a_partial_last_group(oracle, k):
    """The 4- and 5-qubit dense kernels take tuples 16 at a time; with fewer than
    16 tuples (or a count that is not a multiple of 16) the lanes past the end
    hold copies of the last tuple.  Every state size from k to k + 5 qubits,
    targets in random order, against the oracle."""
    rng = np.random.default_rng(50 + k)
    for n in range(k, k + 6):
        dim = 1 << n
        psi = random_state(rng, n)
        ref = psi.copy()
        with DevState.from_host(psi) as s:
            for _ in range(3):
                tg = rng.choice(n, k, replace=False).tolist()
                u = haar_unitary(rng, k)
                oracle.dense(tg, u, ref)
                _lib.call("multi_qubit_dense_matrix_gate_host", u32(tg), k, cmat(u), s.ptr, dim, None, 0)
            assert rel(s.download(), ref) < STATE_TOL, (k, n)


@pytest.mark.parametrize("k", [4, 5])
def test_controlled_dense_on_four_and_five_targets(oracle, k):
    """4- and 5-target dense gates with one or two controls (either control value),
    target sets both scattered and clustered (the 4-target case takes the register
    kernel or the matrix-core kernel by target set), against the oracle."""
    rng = np.random.default_rng(70 + k)
    n = 12
    dim = 1 << n
    psi = random_state(rng, n)
    ref = psi.copy()
    with DevState.from_host(psi) as s:
        for tg in ([0, 1, 2, 3, 4][:k], [6, 7, 8, 9, 10][:k], rng.choice(n, k, replace=False).tolist(),
                   rng.choice(n, k, replace=False).tolist()):
            rest = [q for q in range(n) if q not in tg]
            for nc in (1, 2):
                cl = rng.choice(rest, nc, replace=False).tolist()
                cv = rng.integers(0, 2, nc).tolist()
                u = haar_unitary(rng, k)
                oracle.controlled_dense(cl, cv, tg, u, ref)
                _lib.call("multi_qubit_control_multi_qubit_dense_matrix_gate_host", u32(cl), u32(cv), nc, u32(tg), k,
                          cmat(u), s.ptr, dim, None, 0)
        assert rel(s.download(), ref) < STATE_TOL


@pytest.mark.parametrize("n", [8, 14])
def test_every_target_position_dense_and_named(oracle, n):
    """Each single-qubit kind on every target, each 2q dense on low/low, low/high and high/high pairs."""
    rng = np.random.default_rng(n)
    dim = 1 << n
    psi = random_state(rng, n)
    ref = psi.copy()
    with DevState.from_host(psi) as s:
        for t in range(n):
            u = haar_unitary(rng, 1)
            oracle.dense1(t, u, ref)
            _lib.call("single_qubit_dense_matrix_gate_host", t, cmat(u), s.ptr, dim, None, 0)
            for nm in ("X", "Y", "Z", "H", "P0") if t % 3 == 0 else ("X", "Y", "H"):
                oracle.named(nm, t, ref)
                _lib.call(f"{nm}_gate_host", t, s.ptr, dim, None, 0)
        for a, b in [(0, 1), (1, 0), (0, n - 1), (n - 1, 2), (n - 2, n - 1), (5, 6), (6, 5)]:
            u = haar_unitary(rng, 2)
            oracle.dense2(a, b, u, ref)
            _lib.call("double_qubit_dense_matrix_gate_host", a, b, cmat(u), s.ptr, dim, None, 0)
        assert rel(s.download(), ref) < STATE_TOL


def test_stats_golden(golden):
    g = golden("sv_stats.npz")
    n = int(g["n"])
    dim = 1 << n
    with DevState.from_host(g["psi"]) as s, DevState.from_host(g["bra"]) as b:
        c = _lib.call
        assert abs(c("state_norm_squared_host", s.ptr, dim, None, 0) - g["norm"]) < 1e-12
        assert abs(to_complex(c("inner_product_host", b.ptr, s.ptr, dim, None, 0)) - complex(g["inner"])) < 1e-12
        for t in range(n):
            assert abs(c("M0_prob_host", t, s.ptr, dim, None, 0) - g["M0"][t]) < 1e-12
            assert abs(c("M1_prob_host", t, s.ptr, dim, None, 0) - g["M1"][t]) < 1e-12
        assert abs(c("measurement_distribution_entropy_host", s.ptr, dim, None, 0) - g["entropy"]) < 1e-12
        for k in range(len(g["marginal"])):
            t = [int(x) for x in g["marg_targets"][k] if x >= 0]
            v = [int(x) for x in g["marg_values"][k] if x >= 0]
            assert abs(c("marginal_prob_host", u32(t), u32(v), len(t), s.ptr, dim, None, 0) - g["marginal"][k]) < 1e-12
        for k in range(len(g["expectation"])):
            t = [int(x) for x in g["term_targets"][k] if x >= 0]
            p = [int(x) for x in g["term_paulis"][k] if x >= 0]
            e = c("expectation_value_multi_qubit_Pauli_operator_partial_list_host", u32(t), u32(p), len(t), s.ptr, dim,
                  None, 0)
            assert abs(e - g["expectation"][k]) < 1e-12
            z = c("transition_amplitude_multi_qubit_Pauli_operator_partial_list_host", u32(t), u32(p), len(t), b.ptr,
                  s.ptr, dim, None, 0)
            assert abs(to_complex(z) - g["transition"][k]) < 1e-12


@pytest.mark.parametrize("n", [12, 20])
def test_reductions_match_oracle(oracle, n):
    rng = np.random.default_rng(100 + n)
    dim = 1 << n
    psi = random_state(rng, n) * 0.9
    bra = random_state(rng, n)
    c = _lib.call
    with DevState.from_host(psi) as s, DevState.from_host(bra) as b:
        assert abs(c("state_norm_squared_host", s.ptr, dim, None, 0) - oracle.norm_squared(psi)) < 1e-12
        assert abs(to_complex(c("inner_product_host", b.ptr, s.ptr, dim, None, 0)) - oracle.inner_product(bra, psi)) < 1e-12
        for t in (0, 1, n // 2, n - 1):
            assert abs(c("M0_prob_host", t, s.ptr, dim, None, 0) - oracle.M0(t, psi)) < 1e-12
        terms = random_pauli_terms(rng, n, 16)
        total = 0.0
        for tg, pl, coef in terms:
            e_ref = oracle.expectation(tg, pl, psi)
            total += coef * e_ref
            e = c("expectation_value_multi_qubit_Pauli_operator_partial_list_host", u32(tg), u32(pl), len(tg), s.ptr,
                  dim, None, 0)
            assert abs(e - e_ref) < 1e-11
            bf, pf, rot, piv = pauli_masks(tg, pl)
            if bf:
                e2 = c("expectation_value_multi_qubit_Pauli_operator_XZ_mask_host", bf, pf, rot, piv, s.ptr, dim, None, 0)
            else:
                e2 = c("expectation_value_multi_qubit_Pauli_operator_Z_mask_host", pf, s.ptr, dim, None, 0)
            assert abs(e2 - e_ref) < 1e-11
            z = c("transition_amplitude_multi_qubit_Pauli_operator_partial_list_host", u32(tg), u32(pl), len(tg),
                  b.ptr, s.ptr, dim, None, 0)
            assert abs(to_complex(z) - oracle.transition(tg, pl, bra, psi)) < 1e-11
        # whole observable in one call
        masks = [pauli_masks(tg, pl) for tg, pl, _ in terms]
        bfs = (_lib.C.c_ulonglong * len(terms))(*[m[0] for m in masks])
        pfs = (_lib.C.c_ulonglong * len(terms))(*[m[1] for m in masks])
        rots = u32([m[2] for m in masks])
        coefs = cmat([cf for _, _, cf in terms])
        z = to_complex(c("qm_observable_expectation_host", bfs, pfs, rots, coefs, len(terms), s.ptr, dim, None, 0))
        scale = max(abs(total), sum(abs(cf) for _, _, cf in terms))
        assert abs(z.real - total) <= 1e-10 * scale and abs(z.imag) <= 1e-10 * scale


def test_observable_with_hundreds_of_terms_on_23_qubits(oracle):
    """More than 256 Z-only terms on a state that needs the full 2048 reduction
    blocks: the terms run in chunks of 256 whose partial sums fill the
    reduction workspace exactly."""
    rng = np.random.default_rng(2323)
    n = 23
    psi = random_state(rng, n)
    terms = []
    for _ in range(260):
        tg = sorted(rng.choice(n, int(rng.integers(1, 5)), replace=False).tolist())
        terms.append((tg, [3] * len(tg), float(rng.standard_normal())))
    terms += random_pauli_terms(rng, n, 12)
    masks = [pauli_masks(tg, pl) for tg, pl, _ in terms]
    k = len(terms)
    bfs = (_lib.C.c_ulonglong * k)(*[m[0] for m in masks])
    pfs = (_lib.C.c_ulonglong * k)(*[m[1] for m in masks])
    with DevState.from_host(psi) as s:
        z = to_complex(_lib.call("qm_observable_expectation_host", bfs, pfs, u32([m[2] for m in masks]),
                                 cmat([cf for *_, cf in terms]), k, s.ptr, s.dim, None, 0))
    want = sum(cf * oracle.expectation(tg, pl, psi) for tg, pl, cf in terms)
    scale = max(abs(want), sum(abs(cf) for *_, cf in terms))
    assert abs(z.real - want) <= 1e-10 * scale and abs(z.imag) <= 1e-10 * scale, (z, want)


def _sampling_states(rng, n):
    dim = 1 << n
    yield "random", random_state(rng, n)
    b = np.zeros(dim, dtype=np.complex128)
    b[dim // 3] = 1.0
    yield "basis", b
    yield "uniform", np.full(dim, 1 / np.sqrt(dim), dtype=np.complex128)
    z = random_state(rng, n)
    z[: dim // 4] = 0
    z[dim // 2: dim // 2 + dim // 8] = 0
    yield "zero_runs", z / np.linalg.norm(z)
    w = random_state(rng, n) * np.exp(-np.arange(dim) / (dim / 30))
    yield "decaying", w / np.linalg.norm(w)
    # short-mantissa probabilities: ties in the sequential rounding
    t = np.sqrt(np.resize([3.0, 1.0, 5.0, 7.0], dim) * 2.0 ** -(n + 2)).astype(np.complex128)
    yield "ties", t
    # growing amplitudes: the running sum crosses a binade every few dozen
    # elements over the whole array, not only at its start
    g = random_state(rng, n) * np.exp(np.arange(dim) / (dim / 30))
    yield "growing", g / np.linalg.norm(g)
    # spikes: runs of four consecutive elements each at least twice the running
    # sum in front of it (several binade crossings inside one group of four),
    # on a floor of tiny values with short mantissas (ties)
    sp = np.sqrt(np.resize([3.0, 1.0, 5.0, 7.0], dim) * 2.0 ** -40).astype(np.complex128)
    step = max(dim // 64, 16)
    for k, i in enumerate(range(step // 2, dim - 4, step)):
        i -= i % 4
        sp[i: i + 4] = 2.0 ** (4 * k) * np.array([1.0, 2.0, 4.0, 8.0]) * (1 + 0.3j)
    yield "spikes", sp / np.linalg.norm(sp)
    # a few exact ties per 4096 elements on a random state: probability 2^-54
    # is half a unit of the sum's grid while the sum is in [1/2, 1), 2^-55 while
    # it is in [1/4, 1/2) (not renormalised, which would round them)
    st = random_state(rng, n)
    idx = np.arange(dim)
    pick = (idx % 1024) == 7
    st[pick & (idx >= dim // 2)] = 2.0 ** -27
    st[pick & (idx < dim // 2)] = 2.0 ** -28 * (1 + 1j)
    yield "sparse_ties", st


@pytest.mark.parametrize("n", [6, 12, 18])
def test_sampling_identical_to_sequential_cpu(oracle, n):
    rng = np.random.default_rng(200 + n)
    dim = 1 << n
    for label, psi in _sampling_states(rng, n):
        cdf = oracle.sampling_cdf(psi)
        us = list(rng.random(1000))
        us += [cdf[1], cdf[dim // 2], cdf[-1], float(np.nextafter(cdf[dim // 3], 0)), 0.0, 2.0]
        us = np.array(us)
        want = oracle.sampling(psi, us)
        out = np.zeros(us.size, dtype=np.uint64)
        with DevState.from_host(psi) as s:
            _lib.call("qm_sampling_host", us.ctypes.data_as(_lib.PD), us.size,
                      out.ctypes.data_as(_lib.PI), s.ptr, dim, None, 0)
        assert np.array_equal(out, want), (label, np.nonzero(out != want)[0][:5])


@pytest.mark.parametrize("n", [15, 16, 17, 20, 22, 23])
def test_sampling_cdf_bit_identical(oracle, n):
    """The device CDF (single work-group below 2^16 amplitudes, the multi-block
    chunk walk from 2^16 up) equals the sequential loop's stacked sums bit for bit."""
    rng = np.random.default_rng(300 + n)
    dim = 1 << n
    states = list(_sampling_states(rng, n))
    # a state whose running sum crosses many binades inside single chunks
    g = random_state(rng, n) * (2.0 ** (-np.arange(dim) % 97))
    states.append(("binades", g / np.linalg.norm(g)))
    for label, psi in states:
        want = oracle.sampling_cdf(psi)
        got = np.zeros(dim + 1)
        with DevState.from_host(psi) as s:
            _lib.call("qm_sampling_cdf_host", got.ctypes.data_as(_lib.PD), s.ptr, dim, None, 0)
        bad = np.nonzero(got.view(np.uint64) != want.view(np.uint64))[0]
        assert bad.size == 0, (label, bad[:5])


def test_state_arithmetic_and_init(oracle):
    rng = np.random.default_rng(5)
    n = 11
    dim = 1 << n
    a, b = random_state(rng, n), random_state(rng, n)
    c = _lib.call
    with DevState.from_host(a) as s, DevState.from_host(b) as t:
        c("state_add_host", t.ptr, s.ptr, dim, None, 0)
        assert rel(s.download(), a + b) < 1e-14
        c("state_multiply_host", cval(0.3 - 0.2j), s.ptr, dim, None, 0)
        assert rel(s.download(), (a + b) * (0.3 - 0.2j)) < 1e-14
        c("qm_state_add_with_coef_host", cval(1.5j), t.ptr, s.ptr, dim, None, 0)
        ref = (a + b) * (0.3 - 0.2j) + 1.5j * b
        assert rel(s.download(), ref) < 1e-14
        nrm = c("state_norm_squared_host", s.ptr, dim, None, 0)
        c("normalize_host", nrm, s.ptr, dim, None, 0)
        assert rel(s.download(), ref / np.linalg.norm(ref)) < 1e-13
        c("copy_quantum_state_from_device_to_device", t.ptr, s.ptr, dim, None, 0)
        assert np.array_equal(t.download(), s.download())
        c("set_computational_basis_host", 37, s.ptr, dim, None, 0)
        e = np.zeros(dim, dtype=np.complex128); e[37] = 1
        assert np.array_equal(s.download(), e)
        c("initialize_quantum_state_host", s.ptr, dim, None, 0)
        e = np.zeros(dim, dtype=np.complex128); e[0] = 1
        assert np.array_equal(s.download(), e)
        c("initialize_Haar_random_state_with_seed_host", s.ptr, dim, 7, None, 0)
        h1 = s.download()
        c("initialize_Haar_random_state_with_seed_host", s.ptr, dim, 7, None, 0)
        assert np.array_equal(s.download(), h1)  # same seed, same state
        assert abs(np.linalg.norm(h1) - 1) < 1e-12


def test_streams_and_async_mode(oracle):
    """Stream arrays from allocate_cuda_stream_host, and stream-ordered (async) updates."""
    rng = np.random.default_rng(6)
    n = 14
    dim = 1 << n
    psi = random_state(rng, n)
    ref = psi.copy()
    st = _lib.call("allocate_cuda_stream_host", 2, 0)
    _lib.call("qm_set_synchronous", 0)
    try:
        with DevState.from_host(psi) as s:
            for _ in range(30):
                t = int(rng.integers(n))
                u = haar_unitary(rng, 1)
                oracle.dense1(t, u, ref)
                _lib.call("single_qubit_dense_matrix_gate_host", t, cmat(u), s.ptr, dim, st, 0)
            e = _lib.call("expectation_value_multi_qubit_Pauli_operator_partial_list_host", u32([0]), u32([3]), 1,
                          s.ptr, dim, st, 0)
            assert abs(e - oracle.expectation([0], [3], ref)) < 1e-12
            _lib.call("qm_synchronize", st, 0)
            assert rel(s.download(), ref) < STATE_TOL
    finally:
        _lib.call("qm_set_synchronous", 1)
        _lib.call("release_cuda_stream_host", st, 2, 0)
