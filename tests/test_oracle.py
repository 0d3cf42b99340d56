"""Pins the CPU oracle (oracle/cpu_ref.cpp) before anything is compared with it.

Three independent anchors:
  * golden vectors written by the reference csim compiled from its own sources
    (tests/golden/*.npz, see make_golden.py);
  * the compiled reference itself on fresh random inputs, when it was built here;
  * closed forms built from definitions in tests/golden/gen.py (expanded
    operators, Kronecker products, U rho U^dagger, brute-force lower_bound over
    an exactly rounded CDF), chosen so that a transposed matrix, a swapped bit
    order, a wrong rotation sign or a dropped term fails at least one of them.
"""
from __future__ import annotations

import bisect
from fractions import Fraction

import numpy as np
import pytest

from tests.golden.gen import (PAULI, controlled, expand, haar_unitary, pauli_matrix, random_pauli_terms,
                              random_state)

TOL = 1e-12


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _valid(a):
    return [int(x) for x in a if x >= 0]


NAMED_MATRICES = {
    "X": PAULI[1], "Y": PAULI[2], "Z": PAULI[3],
    "H": np.array([[1, 1], [1, -1]]) / np.sqrt(2),
    "S": np.diag([1, 1j]), "Sdag": np.diag([1, -1j]),
    "T": np.diag([1, (1 + 1j) / np.sqrt(2)]), "Tdag": np.diag([1, (1 - 1j) / np.sqrt(2)]),
    "sqrtX": np.array([[0.5 + 0.5j, 0.5 - 0.5j], [0.5 - 0.5j, 0.5 + 0.5j]]),
    "sqrtXdag": np.array([[0.5 - 0.5j, 0.5 + 0.5j], [0.5 + 0.5j, 0.5 - 0.5j]]),
    "sqrtY": np.array([[0.5 + 0.5j, -0.5 - 0.5j], [0.5 + 0.5j, 0.5 + 0.5j]]),
    "sqrtYdag": np.array([[0.5 - 0.5j, 0.5 - 0.5j], [-0.5 + 0.5j, 0.5 - 0.5j]]),
    "P0": np.diag([1, 0]), "P1": np.diag([0, 1]),
}


def rot_matrix(name, angle):
    """qulacs convention exp(+i angle P / 2) (src/csim/update_ops_pauli_single.cpp:54-116)."""
    p = {"RX": PAULI[1], "RY": PAULI[2], "RZ": PAULI[3]}[name]
    return np.cos(angle / 2) * np.eye(2) + 1j * np.sin(angle / 2) * p


def apply_case_oracle(o, case, psi):
    """Apply one golden SV case with the oracle; returns the updated state."""
    name, kind, tg, cl, cv, param, mat = case
    out = psi.copy()
    if kind == "dense":
        o.dense(tg, mat.reshape(1 << len(tg), 1 << len(tg)), out)
    elif kind == "cdense":
        o.controlled_dense(cl, cv, tg, mat.reshape(1 << len(tg), 1 << len(tg)), out)
    elif kind == "named":
        if name in ("X", "Y", "Z", "H", "P0", "P1"):
            o.named(name, tg[0], out)
        else:
            o.dense1(tg[0], NAMED_MATRICES[name], out)
    elif kind == "rot":
        o.dense1(tg[0], rot_matrix(name, param), out)
    elif kind == "two":
        o.two(name, cl[0], tg[0], out)
    elif kind == "diag1":
        o.diag1(tg[0], mat, out)
    elif kind == "phase1":
        o.phase1(tg[0], complex(mat[0]), out)
    elif kind == "cdiag":
        o.controlled_diag(cl, cv, tg, mat, out)
    elif kind == "pauli":
        o.pauli(tg, cv, out)
    elif kind == "pauli_rot":
        o.pauli_rotation(tg, cv, param, out)
    else:
        raise AssertionError(kind)
    return out


def golden_sv_cases(g):
    off = g["matrix_offset"]
    for i in range(len(g["name"])):
        yield i, (str(g["name"][i]), str(g["kind"][i]), _valid(g["targets"][i]), _valid(g["controls"][i]),
                  _valid(g["cvals"][i]), float(g["param"][i]), g["matrix"][off[i]:off[i + 1]])


# ---------------------------------------------------------------- golden vectors
def test_oracle_matches_reference_golden_sv_gates(oracle, golden):
    g = golden("sv_gates.npz")
    n_cases = 0
    for i, case in golden_sv_cases(g):
        out = apply_case_oracle(oracle, case, g["psi_in"][i])
        assert rel(out, g["psi_out"][i]) < TOL, (i, case[0], case[1], case[2])
        n_cases += 1
    assert n_cases > 100


def test_oracle_matches_reference_golden_stats(oracle, golden):
    g = golden("sv_stats.npz")
    psi, bra, n = g["psi"], g["bra"], int(g["n"])
    assert abs(oracle.norm_squared(psi) - g["norm"]) < TOL
    assert abs(oracle.inner_product(bra, psi) - complex(g["inner"])) < TOL
    for t in range(n):
        assert abs(oracle.M0(t, psi) - g["M0"][t]) < TOL
        assert abs(oracle.M1(t, psi) - g["M1"][t]) < TOL
    assert abs(oracle.entropy(psi) - g["entropy"]) < TOL
    for k in range(len(g["marginal"])):
        assert abs(oracle.marginal(_valid(g["marg_targets"][k]), _valid(g["marg_values"][k]), psi)
                   - g["marginal"][k]) < TOL
    for k in range(len(g["expectation"])):
        t, p = _valid(g["term_targets"][k]), _valid(g["term_paulis"][k])
        assert abs(oracle.expectation(t, p, psi) - g["expectation"][k]) < TOL
        assert abs(oracle.transition(t, p, bra, psi) - g["transition"][k]) < TOL


def apply_dm_case(o, g, i, rho, dim):
    name = str(g["name"][i])
    tg, cl, cv = _valid(g["targets"][i]), _valid(g["controls"][i]), _valid(g["cvals"][i])
    k = len(tg)
    mat = g["matrix"][i][: 1 << k, : 1 << k]
    out = rho.copy()
    if name == "dm_dense1":
        o.dm_dense1(tg[0], mat, out, dim)
    elif name == "dm_dense":
        o.dm_dense(tg, mat, out, dim)
    elif name == "dm_cdense":
        o.dm_controlled_dense(cl, cv, tg, mat, out, dim)
    elif name == "dm_CNOT":
        o.dm_controlled_dense(cl, [1], tg, PAULI[1], out, dim)
    elif name == "dm_CZ":
        o.dm_controlled_dense(cl, [1], tg, PAULI[3], out, dim)
    elif name == "dm_SWAP":
        sw = np.eye(4)[[0, 2, 1, 3]]
        o.dm_dense([cl[0], tg[0]], sw, out, dim)
    elif name in ("dm_RX", "dm_RY", "dm_RZ"):
        o.dm_dense1(tg[0], rot_matrix(name[3:], float(g["param"][i])), out, dim)
    else:
        nm = name[3:]
        m = NAMED_MATRICES[nm]
        if nm == "T":  # the DM path uses T up to a global phase (src/csim/constant.cpp:9-10)
            m = np.diag([np.exp(-1j * np.pi / 8), np.exp(1j * np.pi / 8)])
        o.dm_dense1(tg[0], m, out, dim)
    return out


def test_oracle_matches_reference_golden_dm(oracle, golden):
    g = golden("dm.npz")
    n = int(g["n"])
    dim = 1 << n
    for i in range(len(g["name"])):
        out = apply_dm_case(oracle, g, i, g["rho_in"][i], dim)
        assert rel(out, g["rho_out"][i]) < TOL, str(g["name"][i])
    rho = g["rho"]
    assert abs(oracle.dm_norm(rho, dim) - g["trace"]) < TOL
    for t in range(n):
        assert abs(oracle.dm_M0(t, rho, dim) - g["M0"][t]) < TOL
    assert abs(oracle.dm_marginal([0, 2], [1, 0], rho, dim) - g["marginal"]) < TOL
    for k in range(len(g["expectation"])):
        t, p = _valid(g["term_targets"][k]), _valid(g["term_paulis"][k])
        assert abs(oracle.dm_expectation(t, p, rho, dim) - g["expectation"][k]) < TOL


def test_golden_comparison_detects_planted_mistakes(oracle, golden):
    """A transposed matrix or swapped target order must fail the golden check."""
    g = golden("sv_gates.npz")
    caught = 0
    for i, case in golden_sv_cases(g):
        name, kind, tg, cl, cv, param, mat = case
        if kind != "dense" or len(tg) < 2:
            continue
        k = len(tg)
        m = mat.reshape(1 << k, 1 << k)
        bad_t = apply_case_oracle(oracle, (name, kind, tg, cl, cv, param, m.T.ravel()), g["psi_in"][i])
        bad_o = apply_case_oracle(oracle, (name, kind, tg[::-1], cl, cv, param, mat), g["psi_in"][i])
        assert rel(bad_t, g["psi_out"][i]) > 1e-3
        assert rel(bad_o, g["psi_out"][i]) > 1e-3
        caught += 1
    assert caught >= 5


# ---------------------------------------------------------------- against the compiled reference
@pytest.mark.parametrize("n", [9, 12])
def test_oracle_matches_compiled_reference_random(oracle, reference, n):
    rng = np.random.default_rng(n)
    psi = random_state(rng, n)
    a, b = psi.copy(), psi.copy()
    for step in range(40):
        kind = step % 8
        if kind == 0:
            t = int(rng.integers(n)); u = haar_unitary(rng, 1)
            oracle.dense1(t, u, a); reference.dense1(t, u, b)
        elif kind == 1:
            t = rng.choice(n, 2, replace=False).tolist(); u = haar_unitary(rng, 2)
            oracle.dense(t, u, a); reference.dense(t, u, b)
        elif kind == 2:
            t = rng.choice(n, 5, replace=False).tolist(); u = haar_unitary(rng, 5)
            oracle.dense(t, u, a); reference.dense(t, u, b)
        elif kind == 3:
            c, t = rng.choice(n, 2, replace=False).tolist()
            oracle.two("CNOT", c, t, a); reference.two("CNOT", c, t, b)
        elif kind == 4:
            c, t = rng.choice(n, 2, replace=False).tolist()
            oracle.two("CZ", c, t, a); reference.two("CZ", c, t, b)
        elif kind == 5:
            tg = rng.choice(n, 4, replace=False).tolist(); pl = rng.integers(0, 4, 4).tolist()
            th = float(rng.uniform(0, 6.3))
            oracle.pauli_rotation(tg, pl, th, a); reference.pauli_rotation(tg, pl, th, b)
        elif kind == 6:
            ct = rng.choice(n, 3, replace=False).tolist(); u = haar_unitary(rng, 2)
            oracle.controlled_dense(ct[:1], [1], ct[1:], u, a); reference.controlled_dense(ct[:1], [1], ct[1:], u, b)
        else:
            t = int(rng.integers(n)); th = float(rng.uniform(0, 6.3))
            d = [np.exp(1j * th / 2), np.exp(-1j * th / 2)]
            oracle.diag1(t, d, a); reference.diag1(t, d, b)
    assert rel(a, b) < 1e-13
    for tg, pl, _c in random_pauli_terms(rng, n, 8):
        assert abs(oracle.expectation(tg, pl, a) - reference.expectation(tg, pl, b)) < 1e-12


# ---------------------------------------------------------------- closed forms
def test_dense_gates_equal_expanded_operators(oracle):
    rng = np.random.default_rng(7)
    n = 6
    for targets in ([0], [5], [1, 4], [4, 1], [0, 2, 5], [3, 1, 0, 5], [5, 4, 3, 2, 1]):
        u = haar_unitary(rng, len(targets))
        psi = random_state(rng, n)
        out = psi.copy()
        oracle.dense(targets, u, out)
        assert rel(out, expand(n, targets, u) @ psi) < TOL, targets
    # separable product: U3 (x) U2 (x) U1 on qubits (2, 1, 0) equals three 1q gates
    us = [haar_unitary(rng, 1) for _ in range(3)]
    psi = random_state(rng, 3)
    a = psi.copy()
    for q in range(3):
        oracle.dense1(q, us[q], a)
    assert rel(a, np.kron(us[2], np.kron(us[1], us[0])) @ psi) < TOL


def test_controlled_and_named_two_qubit_gates(oracle):
    rng = np.random.default_rng(8)
    n = 5
    psi = random_state(rng, n)
    for c, t in [(0, 1), (3, 0), (4, 2)]:
        out = psi.copy(); oracle.two("CNOT", c, t, out)
        assert rel(out, controlled(n, [c], [1], [t], PAULI[1]) @ psi) < TOL
        out = psi.copy(); oracle.two("CZ", c, t, out)
        assert rel(out, controlled(n, [c], [1], [t], PAULI[3]) @ psi) < TOL
        out = psi.copy(); oracle.two("SWAP", c, t, out)
        assert rel(out, expand(n, [c, t], np.eye(4)[[0, 2, 1, 3]]) @ psi) < TOL
    u = haar_unitary(rng, 2)
    out = psi.copy(); oracle.controlled_dense([1, 3], [1, 0], [4, 0], u, out)
    assert rel(out, controlled(n, [1, 3], [1, 0], [4, 0], u) @ psi) < TOL


def test_rotations_and_pauli_follow_qulacs_sign(oracle):
    rng = np.random.default_rng(9)
    n = 5
    psi = random_state(rng, n)
    for tg, pl in (([0, 2, 3], [1, 2, 3]), ([4], [2]), ([1, 3], [3, 3]), ([2, 0], [1, 1])):
        th = 0.731
        P = pauli_matrix(n, tg, pl)
        out = psi.copy(); oracle.pauli_rotation(tg, pl, th, out)
        assert rel(out, (np.cos(th / 2) * np.eye(1 << n) + 1j * np.sin(th / 2) * P) @ psi) < TOL
        out = psi.copy(); oracle.pauli(tg, pl, out)
        assert rel(out, P @ psi) < TOL
        assert abs(oracle.expectation(tg, pl, psi) - np.vdot(psi, P @ psi).real) < TOL
        bra = random_state(rng, n)
        assert abs(oracle.transition(tg, pl, bra, psi) - np.vdot(bra, P @ psi)) < TOL


def test_reductions_closed_forms(oracle):
    rng = np.random.default_rng(10)
    n = 6
    psi = random_state(rng, n) * 1.3
    p = np.abs(psi) ** 2
    assert abs(oracle.norm_squared(psi) - p.sum()) < TOL
    idx = np.arange(1 << n)
    for t in range(n):
        assert abs(oracle.M0(t, psi) - p[((idx >> t) & 1) == 0].sum()) < TOL
    sel = (((idx >> 1) & 1) == 1) & (((idx >> 4) & 1) == 0)
    assert abs(oracle.marginal([1, 4], [1, 0], psi) - p[sel].sum()) < TOL
    pc = np.maximum(p, 1e-15)
    assert abs(oracle.entropy(psi) - (-(pc * np.log(pc)).sum())) < 1e-11


def exact_fma_prob(re: float, im: float) -> float:
    """round(re*re + round(im*im)) -- the value of fma(re, re, im*im), exactly."""
    return float(Fraction(re) * Fraction(re) + Fraction(im * im))


def brute_samples(psi, uniforms):
    cdf = [0.0]
    s = 0.0
    for z in psi:
        s = s + exact_fma_prob(z.real, z.imag)  # Python floats add with IEEE rounding
        cdf.append(s)
    return [bisect.bisect_left(cdf, u) - 1 for u in uniforms], cdf


def test_sampling_brute_force_and_known_answers(oracle):
    rng = np.random.default_rng(11)
    psi = random_state(rng, 8)
    us = rng.random(500).tolist()
    want, cdf = brute_samples(psi, us)
    # uniforms placed exactly on CDF values exercise lower_bound's tie rule
    us += [cdf[5], cdf[100], cdf[-1], 0.0, 1.5]
    want = brute_samples(psi, us)[0]
    got = oracle.sampling(psi, us)
    assert [int(x) for x in got[:-2]] == want[:-2]
    assert int(got[-2]) == 2**64 - 1  # u = 0: lower_bound hits cdf[0], index -1 as ITYPE
    assert int(got[-1]) == psi.size   # u above the total: past the end
    assert np.array_equal(oracle.sampling_cdf(psi), np.array(cdf))
    # all samples of basis state 100 are 100 (test/cppsim/test_state.cpp:38-47)
    basis = np.zeros(256, dtype=np.complex128); basis[100] = 1
    assert set(oracle.sampling(basis, rng.random(200) * 0.999 + 1e-9).tolist()) == {100}
    # amplitudes proportional to sqrt(2^i): histogram is monotone (test_state.cpp:49-81)
    amp = np.sqrt(2.0 ** np.arange(10)); amp = (amp / np.linalg.norm(amp)).astype(np.complex128)
    hist = np.bincount(oracle.sampling(amp, rng.random(20000)).astype(np.int64), minlength=10)
    assert all(hist[i] <= hist[i + 1] for i in range(5, 9))


def test_dm_gates_equal_u_rho_udagger(oracle):
    rng = np.random.default_rng(12)
    n = 3
    dim = 1 << n
    psi1, psi2 = random_state(rng, n), random_state(rng, n)
    R = 0.3 * np.outer(psi1, psi1.conj()) + 0.7 * np.outer(psi2, psi2.conj())
    for tg in ([0], [2], [1, 2], [2, 0, 1]):
        u = haar_unitary(rng, len(tg))
        U = expand(n, tg, u)
        rho = R.ravel().copy()
        if len(tg) == 1:
            oracle.dm_dense1(tg[0], u, rho, dim)
        else:
            oracle.dm_dense(tg, u, rho, dim)
        assert rel(rho.reshape(dim, dim), U @ R @ U.conj().T) < TOL
    u = haar_unitary(rng, 1)
    U = controlled(n, [0], [1], [2], u)
    rho = R.ravel().copy()
    oracle.dm_controlled_dense([0], [1], [2], u, rho, dim)
    assert rel(rho.reshape(dim, dim), U @ R @ U.conj().T) < TOL
    for tg, pl in (([0, 2], [1, 2]), ([1], [3])):
        P = pauli_matrix(n, tg, pl)
        assert abs(oracle.dm_expectation(tg, pl, R.ravel(), dim) - np.trace(P @ R).real) < TOL


def test_depolarizing_closed_form(oracle):
    """DepolarizingNoise(q, 3p/4) maps rho to (1-p) rho + p (I/2 (x) tr_q rho)
    (test/cppsim/test_noise_dm.cpp:274-309)."""
    rng = np.random.default_rng(13)
    n = 3
    dim = 1 << n
    psi = random_state(rng, n)
    R = np.outer(psi, psi.conj())
    p = 0.37
    for q in range(n):
        rho = R.ravel().copy()
        oracle.dm_depolarizing(q, 3 * p / 4, rho, dim)
        # partial trace over q, then re-embed with I/2 on q
        T = R.reshape([2] * (2 * n))
        ax_row, ax_col = n - 1 - q, 2 * n - 1 - q
        red = np.trace(T, axis1=ax_row, axis2=ax_col)
        full = np.zeros_like(R)
        for a in range(dim):
            for b in range(dim):
                if ((a >> q) & 1) == ((b >> q) & 1):
                    ar = [(a >> k) & 1 for k in reversed(range(n)) if k != q]
                    br = [(b >> k) & 1 for k in reversed(range(n)) if k != q]
                    full[a, b] = red[tuple(ar + br)] / 2
        assert rel(rho.reshape(dim, dim), (1 - p) * R + p * full) < TOL


# ---------------------------------------------------------------- DM sampling,
# the random stream and the per-amplitude probability (src/cppsim/state_dm.hpp:383-408,
# src/cppsim/utility.hpp:54-102, src/cppsim/state.hpp:915-920)

def _libm_hypot():
    import ctypes

    libm = ctypes.CDLL("libm.so.6")
    libm.hypot.restype = ctypes.c_double
    libm.hypot.argtypes = [ctypes.c_double, ctypes.c_double]
    return libm.hypot


def _hypot_inputs(rng, count):
    """Operand pairs across the whole double range, with equal, widely varying,
    zero and tiny/huge operands."""
    e = rng.integers(-1070, 1020, size=(count, 2))
    v = rng.random((count, 2)) * np.exp2(e.astype(np.float64))
    v[: count // 4, 1] = v[: count // 4, 0] * rng.random(count // 4) * np.exp2(-rng.integers(0, 80, count // 4))
    v[count // 4: count // 4 + 50, 1] = 0.0
    v[count // 4 + 50: count // 4 + 100, 1] = v[count // 4 + 50: count // 4 + 100, 0]
    return [(float(a), float(b)) for a, b in v]


def test_host_hypot_model_equals_libm_and_is_not_correctly_rounded():
    """The model in gen.host_hypot reproduces libm's hypot (what std::abs of a
    complex double calls) on every input tried, and libm's hypot is not the
    correctly rounded value: a device restatement must follow the model, not
    compute |z| exactly."""
    from tests.golden.gen import host_hypot

    hyp = _libm_hypot()
    rng = np.random.default_rng(31)
    pairs = _hypot_inputs(rng, 40000)
    assert all(host_hypot(a, b) == hyp(a, b) for a, b in pairs)
    # correctly rounded sqrt(a^2 + b^2) by exact rational comparison
    import math

    def cr(a, b):
        x = Fraction(a) ** 2 + Fraction(b) ** 2
        d = math.sqrt(float(x))
        cands = (math.nextafter(d, 0), d, math.nextafter(d, math.inf))
        return min(cands, key=lambda c: abs(Fraction(c) ** 2 - x))

    small = [(rng.random(), rng.random() * 2.0 ** -int(rng.integers(0, 30))) for _ in range(20000)]
    diffs = sum(hyp(a, b) != cr(a, b) for a, b in small)
    assert diffs > 0  # glibc 2.35's hypot is faithful, not correctly rounded


def _brute_dm_samples(diag, uniforms):
    from tests.golden.gen import host_hypot

    cdf = [0.0]
    s = 0.0
    for z in diag:
        s = s + (abs(z.real) if z.imag == 0.0 else host_hypot(z.real, z.imag))
        cdf.append(s)
    return [bisect.bisect_left(cdf, u) - 1 for u in uniforms], cdf


def test_dm_sampling_brute_force_complex_diagonal(oracle):
    """oracle_dm_sampling against a brute-force CDF of |rho_ii| (libm-rounded
    modulus, IEEE sequential sums) and bisect_left, on a diagonal with nonzero
    imaginary parts (the hypot path) and uniforms on and beside every step."""
    rng = np.random.default_rng(32)
    n = 6
    dim = 1 << n
    rho = (rng.standard_normal((dim, dim)) + 1j * rng.standard_normal((dim, dim))) * 0.02
    diag = rng.random(dim) * np.exp2(-rng.integers(0, 12, dim)) + 1j * rng.standard_normal(dim) * 1e-3
    diag[5] = 0.0
    diag[9] = -0.0 + 1e-3j  # purely imaginary
    diag[11] = -0.25        # negative real part: |rho_ii|
    rho[np.arange(dim), np.arange(dim)] = diag
    rho = rho.ravel().copy()
    us = rng.random(300) * 1.5
    _, cdf = _brute_dm_samples(diag, [])
    us = np.concatenate([us, cdf, [np.nextafter(c, 2.0) for c in cdf], [np.nextafter(c, -1.0) for c in cdf]])
    want, cdf = _brute_dm_samples(diag, us.tolist())
    got = oracle.dm_sampling(rho, dim, us)
    assert [int(x) if x < 2**63 else int(x) - 2**64 for x in got] == want
    assert np.array_equal(oracle.dm_sampling_cdf(rho, dim), np.array(cdf))


def test_dm_sampling_known_answers(oracle):
    """A diagonal density matrix with exactly representable entries has the exact
    CDF (0, 1/8, 1/2, 1/2, 1): a uniform on a step selects the index below it
    (lower_bound), one just above selects the next non-empty index, the empty
    basis state 2 is never selected, and a basis-state density matrix always
    gives its index (test/cppsim/test_state.cpp:38-47 for the state vector)."""
    dim = 4
    rho = np.zeros((dim, dim), dtype=np.complex128)
    rho[np.arange(dim), np.arange(dim)] = [0.125, 0.375, 0.0, 0.5]
    rho[0, 3] = rho[3, 0] = 0.2  # off-diagonal entries do not enter
    r = rho.ravel().copy()
    assert np.array_equal(oracle.dm_sampling_cdf(r, dim), [0.0, 0.125, 0.5, 0.5, 1.0])
    us = [0.125, np.nextafter(0.125, 1), 0.5, np.nextafter(0.5, 1), 1.0, 1e-300, 0.0, np.nextafter(1.0, 2)]
    got = [int(x) for x in oracle.dm_sampling(r, dim, us)]
    assert got == [0, 1, 1, 3, 3, 0, 2**64 - 1, 4]
    basis = np.zeros((8, 8), dtype=np.complex128)
    basis[6, 6] = 1.0
    assert set(oracle.dm_sampling(basis.ravel().copy(), 8, np.linspace(1e-9, 1.0, 100)).tolist()) == {6}


def test_mt19937_64_check_value_and_canonical_uniforms(oracle):
    """The reference's Random (src/cppsim/utility.hpp:54-102) is std::mt19937_64
    with uniform_real_distribution<double>(0, 1).  The C++ standard fixes the
    engine by its check value ([rand.predef]: the 10000th output of a
    default-constructed mt19937_64, seed 5489, is 9981545732273789042), and
    generate_canonical with a 64-bit engine is one draw x -> RN(x) / 2^64
    (clamped below 1)."""
    raw = oracle.mt64_raw(5489, 10000)
    assert int(raw[-1]) == 9981545732273789042
    for seed in (0, 1, 12345, 2**63 + 7):
        raw = oracle.mt64_raw(seed, 2000)
        us = oracle.mt64_uniforms(seed, 2000)
        want = [min(float(int(x)) / 2.0 ** 64, np.nextafter(1.0, 0.0)) for x in raw]
        assert us.tolist() == want


def test_norm_probe_equals_oracle_probability():
    """The reference's stacked probabilities (sum += std::norm(psi_i),
    src/cppsim/state.hpp:915-920) compiled with its release flags (gnu++14 -O3
    -march=native -mfpmath=both, CMakeLists.txt:36,281,349) equal the oracle's
    fma(re, re, im*im) sums bit for bit.  A plain re*re + im*im or the other
    contraction order fma(im, im, re*re) would differ."""
    from oracle import oracle as O

    if not O.NORM_PROBE_SO.exists():
        pytest.skip("norm probe not built here (make -C oracle norm-probe)")
    o = O.Oracle()
    rng = np.random.default_rng(33)
    diff_plain = diff_swapped = 0
    for n in (6, 10, 14):
        psi = random_state(rng, n) * np.exp2(-rng.integers(0, 8, 1 << n))
        probe = O.norm_probe_stacked_prob(psi)
        assert np.array_equal(probe, o.sampling_cdf(psi))
        re, im = psi.real, psi.imag
        plain = np.concatenate([[0.0], np.cumsum(re * re + im * im)])
        diff_plain += int(not np.array_equal(plain, probe))
        swapped = [0.0]
        s = 0.0
        for a, b in zip(re.tolist(), im.tolist()):
            s = s + float(Fraction(b) * Fraction(b) + Fraction(a * a))
            swapped.append(s)
        diff_swapped += int(swapped != probe.tolist())
    assert diff_plain > 0 and diff_swapped > 0


def test_parameter_shift_closed_forms_and_finite_differences(oracle):
    """oracle.parameter_shift_grad (GradCalculator.cpp:8-44) on circuits with a
    closed-form energy: <Z> after RX(a) on |0> is cos(a) (qulacs RX = exp(+i a X/2)),
    so d/da = -sin(a); <Z0 Z1> after RY(a) (x) RY(b) is cos(a) cos(b).  Then a
    random 4-qubit circuit against central differences (step 1e-6)."""
    from oracle.oracle import parameter_shift_grad

    def e_rx(th):
        psi = np.array([1, 0], dtype=np.complex128)
        oracle.pauli_rotation([0], [1], th[0], psi)
        return oracle.expectation([0], [3], psi)

    for a in (0.3, 1.7, -2.9):
        g = parameter_shift_grad(e_rx, [a])
        assert abs(e_rx([a]) - np.cos(a)) < TOL and abs(g[0] + np.sin(a)) < TOL

    def e_ry2(th):
        psi = np.zeros(4, dtype=np.complex128)
        psi[0] = 1
        oracle.pauli_rotation([0], [2], th[0], psi)
        oracle.pauli_rotation([1], [2], th[1], psi)
        return oracle.expectation([0, 1], [3, 3], psi)

    a, b = 0.4, -1.1
    g = parameter_shift_grad(e_ry2, [a, b])
    assert np.allclose(g, [-np.sin(a) * np.cos(b), -np.cos(a) * np.sin(b)], atol=TOL, rtol=0)

    rng = np.random.default_rng(34)
    n = 4
    rots = [([int(rng.integers(n))], [int(rng.integers(1, 4))]) for _ in range(10)] + [([0, 2], [1, 3]), ([1, 3], [2, 2])]
    ents = [haar_unitary(rng, 2) for _ in range(3)]
    terms = random_pauli_terms(rng, n, 5)

    def energy(th):
        psi = np.zeros(1 << n, dtype=np.complex128)
        psi[0] = 1
        for i, (tq, pl) in enumerate(rots):
            oracle.pauli_rotation(tq, pl, th[i], psi)
            if i % 4 == 3:
                oracle.dense([i % n, (i + 1) % n], ents[i // 4], psi)
        return sum(c * oracle.expectation(t, p, psi) for t, p, c in terms)

    th = rng.uniform(0, 2 * np.pi, len(rots))
    g = parameter_shift_grad(energy, th)
    h = 1e-6
    for k in range(th.size):
        tp, tm = th.copy(), th.copy()
        tp[k] += h
        tm[k] -= h
        assert abs(g[k] - (energy(tp) - energy(tm)) / (2 * h)) < 1e-8
