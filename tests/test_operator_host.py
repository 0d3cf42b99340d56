"""Operator algebra and OpenFermion text (no device needed): Pauli products against
the 2x2 matrix products, operator sums / differences / products against the sums and
products of their matrices, the reference's term-merging rule
(src/cppsim/general_quantum_operator.cpp:547-805), and the H2 Hamiltonian file the
reference's tests load (test/cppsim/H2.txt) against the matrix entries they assert
(test/cppsim/test_hamiltonian.cpp:181-240)."""
from __future__ import annotations

import pickle
from pathlib import Path

import numpy as np
import pytest

q = pytest.importorskip("qulacs_amd")
from qulacs_amd import GeneralQuantumOperator, Observable, PauliOperator, observable, quantum_operator  # noqa: E402

GOLDEN = Path(__file__).parent / "golden"
PAULI = [np.eye(2), np.array([[0, 1], [1, 0]]), np.array([[0, -1j], [1j, 0]]), np.diag([1.0, -1.0])]


def pauli_matrix(p: PauliOperator, n: int) -> np.ndarray:
    """coef * prod_k P_k with qubit q = bit q of the index (kron puts qubit 0 last)."""
    ops = [np.eye(2, dtype=complex) for _ in range(n)]
    for t, pid in zip(p.get_index_list(), p.get_pauli_id_list()):
        ops[t] = ops[t] @ PAULI[pid]
    m = np.eye(1, dtype=complex)
    for k in reversed(range(n)):
        m = np.kron(m, ops[k])
    return p.get_coef() * m


def op_matrix(o, n=None) -> np.ndarray:
    return o.get_matrix().toarray()


def random_op(rng, n, terms, cls=GeneralQuantumOperator, real=False):
    """Distinct Pauli strings: with a string repeated inside one operand the reference's
    sum adds the other operand's coefficient to every copy, which is not the matrix sum."""
    o = cls(n)
    seen = set()
    while len(seen) < terms:
        pl = tuple(rng.integers(0, 4, size=n).tolist())
        if pl in seen:
            continue
        seen.add(pl)
        c = rng.normal() + (0 if real else 1j * rng.normal())
        o.add_operator(PauliOperator(list(range(n)), list(pl), c))
    return o


def test_single_qubit_pauli_products_follow_the_algebra():
    for a in range(4):
        for b in range(4):
            pa = PauliOperator([0], [a], 1.0)
            pb = PauliOperator([0], [b], 1.0)
            got = pa * pb
            np.testing.assert_allclose(pauli_matrix(got, 1), PAULI[a] @ PAULI[b], atol=0)


def test_multi_qubit_pauli_product_with_overlap_and_coefs():
    rng = np.random.default_rng(3)
    n = 4
    for _ in range(20):
        ta = rng.choice(n, size=rng.integers(1, n + 1), replace=False).tolist()
        tb = rng.choice(n, size=rng.integers(1, n + 1), replace=False).tolist()
        pa = PauliOperator(ta, rng.integers(1, 4, size=len(ta)).tolist(), complex(rng.normal(), rng.normal()))
        pb = PauliOperator(tb, rng.integers(1, 4, size=len(tb)).tolist(), complex(rng.normal(), rng.normal()))
        prod = pa * pb
        np.testing.assert_allclose(pauli_matrix(prod, n), pauli_matrix(pa, n) @ pauli_matrix(pb, n), atol=1e-12)
        # result lists its Paulis in ascending qubit order, identities dropped
        assert prod.get_index_list() == sorted(prod.get_index_list())
        assert 0 not in prod.get_pauli_id_list()
        c = complex(0.5, -2.0)
        np.testing.assert_allclose(pauli_matrix(pa * c, n), c * pauli_matrix(pa, n), atol=1e-12)
        pa2 = pa.copy()
        pa2 *= pb
        np.testing.assert_allclose(pauli_matrix(pa2, n), pauli_matrix(prod, n), atol=1e-12)


def test_operator_sum_difference_product_match_matrices():
    rng = np.random.default_rng(5)
    n = 3
    for _ in range(10):
        a = random_op(rng, n, 4)
        b = random_op(rng, n, 3)
        p = PauliOperator([0, 2], [1, 3], complex(0.3, 0.7))
        A, B, P = op_matrix(a), op_matrix(b), pauli_matrix(p, n)
        np.testing.assert_allclose(op_matrix(a + b), A + B, atol=1e-12)
        np.testing.assert_allclose(op_matrix(a - b), A - B, atol=1e-12)
        np.testing.assert_allclose(op_matrix(a * b), A @ B, atol=1e-12)
        np.testing.assert_allclose(op_matrix(a + p), A + P, atol=1e-12)
        np.testing.assert_allclose(op_matrix(a - p), A - P, atol=1e-12)
        np.testing.assert_allclose(op_matrix(a * p), A @ P, atol=1e-12)
        np.testing.assert_allclose(op_matrix(a * 2.5j), 2.5j * A, atol=1e-12)
        c = a.copy()
        c += b
        c -= p
        c *= b
        c *= p
        c *= complex(0, -1)
        np.testing.assert_allclose(op_matrix(c), ((A + B - P) @ B @ P) * -1j, atol=1e-11)
        # the operands are left unchanged
        np.testing.assert_allclose(op_matrix(a), A, atol=0)


def test_sums_merge_equal_pauli_strings():
    a = GeneralQuantumOperator(2)
    a.add_operator(1.0, "X 0 Z 1")
    a.add_operator(2.0, "Y 1")
    b = GeneralQuantumOperator(2)
    b.add_operator(0.5, "Z 1 X 0")  # same string, other order
    b.add_operator(3.0, "Z 0")
    s = a + b
    assert s.get_term_count() == 3
    assert s.get_term(0).get_coef() == 1.5
    assert s.get_term(1).get_coef() == 2.0
    assert s.get_term(2).get_coef() == 3.0 and s.get_term(2).get_pauli_string() == "Z 0"
    d = a - b
    assert d.get_term_count() == 3 and d.get_term(0).get_coef() == 0.5 and d.get_term(2).get_coef() == -3.0
    # X0 X0 = I: the product of a term with itself merges into one identity term
    x = GeneralQuantumOperator(1)
    x.add_operator(1.0, "X 0")
    x.add_operator(1.0, "Z 0")
    sq = x * x
    np.testing.assert_allclose(op_matrix(sq), 2 * np.eye(2), atol=1e-15)
    assert sq.get_term_count() == 2  # I (merged from X·X and Z·Z) and the Y term (iY - iY = 0)


def test_sum_with_repeated_string_follows_reference_rule():
    # every copy of a repeated string in the left operand receives the right operand's
    # coefficient (src/cppsim/general_quantum_operator.cpp:557-581)
    a = GeneralQuantumOperator(1)
    a.add_operator(1.0, "X 0")
    a.add_operator(2.0, "X 0")
    b = GeneralQuantumOperator(1)
    b.add_operator(10.0, "X 0")
    s = a + b
    assert [s.get_term(i).get_coef() for i in range(s.get_term_count())] == [11.0, 12.0]


def test_observable_rejects_complex_products_in_place():
    o = Observable(1)
    o.add_operator(1.0, "X 0")
    with pytest.raises(RuntimeError):
        o.add_operator(1j, "Z 0")
    y = PauliOperator("Y 0", 1.0)
    prod = o * y  # a new GeneralQuantumOperator: complex coefficients allowed
    np.testing.assert_allclose(op_matrix(prod), PAULI[1] @ PAULI[2], atol=0)
    with pytest.raises(RuntimeError):
        o *= y  # X·Y = iZ cannot stay in an Observable


def test_dagger_str_and_pickle():
    o = GeneralQuantumOperator(2)
    o.add_operator(complex(1, 2), "X 0 Y 1")
    o.add_operator(-0.5, "Z 1")
    np.testing.assert_allclose(op_matrix(o.get_dagger()), op_matrix(o).conj().T, atol=0)
    assert str(o) == "(1,2) X 0 Y 1 + (-0.5,0) Z 1"
    ob = Observable(2)
    ob.add_operator(0.25, "Z 0")
    ob.add_operator(-1.0, "X 0 X 1")
    assert str(ob) == "0.25 Z 0 + -1 X 0 X 1"
    for x in (o, ob, PauliOperator("X 0 Z 3", 0.5j)):
        y = pickle.loads(pickle.dumps(x))
        assert type(y) is type(x)
        assert y.to_json() == x.to_json() if hasattr(x, "to_json") else True
    p = pickle.loads(pickle.dumps(PauliOperator("X 0 Z 3", 0.5j)))
    assert p.get_pauli_string() == "X 0 Z 3" and p.get_coef() == 0.5j


def test_identity_entries_kept_in_index_list_not_in_string():
    p = PauliOperator([0, 1, 2], [1, 0, 3], 1.0)
    assert p.get_index_list() == [0, 1, 2]
    assert p.get_pauli_id_list() == [1, 0, 3]
    assert p.get_pauli_string() == "X 0 Z 2"


def test_add_random_operator_seeded():
    a = Observable(4)
    a.add_random_operator(6, 123)
    b = Observable(4)
    b.add_random_operator(6, 123)
    assert a.get_term_count() == 6
    for i in range(6):
        ta, tb = a.get_term(i), b.get_term(i)
        assert ta.get_index_list() == [0, 1, 2, 3]
        assert ta.get_pauli_id_list() == tb.get_pauli_id_list()
        assert ta.get_coef() == tb.get_coef()
        assert -1.0 <= ta.get_coef().real < 1.0 and ta.get_coef().imag == 0
    ids = np.array([a.get_term(i).get_pauli_id_list() for i in range(6)])
    assert len(set(ids.ravel().tolist())) > 1


# test/cppsim/test_hamiltonian.cpp:207-226: non-zero entries of the H2 Hamiltonian
H2_EXPECTED = {
    (0, 0): -4.163336342344337e-16, (1, 1): -1.8304610000000006, (1, 4): 0.181287,
    (2, 2): -1.2462260000000005, (3, 3): -1.0649390000000005, (3, 6): -0.181287,
    (4, 1): 0.181287, (4, 4): -0.25447100000000045, (5, 5): 0.20638199999999932,
    (6, 3): -0.181287, (6, 6): -1.0649390000000005, (7, 7): -1.2462260000000005,
    (8, 8): -0.4759340000000004, (9, 9): -1.1607380000000005, (10, 10): -1.2524770000000005,
    (11, 11): -1.2524770000000005, (12, 12): -0.4759340000000005, (13, 13): -1.1607380000000007,
    (14, 14): -0.3612910000000006, (15, 15): -0.3612910000000006,
}


def _h2_expected():
    m = np.zeros((16, 16), dtype=complex)
    for (i, j), v in H2_EXPECTED.items():
        m[i, j] = v
    return m


def test_openfermion_text_h2_matrix_matches_reference_entries():
    text = (GOLDEN / "openfermion_H2.txt").read_text()
    for make in (observable.create_observable_from_openfermion_text,
                 quantum_operator.create_quantum_operator_from_openfermion_text):
        o = make(text)
        assert o.get_qubit_count() == 4
        assert o.get_term_count() == 15
        np.testing.assert_allclose(op_matrix(o), _h2_expected(), atol=1e-6)
    o = observable.create_observable_from_openfermion_text(text)
    assert o.get_term(0).get_pauli_string() == ""  # "[]" is the identity
    assert o.get_term(1).get_pauli_string() == "X 0 Z 1 X 2"


def test_openfermion_file_and_split():
    path = str(GOLDEN / "openfermion_H2.txt")
    o = observable.create_observable_from_openfermion_file(path)
    g = quantum_operator.create_quantum_operator_from_openfermion_file(path)
    np.testing.assert_allclose(op_matrix(o), _h2_expected(), atol=1e-6)
    np.testing.assert_allclose(op_matrix(g), _h2_expected(), atol=1e-6)
    for diag, nondiag in (observable.create_split_observable(path), quantum_operator.create_split_quantum_operator(path)):
        assert diag.get_term_count() == 11 and nondiag.get_term_count() == 4
        d = op_matrix(diag)
        assert np.count_nonzero(d - np.diag(np.diag(d))) == 0
        np.testing.assert_allclose(d + op_matrix(nondiag), _h2_expected(), atol=1e-6)
    with pytest.raises(RuntimeError):
        observable.create_observable_from_openfermion_file(str(GOLDEN / "no_such_file.txt"))


def test_openfermion_coefficient_forms():
    text = "\n".join([
        "(0.5+0j) [X0 Y1]",
        "(0.25-1.5j) [Z2]",
        "0.75 [Y0]",
        "2j [X1]",
        "",
        "(1e-3+0j) []",
    ])
    o = quantum_operator.create_quantum_operator_from_openfermion_text(text)
    assert o.get_qubit_count() == 3 and o.get_term_count() == 5
    coefs = [o.get_term(i).get_coef() for i in range(5)]
    assert coefs == [0.5, complex(0.25, -1.5), 0.75, 2j, 1e-3]
    strings = [o.get_term(i).get_pauli_string() for i in range(5)]
    assert strings == ["X 0 Y 1", "Z 2", "Y 0", "X 1", ""]
    with pytest.raises(RuntimeError):
        observable.create_observable_from_openfermion_text("(0.25-1.5j) [Z2]")  # not Hermitian
    with pytest.raises(RuntimeError):
        quantum_operator.create_quantum_operator_from_openfermion_text("abc [X0]")


class _QubitOperatorLike:
    """OpenFermion's QubitOperator layout: `terms` maps tuples of (qubit, axis)
    pairs to coefficients."""

    def __init__(self, terms):
        self.terms = terms


def test_convert_openfermion_op():
    # python/tests/single_cpu/test_utils.py:15-27: three terms, any order
    from qulacs_amd.utils import convert_openfermion_op

    op = convert_openfermion_op(_QubitOperatorLike({
        ((0, "X"),): 1.0,
        ((0, "Z"), (1, "Y")): 2.0,
        ((0, "Z"), (3, "Y")): 1.0,
    }))
    text = str(op)
    for part in ("(1,0) X 0", "(2,0) Z 0 Y 1", "(1,0) Z 0 Y 3"):
        text = text.replace(part, "")
    assert text == " +  + "
    assert op.get_qubit_count() == 4
    # only the real part of a coefficient is kept; an explicit size wins; the empty
    # product is the identity
    op = convert_openfermion_op(_QubitOperatorLike({((1, "X"),): 0.5 + 2j, (): -0.25}), n_qubits=3)
    assert op.get_qubit_count() == 3 and op.get_term_count() == 2
    assert op.get_term(0).get_coef() == 0.5
    m = op.get_matrix().toarray()
    x1 = np.kron(np.eye(2), np.kron(np.array([[0, 1], [1, 0]]), np.eye(2)))
    np.testing.assert_allclose(m, 0.5 * x1 - 0.25 * np.eye(8), atol=0)


def test_pauli_string_parsing_edge_cases():
    """test/cppsim/test_pauli_operator.cpp:5-75: trailing whitespace is ignored,
    the empty string is the identity, and a string whose axes and indices do not
    alternate is an error."""
    p = PauliOperator("X 0 ", 1.0)
    assert p.get_index_list() == [0] and p.get_pauli_id_list() == [1]
    assert p.get_pauli_string() == PauliOperator("X 0", 1.0).get_pauli_string()
    e = PauliOperator("", 1.0)
    assert e.get_index_list() == [] and e.get_pauli_id_list() == [] and e.get_pauli_string() == ""
    assert len(PauliOperator("X 0 Y 1 ", 2.0).get_index_list()) == 2
    for broken in ("X 0 X Z 1 Y 2", "X {i}", "X 4x", "4 X", "X 0 Y "):
        with pytest.raises(Exception):
            PauliOperator(broken, 2.0)


@pytest.mark.parametrize("a,b,string,coef", [
    ("X 0", "X 0", "", 4), ("X 0", "Y 0", "Z 0", 4j), ("X 0", "Z 0", "Y 0", -4j),
    ("Y 0", "X 0", "Z 0", -4j), ("Y 0", "Y 0", "", 4), ("Y 0", "Z 0", "X 0", 4j),
    ("Z 0", "X 0", "Y 0", 4j), ("Z 0", "Y 0", "X 0", -4j), ("Z 0", "Z 0", "", 4),
    ("X 0", "Y 1", "X 0 Y 1", 4), ("X 0 Y 1", "Y 0 X 1", "Z 0 Z 1", 4),
])
def test_pauli_products_of_the_reference_table(a, b, string, coef):
    """test/cppsim/test_pauli_operator.cpp:80-140, both operator* and operator*=."""
    r = PauliOperator(a, 2.0) * PauliOperator(b, 2.0)
    assert r.get_pauli_string() == string and r.get_coef() == coef
    r = PauliOperator(a, 2.0)
    r *= PauliOperator(b, 2.0)
    assert r.get_pauli_string() == string and r.get_coef() == coef
