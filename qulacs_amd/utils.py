"""Conversions from other packages' operator types (pysrc/qulacs/utils/).

`convert_openfermion_op` reads any object with OpenFermion's QubitOperator
layout -- a `terms` mapping from tuples of (qubit, 'X'|'Y'|'Z') pairs to
coefficients -- so OpenFermion itself need not be installed to use it."""
from __future__ import annotations

from ._core import GeneralQuantumOperator

__all__ = ["convert_openfermion_op"]


def _qubits_needed(terms) -> int:
    """One more than the highest qubit any term touches; 1 for an operator with no
    qubits (pysrc/qulacs/utils/conversions_openfermion.py:30-42)."""
    top = 0
    for product in terms:
        for qubit, _ in product:
            top = max(top, int(qubit))
    return top + 1


def convert_openfermion_op(openfermion_op, n_qubits=None) -> GeneralQuantumOperator:
    """A GeneralQuantumOperator with one term per entry of `openfermion_op.terms`
    (pysrc/qulacs/utils/conversions_openfermion.py:5-27).  As there, each
    coefficient keeps only its real part, and `n_qubits` defaults to the smallest
    count that holds every term."""
    terms = openfermion_op.terms
    out = GeneralQuantumOperator(_qubits_needed(terms) if n_qubits is None else int(n_qubits))
    for product, coef in terms.items():
        text = " ".join(f"{axis} {int(qubit)}" for qubit, axis in product)
        out.add_operator(float(complex(coef).real), text)
    return out
