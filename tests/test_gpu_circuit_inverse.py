"""QuantumCircuit.get_inverse (test/cppsim/test_circuit.cpp:215-320, CircuitRev):
a circuit of every named gate, the Rot/RotInv rotations, CNOT/CZ/SWAP and a diagonal
observable rotation, followed by its inverse, returns the state it started from.
Sizes on both sides of the fused circuit loop's threshold."""
from __future__ import annotations

import numpy as np
import pytest

import qulacs_amd as q

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [3, 6, 12])
def test_circuit_then_inverse_is_identity(n):
    rng = np.random.default_rng(n)
    c = q.QuantumCircuit(n)
    for _ in range(3):
        t, s = (int(x) for x in rng.permutation(n)[:2])
        a = float(rng.uniform(0, 2 * np.pi))
        for name in ("X", "Y", "Z", "H", "S", "Sdag", "T", "Tdag", "sqrtX", "sqrtXdag", "sqrtY", "sqrtYdag"):
            getattr(c, f"add_{name}_gate")(t)
        for name in ("RX", "RotInvX", "RotX", "RY", "RotInvY", "RotY", "RZ", "RotInvZ", "RotZ"):
            getattr(c, f"add_{name}_gate")(t, a)
        c.add_CNOT_gate(t, s)
        c.add_CZ_gate(t, s)
        c.add_SWAP_gate(t, s)
        c.add_U3_gate(s, a, 0.3, -0.7)
        c.add_random_unitary_gate([t, s], int(rng.integers(1 << 30)))
    obs = q.Observable(n)
    obs.add_operator(1.0, "Z 0")
    obs.add_operator(2.0, "Z 0 Z 1")
    c.add_diagonal_observable_rotation_gate(obs, 0.37)
    st = q.QuantumState(n)
    st.set_Haar_random_state(3)
    v0 = st.get_vector()
    c.update_quantum_state(st)
    assert np.max(np.abs(st.get_vector() - v0)) > 1e-3
    c.get_inverse().update_quantum_state(st)
    np.testing.assert_allclose(st.get_vector(), v0, atol=1e-12)
