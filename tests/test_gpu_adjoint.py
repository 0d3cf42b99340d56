"""Adjoint gradients (qm_adjoint_*) against central finite differences of the
oracle's energies (step 1e-6, |d| <= 1e-8 * scale), the host-side
ParametricQuantumCircuit.backprop and GradCalculator (|d| <= 1e-10 * scale),
and through torch.autograd."""
from __future__ import annotations

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import qulacs_amd as q  # noqa: E402
from qulacs_amd import gate  # noqa: E402
from qulacs_amd.adjoint import AdjointCircuit, expectation, hardware_efficient_ansatz, heisenberg_terms  # noqa: E402
from tests.batch_util import masks_to_lists  # noqa: E402
from oracle.oracle import parameter_shift_grad  # noqa: E402
from tests.golden.gen import haar_unitary  # noqa: E402

pytestmark = pytest.mark.gpu


def _circuit(n, layers, rng):
    gates, P = hardware_efficient_ansatz(n, layers)
    gates.append(("rot", [0, 3], [1, 2], P, 0.5))  # X0 Y3 rotation with coefficient 0.5
    gates.append(("dense", [1, 4], haar_unitary(rng, 2)))
    gates.append(("CNOT", [2, 5], None))
    gates.append(("diag", [3], np.exp(1j * rng.uniform(0, 6, 2))))
    gates.append(("rot", [5], [1], P + 1, 1.0))
    return gates, P + 2


def _oracle_energy(oracle, n, gates, th, terms):
    psi = np.zeros(1 << n, dtype=np.complex128)
    psi[0] = 1
    for g in gates:
        if g[0] == "rot":
            _, qs, ps, p, c = g
            oracle.pauli_rotation(qs, ps, c * th[p], psi)
        elif g[0] == "dense":
            oracle.dense(g[1], g[2], psi)
        elif g[0] == "diag":
            oracle.controlled_diag([], [], g[1], g[2], psi)
        else:
            oracle.two(g[0], g[1][0], g[1][1], psi)
    e = 0.0
    for bf, pf, rot, c in terms:
        t, p = masks_to_lists(bf, pf, n)
        e += c * oracle.expectation(t, p, psi)
    return e


def _oracle_energy_occ(oracle, n, gates, angles, terms):
    """The oracle energy with one angle per rotation occurrence (coefficient included)."""
    it = iter(angles)
    occ = [(g[0], g[1], g[2], None, next(it)) if g[0] == "rot" else g for g in gates]
    psi = np.zeros(1 << n, dtype=np.complex128)
    psi[0] = 1
    for g in occ:
        if g[0] == "rot":
            oracle.pauli_rotation(g[1], g[2], g[4], psi)
        elif g[0] == "dense":
            oracle.dense(g[1], g[2], psi)
        elif g[0] == "diag":
            oracle.controlled_diag([], [], g[1], g[2], psi)
        else:
            oracle.two(g[0], g[1][0], g[1][1], psi)
    e = 0.0
    for bf, pf, rot, c in terms:
        t, p = masks_to_lists(bf, pf, n)
        e += c * oracle.expectation(t, p, psi)
    return e


def _host_circuit(n, gates, th):
    c = q.ParametricQuantumCircuit(n)
    for g in gates:
        if g[0] == "rot":
            _, qs, ps, p, cf = g
            c.add_parametric_multi_Pauli_rotation_gate(qs, ps, cf * th[p])
        elif g[0] == "dense":
            c.add_gate(gate.DenseMatrix(g[1], g[2]))
        elif g[0] == "diag":
            c.add_gate(gate.DiagonalMatrix(g[1], g[2]))
        elif g[0] == "CZ":
            c.add_CZ_gate(*g[1])
        else:
            c.add_CNOT_gate(*g[1])
    return c


def test_adjoint_matches_finite_differences_and_backprop(oracle):
    n, layers, B = 6, 3, 3
    rng = np.random.default_rng(8)
    gates, P = _circuit(n, layers, rng)
    terms = heisenberg_terms(n)
    scale = sum(abs(t[3]) for t in terms)
    circ = AdjointCircuit(n, gates, P)
    theta = rng.uniform(0, 2 * np.pi, (B, P))
    e, g = circ.run(torch.tensor(theta, device="cuda:0"), terms)
    e, g = e.cpu().numpy(), g.cpu().numpy()
    h = 1e-6
    for b in range(B):
        assert abs(e[b] - _oracle_energy(oracle, n, gates, theta[b], terms)) < 1e-10 * scale
        for p in range(P):
            tp, tm = theta[b].copy(), theta[b].copy()
            tp[p] += h
            tm[p] -= h
            fd = (_oracle_energy(oracle, n, gates, tp, terms) - _oracle_energy(oracle, n, gates, tm, terms)) / (2 * h)
            assert abs(g[b, p] - fd) < 1e-8 * scale, (b, p, g[b, p], fd)
    # host backprop and parameter shift on the first parameter set: gradients per
    # rotation occurrence, each parameter used once with coefficient coef
    obs = q.Observable(n)
    for bf, pf, rot, c in terms:
        t, pl = masks_to_lists(bf, pf, n)
        obs.add_operator(q.PauliOperator(t, pl, c))
    hc = _host_circuit(n, gates, theta[0])
    coefs = [gg[4] for gg in gates if gg[0] == "rot"]
    params = [gg[3] for gg in gates if gg[0] == "rot"]
    bp = np.array(hc.backprop(obs))
    shift = np.array(q.GradCalculator().calculate_grad(hc, obs)).real
    # the oracle's parameter shift over per-occurrence angles (GradCalculator.cpp:8-44)
    occ_angles = np.array([cf * theta[0][p] for p, cf in zip(params, coefs)])
    ref_shift = parameter_shift_grad(lambda a: _oracle_energy_occ(oracle, n, gates, a, terms), occ_angles)
    want = np.zeros(P)
    for occ, (p, cf) in enumerate(zip(params, coefs)):
        want[p] += cf * bp[occ]
        assert abs(bp[occ] - shift[occ]) < 1e-9 * scale
        assert abs(shift[occ] - ref_shift[occ]) < 1e-10 * scale, (occ, shift[occ], ref_shift[occ])
        assert abs(bp[occ] - ref_shift[occ]) < 1e-10 * scale, (occ, bp[occ], ref_shift[occ])
    assert np.abs(g[0] - want).max() < 1e-10 * scale


def test_shared_parameters_and_autograd(oracle):
    n = 5
    rng = np.random.default_rng(9)
    gates = [("rot", [0], [2], 0, 1.0), ("rot", [1], [2], 0, -0.5), ("CZ", [0, 1], None),
             ("rot", [0, 1, 2], [3, 3, 1], 1, 2.0), ("rot", [4], [1], 1, 1.0), ("CNOT", [4, 3], None)]
    terms = heisenberg_terms(n)
    scale = sum(abs(t[3]) for t in terms)
    circ = AdjointCircuit(n, gates, 2)
    theta_np = rng.uniform(0, 2 * np.pi, (4, 2))
    theta = torch.tensor(theta_np, device="cuda:0", requires_grad=True)
    w = torch.tensor([1.0, -2.0, 0.5, 3.0], dtype=torch.float64, device="cuda:0")
    e = expectation(circ, theta, terms)
    (e * w).sum().backward()
    _, g = circ.run(theta.detach(), terms)
    assert torch.allclose(theta.grad, w[:, None] * g, rtol=0, atol=1e-14)
    h = 1e-6
    for b in range(4):
        for p in range(2):
            tp, tm = theta_np[b].copy(), theta_np[b].copy()
            tp[p] += h
            tm[p] -= h
            fd = (_oracle_energy(oracle, n, gates, tp, terms) - _oracle_energy(oracle, n, gates, tm, terms)) / (2 * h)
            assert abs(g[b, p].item() - fd) < 1e-8 * scale


def test_adjoint_rejects_bad_gates():
    from qulacs_amd._lib import QulacsMI355Error

    with pytest.raises(QulacsMI355Error):
        AdjointCircuit(3, [("rot", [3], [1], 0, 1.0)], 1)
    with pytest.raises(QulacsMI355Error):
        AdjointCircuit(3, [("rot", [0], [1], 2, 1.0)], 1)  # parameter index out of range
    with pytest.raises(QulacsMI355Error):
        AdjointCircuit(4, [("dense", [0, 1, 2], np.eye(8))], 0)


def test_adjoint_full_ansatz_shape_against_oracle(oracle):
    """The 20-qubit, 12-layer hardware-efficient ansatz (480 RY/RZ parameters)
    with the 57-term Heisenberg chain, two parameter sets: energies against the
    oracle, and three gradient entries of the first set against central
    differences (step 1e-6) and the parameter shift of oracle energies."""
    n, layers, B = 20, 12, 2
    gates, P = hardware_efficient_ansatz(n, layers)
    assert P == 480
    terms = heisenberg_terms(n)
    assert len(terms) == 57
    scale = sum(abs(t[3]) for t in terms)
    rng = np.random.default_rng(21)
    theta = rng.uniform(0, 2 * np.pi, (B, P))
    circ = AdjointCircuit(n, gates, P)
    e, g = circ.run(torch.tensor(theta, device="cuda:0"), terms)
    e, g = e.cpu().numpy(), g.cpu().numpy()
    print("energies", e, flush=True)
    for b in range(B):
        assert abs(e[b] - _oracle_energy(oracle, n, gates, theta[b], terms)) < 1e-10 * scale
    h = 1e-6
    for p in (0, 241, 479):
        tp, tm = theta[0].copy(), theta[0].copy()
        tp[p] += h
        tm[p] -= h
        fd = (_oracle_energy(oracle, n, gates, tp, terms) - _oracle_energy(oracle, n, gates, tm, terms)) / (2 * h)
        tp[p] = theta[0][p] + np.pi / 2
        tm[p] = theta[0][p] - np.pi / 2
        ps = (_oracle_energy(oracle, n, gates, tp, terms) - _oracle_energy(oracle, n, gates, tm, terms)) / 2
        print(f"param {p}: adjoint {g[0, p]:+.15f} fd {fd:+.15f} shift {ps:+.15f}", flush=True)
        assert abs(g[0, p] - fd) < 1e-8 * scale
        assert abs(g[0, p] - ps) < 1e-10 * scale
