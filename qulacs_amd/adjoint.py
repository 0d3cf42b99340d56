"""Adjoint gradients of <H> for parametric circuits, batched over parameter
sets, behind a torch.autograd.Function (qm_adjoint_* entry points).

    circ = AdjointCircuit(n, gates, n_params)
    energy = expectation(circ, theta, terms)     # theta: (B, P) float64 on the GPU
    energy.sum().backward()                      # theta.grad from the adjoint kernel

Gates are ("rot", qubits, paulis, param, coef) for exp(i coef theta[param] P / 2)
(the reference's ParametricRX/RY/RZ/PauliRotation convention,
src/vqcsim/parametric_gate.hpp:31-212) or the fixed batch kinds
("dense" / "diag" on 1..2 qubits, "CZ", "CNOT").  Terms are (bit_flip,
phase_flip, rot90, coef) with real coefficients.  torch only provides device
memory and the stream; every kernel is libqulacs_mi355's.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib

_FIXED = {"dense": _lib.BATCH_DENSE, "diag": _lib.BATCH_DIAG, "CZ": _lib.BATCH_CZ, "CNOT": _lib.BATCH_CNOT}


class AdjointCircuit:
    def __init__(self, n: int, gates, n_params: int):
        gates = list(gates)
        arr = (_lib.QmAdjointGate * max(1, len(gates)))()
        self._keep = []
        for i, g in enumerate(gates):
            d = arr[i]
            if g[0] == "rot":
                _, qubits, paulis, param, coef = g
                d.kind = _lib.ADJ_ROT
                d.param = int(param)
                d.coef = float(coef)
                for j, p in enumerate(paulis):
                    d.paulis[j] = int(p)
            else:
                kind, qubits, matrix = g
                d.kind = _FIXED[kind]
                if matrix is not None:
                    m = np.ascontiguousarray(np.asarray(matrix, dtype=np.complex128)).ravel()
                    self._keep.append(m)
                    d.matrix = C.cast(m.ctypes.data, _lib.PC)
            if len(qubits) > 4:
                raise _lib.QulacsMI355Error(f"gate {i}: at most 4 qubits")
            d.qubit_count = len(qubits)
            for j, q in enumerate(qubits):
                d.qubits[j] = int(q)
        self.n = n
        self.n_params = n_params
        self.h = _lib.call("qm_adjoint_create", int(n), arr, len(gates), int(n_params))
        if not self.h:
            raise _lib.QulacsMI355Error("qm_adjoint_create failed")

    def stats(self) -> dict:
        """Register-block passes per sweep (0: one kernel per gate) and gates."""
        p, g = C.c_ulonglong(), C.c_ulonglong()
        _lib.call("qm_adjoint_stats", self.h, C.byref(p), C.byref(g))
        return {"passes": int(p.value), "gates": int(g.value)}

    def close(self) -> None:
        if self.h:
            _lib.call("qm_adjoint_destroy", self.h)
            self.h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def run(self, theta: torch.Tensor, terms):
        """(energy (B,), grad (B, P)) for theta of shape (B, P), float64 on a GPU."""
        if theta.dtype != torch.float64 or not theta.is_cuda or theta.dim() != 2 or theta.shape[1] != self.n_params:
            raise ValueError("theta must be a (batch, n_params) float64 GPU tensor")
        theta = theta.contiguous()
        B = theta.shape[0]
        dev = theta.device
        energy = torch.empty(B, dtype=torch.float64, device=dev)
        grad = torch.empty(B, self.n_params, dtype=torch.float64, device=dev)
        ws = torch.empty(int(_lib.call("qm_adjoint_workspace_bytes", self.h, B)), dtype=torch.uint8, device=dev)
        t = max(1, len(terms))
        bf = (C.c_ulonglong * t)(*[int(x[0]) for x in terms])
        pf = (C.c_ulonglong * t)(*[int(x[1]) for x in terms])
        rot = (C.c_uint * t)(*[int(x[2]) for x in terms])
        coef = (C.c_double * t)(*[float(x[3]) for x in terms])
        stream = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        _lib.call("qm_adjoint_run", self.h, B, C.c_void_p(theta.data_ptr()), bf, pf, rot, coef, len(terms),
                  C.c_void_p(energy.data_ptr()), C.c_void_p(grad.data_ptr()), C.c_void_p(ws.data_ptr()),
                  C.byref(stream), dev.index or 0)
        return energy, grad


class _Expectation(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, circuit, terms):
        energy, grad = circuit.run(theta.detach(), terms)
        ctx.save_for_backward(grad)
        return energy

    @staticmethod
    def backward(ctx, grad_out):
        (grad,) = ctx.saved_tensors
        return grad_out[:, None] * grad, None, None


def expectation(circuit: AdjointCircuit, theta: torch.Tensor, terms) -> torch.Tensor:
    """<H> for every row of theta; differentiable through the adjoint kernel."""
    return _Expectation.apply(theta, circuit, terms)


def hardware_efficient_ansatz(n: int, layers: int):
    """Per layer RY(theta) then RZ(theta') on every qubit, then CZ(i, i+1):
    returns (gates, n_params) with parameters numbered in gate order."""
    gates = []
    p = 0
    for _ in range(layers):
        for q in range(n):
            gates.append(("rot", [q], [2], p, 1.0))
            gates.append(("rot", [q], [3], p + 1, 1.0))
            p += 2
        for q in range(n - 1):
            gates.append(("CZ", [q, q + 1], None))
    return gates, p


def heisenberg_terms(n: int) -> list:
    """sum_{i<n-1} X_i X_{i+1} + Y_i Y_{i+1} + Z_i Z_{i+1} in mask form."""
    terms = []
    for i in range(n - 1):
        m = (1 << i) | (1 << (i + 1))
        terms += [(m, 0, 0, 1.0), (m, m, 2, 1.0), (0, m, 0, 1.0)]
    return terms
