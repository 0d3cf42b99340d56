"""Device-side helpers for the GPU tests: every call goes through the C ABI
(qulacs_amd._lib), never through PyTorch or the oracle."""
from __future__ import annotations

import ctypes as C

import numpy as np

from qulacs_amd import _lib


def u32(xs) -> C.Array:
    xs = [int(x) for x in xs]
    return (C.c_uint * max(1, len(xs)))(*xs)


def cmat(m) -> C.Array:
    a = np.ascontiguousarray(np.asarray(m, dtype=np.complex128)).ravel()
    arr = (_lib.QmComplex * a.size)()
    for i, z in enumerate(a):
        arr[i].re = z.real
        arr[i].im = z.imag
    return arr


def cval(z: complex) -> _lib.QmComplex:
    return _lib.QmComplex(float(np.real(z)), float(np.imag(z)))


def to_complex(q: _lib.QmComplex) -> complex:
    return complex(q.re, q.im)


class DevState:
    """A device buffer of `dim` complex128 values (a state vector, or a flat DM)."""

    def __init__(self, dim: int, device: int = 0):
        self.dim = dim
        self.device = device
        self.ptr = _lib.call("allocate_quantum_state_host", dim, device)
        assert self.ptr

    @classmethod
    def from_host(cls, host: np.ndarray) -> "DevState":
        h = np.ascontiguousarray(host, dtype=np.complex128)
        s = cls(h.size)
        s.upload(h)
        return s

    def upload(self, host: np.ndarray) -> None:
        h = np.ascontiguousarray(host, dtype=np.complex128)
        assert h.size == self.dim
        _lib.call("copy_quantum_state_from_host_to_device", self.ptr, h.ctypes.data, self.dim, None, self.device)

    def download(self) -> np.ndarray:
        out = np.empty(self.dim, dtype=np.complex128)
        _lib.call("copy_quantum_state_from_device_to_host", out.ctypes.data, self.ptr, self.dim, None, self.device)
        return out

    def free(self) -> None:
        if self.ptr:
            _lib.call("release_quantum_state_host", self.ptr, self.device)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def rel(a, b) -> float:
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def apply_golden_case(case, s: DevState, dim: int) -> None:
    """Apply one golden state-vector case (see tests/test_oracle.py) through the C ABI."""
    name, kind, tg, cl, cv, param, mat = case
    k = len(tg)
    call = _lib.call
    if kind == "dense":
        m = cmat(mat)
        if k == 1:
            call("single_qubit_dense_matrix_gate_host", tg[0], m, s.ptr, dim, None, 0)
        elif k == 2:
            call("double_qubit_dense_matrix_gate_host", tg[0], tg[1], m, s.ptr, dim, None, 0)
        elif k == 3:
            call("triple_qubit_dense_matrix_gate_host", tg[0], tg[1], tg[2], m, s.ptr, dim, None, 0)
        elif k == 4:
            call("quad_qubit_dense_matrix_gate_host", u32(tg), m, s.ptr, dim, None, 0)
        elif k == 5:
            call("penta_qubit_dense_matrix_gate_host", u32(tg), m, s.ptr, dim, None, 0)
        else:
            call("multi_qubit_dense_matrix_gate_host", u32(tg), k, m, s.ptr, dim, None, 0)
    elif kind == "cdense":
        m = cmat(mat)
        if k == 1 and len(cl) == 1:
            call("single_qubit_control_single_qubit_dense_matrix_gate_host", cl[0], cv[0], tg[0], m, s.ptr, dim, None, 0)
        elif k == 1:
            call("multi_qubit_control_single_qubit_dense_matrix_gate_host", u32(cl), u32(cv), len(cl), tg[0], m,
                 s.ptr, dim, None, 0)
        elif len(cl) == 1:
            call("single_qubit_control_multi_qubit_dense_matrix_gate_host", cl[0], cv[0], u32(tg), k, m, s.ptr, dim,
                 None, 0)
        else:
            call("multi_qubit_control_multi_qubit_dense_matrix_gate_host", u32(cl), u32(cv), len(cl), u32(tg), k, m,
                 s.ptr, dim, None, 0)
    elif kind == "named":
        call(f"{name}_gate_host", tg[0], s.ptr, dim, None, 0)
    elif kind == "rot":
        call(f"{name}_gate_host", tg[0], param, s.ptr, dim, None, 0)
    elif kind == "two":
        call(f"{name}_gate_host", cl[0], tg[0], s.ptr, dim, None, 0)
    elif kind == "diag1":
        call("single_qubit_diagonal_matrix_gate_host", tg[0], cmat(mat), s.ptr, dim, None, 0)
    elif kind == "phase1":
        call("single_qubit_phase_gate_host", tg[0], cval(complex(mat[0])), s.ptr, dim, None, 0)
    elif kind == "cdiag":
        call("qm_multi_qubit_control_multi_qubit_diagonal_matrix_gate_host", u32(cl), u32(cv), len(cl), u32(tg), k,
             cmat(mat), s.ptr, dim, None, 0)
    elif kind == "pauli":
        call("multi_qubit_Pauli_gate_partial_list_host", u32(tg), u32(cv), k, s.ptr, dim, None, 0)
    elif kind == "pauli_rot":
        call("multi_qubit_Pauli_rotation_gate_partial_list_host", u32(tg), u32(cv), k, param, s.ptr, dim, None, 0)
    else:
        raise AssertionError(kind)
