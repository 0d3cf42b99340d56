"""Python access to the CPU checker (TEST INFRASTRUCTURE ONLY).

Loads oracle/liboracle.so (the in-repo restatement of the reference csim loops,
cpu_ref.cpp) and, when it was built, oracle/_ref/libcsim_ref.so (the reference
csim compiled from its own sources plus ref_bridge.cpp).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.

All arrays are numpy complex128 (states, matrices) / uint32 (qubit lists) /
float64 (uniforms); functions update states in place like csim does.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "liboracle.so"
REF_SO = HERE / "_ref" / "libcsim_ref.so"
NORM_PROBE_SO = HERE / "_ref" / "libnorm_probe.so"

_u = C.c_uint
_ull = C.c_ulonglong
_d = C.c_double
_p = C.c_void_p


def build(ref: bool = True) -> None:
    """make -C oracle (and the reference build when its sources are present)."""
    targets = ["all", "norm-probe"]
    if ref and Path("/root/reference/src/csim").is_dir():
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", str(HERE), "-j8", *targets], check=True)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_p)


def _u32(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint32))


def _c128(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.complex128))


class _Kernels:
    """Common wrapper: the same function names exist in both libraries, with
    prefix `oracle_` (restatement) or `csimref_` (compiled reference)."""

    prefix = ""

    def __init__(self, path: Path):
        if not path.exists():
            raise FileNotFoundError(f"{path} not built (run oracle.build())")
        self.lib = C.CDLL(str(path))

    def f(self, name: str):
        return getattr(self.lib, self.prefix + name)

    def has(self, name: str) -> bool:
        return hasattr(self.lib, self.prefix + name)

    # ------------------------------------------------------------ threads
    def set_num_threads(self, n: int) -> None:
        self.f("set_num_threads")(C.c_int(n))

    def num_threads(self) -> int:
        fn = self.f("num_threads")
        fn.restype = C.c_int
        return fn()

    # ------------------------------------------------------------ gates
    def named(self, name: str, target: int, psi: np.ndarray) -> None:
        self.f(f"{name}_gate")(_u(target), _ptr(psi), _ull(psi.size))

    def rotation(self, name: str, target: int, angle: float, psi: np.ndarray) -> None:
        self.f(f"{name}_gate")(_u(target), _d(angle), _ptr(psi), _ull(psi.size))

    def two(self, name: str, a: int, b: int, psi: np.ndarray) -> None:
        self.f(f"{name}_gate")(_u(a), _u(b), _ptr(psi), _ull(psi.size))

    def dense(self, targets, matrix, psi: np.ndarray) -> None:
        t = _u32(targets)
        m = _c128(matrix)
        self.f("multi_qubit_dense_matrix_gate")(_ptr(t), _u(t.size), _ptr(m), _ptr(psi), _ull(psi.size))

    def dense1(self, target: int, matrix, psi: np.ndarray) -> None:
        m = _c128(matrix)
        self.f("single_qubit_dense_matrix_gate")(_u(target), _ptr(m), _ptr(psi), _ull(psi.size))

    def dense2(self, t0: int, t1: int, matrix, psi: np.ndarray) -> None:
        m = _c128(matrix)
        self.f("double_qubit_dense_matrix_gate")(_u(t0), _u(t1), _ptr(m), _ptr(psi), _ull(psi.size))

    def controlled_dense(self, controls, cvals, targets, matrix, psi: np.ndarray) -> None:
        c, v, t, m = _u32(controls), _u32(cvals), _u32(targets), _c128(matrix)
        self.f("multi_qubit_control_multi_qubit_dense_matrix_gate")(
            _ptr(c), _ptr(v), _u(c.size), _ptr(t), _u(t.size), _ptr(m), _ptr(psi), _ull(psi.size))

    def diag1(self, target: int, diag, psi: np.ndarray) -> None:
        d = _c128(diag)
        self.f("single_qubit_diagonal_matrix_gate")(_u(target), _ptr(d), _ptr(psi), _ull(psi.size))

    def phase1(self, target: int, phase: complex, psi: np.ndarray) -> None:
        p = _c128([phase])
        self.f("single_qubit_phase_gate")(_u(target), _ptr(p), _ptr(psi), _ull(psi.size))

    def controlled_diag(self, controls, cvals, targets, diag, psi: np.ndarray) -> None:
        c, v, t, d = _u32(controls), _u32(cvals), _u32(targets), _c128(diag)
        self.f("multi_qubit_control_multi_qubit_diagonal_matrix_gate")(
            _ptr(c), _ptr(v), _u(c.size), _ptr(t), _u(t.size), _ptr(d), _ptr(psi), _ull(psi.size))

    def pauli(self, targets, paulis, psi: np.ndarray) -> None:
        t, p = _u32(targets), _u32(paulis)
        self.f("multi_qubit_Pauli_gate_partial_list")(_ptr(t), _ptr(p), _u(t.size), _ptr(psi), _ull(psi.size))

    def pauli_rotation(self, targets, paulis, angle: float, psi: np.ndarray) -> None:
        t, p = _u32(targets), _u32(paulis)
        self.f("multi_qubit_Pauli_rotation_gate_partial_list")(
            _ptr(t), _ptr(p), _u(t.size), _d(angle), _ptr(psi), _ull(psi.size))

    def normalize(self, sq: float, psi: np.ndarray) -> None:
        self.f("normalize")(_d(sq), _ptr(psi), _ull(psi.size))

    def add_with_coef(self, coef: complex, added: np.ndarray, psi: np.ndarray) -> None:
        c = _c128([coef])
        self.f("state_add_with_coef")(_ptr(c), _ptr(added), _ptr(psi), _ull(psi.size))

    def multiply(self, coef: complex, psi: np.ndarray) -> None:
        c = _c128([coef])
        self.f("state_multiply")(_ptr(c), _ptr(psi), _ull(psi.size))

    # ------------------------------------------------------------ reductions
    def _dret(self, name):
        fn = self.f(name)
        fn.restype = C.c_double
        return fn

    def norm_squared(self, psi: np.ndarray) -> float:
        return self._dret("state_norm_squared")(_ptr(psi), _ull(psi.size))

    def inner_product(self, bra: np.ndarray, ket: np.ndarray) -> complex:
        out = np.zeros(2)
        self.f("state_inner_product")(_ptr(bra), _ptr(ket), _ull(ket.size), _ptr(out))
        return complex(out[0], out[1])

    def M0(self, t: int, psi: np.ndarray) -> float:
        return self._dret("M0_prob")(_u(t), _ptr(psi), _ull(psi.size))

    def M1(self, t: int, psi: np.ndarray) -> float:
        return self._dret("M1_prob")(_u(t), _ptr(psi), _ull(psi.size))

    def marginal(self, sorted_targets, values, psi: np.ndarray) -> float:
        t, v = _u32(sorted_targets), _u32(values)
        return self._dret("marginal_prob")(_ptr(t), _ptr(v), _u(t.size), _ptr(psi), _ull(psi.size))

    def entropy(self, psi: np.ndarray) -> float:
        return self._dret("measurement_distribution_entropy")(_ptr(psi), _ull(psi.size))

    def expectation(self, targets, paulis, psi: np.ndarray) -> float:
        t, p = _u32(targets), _u32(paulis)
        return self._dret("expectation_value_multi_qubit_Pauli_operator_partial_list")(
            _ptr(t), _ptr(p), _u(t.size), _ptr(psi), _ull(psi.size))

    def transition(self, targets, paulis, bra: np.ndarray, ket: np.ndarray) -> complex:
        t, p = _u32(targets), _u32(paulis)
        out = np.zeros(2)
        self.f("transition_amplitude_multi_qubit_Pauli_operator_partial_list")(
            _ptr(t), _ptr(p), _u(t.size), _ptr(bra), _ptr(ket), _ull(ket.size), _ptr(out))
        return complex(out[0], out[1])

    # ------------------------------------------------------------ density matrix (dim = 2^n)
    def dm_from_pure(self, psi: np.ndarray) -> np.ndarray:
        rho = np.zeros(psi.size * psi.size, dtype=np.complex128)
        self.f("dm_initialize_with_pure_state")(_ptr(rho), _ptr(psi), _ull(psi.size))
        return rho

    def dm_dense1(self, target: int, matrix, rho: np.ndarray, dim: int) -> None:
        m = _c128(matrix)
        self.f("dm_single_qubit_dense_matrix_gate")(_u(target), _ptr(m), _ptr(rho), _ull(dim))

    def dm_dense(self, targets, matrix, rho: np.ndarray, dim: int) -> None:
        t, m = _u32(targets), _c128(matrix)
        self.f("dm_multi_qubit_dense_matrix_gate")(_ptr(t), _u(t.size), _ptr(m), _ptr(rho), _ull(dim))

    def dm_controlled_dense(self, controls, cvals, targets, matrix, rho: np.ndarray, dim: int) -> None:
        c, v, t, m = _u32(controls), _u32(cvals), _u32(targets), _c128(matrix)
        self.f("dm_multi_qubit_control_multi_qubit_dense_matrix_gate")(
            _ptr(c), _ptr(v), _u(c.size), _ptr(t), _u(t.size), _ptr(m), _ptr(rho), _ull(dim))

    def dm_norm(self, rho: np.ndarray, dim: int) -> float:
        return self._dret("dm_state_norm_squared")(_ptr(rho), _ull(dim))

    def dm_M0(self, t: int, rho: np.ndarray, dim: int) -> float:
        return self._dret("dm_M0_prob")(_u(t), _ptr(rho), _ull(dim))

    def dm_marginal(self, sorted_targets, values, rho: np.ndarray, dim: int) -> float:
        t, v = _u32(sorted_targets), _u32(values)
        return self._dret("dm_marginal_prob")(_ptr(t), _ptr(v), _u(t.size), _ptr(rho), _ull(dim))

    def dm_expectation(self, targets, paulis, rho: np.ndarray, dim: int) -> float:
        t, p = _u32(targets), _u32(paulis)
        return self._dret("dm_expectation_value_multi_qubit_Pauli_operator_partial_list")(
            _ptr(t), _ptr(p), _u(t.size), _ptr(rho), _ull(dim))


class Oracle(_Kernels):
    """The in-repo restatement (oracle/cpu_ref.cpp)."""

    prefix = "oracle_"

    def __init__(self):
        super().__init__(ORACLE_SO)

    def parallel_fill(self, psi: np.ndarray, value: complex) -> None:
        """First touch of a new buffer by the OpenMP threads (static schedule);
        the compiled reference in the same process shares the thread pool."""
        self.f("parallel_fill")(_ptr(psi), _ull(psi.size), _d(complex(value).real), _d(complex(value).imag))

    def sampling(self, psi: np.ndarray, uniforms) -> np.ndarray:
        u = np.ascontiguousarray(np.asarray(uniforms, dtype=np.float64))
        out = np.zeros(u.size, dtype=np.uint64)
        self.f("sampling")(_ptr(psi), _ull(psi.size), _ptr(u), _u(u.size), _ptr(out))
        return out

    def sampling_cdf(self, psi: np.ndarray) -> np.ndarray:
        cdf = np.zeros(psi.size + 1)
        self.f("sampling_cdf")(_ptr(psi), _ull(psi.size), _ptr(cdf))
        return cdf

    def dm_sampling(self, rho: np.ndarray, dim: int, uniforms) -> np.ndarray:
        u = np.ascontiguousarray(np.asarray(uniforms, dtype=np.float64))
        out = np.zeros(u.size, dtype=np.uint64)
        self.f("dm_sampling")(_ptr(rho), _ull(dim), _ptr(u), _u(u.size), _ptr(out))
        return out

    def dm_sampling_cdf(self, rho: np.ndarray, dim: int) -> np.ndarray:
        cdf = np.zeros(dim + 1)
        self.f("dm_sampling_cdf")(_ptr(rho), _ull(dim), _ptr(cdf))
        return cdf

    def dm_depolarizing(self, t: int, p: float, rho: np.ndarray, dim: int) -> None:
        self.f("dm_depolarizing")(_u(t), _d(p), _ptr(rho), _ull(dim))

    def mt64_uniforms(self, seed: int, count: int) -> np.ndarray:
        out = np.zeros(count, dtype=np.float64)
        self.f("mt64_uniforms")(_ull(seed), _u(count), _ptr(out))
        return out

    def mt64_raw(self, seed: int, count: int) -> np.ndarray:
        out = np.zeros(count, dtype=np.uint64)
        self.f("mt64_raw")(_ull(seed), _u(count), _ptr(out))
        return out

    def whole_diagonal(self, diag, psi: np.ndarray) -> None:
        d = _c128(diag)
        self.f("whole_state_diagonal_gate")(_ptr(d), _ptr(psi), _ull(psi.size))


class Reference(_Kernels):
    """The reference csim, compiled from its sources (oracle/_ref/libcsim_ref.so)."""

    prefix = "csimref_"

    def __init__(self):
        super().__init__(REF_SO)

    def dm_named(self, name: str, target: int, rho: np.ndarray, dim: int) -> None:
        self.f(f"dm_{name}_gate")(_u(target), _ptr(rho), _ull(dim))

    def dm_two(self, name: str, a: int, b: int, rho: np.ndarray, dim: int) -> None:
        self.f(f"dm_{name}_gate")(_u(a), _u(b), _ptr(rho), _ull(dim))

    def dm_rotation(self, name: str, target: int, angle: float, rho: np.ndarray, dim: int) -> None:
        self.f(f"dm_{name}_gate")(_u(target), _d(angle), _ptr(rho), _ull(dim))


def reference_available() -> bool:
    return REF_SO.exists()


def norm_probe_stacked_prob(psi: np.ndarray) -> np.ndarray:
    """The reference's sampling prefix sum (sum += std::norm(psi_i)) compiled from
    oracle/norm_probe.cpp with the reference's release flags (built here, used on
    this machine only: -march=native)."""
    lib = C.CDLL(str(NORM_PROBE_SO))
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    out = np.zeros(psi.size + 1)
    lib.norm_probe_stacked_prob.argtypes = [_p, C.c_size_t, _p]
    lib.norm_probe_stacked_prob(_ptr(psi), psi.size, _ptr(out))
    return out


def parameter_shift_grad(energy, theta) -> np.ndarray:
    """GradCalculator::calculate_grad (src/vqcsim/GradCalculator.cpp:8-44): for
    each parameter k, the energy with theta_k + pi/2 and with theta_k - pi/2
    (every other parameter at theta), and grad_k = (E+ - E-) / 2.  `energy` maps
    a parameter vector (one angle per parametric gate, in circuit order) to
    <H>; the reference evaluates it with CausalConeSimulator, whose value is the
    full simulation's, so any exact evaluation (here the oracle's kernels) gives
    the same gradient up to rounding."""
    theta = np.asarray(theta, dtype=np.float64)
    grad = np.zeros(theta.size)
    for k in range(theta.size):
        tp = theta.copy()
        tm = theta.copy()
        tp[k] = theta[k] + np.pi / 2
        tm[k] = theta[k] - np.pi / 2
        grad[k] = (energy(tp) - energy(tm)) / 2.0
    return grad
