"""qulacs_amd: qulacs's state-vector / density-matrix simulator API with every
state resident in MI355X HBM and every update run by HIP kernels (complex128).

Drop-in names (python/cppsim_wrapper.cpp in the reference):

    from qulacs_amd import QuantumState, DensityMatrix, QuantumCircuit, Observable
    from qulacs_amd import gate
    from qulacs_amd.state import inner_product

``QuantumState`` and ``QuantumStateGpu`` both name the GPU state vector
(``QuantumStateMI355``); ``DensityMatrix`` names ``DensityMatrixMI355``.
Importing the package loads the compiled module ``_core`` and through it
``libqulacs_mi355.so``; there is no CPU fallback, so a missing build raises
ImportError here.
"""
from __future__ import annotations

import sys as _sys

# One HIP runtime per process.  PyTorch ships its own HIP/HSA runtime; when torch
# is installed it is loaded first, so this library binds to that same runtime
# (by soname) instead of bringing a second HIP/HSA stack into the process, which
# cannot open the device once the other stack has.
try:
    import torch as _torch  # noqa: F401
except ImportError:  # pragma: no cover - torch is optional for the library itself
    _torch = None

try:
    from . import _core
except ImportError as e:  # pragma: no cover - exercised only without a build
    raise ImportError(
        "qulacs_amd: the compiled module is missing; run qulacs_amd/build.py "
        f"(or __graft_entry__.build()) first ({e})"
    ) from e

from ._core import (  # noqa: F401
    CausalConeSimulator,
    ClsNoisyEvolution,
    ClsNoisyEvolution_fast,
    ClsNpairQubitGate,
    ClsPauliGate,
    ClsStateReflectionGate,
    DensityMatrixMI355,
    GeneralQuantumOperator,
    GradCalculator,
    NoiseSimulator,
    Observable,
    ParametricQuantumCircuit,
    PauliOperator,
    QuantumCircuit,
    QuantumCircuitSimulator,
    QuantumGate_Adaptive,
    QuantumGate_CP,
    QuantumGate_CPTP,
    QuantumGate_LinearCombination,
    QuantumGate_Probabilistic,
    QuantumGate_SingleParameter,
    QuantumGateBase,
    QuantumGateDiagonalMatrix,
    QuantumGateMatrix,
    QuantumStateBase,
    QuantumStateMI355,
    QulacsError,
    SimulationResult,
    check_build_for_mpi,
    circuit,
    gate,
    get_circuit_fusion,
    get_num_device,
    observable,
    quantum_operator,
    get_synchronous,
    sampling_margin,
    set_circuit_fusion,
    set_synchronous,
    state,
    synchronize,
    to_general_quantum_operator,
)

QuantumState = QuantumStateMI355
QuantumStateGpu = QuantumStateMI355
StateVector = QuantumStateMI355
StateVectorGpu = QuantumStateMI355
DensityMatrix = DensityMatrixMI355

# The reference's per-family gate classes.  Here every named, rotation, CNOT/CZ/SWAP
# and sparse gate is a QuantumGateMatrix that dispatches on its kind, and a Pauli
# rotation is a ClsPauliGate, so the family names refer to those classes.
ClsOneQubitGate = QuantumGateMatrix
ClsOneQubitRotationGate = QuantumGateMatrix
ClsOneControlOneTargetGate = QuantumGateMatrix
ClsTwoQubitGate = QuantumGateMatrix
QuantumGateSparseMatrix = QuantumGateMatrix
ClsPauliRotationGate = ClsPauliGate
QuantumGate_Instrument = QuantumGate_CPTP
ClsReversibleBooleanGate = QuantumGateMatrix

# `import qulacs_amd.gate` / `from qulacs_amd.state import inner_product`
_sys.modules[__name__ + ".gate"] = gate
_sys.modules[__name__ + ".state"] = state
_sys.modules[__name__ + ".circuit"] = circuit
_sys.modules[__name__ + ".quantum_operator"] = quantum_operator
_sys.modules[__name__ + ".observable"] = observable

__version__ = _core.__version__
