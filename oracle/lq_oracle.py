"""CPU oracle for the lightning.qubit hot path -- TEST INFRASTRUCTURE ONLY.

This module is a numpy restatement of the reference algorithm. Only tests/,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
it, and only as the checker. The product (``liblq_amd.so`` and the
``lightning_amd`` Python package) never imports or calls it.

Everything is computed in complex128 / float64 whatever the precision of the
engine under test. Reference paths are relative to
``pennylane_lightning/core/`` (``lq/`` = ``simulators/lightning_qubit/``).

Conventions restated (all pinned by tests/test_oracle.py):
  * wire ``w`` is amplitude-index bit ``n-1-w``: wire 0 is the most
    significant bit (lq/StateVectorLQubit.hpp:950,
    lq/gates/cpu_kernels/GateImplementationsLM.hpp:690). With a C-order
    reshape of the state to ``(2,)*n`` axis ``w`` is wire ``w``.
  * gate matrices are row-major and ``wires[0]`` is the most significant bit
    of the local index (GateImplementationsLM.hpp:336-381, 1363-1381).
  * ``inverse`` applies U^dagger (GateImplementationsLM.hpp:287-293 for
    matrices; per-gate angle negation for named gates, gates/Gates.hpp:1249-1330).
  * controlled ops act on the subspace where every control wire equals its
    control value (GateImplementationsLM.hpp:407-440).

Pinning (what checks each function; see tests/test_oracle.py):
  gate_matrix        closed forms from the reference's own known-answer tests
                     (lq/gates/tests/Test_GateImplementations_Param.cpp:146-166),
                     unitarity, inverse, and the compiled reference
                     (oracle/_ref) over every gate name, random states,
                     controls and inverses.
  generator          the identity i*s*G|psi> = dU/dtheta|psi> at theta=0
                     (lq/gates/tests/Test_GateImplementations_Generator.cpp),
                     and oracle/_ref's applyGenerator (state and scale).
  expval/probs       Test_MeasurementsLQubit.cpp:71-177 known answers
                     (0.49272486, 0.42073549, 0.28232124), sum(probs)=1,
                     <Z_i> = sum(probs*(+-1)), oracle/_ref.
  adjoint_jacobian   closed forms of core/algorithms/tests/Test_AdjointJacobian.cpp
                     (RX/Z -> -sin p, PhaseShift/Y -> cos p), central finite
                     differences, and oracle/_ref's AdjointJacobian.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np

# ---------------------------------------------------------------------------
# Gate tables: names, wire counts, parameter counts (gates/Constant.hpp:60-431)
# ---------------------------------------------------------------------------
GATE_WIRES = {
    "Identity": 1, "PauliX": 1, "PauliY": 1, "PauliZ": 1, "Hadamard": 1,
    "S": 1, "SX": 1, "T": 1, "PhaseShift": 1, "RX": 1, "RY": 1, "RZ": 1,
    "Rot": 1, "CNOT": 2, "CY": 2, "CZ": 2, "SWAP": 2, "IsingXX": 2,
    "IsingXY": 2, "IsingYY": 2, "IsingZZ": 2, "ControlledPhaseShift": 2,
    "CRX": 2, "CRY": 2, "CRZ": 2, "CRot": 2, "SingleExcitation": 2,
    "SingleExcitationMinus": 2, "SingleExcitationPlus": 2, "Toffoli": 3,
    "CSWAP": 3, "DoubleExcitation": 4, "DoubleExcitationMinus": 4,
    "DoubleExcitationPlus": 4, "GlobalPhase": 1, "PSWAP": 2,
    # multi-qubit gates take any number of wires
    "MultiRZ": -1, "PCPhase": -1,
}
GATE_NPARAMS = {
    "Identity": 0, "PauliX": 0, "PauliY": 0, "PauliZ": 0, "Hadamard": 0,
    "S": 0, "SX": 0, "T": 0, "PhaseShift": 1, "RX": 1, "RY": 1, "RZ": 1,
    "Rot": 3, "CNOT": 0, "CY": 0, "CZ": 0, "SWAP": 0, "IsingXX": 1,
    "IsingXY": 1, "IsingYY": 1, "IsingZZ": 1, "ControlledPhaseShift": 1,
    "CRX": 1, "CRY": 1, "CRZ": 1, "CRot": 3, "SingleExcitation": 1,
    "SingleExcitationMinus": 1, "SingleExcitationPlus": 1, "Toffoli": 0,
    "CSWAP": 0, "DoubleExcitation": 1, "DoubleExcitationMinus": 1,
    "DoubleExcitationPlus": 1, "GlobalPhase": 1, "PSWAP": 1, "MultiRZ": 1,
    "PCPhase": 2,
}
# Gates accepted in the controlled form (gates/Constant.hpp:101-129).
CONTROLLED_GATES = {
    "PauliX", "PauliY", "PauliZ", "Hadamard", "S", "SX", "T", "PhaseShift",
    "RX", "RY", "RZ", "Rot", "SWAP", "IsingXX", "IsingXY", "IsingYY",
    "IsingZZ", "SingleExcitation", "SingleExcitationMinus",
    "SingleExcitationPlus", "DoubleExcitation", "DoubleExcitationMinus",
    "DoubleExcitationPlus", "MultiRZ", "GlobalPhase", "PSWAP", "PCPhase",
}

_I2 = np.eye(2, dtype=np.complex128)
_X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
_Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
_Z = np.array([[1, 0], [0, -1]], dtype=np.complex128)
_H = np.array([[1, 1], [1, -1]], dtype=np.complex128) / math.sqrt(2.0)
_P1 = np.array([[0, 0], [0, 1]], dtype=np.complex128)  # |1><1|


def _ctrl_block(u: np.ndarray) -> np.ndarray:
    """|0><0| (x) I + |1><1| (x) U: control on wires[0] (Gates.hpp:402-521)."""
    d = u.shape[0]
    m = np.eye(2 * d, dtype=np.complex128)
    m[d:, d:] = u
    return m


def _rot(phi: float, theta: float, omega: float) -> np.ndarray:
    """Rot(phi,theta,omega) = RZ(omega) RY(theta) RZ(phi) (Gates.hpp:389-399)."""
    c, s = math.cos(theta / 2), math.sin(theta / 2)
    p, m = phi + omega, phi - omega
    return np.array(
        [[complex(math.cos(p / 2) * c, -math.sin(p / 2) * c),
          complex(-math.cos(m / 2) * s, -math.sin(m / 2) * s)],
         [complex(math.cos(m / 2) * s, -math.sin(m / 2) * s),
          complex(math.cos(p / 2) * c, math.sin(p / 2) * c)]])


def gate_matrix(name: str, params: Sequence[float] = (), inverse: bool = False,
                num_wires: int | None = None) -> np.ndarray:
    """Row-major local matrix of a named gate (gates/Gates.hpp:29-1330).

    ``inverse`` returns U^dagger, which for every gate equals the reference's
    per-gate rule (negated angles, Rot(-w,-t,-p), conjugated S/T/SX;
    Gates.hpp:1249-1330, GateImplementationsLM.hpp:1063-1065,1194-1196).
    ``num_wires`` sizes MultiRZ / PCPhase.
    """
    p = [float(x) for x in params]
    if name == "Identity":
        u = _I2
    elif name == "PauliX":
        u = _X
    elif name == "PauliY":
        u = _Y
    elif name == "PauliZ":
        u = _Z
    elif name == "Hadamard":
        u = _H
    elif name == "S":
        u = np.diag([1, 1j])
    elif name == "T":
        u = np.diag([1, np.exp(1j * math.pi / 4)])
    elif name == "SX":
        u = 0.5 * np.array([[1 + 1j, 1 - 1j], [1 - 1j, 1 + 1j]])
    elif name == "PhaseShift":
        u = np.diag([1, np.exp(1j * p[0])])
    elif name == "RX":
        c, s = math.cos(p[0] / 2), math.sin(p[0] / 2)
        u = np.array([[c, -1j * s], [-1j * s, c]])
    elif name == "RY":
        c, s = math.cos(p[0] / 2), math.sin(p[0] / 2)
        u = np.array([[c, -s], [s, c]], dtype=np.complex128)
    elif name == "RZ":
        u = np.diag([np.exp(-0.5j * p[0]), np.exp(0.5j * p[0])])
    elif name == "Rot":
        u = _rot(p[0], p[1], p[2])
    elif name == "GlobalPhase":
        # e^{-i theta} on the target (GateImplementationsLM.hpp:2044-2111)
        u = np.exp(-1j * p[0]) * _I2
    elif name == "CNOT":
        u = _ctrl_block(_X)
    elif name == "CY":
        u = _ctrl_block(_Y)
    elif name == "CZ":
        u = _ctrl_block(_Z)
    elif name == "SWAP":
        u = np.eye(4, dtype=np.complex128)[[0, 2, 1, 3]]
    elif name == "ControlledPhaseShift":
        u = np.diag([1, 1, 1, np.exp(1j * p[0])])
    elif name == "CRX":
        u = _ctrl_block(gate_matrix("RX", p))
    elif name == "CRY":
        u = _ctrl_block(gate_matrix("RY", p))
    elif name == "CRZ":
        u = _ctrl_block(gate_matrix("RZ", p))
    elif name == "CRot":
        u = _ctrl_block(_rot(*p))
    elif name == "IsingXX":
        c, s = math.cos(p[0] / 2), math.sin(p[0] / 2)
        u = c * np.eye(4) - 1j * s * np.kron(_X, _X)
    elif name == "IsingYY":
        c, s = math.cos(p[0] / 2), math.sin(p[0] / 2)
        u = c * np.eye(4) - 1j * s * np.kron(_Y, _Y)
    elif name == "IsingZZ":
        e = np.exp(-0.5j * p[0])
        u = np.diag([e, np.conj(e), np.conj(e), e])
    elif name == "IsingXY":
        c, s = math.cos(p[0] / 2), math.sin(p[0] / 2)
        u = np.eye(4, dtype=np.complex128)
        u[1, 1] = u[2, 2] = c
        u[1, 2] = u[2, 1] = 1j * s
    elif name in ("SingleExcitation", "SingleExcitationMinus", "SingleExcitationPlus"):
        c, s = math.cos(p[0] / 2), math.sin(p[0] / 2)
        e = {"SingleExcitation": 1.0, "SingleExcitationMinus": np.exp(-0.5j * p[0]),
             "SingleExcitationPlus": np.exp(0.5j * p[0])}[name]
        u = np.diag([e, c, c, e]).astype(np.complex128)
        u[1, 2], u[2, 1] = -s, s
    elif name == "PSWAP":
        e = np.exp(1j * p[0])
        u = np.zeros((4, 4), dtype=np.complex128)
        u[0, 0] = u[3, 3] = 1
        u[1, 2] = u[2, 1] = e
    elif name == "Toffoli":
        u = _ctrl_block(_ctrl_block(_X))
    elif name == "CSWAP":
        u = _ctrl_block(gate_matrix("SWAP"))
    elif name in ("DoubleExcitation", "DoubleExcitationMinus", "DoubleExcitationPlus"):
        # Gates.hpp:741-930: rotation on |0011>,|1100>, phase elsewhere
        c, s = math.cos(p[0] / 2), math.sin(p[0] / 2)
        e = {"DoubleExcitation": 1.0, "DoubleExcitationMinus": np.exp(-0.5j * p[0]),
             "DoubleExcitationPlus": np.exp(0.5j * p[0])}[name]
        u = e * np.eye(16, dtype=np.complex128)
        u[3, 3] = u[12, 12] = c
        u[3, 12], u[12, 3] = -s, s
    elif name == "MultiRZ":
        # GateImplementationsLM.hpp:1988-2010: e^{-i t/2} on even parity
        k = int(num_wires)
        par = np.array([bin(i).count("1") & 1 for i in range(1 << k)])
        u = np.diag(np.where(par == 0, np.exp(-0.5j * p[0]), np.exp(0.5j * p[0])))
    elif name == "PCPhase":
        # GateImplementationsLM.hpp:2113-2162: e^{i phi} on the first `dim`
        # local basis states, e^{-i phi} on the rest
        k = int(num_wires)
        dim = int(round(p[1]))
        d = np.full(1 << k, np.exp(-1j * p[0]))
        d[:dim] = np.exp(1j * p[0])
        u = np.diag(d)
    else:
        raise ValueError(f"unknown gate {name}")
    u = np.asarray(u, dtype=np.complex128)
    return u.conj().T.copy() if inverse else u


# Generators: (local Hermitian G, scale s) with U(theta) = exp(i s theta G)
# (lq/gates/cpu_kernels/PauliGenerator.hpp:31-60,
#  GateImplementationsLM.hpp:2181-2952). Controlled generators are P_ctrl (x) G:
#  the non-control subspace is zeroed (GateImplementationsLM.hpp:2181-2228).
def generator(name: str, num_wires: int | None = None) -> tuple[np.ndarray, float]:
    if name == "RX":
        return _X.copy(), -0.5
    if name == "RY":
        return _Y.copy(), -0.5
    if name == "RZ":
        return _Z.copy(), -0.5
    if name == "PhaseShift":
        return _P1.copy(), 1.0
    if name == "GlobalPhase":
        return _I2.copy(), -1.0
    if name == "CRX":
        return np.kron(_P1, _X), -0.5
    if name == "CRY":
        return np.kron(_P1, _Y), -0.5
    if name == "CRZ":
        return np.kron(_P1, _Z), -0.5
    if name == "ControlledPhaseShift":
        return np.kron(_P1, _P1), 1.0
    if name == "IsingXX":
        return np.kron(_X, _X), -0.5
    if name == "IsingYY":
        return np.kron(_Y, _Y), -0.5
    if name == "IsingZZ":
        return np.kron(_Z, _Z), -0.5
    if name == "IsingXY":
        g = np.zeros((4, 4), dtype=np.complex128)
        g[1, 2] = g[2, 1] = 1
        return g, 0.5
    if name in ("SingleExcitation", "SingleExcitationMinus", "SingleExcitationPlus"):
        d = {"SingleExcitation": 0.0, "SingleExcitationMinus": 1.0,
             "SingleExcitationPlus": -1.0}[name]
        g = np.diag([d, 0, 0, d]).astype(np.complex128)
        g[1, 2], g[2, 1] = -1j, 1j
        return g, -0.5
    if name == "PSWAP":
        g = np.zeros((4, 4), dtype=np.complex128)
        g[1, 2] = g[2, 1] = 1
        return g, 1.0
    if name == "DoubleExcitation":
        g = np.zeros((16, 16), dtype=np.complex128)
        g[3, 12], g[12, 3] = -1j, 1j
        return g, -0.5
    if name == "DoubleExcitationMinus":
        g = np.eye(16, dtype=np.complex128)
        g[3, 3] = g[12, 12] = 0
        g[3, 12], g[12, 3] = -1j, 1j
        return g, -0.5
    if name == "DoubleExcitationPlus":
        # LM applies -G of Gates.hpp:907-930 and returns +1/2 (:2855-2882)
        g = np.eye(16, dtype=np.complex128)
        g[3, 3] = g[12, 12] = 0
        g[3, 12], g[12, 3] = 1j, -1j
        return g, 0.5
    if name == "MultiRZ":
        k = int(num_wires)
        par = np.array([bin(i).count("1") & 1 for i in range(1 << k)])
        return np.diag(1.0 - 2.0 * par).astype(np.complex128), -0.5
    raise ValueError(f"gate {name} has no generator")


# ---------------------------------------------------------------------------
# State-vector primitives
# ---------------------------------------------------------------------------
def zero_state(n: int) -> np.ndarray:
    """|0...0> (lq/StateVectorLQubitManaged.hpp:75-82)."""
    s = np.zeros(1 << n, dtype=np.complex128)
    s[0] = 1.0
    return s


def apply_matrix(state: np.ndarray, n: int, mat: np.ndarray, wires: Sequence[int],
                 ctrls: Sequence[int] = (), cvals: Sequence[int] = (),
                 inverse: bool = False) -> np.ndarray:
    """In-place U (or U^dagger) on ``wires`` restricted to the control subspace.

    Restates applyNCN / applyNCMultiQubitOp (GateImplementationsLM.hpp:407-498)
    and applyNCSingleQubitOp/TwoQubitOp (:276-381): wires[0] is the MSB of the
    local index, controls select the subspace ctrl==cval.
    """
    wires, ctrls = list(wires), list(ctrls)
    cvals = [int(bool(v)) for v in (cvals if len(cvals) else [1] * len(ctrls))]
    k = len(wires)
    m = np.asarray(mat, dtype=np.complex128).reshape(1 << k, 1 << k)
    if inverse:
        m = m.conj().T
    if len(set(wires) | set(ctrls)) != k + len(ctrls):
        raise ValueError("wires and controls must be disjoint and unique")
    psi = state.reshape((2,) * n)
    axes = ctrls + wires
    moved = np.moveaxis(psi, axes, list(range(len(axes))))
    sub = moved[tuple(cvals)] if ctrls else moved
    shape = sub.shape
    flat = sub.reshape(1 << k, -1)
    new = m @ flat
    if ctrls:
        moved[tuple(cvals)] = new.reshape(shape)
    else:
        moved[...] = new.reshape(shape)
    return state


def apply_op(state: np.ndarray, n: int, name: str, wires: Sequence[int],
             inverse: bool = False, params: Sequence[float] = (),
             ctrls: Sequence[int] = (), cvals: Sequence[int] = ()) -> np.ndarray:
    """applyOperation(name, [ctrls, cvals,] wires, inverse, params)
    (lq/StateVectorLQubit.hpp:378-423). Identity is skipped (:380-382)."""
    if name == "Identity":
        return state
    if name == "GlobalPhase" and not ctrls:
        # uncontrolled GlobalPhase multiplies the whole state (:2044-2111)
        state *= np.exp((1j if inverse else -1j) * float(params[0]))
        return state
    u = gate_matrix(name, params, inverse, num_wires=len(wires))
    return apply_matrix(state, n, u, wires, ctrls, cvals)


def apply_generator(state: np.ndarray, n: int, name: str, wires: Sequence[int],
                    ctrls: Sequence[int] = (), cvals: Sequence[int] = ()) -> float:
    """applyGenerator (lq/StateVectorLQubit.hpp:527-560): state <- G state
    (non-control subspace zeroed when controlled), returns the scale."""
    g, s = generator(name, num_wires=len(wires))
    if name == "GlobalPhase" and not ctrls:
        return s
    if ctrls:
        cvals = [int(bool(v)) for v in (cvals if len(cvals) else [1] * len(ctrls))]
        psi = state.reshape((2,) * n)
        moved = np.moveaxis(psi, list(ctrls), list(range(len(ctrls))))
        keep = moved[tuple(cvals)].copy()
        moved[...] = 0
        moved[tuple(cvals)] = keep
    apply_matrix(state, n, g, wires, ctrls, cvals)
    return s


def apply_pauli_rot(state: np.ndarray, n: int, wires: Sequence[int], theta: float,
                    word: str, inverse: bool = False) -> np.ndarray:
    """exp(-i theta/2 P_word) (GateImplementationsLM.hpp:575-629).

    Reproduces the reference as it is: its ``mask_xy`` takes every letter
    that is not 'Z' (:603-604), so an 'I' in the word acts as 'X'. An all-'Z'
    word goes to MultiRZ (:583-587), which is the same operator.
    """
    mats = {"X": _X, "Y": _Y, "Z": _Z}
    p = np.array([[1.0]], dtype=np.complex128)
    for ch in word:
        p = np.kron(p, mats.get(ch, _X))
    t = -theta if inverse else theta
    u = math.cos(t / 2) * np.eye(p.shape[0]) - 1j * math.sin(t / 2) * p
    return apply_matrix(state, n, u, wires)


# ---------------------------------------------------------------------------
# Measurements (lq/measurements/MeasurementsLQubit.hpp, ExpValFunc.hpp)
# ---------------------------------------------------------------------------
def probs(state: np.ndarray, n: int, wires: Sequence[int] | None = None) -> np.ndarray:
    """|psi|^2 marginalised onto ``wires``; wires[0] is the output MSB
    (MeasurementsLQubit.hpp:90-163)."""
    p = np.abs(state) ** 2
    if wires is None:
        return p
    wires = list(wires)
    if len(wires) == 0:
        return np.array([1.0])
    t = p.reshape((2,) * n)
    rest = [a for a in range(n) if a not in wires]
    t = t.sum(axis=tuple(rest)) if rest else t
    # t's axes are the kept wires in increasing order; reorder to `wires`
    order = sorted(wires)
    t = np.transpose(t, [order.index(w) for w in wires])
    return t.reshape(-1)


_NAMED_OBS = {"PauliX": _X, "PauliY": _Y, "PauliZ": _Z, "Hadamard": _H, "Identity": _I2}


def expval_matrix(state: np.ndarray, n: int, mat: np.ndarray, wires: Sequence[int]) -> float:
    """Re <psi|M|psi> over the local block (ExpValFunc.hpp:106-333)."""
    phi = apply_matrix(state.copy(), n, mat, wires)
    return float(np.real(np.vdot(state, phi)))


def expval_named(state: np.ndarray, n: int, name: str, wires: Sequence[int]) -> float:
    """Measurements::expval(name, wires) (MeasurementsLQubit.hpp:247-305,
    applyExpVal1 ExpValFunc.hpp:68-93; Identity returns 1 without reading)."""
    if name == "Identity":
        return 1.0
    if name in _NAMED_OBS:
        return expval_matrix(state, n, _NAMED_OBS[name], wires)
    return expval_matrix(state, n, gate_matrix(name, (), False, len(wires)), wires)


# Observables are nested tuples mirroring NamedObs / TensorProdObs /
# HermitianObs / Hamiltonian (core/observables/Observables.hpp:128-584):
#   ("N", name, wire) | ("T", [(name, wire), ...]) | ("M", matrix, wires)
#   | ("H", coeffs, [obs, ...])
def obs_apply(obs, state: np.ndarray, n: int) -> np.ndarray:
    """Observable::applyInPlace (lq/observables/ObservablesLQubit.hpp:52-366)."""
    tag = obs[0]
    if tag == "N":
        _, name, w = obs
        if name == "Identity":
            return state
        return apply_matrix(state, n, _NAMED_OBS[name], [w])
    if tag == "T":
        for name, w in obs[1]:
            if name != "Identity":
                apply_matrix(state, n, _NAMED_OBS[name], [w])
        return state
    if tag == "M":
        return apply_matrix(state, n, obs[1], obs[2])
    if tag == "S":  # ("S", data, indices, indptr, wires): CSR over the full state index
        if len(obs[4]) != n:
            raise ValueError("SparseH wire count does not match state-vector size")
        state[...] = _sparse_apply(obs[3], obs[2], obs[1], state)
        return state
    if tag == "H":
        acc = np.zeros_like(state)
        for c, o in zip(obs[1], obs[2]):
            acc += float(c) * obs_apply(o, state.copy(), n)
        state[...] = acc
        return state
    raise ValueError(f"bad observable {obs!r}")


def expval_obs(state: np.ndarray, n: int, obs) -> float:
    """Measurements::expval(Observable) = Re <psi|O psi>
    (MeasurementsLQubit.hpp:372-394, calculateObsExpval :698-704)."""
    phi = obs_apply(obs, state.copy(), n)
    return float(np.real(np.vdot(state, phi)))


def _sparse_apply(indptr, indices, data, state: np.ndarray) -> np.ndarray:
    """apply_Sparse_Matrix (lq/utils/SparseLinAlg.hpp:42-63): row i of the
    CSR matrix times the state, rows over the full state index."""
    indptr = np.asarray(indptr, dtype=np.int64)
    indices = np.asarray(indices, dtype=np.int64)
    data = np.asarray(data, dtype=np.complex128)
    out = np.zeros(len(indptr) - 1, dtype=np.complex128)
    rows = np.repeat(np.arange(len(indptr) - 1), np.diff(indptr))
    np.add.at(out, rows, data * state[indices])
    return out


def sparse_to_dense(indptr, indices, data, dim: int) -> np.ndarray:
    """Row-major dense form of a CSR operator (duplicates summed)."""
    m = np.zeros((dim, dim), dtype=np.complex128)
    indptr = np.asarray(indptr, dtype=np.int64)
    for r in range(dim):
        for k in range(indptr[r], indptr[r + 1]):
            m[r, int(indices[k])] += data[k]
    return m


def apply_sparse_matrix(state: np.ndarray, n: int, indptr, indices, data, wires: Sequence[int],
                        ctrls: Sequence[int] = (), cvals: Sequence[int] = ()) -> np.ndarray:
    """applyControlledSparseMatrix / applySparseMatrix
    (GateImplementationsLM.hpp:516-574): the CSR rows over the local index of
    `wires` (wires[0] the MSB), on the control subspace; no inverse."""
    m = sparse_to_dense(indptr, indices, data, 1 << len(wires))
    return apply_matrix(state, n, m, wires, ctrls, cvals)


def var_named(state: np.ndarray, n: int, name: str, wires: Sequence[int]) -> float:
    """Measurements::var(name, wires) (MeasurementsLQubit.hpp:461-479): the
    operation applied by name to a copy (Identity skipped),
    <O psi|O psi> - Re<psi|O psi>^2."""
    if name == "Identity":
        phi = state.copy()
    else:
        mat = _NAMED_OBS[name] if name in _NAMED_OBS else gate_matrix(name, (), False, len(wires))
        phi = apply_matrix(state.copy(), n, mat, wires)
    return float(np.real(np.vdot(phi, phi)) - np.real(np.vdot(state, phi)) ** 2)


def var_matrix(state: np.ndarray, n: int, mat: np.ndarray, wires: Sequence[int]) -> float:
    """Measurements::var(matrix, wires) (MeasurementsLQubit.hpp:488-506)."""
    phi = apply_matrix(state.copy(), n, np.asarray(mat, dtype=np.complex128), wires)
    return float(np.real(np.vdot(phi, phi)) - np.real(np.vdot(state, phi)) ** 2)


def var_obs(state: np.ndarray, n: int, obs) -> float:
    """Measurements::var(Observable) (MeasurementsLQubit.hpp:430-452,
    calculateObsVar :716-727): bra = O psi; <bra|bra> - Re<bra|psi>^2."""
    bra = obs_apply(obs, state.copy(), n)
    return float(np.real(np.vdot(bra, bra)) - np.real(np.vdot(bra, state)) ** 2)


def expval_sparse(state: np.ndarray, indptr, indices, data) -> float:
    """Measurements::expval(row_map, col_idx, values) (MeasurementsLQubit.hpp:323-337)."""
    if len(indptr) - 1 != state.size:
        raise ValueError("Statevector and Hamiltonian have incompatible sizes.")
    return float(np.real(np.vdot(state, _sparse_apply(indptr, indices, data, state))))


def var_sparse(state: np.ndarray, indptr, indices, data) -> float:
    """Measurements::var(row_map, col_idx, values) (MeasurementsLQubit.hpp:616-636)."""
    if len(indptr) - 1 != state.size:
        raise ValueError("Statevector and Hamiltonian have incompatible sizes.")
    v = _sparse_apply(indptr, indices, data, state)
    return float(np.real(np.vdot(v, v)) - np.real(np.vdot(v, state)) ** 2)


# Samples. The reference draws with std::mt19937 seeded by the Measurements
# seed (MeasurementsBase.hpp:77-84) through std::uniform_real_distribution
# (libstdc++'s generate_canonical: enough 32-bit draws for the mantissa,
# sum_k g_k 2^(32k) / 2^(32m), clamped below 1) and the alias table of
# DiscreteRandomVariable (MeasurementKernels.hpp:308-381). numpy's legacy
# MT19937 seeding is the same init_genrand as std::mt19937(seed).
class _StdMt19937:
    def __init__(self, seed: int):
        self._bg = np.random.MT19937(0)
        self._bg._legacy_seeding(int(seed) & 0xFFFFFFFF)

    def draw(self) -> int:
        return int(self._bg.random_raw())

    def canonical(self, single: bool) -> float:
        if single:  # float: one draw, rounded to float, / 2^32
            v = np.float32(self.draw()) / np.float32(2.0 ** 32)
            return float(np.nextafter(np.float32(1), np.float32(0))) if v >= 1 else float(v)
        s = float(self.draw()) + float(self.draw()) * 2.0 ** 32
        v = s / 2.0 ** 64
        return np.nextafter(1.0, 0.0) if v >= 1 else v


def _alias_table(p: np.ndarray, single: bool):
    nprob = len(p)
    first = [float(np.float32(nprob) * np.float32(x)) if single else float(nprob * float(x)) for x in p]
    second = [-1] * nprob
    under = [i for i in range(nprob) if first[i] < 1.0]
    over = [i for i in range(nprob) if first[i] >= 1.0]
    while under and over:
        i = over.pop()
        j = under.pop()
        second[j] = i
        first[i] += first[j] - 1.0
        (under if first[i] < 1.0 else over).append(i)
    return first, second


def generate_samples(state: np.ndarray, n: int, wires: Sequence[int] | None, shots: int, seed: int,
                     single: bool = False) -> np.ndarray:
    """Measurements::generate_samples(wires, shots) with a set seed
    (MeasurementsLQubit.hpp:646-677): shots x len(wires) of 0/1, column j =
    wires[j]. ``single`` follows the complex64 engine (float probabilities
    and draws). Pure-Python loops: small cases only."""
    w = list(range(n)) if wires is None else list(wires)
    p = probs(state, n, w)
    if single:
        p = p.astype(np.float32)
    first, second = _alias_table(p, single)
    rng = _StdMt19937(seed)
    k = len(w)
    out = np.zeros((shots, k), dtype=np.uint64)
    nprob = len(p)
    for s in range(shots):
        u = rng.canonical(single)
        idx = int(np.float32(u) * np.float32(nprob)) if single else int(u * nprob)
        u2 = rng.canonical(single)
        if u2 >= first[idx] and second[idx] >= 0:
            idx = second[idx]
        for j in range(k):
            out[s, k - 1 - j] = (idx >> j) & 1
    return out


def obs_to_spec(obs) -> str:
    """Text form read by oracle/ref_driver.cpp's parse_obs."""
    tag = obs[0]
    if tag == "N":
        return f"N {obs[1]} {int(obs[2])}"
    if tag == "T":
        return f"T {len(obs[1])} " + " ".join(f"{nm} {int(w)}" for nm, w in obs[1])
    if tag == "M":
        m = np.asarray(obs[1], dtype=np.complex128).reshape(-1)
        ws = " ".join(str(int(w)) for w in obs[2])
        vals = " ".join(f"{v.real:.17g} {v.imag:.17g}" for v in m)
        return f"M {len(obs[2])} {ws} {vals}"
    if tag == "H":
        parts = [f"H {len(obs[1])}"]
        for c, o in zip(obs[1], obs[2]):
            parts.append(f"{float(c):.17g} {obs_to_spec(o)}")
        return " ".join(parts)
    raise ValueError(obs)


# ---------------------------------------------------------------------------
# Ops list and adjoint Jacobian (core/algorithms/JacobianData.hpp:39-253,
# lq/algorithms/AdjointJacobianLQubit.hpp:225-491)
# ---------------------------------------------------------------------------
@dataclass
class Op:
    name: str
    wires: list
    params: list = field(default_factory=list)
    inverse: bool = False
    matrix: np.ndarray | None = None
    ctrls: list = field(default_factory=list)
    cvals: list = field(default_factory=list)


def apply_ops(state: np.ndarray, n: int, ops: Sequence[Op], adj: bool = False) -> np.ndarray:
    """AdjointJacobianBase::applyOperations (core/algorithms/AdjointJacobianBase.hpp:75-100):
    named gate if known, else its matrix (lq/StateVectorLQubit.hpp:433-482)."""
    for op in ops:
        inv = bool(op.inverse) ^ adj
        if op.name in GATE_NPARAMS:
            apply_op(state, n, op.name, op.wires, inv, op.params, op.ctrls, op.cvals)
        else:
            apply_matrix(state, n, op.matrix, op.wires, op.ctrls, op.cvals, inverse=inv)
    return state


def _apply_op_adj(state: np.ndarray, n: int, op: Op) -> None:
    """applyOperationAdj (core/algorithms/AdjointJacobianBase.hpp:108-127)."""
    apply_ops(state, n, [op], adj=True)


def adjoint_jacobian(final_state: np.ndarray, n: int, ops: Sequence[Op],
                     observables: Sequence, trainable: Sequence[int]) -> np.ndarray:
    """AdjointJacobian::adjointJacobian with apply_operations=false.

    Backward sweep over ``ops`` (AdjointJacobianLQubit.hpp:394-486): for each
    op from last to first, if it is the next trainable parametric op,
    jac[o][t] = -2 s Im <H_o lambda | G lambda> with s negated for inverse ops
    (:85-93, :283-309); then lambda and every H_o lambda get U^dagger.
    ``trainable`` indexes parametric ops (ops with non-empty params). Returns
    [obs][param] row-major as a (len(obs), len(trainable)) array (:315-318).
    """
    tp = sorted(int(t) for t in trainable)
    nobs = len(observables)
    jac = np.zeros((nobs, len(tp)))
    if not tp:
        return jac
    lam = final_state.astype(np.complex128).copy()
    hlam = [obs_apply(o, lam.copy(), n) for o in observables]
    num_par = sum(1 for op in ops if len(op.params) > 0)
    cur = num_par - 1
    tnum = len(tp) - 1
    tp_it = len(tp) - 1
    for op in reversed(ops):
        if len(op.params) > 1:
            raise RuntimeError("The operation is not supported using the adjoint differentiation method")
        if op.name in ("StatePrep", "BasisState"):
            continue
        if tp_it < 0:
            break
        if len(op.params) > 0:
            if cur == tp[tp_it]:
                mu = lam.copy()
                s = apply_generator(mu, n, op.name, op.wires, op.ctrls, op.cvals)
                s = s * (-1.0 if op.inverse else 1.0)
                for o in range(nobs):
                    jac[o, tnum] = -2.0 * s * np.imag(np.vdot(hlam[o], mu))
                tnum -= 1
                tp_it -= 1
            cur -= 1
        _apply_op_adj(lam, n, op)
        for h in hlam:
            _apply_op_adj(h, n, op)
    return jac


def vector_jacobian_product(final_state: np.ndarray, n: int, ops: Sequence[Op], dy: np.ndarray,
                            trainable: Sequence[int]) -> np.ndarray:
    """VectorJacobianProduct::operator() with apply_operations=false
    (lq/algorithms/VectorJacobianProduct.hpp:84-157).

    mu starts as the state cotangent dy; for each op from the last, if it is
    the next trainable parametric op, w[t] = i s <G mu | lambda> (:128-146,
    s negated for inverse ops), then lambda and mu get U^dagger (:152-155).
    Returns len(trainable) complex values.
    """
    tp = sorted(int(t) for t in trainable)
    w = np.zeros(len(tp), dtype=np.complex128)
    if not tp:
        return w
    lam = final_state.astype(np.complex128).copy()
    mu = np.asarray(dy, dtype=np.complex128).copy()
    cur = sum(1 for op in ops if len(op.params) > 0) - 1
    tnum = len(tp) - 1
    for op in reversed(ops):
        if len(op.params) > 1:
            raise RuntimeError("The operation is not supported using the adjoint differentiation method")
        if op.name in ("StatePrep", "BasisState"):
            continue
        if tnum < 0:
            break
        if len(op.params) > 0:
            if cur == tp[tnum]:
                mu_d = mu.copy()
                s = apply_generator(mu_d, n, op.name, op.wires, op.ctrls, op.cvals)
                s = s * (-1.0 if op.inverse else 1.0)
                w[tnum] = 1j * s * np.vdot(mu_d, lam)
                tnum -= 1
            cur -= 1
        _apply_op_adj(lam, n, op)
        _apply_op_adj(mu, n, op)
    return w


# ---------------------------------------------------------------------------
# Seeded circuits of the measured workloads (identical ops lists are fed to the
# reference, the oracle and the engine)
# ---------------------------------------------------------------------------
def circuit_layers_rxyz_cnot(n: int = 20, layers: int = 10, seed: int = 0) -> list[Op]:
    """RX,RY,RZ on every wire (angles U[0,2pi), seed) then CNOT(i,i+1) ladder."""
    rng = np.random.default_rng(seed)
    ang = rng.uniform(0.0, 2 * math.pi, size=(layers, n, 3))
    ops: list[Op] = []
    for layer in range(layers):
        for w in range(n):
            for g, name in enumerate(("RX", "RY", "RZ")):
                ops.append(Op(name, [w], [float(ang[layer, w, g])]))
        for w in range(n - 1):
            ops.append(Op("CNOT", [w, w + 1]))
    return ops


def circuit_hea_rot_cz(n: int = 30, layers: int = 20, seed: int = 1) -> list[Op]:
    """Hardware-efficient ansatz: Rot(phi,theta,omega) on every wire, then CZ
    on (i,i+1) for even i, then for odd i."""
    rng = np.random.default_rng(seed)
    ang = rng.uniform(0.0, 2 * math.pi, size=(layers, n, 3))
    ops: list[Op] = []
    for layer in range(layers):
        for w in range(n):
            ops.append(Op("Rot", [w], [float(x) for x in ang[layer, w]]))
        for w in range(0, n - 1, 2):
            ops.append(Op("CZ", [w, w + 1]))
        for w in range(1, n - 1, 2):
            ops.append(Op("CZ", [w, w + 1]))
    return ops


def circuit_strongly_entangling(n: int = 24, layers: int = 6, seed: int = 2) -> list[Op]:
    """Strongly entangling layers in the adjoint-ready form: each Rot is
    decomposed as RZ(phi) RY(theta) RZ(omega) (the serializer's rule,
    pennylane_lightning/lightning_base/_serialize.py:526-529), followed by
    CNOT(i, (i+r) mod n) with r = layer mod (n-1) + 1."""
    rng = np.random.default_rng(seed)
    ang = rng.uniform(0.0, 2 * math.pi, size=(layers, n, 3))
    ops: list[Op] = []
    for layer in range(layers):
        for w in range(n):
            phi, theta, omega = (float(x) for x in ang[layer, w])
            ops.append(Op("RZ", [w], [phi]))
            ops.append(Op("RY", [w], [theta]))
            ops.append(Op("RZ", [w], [omega]))
        r = layer % (n - 1) + 1
        for w in range(n):
            ops.append(Op("CNOT", [w, (w + r) % n]))
    return ops


def random_pauli_hamiltonian(n: int, terms: int, seed: int,
                             probs_ixyz=(0.7, 0.1, 0.1, 0.1)):
    """Random Pauli sum: each wire I/X/Y/Z with the given law, coefficients
    N(0,1). Returned as ("H", coeffs, [("T", [(name, wire), ...]), ...]), the
    shape the serializer emits for a Pauli sentence (_serialize.py:358-375)."""
    rng = np.random.default_rng(seed)
    names = ("Identity", "PauliX", "PauliY", "PauliZ")
    coeffs = rng.standard_normal(terms)
    obs = []
    for _ in range(terms):
        letters = rng.choice(4, size=n, p=list(probs_ixyz))
        factors = [(names[c], w) for w, c in enumerate(letters) if c != 0]
        if not factors:
            factors = [("Identity", 0)]
        obs.append(("T", factors))
    return ("H", [float(c) for c in coeffs], obs)
