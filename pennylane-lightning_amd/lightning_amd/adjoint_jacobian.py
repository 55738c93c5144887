"""Tape-level adjoint Jacobian: the Python layer between a circuit description
and ``AdjointJacobianC*``.

Mirrors the reference's device glue for the adjoint method
(pennylane_lightning/lightning_base/_adjoint_jacobian.py:44-288 and
pennylane_lightning/lightning_qubit/_adjoint_jacobian.py:40-123), with the
serializer steps it relies on (pennylane_lightning/lightning_base/_serialize.py:
273-285, 402-563) and the final-state preparation
(pennylane_lightning/lightning_base/_state_vector.py:105-260):

* ``QuantumScript`` holds what the glue reads from a PennyLane tape:
  ``operations`` (``CircuitOp`` records), ``measurements`` (``expval``,
  ``var``, ``probs``, ``state`` records), ``trainable_params`` (indices into
  the operations' parameters, all by default) and ``shots``. A
  ``StatePrep`` / ``BasisState`` first op takes the state / the basis bits as
  its one parameter; a matrix op (``QubitUnitary``) counts its matrix as one
  parameter, as PennyLane does.
* ``LightningStateVector.get_final_state(tape)`` prepares the state (a
  leading state preparation through ``setStateVector`` / ``setBasisState``)
  and applies the rest of the operations.
* ``LightningAdjointJacobian.calculate_jacobian(tape)`` serializes the
  operations (state preparations left out, a plain ``Rot`` as RZ RY RZ, a
  matrix op with a dummy 0.0 parameter) and the observables, maps the
  tape's trainable parameters to operation parameters (``tp_shift``: one
  less after a leading state preparation) and to Jacobian columns
  (``record_tp_rows``: state-preparation parameters get zero columns), runs
  the adjoint sweep on the device and sums the rows of split observables
  back to one row per measurement. ``calculate_vjp(tape, dy)`` folds the
  cotangent into one Hamiltonian ``sum_k dy_k O_k`` and differentiates that.

Observables are spec tuples: ``("N", name, wire)``, ``("T", [(name, wire),
...])``, ``("M", matrix, wires)`` (Hermitian) and ``("H", coeffs, [terms])``.
Parameters of observables are not tape parameters here.

With ``batch_obs`` set to an integer k > 1, a Hamiltonian measured together
with other measurements is split into k sub-Hamiltonians (terms sorted by
wire count and dealt out round-robin, as the reference's
``_chunk_ham_terms``); on the device this costs more than it saves (each
part is one more state carried through the backward sweep), so the default
is no split.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import lightning_qubit_ops as L
from .workloads import CircuitOp

try:  # the PennyLane exception type when PennyLane is installed
    from pennylane import QuantumFunctionError  # type: ignore
except Exception:  # pragma: no cover - PennyLane is optional
    class QuantumFunctionError(Exception):
        """Raised for tapes the adjoint method cannot differentiate."""

_STATE_PREPS = ("StatePrep", "BasisState")


# ---------------------------------------------------------------- measurements

@dataclass(frozen=True)
class Measurement:
    kind: str  # "expval" | "var" | "probs" | "state"
    obs: tuple | None = None
    wires: tuple | None = None


def expval(obs) -> Measurement:
    return Measurement("expval", obs=obs)


def var(obs) -> Measurement:
    return Measurement("var", obs=obs)


def probs(wires=None) -> Measurement:
    return Measurement("probs", wires=None if wires is None else tuple(int(w) for w in wires))


def state() -> Measurement:
    return Measurement("state")


# ---------------------------------------------------------------------- tape

def op_num_params(op: CircuitOp) -> int:
    """Parameters an operation contributes to its tape: a state preparation
    and a matrix op have one (the state, the basis bits, the matrix)."""
    if op.name in _STATE_PREPS or op.matrix is not None:
        return 1
    return len(op.params)


@dataclass
class QuantumScript:
    operations: list
    measurements: list
    trainable_params: list | None = None
    shots: int | None = None
    _trainable: list = field(init=False, repr=False)

    def __post_init__(self):
        n = self.num_params
        tp = list(range(n)) if self.trainable_params is None else sorted(int(t) for t in self.trainable_params)
        if any(t < 0 or t >= n for t in tp) or len(set(tp)) != len(tp):
            raise ValueError(f"trainable parameter indices must be distinct and in [0, {n})")
        self._trainable = tp
        self.trainable_params = tp

    @property
    def num_params(self) -> int:
        return sum(op_num_params(op) for op in self.operations)

    @property
    def observables(self) -> list:
        return [m.obs for m in self.measurements]

    def get_operation(self, idx: int):
        """(operation, its index in ``operations``, parameter index within it)
        of the ``idx``-th trainable parameter."""
        p = self._trainable[idx]
        for k, op in enumerate(self.operations):
            c = op_num_params(op)
            if p < c:
                return op, k, p
            p -= c
        raise IndexError(idx)


# -------------------------------------------------------------- observables

def observable(spec, bits: int = 128):
    """Binding observable (``NamedObsC*`` ...) of a spec tuple."""
    sfx = f"C{bits}"
    tag = spec[0]
    if tag == "N":
        return getattr(L, "NamedObs" + sfx)(spec[1], [int(spec[2])])
    if tag == "T":
        return getattr(L, "TensorProdObs" + sfx)([getattr(L, "NamedObs" + sfx)(nm, [int(w)]) for nm, w in spec[1]])
    if tag == "M":
        return getattr(L, "HermitianObs" + sfx)(np.asarray(spec[1]), [int(w) for w in spec[2]])
    if tag == "H":
        if len(spec[2]) == 1 and float(spec[1][0]) == 1.0:
            # a single unit-coefficient term is the term itself (_serialize.py:372-373)
            return observable(spec[2][0], bits)
        return getattr(L, "Hamiltonian" + sfx)([float(c) for c in spec[1]], [observable(t, bits) for t in spec[2]])
    raise ValueError(f"bad observable {spec!r}")


def spec_wires(spec) -> list:
    tag = spec[0]
    if tag == "N":
        return [int(spec[2])]
    if tag == "T":
        return [int(w) for _, w in spec[1]]
    if tag == "M":
        return [int(w) for w in spec[2]]
    if tag == "H":
        out = []
        for t in spec[2]:
            out += [w for w in spec_wires(t) if w not in out]
        return out
    raise ValueError(f"bad observable {spec!r}")


def weighted_terms(spec, w: float) -> list:
    """``w * spec`` as (coefficient, term) pairs with no Hamiltonian term."""
    if spec[0] == "H":
        out = []
        for c, t in zip(spec[1], spec[2]):
            out += weighted_terms(t, w * float(c))
        return out
    return [(w, spec)]


def cotangent_observable(observables, dy):
    """``sum_k dy_k O_k`` as one Hamiltonian spec (None when every dy_k is 0)."""
    terms = []
    for spec, w in zip(observables, dy):
        if w != 0.0:
            terms += weighted_terms(spec, float(w))
    if not terms:
        return None
    return ("H", [c for c, _ in terms], [t for _, t in terms])


def chunk_hamiltonian(spec, parts: int) -> list:
    """Split a Hamiltonian into at most ``parts`` sub-Hamiltonians: terms
    ordered by wire count (stable), then dealt out round-robin."""
    coeffs, terms = list(spec[1]), list(spec[2])
    order = np.argsort(np.array([len(spec_wires(t)) for t in terms]), kind="stable")
    coeffs = [coeffs[i] for i in order]
    terms = [terms[i] for i in order]
    k = min(len(terms), parts)
    return [("H", coeffs[i::parts], terms[i::parts]) for i in range(k)]


# ------------------------------------------------------------- serialization

def serialize_ops(tape: QuantumScript, bits: int = 128):
    """(``OpsStructC*``, uses a state preparation). State preparations are
    left out; a plain ``Rot`` becomes RZ(phi) RY(theta) RZ(omega); a matrix op
    carries a dummy 0.0 parameter so parameter positions match the tape."""
    names, params, wires, inverses, mats, ctrls, cvals = [], [], [], [], [], [], []
    uses_stateprep = False
    for op in tape.operations:
        if op.name in _STATE_PREPS:
            uses_stateprep = True
            continue
        if op.name == "Rot" and op.matrix is None and not op.inverse and not op.ctrls:
            phi, theta, omega = op.params
            seq = [("RZ", [phi]), ("RY", [theta]), ("RZ", [omega])]
            for nm, p in seq:
                names.append(nm)
                params.append(p)
                wires.append(list(op.wires))
                inverses.append(False)
                mats.append([])
                ctrls.append([])
                cvals.append([])
            continue
        names.append(op.name if op.matrix is None else "QubitUnitary")
        params.append([0.0] if op.matrix is not None else list(op.params))
        wires.append(list(op.wires))
        inverses.append(bool(op.inverse))
        mats.append(op.matrix if op.matrix is not None else [])
        ctrls.append(list(op.ctrls))
        cvals.append(list(op.cvals) if op.cvals else [True] * len(op.ctrls))
    cls = L.OpsStructC128 if bits == 128 else L.OpsStructC64
    return cls(names, params, wires, inverses, mats, ctrls, cvals), uses_stateprep


def serialize_observables(tape: QuantumScript, bits: int = 128, split_obs=False):
    """(binding observables, measurement index of each). With ``split_obs``
    an integer, each Hamiltonian becomes that many sub-Hamiltonians."""
    out, idx = [], []
    for i, spec in enumerate(tape.observables):
        if split_obs and spec[0] == "H":
            parts = chunk_hamiltonian(spec, int(split_obs))
            out += [observable(p, bits) for p in parts]
            idx += [i] * len(parts)
        else:
            out.append(observable(spec, bits))
            idx.append(i)
    return out, idx


# ---------------------------------------------------------- state and sweep

class LightningStateVector:
    """The device state of one tape: |0...0>, a leading state preparation,
    then the operations."""

    def __init__(self, num_wires: int, dtype=np.complex128):
        self._dtype = np.dtype(dtype)
        if self._dtype not in (np.dtype(np.complex64), np.dtype(np.complex128)):
            raise TypeError(f"Unsupported complex type: {dtype}")
        self._bits = 64 if self._dtype == np.complex64 else 128
        self._num_wires = int(num_wires)
        self._qubit_state = (L.StateVectorC64 if self._bits == 64 else L.StateVectorC128)(self._num_wires)

    @property
    def dtype(self):
        return self._dtype

    @property
    def num_wires(self) -> int:
        return self._num_wires

    @property
    def state_vector(self):
        return self._qubit_state

    @property
    def state(self) -> np.ndarray:
        return self._qubit_state.state()

    def reset_state(self):
        self._qubit_state.resetStateVector()

    def _apply_basis_state(self, bits, wires):
        bits = np.asarray(bits).reshape(-1)
        if not set(bits.tolist()).issubset({0, 1}):
            raise ValueError("BasisState parameter must consist of 0 or 1 integers.")
        if len(bits) != len(wires):
            raise ValueError("BasisState parameter and wires must be of equal length.")
        self._qubit_state.setBasisState([int(b) for b in bits], list(wires))

    def apply_operations(self, operations):
        from .workloads import apply_circuit
        ops = list(operations)
        if ops and ops[0].name == "StatePrep":
            self._qubit_state.setStateVector(np.asarray(ops[0].params[0], dtype=self._dtype), list(ops[0].wires))
            ops = ops[1:]
        elif ops and ops[0].name == "BasisState":
            self._apply_basis_state(ops[0].params[0], ops[0].wires)
            ops = ops[1:]
        if any(op.name in _STATE_PREPS for op in ops):
            raise ValueError("a state preparation must be the first operation")
        apply_circuit(self._qubit_state, ops)

    def get_final_state(self, circuit: QuantumScript):
        self.apply_operations(circuit.operations)
        return self


class LightningAdjointJacobian:
    """Adjoint Jacobian and vector-Jacobian product of a tape on a prepared
    ``LightningStateVector`` (its final state)."""

    def __init__(self, qubit_state: LightningStateVector, batch_obs=False):
        self._qubit_state = qubit_state
        self._batch_obs = batch_obs
        bits = qubit_state._bits
        self._jacobian_lightning = (L.algorithms.AdjointJacobianC64 if bits == 64 else
                                    L.algorithms.AdjointJacobianC128)()

    @property
    def qubit_state(self):
        return self._qubit_state

    @property
    def state(self):
        return self._qubit_state.state_vector

    @property
    def dtype(self):
        return self._qubit_state.dtype

    @staticmethod
    def _get_return_type(measurements):
        if not measurements:
            return None
        if len(measurements) == 1 and measurements[0].kind == "state":
            return "state"
        return "expval"

    def _handle_raises(self, tape: QuantumScript, is_jacobian: bool, grad_vec=None) -> bool:
        """True when the result is empty; raises for what the method cannot do."""
        if tape.shots:
            raise QuantumFunctionError(
                "Requested adjoint differentiation to be computed with finite shots. "
                "The derivative is always exact when using the adjoint differentiation method.")
        rtype = self._get_return_type(tape.measurements)
        if is_jacobian:
            if not rtype:
                return True
            if rtype == "state":
                raise QuantumFunctionError("Adjoint differentiation method does not support measurement StateMP.")
        else:
            if not rtype or np.allclose(np.asarray(grad_vec, dtype=np.complex128), 0.0):
                return True
            if rtype == "state":
                raise QuantumFunctionError("Adjoint differentiation does not support State measurements.")
        if any(m.kind != "expval" for m in tape.measurements):
            raise QuantumFunctionError(
                "Adjoint differentiation method does not support expectation return type mixed with other "
                "return types")
        return False

    def _process_jacobian_tape(self, tape: QuantumScript, split_obs=False):
        bits = self._qubit_state._bits
        obs_serialized, obs_indices = serialize_observables(tape, bits, split_obs)
        ops_serialized, use_sp = serialize_ops(tape, bits)
        trainable = sorted(tape.trainable_params)
        if not trainable:
            return None
        tp_shift, record_tp_rows = [], []
        for row, param in enumerate(trainable):
            op, _, _ = tape.get_operation(row)
            if op.name not in _STATE_PREPS:
                tp_shift.append(param)
                record_tp_rows.append(row)
        if use_sp:
            tp_shift = [t - 1 for t in tp_shift]
        return {"state_vector": self.state, "obs_serialized": obs_serialized, "ops_serialized": ops_serialized,
                "tp_shift": tp_shift, "record_tp_rows": record_tp_rows, "all_params": len(trainable),
                "obs_indices": obs_indices}

    @staticmethod
    def _adjoint_jacobian_processing(jac):
        """One measurement and one parameter -> 0-d array; one of either ->
        tuple of arrays; otherwise a tuple (per measurement) of tuples."""
        jac = np.squeeze(jac)
        if jac.ndim == 0:
            return np.array(jac)
        if jac.ndim == 1:
            return tuple(np.array(j) for j in jac)
        return tuple(tuple(np.array(j_) for j_ in j) for j in jac)

    def calculate_jacobian(self, tape: QuantumScript):
        if self._handle_raises(tape, is_jacobian=True):
            return np.array([], dtype=self.dtype)
        split_obs = False
        if len(tape.measurements) > 1 and self._batch_obs and int(self._batch_obs) > 1:
            split_obs = int(self._batch_obs)
        data = self._process_jacobian_tape(tape, split_obs=split_obs)
        if not data:
            return np.array([], dtype=self.dtype)
        tp = data["tp_shift"]
        rows = data["obs_indices"]
        nmeas = len(tape.measurements)
        jac = np.zeros((nmeas, len(tp)), dtype=np.float64)
        if tp:
            raw = np.asarray(self._jacobian_lightning(data["state_vector"], data["obs_serialized"],
                                                      data["ops_serialized"], tp), dtype=np.float64)
            np.add.at(jac, np.asarray(rows, dtype=np.int64), raw.reshape(len(rows), len(tp)))
        jac_r = np.zeros((nmeas, data["all_params"]), dtype=np.float64)
        jac_r[:, data["record_tp_rows"]] = jac
        return self._adjoint_jacobian_processing(jac_r)

    def calculate_vjp(self, tape: QuantumScript, grad_vec):
        if self._handle_raises(tape, is_jacobian=False, grad_vec=grad_vec):
            return np.zeros(len(tape.trainable_params))
        if np.ndim(grad_vec) == 0:
            grad_vec = (grad_vec,)
        if len(grad_vec) != len(tape.measurements):
            raise ValueError("Number of observables in the tape must be the same as the length of grad_vec in "
                             "the vjp method")
        if np.iscomplexobj(grad_vec):
            raise ValueError("The vjp method only works with a real-valued grad_vec when the tape is returning an "
                             "expectation value")
        if len(tape.trainable_params) == 0:
            return np.array([], dtype=self.dtype)
        ham = cotangent_observable(tape.observables, [float(g) for g in grad_vec])
        if ham is None:
            return np.zeros(len(tape.trainable_params))
        new_tape = QuantumScript(tape.operations, [expval(ham)], trainable_params=tape.trainable_params,
                                 shots=tape.shots)
        return self.calculate_jacobian(new_tape)
