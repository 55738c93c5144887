"""Tape-level adjoint Jacobian on the device (lightning_amd/adjoint_jacobian.py):
final state after a leading StatePrep / BasisState, Jacobian columns for
every trainable tape parameter (zero for the state preparation's, as
pennylane_lightning/lightning_base/_adjoint_jacobian.py:125-143 gives), split
Hamiltonians summed back to one row, and the vector-Jacobian product.

Checked against the oracle's adjoint Jacobian on the serialized operations
(complex128 within 1e-10; complex64 within 5e-4) and, independently of any
adjoint code, against central finite differences of the oracle's expectation
values (h = 1e-6, within 1e-7)."""
import numpy as np
import pytest

from helpers import rand_state, to_oracle_ops
from oracle import lq_oracle as O

pytestmark = pytest.mark.gpu

A = pytest.importorskip("lightning_amd.adjoint_jacobian")
W = pytest.importorskip("lightning_amd.workloads")

N = 6


def build(prep, trainable):
    rng = np.random.default_rng(11)
    op = W.CircuitOp
    ops = []
    if prep == "state":
        ops.append(op("StatePrep", list(range(N)), [rand_state(N, rng)]))
    elif prep == "basis":
        ops.append(op("BasisState", [0, 2, 3], [np.array([1, 0, 1])]))
    ops += [op("RX", [0], [0.3]), op("Rot", [1], [0.4, 1.1, -0.7]), op("CNOT", [0, 1]),
            op("QubitUnitary", [2, 3], matrix=W.haar_unitary(4, rng)), op("CRY", [4, 5], [0.9]),
            op("IsingXX", [1, 4], [0.35]), op("PhaseShift", [3], [0.8], inverse=True), op("RY", [5], [-1.2]),
            op("CZ", [2, 5]), op("RZ", [2], [0.25])]
    coeffs, words = W.random_pauli_hamiltonian(N, 9, seed=5, law=(0.4, 0.2, 0.2, 0.2))
    ham = ("H", list(coeffs), [("T", list(w)) for w in words])
    meas = [A.expval(ham), A.expval(("N", "PauliZ", 3)), A.expval(("T", [("PauliX", 0), ("PauliY", 5)]))]
    return A.QuantumScript(ops, meas, trainable_params=trainable)


def oracle_initial(tape):
    first = tape.operations[0]
    if first.name == "StatePrep":
        return np.asarray(first.params[0], np.complex128).copy(), tape.operations[1:]
    s = O.zero_state(N)
    if first.name == "BasisState":
        idx = 0
        bits = dict(zip(first.wires, first.params[0]))
        for w in range(N):
            idx = (idx << 1) | int(bits.get(w, 0))
        s[:] = 0
        s[idx] = 1.0
        return s, tape.operations[1:]
    return s, tape.operations


def oracle_expvals(tape, ops_override=None):
    s, ops = oracle_initial(tape)
    st = O.apply_ops(s, N, to_oracle_ops(ops_override if ops_override is not None else ops))
    return st, np.array([O.expval_obs(st, N, m.obs) for m in tape.measurements])


def finite_differences(tape, h=1e-6):
    """d<O_k>/d(tape parameter p) for every trainable gate parameter (state
    preparation columns stay 0), from the oracle's expectation values."""
    _, rest = oracle_initial(tape)
    offset = 1 if rest is not tape.operations else 0
    out = np.zeros((len(tape.measurements), len(tape.trainable_params)))
    for col, p in enumerate(tape.trainable_params):
        o, k, j = tape.get_operation(col)
        if o.name in ("StatePrep", "BasisState"):
            continue
        vals = []
        for sgn in (1.0, -1.0):
            ops = [W.CircuitOp(x.name, x.wires, list(x.params), x.inverse, x.matrix, x.ctrls, x.cvals)
                   for x in rest]
            ops[k - offset].params[j] += sgn * h
            vals.append(oracle_expvals(tape, ops)[1])
        out[:, col] = (vals[0] - vals[1]) / (2 * h)
    return out


def oracle_jacobian(tape):
    """The oracle's adjoint Jacobian on the serialized operation list, placed
    into the tape's columns."""
    st, _ = oracle_expvals(tape)
    ser = []
    for o in tape.operations:
        if o.name in ("StatePrep", "BasisState"):
            continue
        if o.name == "Rot":
            ser += [W.CircuitOp(nm, o.wires, [p]) for nm, p in zip(("RZ", "RY", "RZ"), o.params)]
        elif o.matrix is not None:
            ser.append(W.CircuitOp("QubitUnitary", o.wires, [0.0], o.inverse, o.matrix))
        else:
            ser.append(o)
    shift = 1 if tape.operations[0].name in ("StatePrep", "BasisState") else 0
    cols, tp = [], []
    for col, p in enumerate(tape.trainable_params):
        if tape.get_operation(col)[0].name not in ("StatePrep", "BasisState"):
            cols.append(col)
            tp.append(p - shift)
    out = np.zeros((len(tape.measurements), len(tape.trainable_params)))
    if tp:
        jac = O.adjoint_jacobian(st, N, to_oracle_ops(ser), [m.obs for m in tape.measurements], tp)
        out[:, cols] = np.asarray(jac).reshape(len(tape.measurements), len(tp))
    return out


# tape parameters (with a state preparation): 0 the state, 1 RX, 2-4 Rot, 5 QubitUnitary (not
# differentiable), 6 CRY, 7 IsingXX, 8 PhaseShift, 9 RY, 10 RZ; one less without it
CASES = [("none", [0, 1, 2, 3, 5, 6, 7, 8, 9]), ("state", [0, 1, 2, 3, 4, 6, 7, 9, 10]), ("state", [0, 4, 9]),
         ("basis", [1, 2, 4, 8, 10]), ("basis", [0, 3, 6, 7, 9])]


@pytest.mark.parametrize("bits", [128, 64])
@pytest.mark.parametrize("prep,trainable", CASES)
def test_tape_jacobian(bits, prep, trainable):
    tape = build(prep, trainable)
    dtype = np.complex128 if bits == 128 else np.complex64
    sv = A.LightningStateVector(N, dtype).get_final_state(tape)
    st, _ = oracle_expvals(tape)
    tol = 1e-10 if bits == 128 else 5e-4
    np.testing.assert_allclose(sv.state, st, atol=1e-12 if bits == 128 else 5e-6)
    jac = np.array(A.LightningAdjointJacobian(sv).calculate_jacobian(tape), dtype=np.float64)
    jac = jac.reshape(len(tape.measurements), len(tape.trainable_params))
    np.testing.assert_allclose(jac, oracle_jacobian(tape), atol=tol)
    if bits == 128:
        np.testing.assert_allclose(jac, finite_differences(tape), atol=1e-7)


def test_split_hamiltonian_matches_unsplit():
    tape = build("state", [1, 2, 3, 6, 10])
    sv = A.LightningStateVector(N).get_final_state(tape)
    one = np.array(A.LightningAdjointJacobian(sv).calculate_jacobian(tape), dtype=np.float64)
    split = np.array(A.LightningAdjointJacobian(sv, batch_obs=3).calculate_jacobian(tape), dtype=np.float64)
    np.testing.assert_allclose(split, one, atol=1e-10)


def test_vjp_against_jacobian():
    tape = build("basis", [1, 2, 4, 8, 10])
    sv = A.LightningStateVector(N).get_final_state(tape)
    g = A.LightningAdjointJacobian(sv)
    jac = np.array(g.calculate_jacobian(tape), dtype=np.float64)
    dy = np.array([0.7, -1.3, 0.4])
    np.testing.assert_allclose(np.array(g.calculate_vjp(tape, dy), dtype=np.float64), dy @ jac, atol=1e-10)


def test_trainable_matrix_op_is_rejected():
    tape = build("state", [1, 5])
    sv = A.LightningStateVector(N).get_final_state(tape)
    from lightning_amd._capi import LightningException
    with pytest.raises(LightningException, match="not supported using the adjoint differentiation method"):
        A.LightningAdjointJacobian(sv).calculate_jacobian(tape)
