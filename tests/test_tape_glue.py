"""Host logic of the tape-level adjoint glue (lightning_amd/adjoint_jacobian.py),
on the CPU: parameter bookkeeping around state preparations
(pennylane_lightning/lightning_base/_adjoint_jacobian.py:91-152), operation
serialization (_serialize.py:522-563), Hamiltonian splitting
(_serialize.py:273-285), error paths (_adjoint_jacobian.py:171-209) and the
row reduction / column placement of the Jacobian
(lightning_qubit/_adjoint_jacobian.py:103-123). The device sweep is replaced by a
stand-in that records its inputs and returns a known matrix; the GPU tests of
the same module compare whole Jacobians with the oracle and with finite
differences."""
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "pennylane-lightning_amd", "lightning_amd", "_lib", "liblq_amd.so")
pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="liblq_amd.so not built")


@pytest.fixture(scope="module")
def A():
    from lightning_amd import adjoint_jacobian
    return adjoint_jacobian


def op(name, wires, params=(), **kw):
    from lightning_amd.workloads import CircuitOp
    return CircuitOp(name, list(wires), list(params), **kw)


def tape_with_stateprep(A, trainable):
    psi = np.ones(8, np.complex128) / np.sqrt(8)
    ops = [op("StatePrep", [0, 1, 2], [psi]),                  # parameter 0
           op("RX", [0], [0.1]),                                # 1
           op("CNOT", [0, 1]),
           op("QubitUnitary", [1], matrix=np.eye(2)),           # 2
           op("Rot", [2], [0.2, 0.3, 0.4]),                     # 3, 4, 5
           op("RZ", [1], [0.5])]                                # 6
    return A.QuantumScript(ops, [A.expval(("N", "PauliZ", 0))], trainable_params=trainable)


class _StubState:
    _bits = 128
    dtype = np.dtype(np.complex128)
    state_vector = "sv"


class _StubSweep:
    def __init__(self, raw):
        self.raw = raw
        self.calls = []

    def __call__(self, sv, obs, ops, tp):
        self.calls.append((sv, list(obs), ops, list(tp)))
        return self.raw


def glue(A, raw=None, batch_obs=False):
    g = A.LightningAdjointJacobian.__new__(A.LightningAdjointJacobian)
    g._qubit_state = _StubState()
    g._batch_obs = batch_obs
    g._jacobian_lightning = _StubSweep(raw)
    return g


def test_get_operation_counts_state_preps_and_matrices(A):
    t = tape_with_stateprep(A, None)
    assert t.num_params == 7 and t.trainable_params == list(range(7))
    got = [(o.name, k, p) for o, k, p in (t.get_operation(i) for i in range(7))]
    assert got == [("StatePrep", 0, 0), ("RX", 1, 0), ("QubitUnitary", 3, 0), ("Rot", 4, 0), ("Rot", 4, 1),
                   ("Rot", 4, 2), ("RZ", 5, 0)]
    with pytest.raises(ValueError):
        A.QuantumScript(t.operations, t.measurements, trainable_params=[7])


def test_trainable_bookkeeping_after_a_state_preparation(A):
    # trainable tape parameters 0 (the state), 1 (RX), 4 (Rot theta), 6 (RZ): the state's column stays
    # zero; the others map to operation parameters 0, 3, 5 of the serialized list (RX, RZ RY RZ, RZ
    # without the state preparation) and to Jacobian columns 1, 2, 3
    t = tape_with_stateprep(A, [6, 0, 4, 1])
    d = glue(A)._process_jacobian_tape(t)
    assert d["tp_shift"] == [0, 3, 5]
    assert d["record_tp_rows"] == [1, 2, 3]
    assert d["all_params"] == 4
    assert d["obs_indices"] == [0]
    ops = d["ops_serialized"]
    assert ops.names == ["RX", "CNOT", "QubitUnitary", "RZ", "RY", "RZ", "RZ"]
    assert ops.params == [[0.1], [], [0.0], [0.2], [0.3], [0.4], [0.5]]


def test_only_state_preparation_trainable_gives_zero_columns(A):
    t = tape_with_stateprep(A, [0])
    g = glue(A)
    d = g._process_jacobian_tape(t)
    assert d["tp_shift"] == [] and d["record_tp_rows"] == [] and d["all_params"] == 1
    out = g.calculate_jacobian(t)
    assert np.asarray(out).shape == () and float(out) == 0.0
    assert g._jacobian_lightning.calls == []   # nothing to sweep


def test_columns_and_split_rows(A):
    # two measurements, the first a 5-term Hamiltonian split three ways: serialized rows 0,1,2 -> 0, row 3 -> 1
    ham = ("H", [1.0, 2.0, 3.0, 4.0, 5.0],
           [("T", [("PauliZ", 0), ("PauliZ", 1)]), ("N", "PauliX", 2), ("T", [("PauliY", 0), ("PauliY", 2)]),
            ("N", "PauliZ", 1), ("M", np.eye(2), [0])])
    ops = [op("RX", [0], [0.1]), op("RY", [1], [0.2]), op("RZ", [2], [0.3])]
    t = A.QuantumScript(ops, [A.expval(ham), A.expval(("N", "PauliZ", 2))], trainable_params=[0, 2])
    raw = np.arange(8, dtype=np.float64)          # 4 serialized observables x 2 trainable
    g = glue(A, raw, batch_obs=3)
    jac = g.calculate_jacobian(t)
    sv, obs, _, tp = g._jacobian_lightning.calls[0]
    assert sv == "sv" and len(obs) == 4 and tp == [0, 2]
    r = raw.reshape(4, 2)
    want = np.array([r[0] + r[1] + r[2], r[3]])
    np.testing.assert_array_equal(np.array(jac, dtype=float), want)


def test_chunk_hamiltonian_orders_by_wire_count_then_deals_round_robin(A):
    terms = [("T", [("PauliZ", 0), ("PauliZ", 1)]), ("N", "PauliX", 2), ("T", [("PauliY", 0), ("PauliY", 2)]),
             ("N", "PauliZ", 1), ("M", np.eye(4), [0, 1])]
    parts = A.chunk_hamiltonian(("H", [1.0, 2.0, 3.0, 4.0, 5.0], terms), 3)
    # wire counts 2,1,2,1,2 -> stable order: terms 1, 3, 0, 2, 4 (coefficients 2, 4, 1, 3, 5)
    assert [p[1] for p in parts] == [[2.0, 3.0], [4.0, 5.0], [1.0]]
    assert A.chunk_hamiltonian(("H", [1.0], [terms[1]]), 3) == [("H", [1.0], [terms[1]])]


def test_error_paths(A):
    ops = [op("RX", [0], [0.1])]
    g = glue(A)
    with pytest.raises(A.QuantumFunctionError, match="finite shots"):
        g.calculate_jacobian(A.QuantumScript(ops, [A.expval(("N", "PauliZ", 0))], shots=100))
    with pytest.raises(A.QuantumFunctionError, match="StateMP"):
        g.calculate_jacobian(A.QuantumScript(ops, [A.state()]))
    with pytest.raises(A.QuantumFunctionError, match="mixed with other"):
        g.calculate_jacobian(A.QuantumScript(ops, [A.expval(("N", "PauliZ", 0)), A.probs([0])]))
    assert g.calculate_jacobian(A.QuantumScript(ops, [])).size == 0
    assert g.calculate_jacobian(A.QuantumScript(ops, [A.expval(("N", "PauliZ", 0))],
                                                trainable_params=[])).size == 0
    t = A.QuantumScript(ops, [A.expval(("N", "PauliZ", 0)), A.expval(("N", "PauliX", 0))])
    with pytest.raises(ValueError, match="length of grad_vec"):
        g.calculate_vjp(t, [1.0])
    with pytest.raises(ValueError, match="real-valued"):
        g.calculate_vjp(t, [1.0 + 1j, 0.0])
    with pytest.raises(A.QuantumFunctionError, match="State measurements"):
        g.calculate_vjp(A.QuantumScript(ops, [A.state()]), [1.0])
    np.testing.assert_array_equal(g.calculate_vjp(t, [0.0, 0.0]), [0.0])


def test_vjp_folds_the_cotangent_into_one_hamiltonian(A):
    ops = [op("RX", [0], [0.1]), op("RY", [1], [0.2])]
    t = A.QuantumScript(ops, [A.expval(("N", "PauliZ", 0)),
                              A.expval(("H", [0.5], [("N", "PauliX", 1)]))])
    g = glue(A, np.array([3.0, 4.0]))
    out = g.calculate_vjp(t, [2.0, -1.0])
    _, obs, _, tp = g._jacobian_lightning.calls[0]
    assert len(obs) == 1 and tp == [0, 1]
    assert obs[0].get_coeffs() == [2.0, -0.5]
    np.testing.assert_array_equal(np.array(out, dtype=float), [3.0, 4.0])


def test_processing_shapes(A):
    P = A.LightningAdjointJacobian._adjoint_jacobian_processing
    assert np.asarray(P(np.array([[1.5]]))).shape == ()
    assert len(P(np.array([[1.0, 2.0]]))) == 2
    two = P(np.arange(6.0).reshape(2, 3))
    assert len(two) == 2 and all(len(r) == 3 for r in two)


def test_single_unit_term_hamiltonian_is_the_term(A):
    from lightning_amd import lightning_qubit_ops as L
    o = A.observable(("H", [1.0], [("N", "PauliZ", 2)]))
    assert isinstance(o, L.NamedObsC128)
    h = A.observable(("H", [0.5], [("N", "PauliZ", 2)]))
    assert isinstance(h, L.HamiltonianC128)
