"""Full-size checks through properties the mathematics fixes (no oracle run at
these sizes):

* 33 qubits, complex128 (128 GiB): the QFT of |0..0> is |+>^n, so every
  <X_i> is 1; the QFT + Haar brickwork + 3-controlled X circuit keeps the norm;
  applying the circuit's inverse (ops reversed, each inverted) returns to
  |+>^n. Tolerance 1e-10.
* 30 qubits: the complex64 run of the depth-20 HEA against the complex128 run:
  fidelity and expvals within the stated complex64 bounds (1e-5).
* 33 qubits: the complex64 run (64 GiB) of the QFT + Haar brickwork +
  3-controlled X circuit against the complex128 run (128 GiB, both resident):
  fidelity, every <Z_i> and the marginal probabilities of wires 0-9 within
  the stated complex64 bound of that configuration (1e-5).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

L = pytest.importorskip("lightning_amd.lightning_qubit_ops")
W = pytest.importorskip("lightning_amd.workloads")


def inverse_ops(ops):
    return [W.CircuitOp(o.name, o.wires, o.params, not o.inverse, o.matrix, o.ctrls, o.cvals) for o in reversed(ops)]


def test_qft33_round_trip():
    n = 33
    ops = W.qft_brickwork_mcx(n, layers=8, mcx_per_layer=1, seed=5)
    n_qft = n + n * (n - 1) // 2 + n // 2
    assert all(o.name in ("Hadamard", "ControlledPhaseShift", "SWAP") for o in ops[:n_qft])
    sv = L.StateVectorC128(n)
    m = L.MeasurementsC128(sv)
    W.apply_circuit(sv, ops[:n_qft])
    x = m.expval_named_batch(["PauliX"] * n, list(range(n)))
    np.testing.assert_allclose(x, np.ones(n), atol=1e-10)
    W.apply_circuit(sv, ops[n_qft:])
    p = m.probs(list(range(10)))
    assert abs(p.sum() - 1.0) < 1e-10
    z = m.expval_named_batch(["PauliZ"] * n, list(range(n)))
    assert np.abs(z).max() <= 1.0 + 1e-12
    W.apply_circuit(sv, inverse_ops(ops[n_qft:]))
    x = m.expval_named_batch(["PauliX"] * n, list(range(n)))
    np.testing.assert_allclose(x, np.ones(n), atol=1e-10)
    del m, sv


def test_hea30_complex64_against_complex128():
    n = 30
    ops = W.hea_rot_cz(n, 20, seed=1)
    a = L.StateVectorC128(n)
    W.apply_circuit(a, ops)
    b = L.StateVectorC64(n)
    W.apply_circuit(b, ops)
    fid = abs(a.overlap(b)) ** 2 / b.overlap(b).real
    assert fid > 1 - 1e-5
    ma, mb = L.MeasurementsC128(a), L.MeasurementsC64(b)
    assert abs(ma.expval("PauliX", [29]) - mb.expval("PauliX", [29])) < 1e-5
    assert abs(ma.expval("PauliZ", [0]) - mb.expval("PauliZ", [0])) < 1e-5
    del ma, mb, a, b


def test_qft33_complex64_against_complex128():
    n = 33
    ops = W.qft_brickwork_mcx(n, layers=8, mcx_per_layer=1, seed=5)
    a = L.StateVectorC128(n)
    W.apply_circuit(a, ops)
    b = L.StateVectorC64(n)
    W.apply_circuit(b, ops)
    fid = abs(a.overlap(b)) ** 2 / b.overlap(b).real
    assert fid > 1 - 1e-5
    ma, mb = L.MeasurementsC128(a), L.MeasurementsC64(b)
    za = np.asarray(ma.expval_named_batch(["PauliZ"] * n, list(range(n))))
    zb = np.asarray(mb.expval_named_batch(["PauliZ"] * n, list(range(n))))
    np.testing.assert_allclose(zb, za, atol=1e-5)
    np.testing.assert_allclose(mb.probs(list(range(10))), ma.probs(list(range(10))), atol=1e-5)
    del ma, mb, a, b
