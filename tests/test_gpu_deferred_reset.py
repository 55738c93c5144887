"""Deferred resets on the GPU.

resetStateVector / setBasisState record the basis state instead of writing
it; the next program's first specialised pass forms it (amplitude 1 at the
basis index) without reading memory, and every other use of the state
(measurements, getState, copies, the device pointer, generic passes) writes
it first. Each case first fills the memory with an unrelated state, so a pass
or measurement that read the memory instead of the recorded basis state
would fail, and compares with the oracle (complex128 within 1e-12).
"""
import numpy as np
import pytest

from helpers import TOL_C128, oracle_run, rand_state
from oracle import lq_oracle as O

pytestmark = pytest.mark.gpu

L = pytest.importorskip("lightning_amd.lightning_qubit_ops")
W = pytest.importorskip("lightning_amd.workloads")

N = 16  # specialised passes are used from 16 qubits


def basis_state(n, bits):
    idx = 0
    for w, b in enumerate(bits):
        idx |= int(b) << (n - 1 - w)
    s = np.zeros(1 << n, dtype=np.complex128)
    s[idx] = 1.0
    return s


def scrambled(n, seed):
    sv = L.StateVectorC128(n)
    sv.updateData(rand_state(n, np.random.default_rng(seed)))
    return sv


def test_basis_state_then_circuit_generic_and_specialised():
    bits = [int(x) for x in np.random.default_rng(1).integers(0, 2, N)]
    ops = W.hea_rot_cz(N, 3, seed=4)
    exp = oracle_run(N, ops, basis_state(N, bits))
    sv = scrambled(N, 2)
    counts = []
    for attempt in range(2):
        sv.updateData(rand_state(N, np.random.default_rng(10 + attempt)))
        sv.setBasisState(bits, list(range(N)))
        sv.pass_counts(reset=True)
        W.apply_circuit(sv, ops)
        np.testing.assert_allclose(sv.state(), exp, atol=TOL_C128)
        counts.append(sv.pass_counts(reset=True))
        if attempt == 0:
            L.jit_wait()  # the second run uses the specialised kernels, whose first pass forms the state
    assert counts[1]["specialised"] > 0, counts


def test_reset_then_measurements_without_gates():
    sv = scrambled(N, 3)
    sv.resetStateVector()
    np.testing.assert_allclose(sv.state(), basis_state(N, [0] * N), atol=0)
    sv.updateData(rand_state(N, np.random.default_rng(4)))
    bits = [1, 0, 1] + [0] * (N - 3)
    sv.setBasisState(bits, list(range(N)))
    m = L.MeasurementsC128(sv)
    assert m.expval("PauliZ", [0]) == pytest.approx(-1.0, abs=TOL_C128)
    assert m.expval("PauliZ", [1]) == pytest.approx(1.0, abs=TOL_C128)
    p = np.asarray(m.probs([0, 1, 2]))
    exp = np.zeros(8)
    exp[0b101] = 1.0
    np.testing.assert_allclose(p, exp, atol=TOL_C128)


def test_reset_then_copy_and_overlap():
    src = scrambled(N, 5)
    src.setBasisState([0, 1] + [0] * (N - 2), list(range(N)))
    dst = scrambled(N, 6)
    dst.copy_from(src)
    exp = basis_state(N, [0, 1] + [0] * (N - 2))
    np.testing.assert_allclose(dst.state(), exp, atol=0)
    other = scrambled(N, 7)
    other.resetStateVector()
    assert abs(src.overlap(other)) == pytest.approx(0.0, abs=TOL_C128)
    other.setBasisState([0, 1] + [0] * (N - 2), list(range(N)))
    assert abs(src.overlap(other)) == pytest.approx(1.0, abs=TOL_C128)


def test_reset_then_matrix_and_adjoint():
    # a dense 4-qubit matrix runs in the LDS pass kernel (not a specialised
    # register pass), so the reset is written before it
    rng = np.random.default_rng(8)
    U = np.linalg.qr(rng.standard_normal((16, 16)) + 1j * rng.standard_normal((16, 16)))[0]
    sv = scrambled(N, 9)
    sv.resetStateVector()
    sv.applyMatrix(U, [3, 7, 11, 15])
    exp = O.apply_ops(O.zero_state(N), N, [O.Op("Matrix", [3, 7, 11, 15], [], False, U, [], [])])
    np.testing.assert_allclose(sv.state(), exp, atol=TOL_C128)
    # adjoint Jacobian right after a reset and a short circuit
    ops = W.strongly_entangling(N, 1, seed=3)
    sv.updateData(rand_state(N, rng))
    sv.resetStateVector()
    W.apply_circuit(sv, ops)
    coeffs, words = W.random_pauli_hamiltonian(N, 6, seed=4, law=(0.5, 0.2, 0.15, 0.15))
    H = W.hamiltonian_obs(coeffs, words, 128)
    tp = list(range(0, sum(1 for op in ops if op.params), 3))
    jac = np.asarray(L.algorithms.AdjointJacobianC128()(sv, [H], W.to_ops(ops, 128), tp)).reshape(-1)
    st = oracle_run(N, ops)
    from helpers import ham_to_oracle, to_oracle_ops
    jexp = np.asarray(O.adjoint_jacobian(st, N, to_oracle_ops(ops), [ham_to_oracle(coeffs, words)], tp)).reshape(-1)
    np.testing.assert_allclose(jac, jexp, atol=1e-10)


def test_reset_then_read_only_pass_leaves_state_written():
    # an expval right after a reset runs a read-only dot pass; once its
    # specialised kernel exists, the reset must still be written to memory
    bits = [1, 1, 0, 1] + [0] * (N - 4)
    exp = basis_state(N, bits)
    sv = scrambled(N, 11)
    m = L.MeasurementsC128(sv)
    counts = []
    for attempt in range(2):
        sv.updateData(rand_state(N, np.random.default_rng(20 + attempt)))
        sv.setBasisState(bits, list(range(N)))
        sv.pass_counts(reset=True)
        assert m.expval("PauliX", [5]) == pytest.approx(0.0, abs=TOL_C128)
        counts.append(sv.pass_counts(reset=True))
        np.testing.assert_allclose(sv.state(), exp, atol=0)
        if attempt == 0:
            L.jit_wait()
    assert counts[1]["specialised"] > 0, counts
