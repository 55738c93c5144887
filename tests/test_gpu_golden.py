"""The engine (HIP, through the C-ABI) against the golden fixtures written by
the reference lightning.qubit, at the workloads' real sizes: the 20-qubit
RX/RY/RZ + CNOT circuit, the 24-qubit strongly-entangling adjoint Jacobian
with a 200-term Hamiltonian (432 parameters) and the 30-qubit depth-20
hardware-efficient ansatz (complex128, 16 GiB state).

Tolerances: amplitudes 1e-12, expectations and Jacobians 1e-10 (complex128).
"""
import os

import numpy as np
import pytest

import golden_circuits as G

pytestmark = pytest.mark.gpu

L = pytest.importorskip("lightning_amd.lightning_qubit_ops")
W = pytest.importorskip("lightning_amd.workloads")

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    path = os.path.join(GOLD, name)
    if not os.path.exists(path):
        pytest.skip(f"fixture {name} not generated")
    return np.load(path)


def run(n, ops, bits=128):
    sv = (L.StateVectorC128 if bits == 128 else L.StateVectorC64)(n)
    W.apply_circuit(sv, G.to_circuit(ops))
    return sv


def test_rxyz_cnot_n20():
    f = load("rxyz_cnot_n20_l10.npz")
    n = 20
    sv = run(n, G.rxyz_cnot_ops(f["angles"]))
    m = L.MeasurementsC128(sv)
    np.testing.assert_allclose([m.expval("PauliZ", [w]) for w in range(n)], f["expval_z"], atol=1e-10)
    pr = m.probs()
    np.testing.assert_allclose(pr.reshape(1024, -1).sum(1), f["probs_block_sums"], atol=1e-12)
    np.testing.assert_allclose(pr[f["sample_idx"]], f["sample_probs"], atol=1e-14)
    st = sv.state()
    np.testing.assert_allclose(st[f["sample_idx"]], f["sample_amps"], atol=1e-12)


@pytest.mark.parametrize("bits", [128, 64])
def test_rxyz_cnot_n20_full_vectors(bits):
    """The whole 2^20 state and probability vector against the oracle (pinned
    to the compiled reference; its block sums and sampled amplitudes match the
    fixture above): complex128 within 1e-12, complex64 within its stated
    bound."""
    from helpers import oracle_run

    f = load("rxyz_cnot_n20_l10.npz")
    n = 20
    ops = G.to_circuit(G.rxyz_cnot_ops(f["angles"]))
    exp = oracle_run(n, ops)
    np.testing.assert_allclose((np.abs(exp) ** 2).reshape(1024, -1).sum(1), f["probs_block_sums"], atol=1e-14)
    sv = run(n, G.rxyz_cnot_ops(f["angles"]), bits)
    m = (L.MeasurementsC128 if bits == 128 else L.MeasurementsC64)(sv)
    tol_a, tol_p = (1e-12, 1e-12) if bits == 128 else (5e-6, 1e-9)
    np.testing.assert_allclose(m.probs(), np.abs(exp) ** 2, atol=tol_p)
    np.testing.assert_allclose(sv.state(), exp, atol=tol_a)


def test_hea_n12_both_precisions():
    f = load("hea_rot_cz_n12_l4.npz")
    n = 12
    ops = G.hea_ops(f["angles"])
    sv = run(n, ops)
    np.testing.assert_allclose(sv.state(), f["state"], atol=1e-12)
    sv64 = run(n, ops, 64)
    fid = abs(np.vdot(f["state"], sv64.state().astype(np.complex128))) ** 2
    assert fid > 1 - 1e-6


def test_qft_haar_mcx_n12():
    f = load("qft_haar_mcx_n12_l8.npz")
    n = 12
    sv = run(n, G.qft_brickwork_ops(n, f["mats"], f["mcx"]))
    m = L.MeasurementsC128(sv)
    np.testing.assert_allclose([m.expval("PauliZ", [w]) for w in range(n)], f["expval_z"], atol=1e-10)
    np.testing.assert_allclose(m.probs(list(range(10))), f["probs_w0_9"], atol=1e-12)


@pytest.mark.parametrize("name,n", [("strongly_entangling_n10_l2_h30.npz", 10),
                                    ("strongly_entangling_n24_l6_h200.npz", 24)])
def test_adjoint_jacobian(name, n):
    f = load(name)
    ops = G.strongly_entangling_ops(f["angles"])
    sv = run(n, ops)
    H = W.hamiltonian_obs(list(f["coeffs"]), G.ham_words(f["letters"]), 128)
    Z = L.NamedObsC128("PauliZ", [n // 2])
    tp = list(range(sum(1 for op in ops if op.params)))
    adj = L.algorithms.AdjointJacobianC128()
    jac, ev = adj(sv, [H, Z], W.to_ops(G.to_circuit(ops), 128), tp, return_expvals=True)
    np.testing.assert_allclose(np.asarray(jac).reshape(2, -1), f["jacobian"], atol=1e-10)
    np.testing.assert_allclose(ev, f["expvals"], atol=1e-10)


def test_hea_n30_full_size():
    f = load("hea_rot_cz_n30_l20.npz")
    n = 30
    sv = run(n, G.hea_ops(f["angles"]))
    m = L.MeasurementsC128(sv)
    zz = m.expval(L.TensorProdObsC128([L.NamedObsC128("PauliZ", [0]), L.NamedObsC128("PauliZ", [1])]))
    assert abs(zz - f["expval_z0z1"]) < 1e-10
    assert abs(m.expval("PauliX", [n - 1]) - f["expval_x_last"]) < 1e-10
    np.testing.assert_allclose([m.expval("PauliZ", [int(w)]) for w in f["z_wires"]], f["expval_z"], atol=1e-10)
    # size-independent properties at full size: unit norm, probs of one wire sum to 1
    p = m.probs([29])
    assert abs(p.sum() - 1.0) < 1e-12
    assert abs((p[0] - p[1]) - m.expval("PauliZ", [29])) < 1e-12
    del m, sv
