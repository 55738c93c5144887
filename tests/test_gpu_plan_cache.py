"""Plan reuse on the GPU: circuits of one structure run with several angle
sets (the second and later runs reuse the first plan with refilled
coefficients, and once compiled the specialised kernels); every run is
checked against the oracle, for states below and above the size where
specialised kernels start, and for the adjoint Jacobian."""
import numpy as np
import pytest

from helpers import TOL_C128, TOL_C64, ham_to_oracle, oracle_run, to_oracle_ops
from oracle import lq_oracle as O

pytestmark = pytest.mark.gpu

L = pytest.importorskip("lightning_amd.lightning_qubit_ops")
W = pytest.importorskip("lightning_amd.workloads")

SV = {128: L.StateVectorC128, 64: L.StateVectorC64}
TOL = {128: TOL_C128, 64: TOL_C64 * 10}


@pytest.mark.parametrize("bits", [128, 64])
@pytest.mark.parametrize("n", [12, 17])
def test_same_structure_new_angles(bits, n):
    for seed in (1, 2, 3, 3, 4):
        if seed == 3:
            L.jit_wait()  # later runs use the specialised kernels where they apply
        ops = W.hea_rot_cz(n, 4, seed=seed)
        sv = SV[bits](n)
        W.apply_circuit(sv, ops)
        np.testing.assert_allclose(sv.state(), oracle_run(n, ops), atol=TOL[bits], err_msg=f"seed {seed}")


@pytest.mark.parametrize("bits", [128, 64])
def test_adjoint_same_structure_new_angles(bits):
    n = 16
    coeffs, words = W.random_pauli_hamiltonian(n, 10, seed=4, law=(0.5, 0.2, 0.15, 0.15))
    tol = 1e-10 if bits == 128 else 5e-4
    sfx = f"C{bits}"
    for seed in (5, 6, 6, 7):
        if seed == 6:
            L.jit_wait()
        ops = W.strongly_entangling(n, 2, seed=seed)
        npar = sum(1 for op in ops if op.params)
        tp = list(range(npar))
        st = oracle_run(n, ops)
        jexp = O.adjoint_jacobian(st, n, to_oracle_ops(ops), [ham_to_oracle(coeffs, words)], tp)
        sv = SV[bits](n)
        W.apply_circuit(sv, ops)
        H = W.hamiltonian_obs(coeffs, words, bits)
        adj = getattr(L.algorithms, "AdjointJacobian" + sfx)()
        jac = np.asarray(adj(sv, [H], W.to_ops(ops, bits), tp)).reshape(1, -1)
        np.testing.assert_allclose(jac, jexp, atol=tol, err_msg=f"seed {seed}")
