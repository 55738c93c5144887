"""The differentiable torch function (adjoint backward on the device) against
the oracle's adjoint Jacobian, central finite differences (h = 1e-6, FP64) and,
at the full 24-qubit / 432-parameter / 200-term size, the reference's
AdjointJacobian (golden fixture).

Tolerances: gradients vs adjoint Jacobians 1e-10 (complex128); vs central
differences 1e-7 (the difference quotient's own rounding error is about
1e-16 * |<H>| / h)."""
import os

import numpy as np
import pytest

import golden_circuits as G
from helpers import to_oracle_ops
from oracle import lq_oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
T = pytest.importorskip("lightning_amd.torch_interface")
W = pytest.importorskip("lightning_amd.workloads")

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def ham_spec(coeffs, words):
    return ("H", list(coeffs), [("T", list(w)) for w in words])


@pytest.mark.parametrize("bits", [128, 64])
def test_grad_small(bits):
    n = 6
    ops = W.hea_rot_cz(n, 2, seed=3) + [W.CircuitOp("CRY", [0, 4], [0.5]), W.CircuitOp("MultiRZ", [1, 2, 5], [0.7])]
    coeffs, words = W.random_pauli_hamiltonian(n, 12, seed=4, law=(0.4, 0.2, 0.2, 0.2))
    obs = [ham_spec(coeffs, words), ("N", "PauliZ", 2), ("T", [("PauliX", 0), ("PauliY", 3)])]
    aops = T.adjoint_form(ops)
    npar = T.num_params(aops)
    rng = np.random.default_rng(5)
    p0 = rng.uniform(0, 2 * np.pi, npar)
    params = torch.tensor(p0, dtype=torch.float64, requires_grad=True)
    ev = T.circuit_expvals(params, n, ops, obs, bits)
    wts = torch.tensor([0.7, -1.3, 0.4], dtype=torch.float64)
    (ev * wts).sum().backward()

    bound = to_oracle_ops(T.bind(aops, p0))
    st = O.apply_ops(O.zero_state(n), n, bound)
    jac = O.adjoint_jacobian(st, n, bound, obs, list(range(npar)))
    tol = 1e-10 if bits == 128 else 5e-4
    np.testing.assert_allclose(ev.detach().numpy(), [O.expval_obs(st, n, o) for o in obs], atol=tol)
    np.testing.assert_allclose(params.grad.numpy(), wts.numpy() @ jac, atol=tol)

    if bits == 128:
        h = 1e-6
        for k in (0, 5, npar - 1):
            pp, pm = p0.copy(), p0.copy()
            pp[k] += h
            pm[k] -= h
            fp = (T.circuit_expvals(torch.tensor(pp), n, ops, obs) * wts).sum().item()
            fm = (T.circuit_expvals(torch.tensor(pm), n, ops, obs) * wts).sum().item()
            assert abs((fp - fm) / (2 * h) - params.grad[k].item()) < 1e-7


def test_grad_config3_full_size():
    path = os.path.join(GOLD, "strongly_entangling_n24_l6_h200.npz")
    if not os.path.exists(path):
        pytest.skip("fixture not generated")
    f = np.load(path)
    n = 24
    ops = G.to_circuit(G.strongly_entangling_ops(f["angles"]))
    obs = [("H", [float(c) for c in f["coeffs"]], [("T", w) for w in G.ham_words(f["letters"])])]
    p0 = np.array([op.params[0] for op in ops if op.params])
    assert p0.size == 432
    params = torch.tensor(p0, dtype=torch.float64, requires_grad=True)
    ev = T.circuit_expvals(params, n, ops, obs)
    ev.sum().backward()
    assert abs(ev.item() - f["expvals"][0]) < 1e-10
    np.testing.assert_allclose(params.grad.numpy(), f["jacobian"][0], atol=1e-10)
    h = 1e-6
    for k in (0, 217, 431):
        pp, pm = p0.copy(), p0.copy()
        pp[k] += h
        pm[k] -= h
        fp = T.circuit_expvals(torch.tensor(pp), n, ops, obs).item()
        fm = T.circuit_expvals(torch.tensor(pm), n, ops, obs).item()
        assert abs((fp - fm) / (2 * h) - params.grad[k].item()) < 1e-7
