"""The oracle against the golden fixtures written by the reference
lightning.qubit (tests/golden/make_golden.py). CPU only; the fixtures that
need a large state (n = 24, 30) are checked against the engine in
test_gpu_golden.py."""
import os

import numpy as np
import pytest

from oracle import lq_oracle as O

import golden_circuits as G

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    path = os.path.join(GOLD, name)
    if not os.path.exists(path):
        pytest.skip(f"fixture {name} not generated")
    return np.load(path)


def test_rxyz_cnot_n20():
    f = load("rxyz_cnot_n20_l10.npz")
    n = 20
    st = O.apply_ops(O.zero_state(n), n, G.rxyz_cnot_ops(f["angles"]))
    np.testing.assert_allclose([O.expval_named(st, n, "PauliZ", [w]) for w in range(n)], f["expval_z"], atol=1e-12)
    pr = O.probs(st, n)
    np.testing.assert_allclose(pr.reshape(1024, -1).sum(1), f["probs_block_sums"], atol=1e-12)
    np.testing.assert_allclose(pr[f["sample_idx"]], f["sample_probs"], atol=1e-14)
    np.testing.assert_allclose(st[f["sample_idx"]], f["sample_amps"], atol=1e-12)


def test_hea_n12():
    f = load("hea_rot_cz_n12_l4.npz")
    n = 12
    st = O.apply_ops(O.zero_state(n), n, G.hea_ops(f["angles"]))
    np.testing.assert_allclose(st, f["state"], atol=1e-13)
    assert abs(O.expval_obs(st, n, ("T", [("PauliZ", 0), ("PauliZ", 1)])) - f["expval_z0z1"]) < 1e-12
    assert abs(O.expval_named(st, n, "PauliX", [n - 1]) - f["expval_x_last"]) < 1e-12


def test_qft_haar_mcx_n12():
    f = load("qft_haar_mcx_n12_l8.npz")
    n = 12
    st = O.apply_ops(O.zero_state(n), n, G.qft_brickwork_ops(n, f["mats"], f["mcx"]))
    np.testing.assert_allclose([O.expval_named(st, n, "PauliZ", [w]) for w in range(n)], f["expval_z"], atol=1e-12)
    np.testing.assert_allclose(O.probs(st, n, list(range(10))), f["probs_w0_9"], atol=1e-13)


def test_adjoint_n10():
    f = load("strongly_entangling_n10_l2_h30.npz")
    n = 10
    ops = G.strongly_entangling_ops(f["angles"])
    obs = [G.ham_spec(f["coeffs"], f["letters"]), ("N", "PauliZ", n // 2)]
    st = O.apply_ops(O.zero_state(n), n, ops)
    tp = list(range(sum(1 for op in ops if op.params)))
    np.testing.assert_allclose(O.adjoint_jacobian(st, n, ops, obs, tp), f["jacobian"], atol=1e-11)
    np.testing.assert_allclose([O.expval_obs(st, n, o) for o in obs], f["expvals"], atol=1e-12)


def test_planted_mistake_is_caught():
    """A wrong wire order in the oracle (wires[0] as the local LSB instead of
    the MSB) must fail the 12-qubit fixture."""
    f = load("hea_rot_cz_n12_l4.npz")
    n = 12
    ops = G.hea_ops(f["angles"])
    ops = [O.Op(o.name, list(reversed(o.wires)), o.params) if len(o.wires) == 2 else o for o in ops]
    # CZ is symmetric, so plant the mistake in the single-qubit targets instead
    ops = [O.Op(o.name, [n - 1 - o.wires[0]], o.params) if len(o.wires) == 1 else o for o in ops]
    st = O.apply_ops(O.zero_state(n), n, ops)
    assert np.abs(st - f["state"]).max() > 1e-3
