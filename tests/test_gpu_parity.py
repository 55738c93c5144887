"""Engine (HIP, through the C-ABI) vs the oracle on seeded inputs.

complex128: amplitudes and expectations within 1e-12 absolute; complex64:
within 5e-5 of the FP64 oracle for these small circuits.
"""
import math

import numpy as np
import pytest

from helpers import TOL_C128, TOL_C64, ham_to_oracle, oracle_run, rand_state, to_oracle_ops
from oracle import lq_oracle as O

pytestmark = pytest.mark.gpu

L = pytest.importorskip("lightning_amd.lightning_qubit_ops")
W = pytest.importorskip("lightning_amd.workloads")

SV = {128: L.StateVectorC128, 64: L.StateVectorC64}
MEAS = {128: L.MeasurementsC128, 64: L.MeasurementsC64}
TOL = {128: TOL_C128, 64: TOL_C64}
CDT = {128: np.complex128, 64: np.complex64}

ALL_GATES = [g for g in O.GATE_NPARAMS if g != "Identity"]


def gate_args(name, rng, k_multi=3):
    nw = O.GATE_WIRES[name] if O.GATE_WIRES[name] > 0 else k_multi
    params = list(rng.uniform(-math.pi, math.pi, size=O.GATE_NPARAMS[name]))
    if name == "PCPhase":
        params[1] = float(rng.integers(0, 1 << nw))
    return nw, params


def make_sv(bits, n, psi):
    sv = SV[bits](n)
    sv.updateData(np.asarray(psi, dtype=CDT[bits]))
    return sv


@pytest.mark.parametrize("n", [7, 12])
@pytest.mark.parametrize("bits", [128, 64])
@pytest.mark.parametrize("name", ALL_GATES)
def test_named_gates(bits, name, n):
    rng = np.random.default_rng(abs(hash((name, bits, n))) % 2**32)
    for inverse in (False, True):
        nw, params = gate_args(name, rng)
        wires = [int(w) for w in rng.permutation(n)[:nw]]
        psi = rand_state(n, rng)
        sv = make_sv(bits, n, psi)
        getattr(sv, name)(wires, inverse, params)
        exp = O.apply_op(psi.copy(), n, name, wires, inverse, params)
        np.testing.assert_allclose(sv.state(), exp, atol=TOL[bits], err_msg=f"{name} inv={inverse}")


@pytest.mark.parametrize("n", [8, 13])
@pytest.mark.parametrize("name", sorted(O.CONTROLLED_GATES))
def test_controlled_gates(name, n):
    rng = np.random.default_rng(abs(hash((name, n))) % 2**32)
    for nc in (1, 2, 3):
        nw, params = gate_args(name, rng, k_multi=2)
        perm = [int(w) for w in rng.permutation(n)]
        wires, ctrls = perm[:nw], perm[nw:nw + nc]
        cvals = [bool(v) for v in rng.integers(0, 2, size=nc)]
        inv = bool(rng.integers(0, 2))
        psi = rand_state(n, rng)
        sv = make_sv(128, n, psi)
        getattr(sv, name)(ctrls, cvals, wires, inv, params)
        exp = O.apply_op(psi.copy(), n, name, wires, inv, params, ctrls, cvals)
        np.testing.assert_allclose(sv.state(), exp, atol=TOL_C128, err_msg=f"nc={nc}")


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
def test_matrices(k):
    rng = np.random.default_rng(k)
    n = 12
    for nc in (0, 1, 2):
        perm = [int(w) for w in rng.permutation(n)]
        wires, ctrls = perm[:k], perm[k:k + nc]
        cvals = [bool(v) for v in rng.integers(0, 2, size=nc)]
        m = rng.standard_normal((1 << k, 1 << k)) + 1j * rng.standard_normal((1 << k, 1 << k))
        for inv in (False, True):
            psi = rand_state(n, rng)
            sv = make_sv(128, n, psi)
            if nc:
                sv.applyControlledMatrix(m, ctrls, cvals, wires, inv)
            else:
                sv.applyMatrix(m, wires, inv)
            exp = O.apply_matrix(psi.copy(), n, m, wires, ctrls, cvals, inverse=inv)
            np.testing.assert_allclose(sv.state(), exp, atol=1e-11)


@pytest.mark.parametrize("bits", [128, 64])
@pytest.mark.parametrize("n", [3, 9, 14, 16])
def test_layered_circuits_multi_tile(bits, n):
    # several tiles per pass at n >= 13, runs of fused gates
    ops = W.layers_rxyz_cnot(n, 4, seed=n) + W.hea_rot_cz(n, 3, seed=n + 1)
    sv = SV[bits](n)
    W.apply_circuit(sv, ops)
    exp = oracle_run(n, ops)
    tol = TOL[bits] * (10 if bits == 64 else 1)
    np.testing.assert_allclose(sv.state(), exp, atol=tol)


def test_random_mixed_circuit():
    rng = np.random.default_rng(77)
    n = 15
    ops = []
    for _ in range(300):
        name = ALL_GATES[int(rng.integers(len(ALL_GATES)))]
        nw, params = gate_args(name, rng)
        wires = [int(w) for w in rng.permutation(n)[:nw]]
        ops.append(W.CircuitOp(name, wires, params, bool(rng.integers(0, 2))))
    sv = SV[128](n)
    W.apply_circuit(sv, ops)
    np.testing.assert_allclose(sv.state(), oracle_run(n, ops), atol=1e-11)


@pytest.mark.parametrize("bits", [128, 64])
def test_measurements(bits):
    rng = np.random.default_rng(5)
    n = 15
    psi = rand_state(n, rng)
    sv = make_sv(bits, n, psi)
    M = MEAS[bits](sv)
    tol = TOL[bits]
    np.testing.assert_allclose(M.probs(), O.probs(psi, n), atol=tol)
    for wires in ([0], [14], [3, 9], [12, 1, 5], list(range(n)), [4, 2, 0, 14, 7, 9, 11, 13, 1, 3, 5, 6, 8]):
        np.testing.assert_allclose(M.probs(wires), O.probs(psi, n, wires), atol=tol, err_msg=str(wires))
    for name in ("PauliX", "PauliY", "PauliZ", "Hadamard", "Identity"):
        for w in (0, 7, 14):
            assert abs(M.expval(name, [w]) - O.expval_named(psi, n, name, [w])) < tol * 10
    z = M.expval_named_batch(["PauliZ"] * n, list(range(n)))
    np.testing.assert_allclose(z, [O.expval_named(psi, n, "PauliZ", [w]) for w in range(n)], atol=tol * 10)
    for k in (1, 2, 3):
        a = rng.standard_normal((1 << k, 1 << k)) + 1j * rng.standard_normal((1 << k, 1 << k))
        h = a + a.conj().T
        wires = [int(w) for w in rng.permutation(n)[:k]]
        assert abs(M.expval(h, wires) - O.expval_matrix(psi, n, h, wires)) < tol * 100


@pytest.mark.parametrize("bits", [128, 64])
def test_observables(bits):
    rng = np.random.default_rng(6)
    n = 14
    psi = rand_state(n, rng)
    sv = make_sv(bits, n, psi)
    M = MEAS[bits](sv)
    sfx = f"C{bits}"
    named = getattr(L, "NamedObs" + sfx)
    tensor = getattr(L, "TensorProdObs" + sfx)
    herm = getattr(L, "HermitianObs" + sfx)
    ham = getattr(L, "Hamiltonian" + sfx)
    tol = TOL[bits] * 100
    # small and large Pauli sums take different paths
    for terms in (3, 60):
        coeffs, words = W.random_pauli_hamiltonian(n, terms, seed=terms, law=(0.5, 0.2, 0.15, 0.15))
        H = W.hamiltonian_obs(coeffs, words, bits)
        assert abs(M.expval(H) - O.expval_obs(psi, n, ham_to_oracle(coeffs, words))) < tol
    t = tensor([named("PauliZ", [0]), named("PauliX", [13]), named("Hadamard", [5])])
    ot = ("T", [("PauliZ", 0), ("PauliX", 13), ("Hadamard", 5)])
    assert abs(M.expval(t) - O.expval_obs(psi, n, ot)) < tol
    a = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
    hm = a + a.conj().T
    H2 = ham([0.3, -1.2], [herm(hm, [2, 9]), t])
    oh = ("H", [0.3, -1.2], [("M", hm, [2, 9]), ot])
    assert abs(M.expval(H2) - O.expval_obs(psi, n, oh)) < tol


def test_state_prep_and_misc():
    rng = np.random.default_rng(8)
    n = 10
    sv = SV[128](n)
    sv.setBasisState([1, 0, 1], [2, 5, 9])
    exp = np.zeros(1 << n, complex)
    exp[(1 << (n - 1 - 2)) | (1 << (n - 1 - 9))] = 1
    np.testing.assert_allclose(sv.state(), exp, atol=0)
    sub = rand_state(3, rng)
    sv.resetStateVector()
    sv.setStateVector(sub, [4, 0, 7])
    e = O.zero_state(n)
    for i in range(8):
        idx = 0
        for j, w in enumerate([4, 0, 7]):
            idx |= ((i >> (2 - j)) & 1) << (n - 1 - w)
        e[idx] = sub[i]
    np.testing.assert_allclose(sv.state(), e, atol=1e-15)
    psi = rand_state(n, rng) * 3.0
    sv.updateData(psi)
    sv.normalize()
    np.testing.assert_allclose(sv.state(), psi / np.linalg.norm(psi), atol=1e-14)
    sv.collapse(3, True)
    c = psi.copy()
    for i in range(1 << n):
        if not (i >> (n - 1 - 3)) & 1:
            c[i] = 0
    np.testing.assert_allclose(sv.state(), c / np.linalg.norm(c), atol=1e-14)
    with pytest.raises(RuntimeError, match="Error in PennyLane Lightning"):
        sv.RX([n], False, [0.1])
    with pytest.raises(RuntimeError, match="Operation does not exist"):
        sv._apply("NotAGate", [0])


@pytest.mark.parametrize("name", ["RX", "RY", "RZ", "PhaseShift", "CRX", "CRY", "CRZ", "ControlledPhaseShift",
                                  "IsingXX", "IsingXY", "IsingYY", "IsingZZ", "SingleExcitation",
                                  "SingleExcitationMinus", "SingleExcitationPlus", "PSWAP", "DoubleExcitation",
                                  "DoubleExcitationMinus", "DoubleExcitationPlus", "MultiRZ", "GlobalPhase"])
def test_generators(name):
    rng = np.random.default_rng(abs(hash(name)) % 2**32)
    n = 12
    nw = O.GATE_WIRES[name] if O.GATE_WIRES[name] > 0 else 3
    for nc in (0, 1):
        if nc and name in ("CRX", "CRY", "CRZ", "ControlledPhaseShift"):
            continue
        perm = [int(w) for w in rng.permutation(n)]
        wires, ctrls = perm[:nw], perm[nw:nw + nc]
        cvals = [bool(v) for v in rng.integers(0, 2, size=nc)]
        psi = rand_state(n, rng)
        sv = make_sv(128, n, psi)
        s = sv.applyGenerator(name, wires, False, ctrls, cvals)
        mine = psi.copy()
        s_exp = O.apply_generator(mine, n, name, wires, ctrls, cvals)
        assert s == s_exp
        np.testing.assert_allclose(sv.state(), mine, atol=TOL_C128)


@pytest.mark.parametrize("bits", [128, 64])
def test_adjoint_jacobian(bits):
    rng = np.random.default_rng(12)
    n = 13
    ops = W.strongly_entangling(n, 2, seed=3)
    ops += [W.CircuitOp("CRX", [0, 3], [0.4]), W.CircuitOp("IsingZZ", [1, 12], [0.9], inverse=True),
            W.CircuitOp("RY", [4], [0.2], ctrls=[5], cvals=[False]),
            W.CircuitOp("QubitUnitary", [2, 5], matrix=O.gate_matrix("CNOT")),
            W.CircuitOp("MultiRZ", [0, 6, 11], [0.7]), W.CircuitOp("PhaseShift", [9], [1.3])]
    coeffs, words = W.random_pauli_hamiltonian(n, 30, seed=4, law=(0.5, 0.2, 0.15, 0.15))
    npar = sum(1 for op in ops if op.params)
    tp = sorted(int(t) for t in rng.choice(npar, size=npar - 5, replace=False))
    sv = SV[bits](n)
    W.apply_circuit(sv, ops)
    H = W.hamiltonian_obs(coeffs, words, bits)
    sfx = f"C{bits}"
    Z = getattr(L, "NamedObs" + sfx)("PauliZ", [7])
    adj = getattr(L.algorithms, "AdjointJacobian" + sfx)()
    jac, ev = adj(sv, [H, Z], W.to_ops(ops, bits), tp, return_expvals=True)
    st = oracle_run(n, ops)
    oobs = [ham_to_oracle(coeffs, words), ("N", "PauliZ", 7)]
    jexp = O.adjoint_jacobian(st, n, to_oracle_ops(ops), oobs, tp)
    tol = 1e-10 if bits == 128 else 5e-4
    np.testing.assert_allclose(np.asarray(jac).reshape(2, -1), jexp, atol=tol)
    np.testing.assert_allclose(ev, [O.expval_obs(st, n, o) for o in oobs], atol=tol)
    # the state is unchanged by the adjoint sweep
    np.testing.assert_allclose(sv.state(), st, atol=TOL[bits] * 10)


def test_adjoint_apply_operations_and_edge_cases():
    n = 5
    ops = W.layers_rxyz_cnot(n, 2, seed=9)
    sv = SV[128](n)
    Z = L.NamedObsC128("PauliZ", [0])
    adj = L.AdjointJacobianC128()
    jac = adj(sv, [Z], W.to_ops(ops), list(range(6)), apply_operations=True)
    st = oracle_run(n, ops)
    jexp = O.adjoint_jacobian(st, n, to_oracle_ops(ops), [("N", "PauliZ", 0)], list(range(6)))
    np.testing.assert_allclose(jac, jexp[0], atol=1e-12)
    assert adj(sv, [Z], W.to_ops(ops), []).size == 0
    bad = [W.CircuitOp("Rot", [0], [0.1, 0.2, 0.3])]
    with pytest.raises(RuntimeError, match="adjoint differentiation"):
        adj(sv, [Z], W.to_ops(bad), [0])


@pytest.mark.parametrize("bits", [128, 64])
@pytest.mark.parametrize("n", [6, 13])
def test_vector_jacobian_product(bits, n):
    rng = np.random.default_rng(60 + n)
    ops = W.strongly_entangling(n, 2, seed=n)
    ops += [W.CircuitOp("CRZ", [0, 3], [0.4]), W.CircuitOp("IsingXX", [1, 2], [0.9], inverse=True),
            W.CircuitOp("RX", [4], [0.2], ctrls=[5], cvals=[False]),
            W.CircuitOp("QubitUnitary", [2, 4], matrix=O.gate_matrix("CNOT")),
            W.CircuitOp("PhaseShift", [3], [1.3])]
    npar = sum(1 for op in ops if op.params)
    tp = sorted(int(t) for t in rng.choice(npar, size=npar - 4, replace=False))
    dy = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    dy /= np.linalg.norm(dy)
    sv = SV[bits](n)
    W.apply_circuit(sv, ops)
    vjp = getattr(L.algorithms, f"VectorJacobianProductC{bits}")()
    w = vjp(sv, W.to_ops(ops, bits), dy.astype(CDT[bits]), tp)
    assert w.dtype == CDT[bits] and w.shape == (len(tp),)
    st = oracle_run(n, ops)
    exp = O.vector_jacobian_product(st, n, to_oracle_ops(ops), dy.astype(CDT[bits]).astype(np.complex128), tp)
    np.testing.assert_allclose(w, exp, atol=1e-10 if bits == 128 else 5e-4)
    # apply_operations=True starts from the initial state
    sv0 = SV[bits](n)
    w0 = vjp(sv0, W.to_ops(ops, bits), dy.astype(CDT[bits]), tp, apply_operations=True)
    np.testing.assert_allclose(w0, exp, atol=1e-10 if bits == 128 else 5e-4)
    np.testing.assert_allclose(sv.state(), st, atol=TOL[bits] * 10)
    with pytest.raises(L.LightningException):
        vjp(sv, W.to_ops(ops, bits), dy[:-1].astype(CDT[bits]), tp)
