"""Pins oracle/lq_oracle.py (CPU only).

Each check is chosen so that a plausible mistake in the oracle (transposed
matrix, wrong wire order, wrong sign or scale, dropped control) fails at least
one of: the reference's own known answers, a mathematical identity, a
brute-force Kronecker construction, or the compiled reference (oracle/_ref).
"""
import math

import numpy as np
import pytest

from oracle import lq_oracle as O
from oracle import ref

rng_global = np.random.default_rng(1234)

ALL_GATES = [g for g in O.GATE_NPARAMS if g != "Identity"]


def rand_state(n, rng):
    s = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    return s / np.linalg.norm(s)


def gate_args(name, rng, k_multi=3):
    nw = O.GATE_WIRES[name]
    if nw < 0:
        nw = k_multi
    npar = O.GATE_NPARAMS[name]
    params = list(rng.uniform(-math.pi, math.pi, size=npar))
    if name == "PCPhase":
        params[1] = float(rng.integers(0, 1 << nw))
    return nw, params


def kron_full(n, mat, wires, ctrls=(), cvals=()):
    """Brute-force 2^n x 2^n operator: sum over basis states, wire w = bit n-1-w."""
    dim = 1 << n
    k = len(wires)
    full = np.zeros((dim, dim), dtype=np.complex128)
    for col in range(dim):
        bits = [(col >> (n - 1 - w)) & 1 for w in range(n)]
        if any(bits[c] != v for c, v in zip(ctrls, cvals)):
            full[col, col] = 1
            continue
        loc = 0
        for w in wires:
            loc = (loc << 1) | bits[w]
        for r in range(1 << k):
            amp = mat[r, loc]
            if amp == 0:
                continue
            nb = list(bits)
            for j, w in enumerate(wires):
                nb[w] = (r >> (k - 1 - j)) & 1
            row = 0
            for b in nb:
                row = (row << 1) | b
            full[row, col] += amp
    return full


# --------------------------------------------------------------------------
# Known answers held by the reference's own tests
# --------------------------------------------------------------------------
def test_rx_known_answers():
    # lq/gates/tests/Test_GateImplementations_Param.cpp:146-166
    exp = {0.1: [0.9987502603949663, -0.04997916927067834],
           0.6: [0.9553364891256061, -0.2955202066613395]}
    for th, (re0, im1) in exp.items():
        s = O.zero_state(1)
        O.apply_op(s, 1, "RX", [0], False, [th])
        assert abs(s[0] - re0) < 1e-15 and abs(s[1] - 1j * im1) < 1e-15


def nontrivial_state(n=3):
    # core/utils/TestHelpers.hpp:436-470: RX(p),RY(p) per wire, p=0.7,0.5,...
    s = O.zero_state(n)
    p = 0.7
    for w in range(n):
        O.apply_op(s, n, "RX", [w], False, [p])
        O.apply_op(s, n, "RY", [w], False, [p])
        p -= 0.2
    return s


def test_expval_known_answers():
    # lq/measurements/tests/Test_MeasurementsLQubit.cpp:71-177
    s = nontrivial_state(3)
    table = {"PauliX": [0.49272486, 0.42073549, 0.28232124],
             "PauliY": [-0.64421768, -0.47942553, -0.29552020],
             "PauliZ": [0.58498357, 0.77015115, 0.91266780],
             "Hadamard": [0.7620549436, 0.8420840225, 0.8449848566]}
    for name, vals in table.items():
        got = [O.expval_named(s, 3, name, [w]) for w in range(3)]
        np.testing.assert_allclose(got, vals, atol=1e-6)


def test_adjoint_closed_forms():
    # core/algorithms/tests/Test_AdjointJacobian.cpp:175-201 (RX, Z -> -sin p)
    for p in (-2.3, -0.1, 0.4, 1.7):
        ops = [O.Op("RX", [0], [p])]
        st = O.apply_ops(O.zero_state(1), 1, ops)
        jac = O.adjoint_jacobian(st, 1, ops, [("N", "PauliZ", 0)], [0])
        assert abs(jac[0, 0] + math.sin(p)) < 1e-14
    # :138-173 controlled PhaseShift with Y on the target -> cos p
    for n in (2, 3, 4):
        for p in (-0.9, 0.3, 2.1):
            ops = [O.Op("PhaseShift", [n - 1], [p], ctrls=list(range(n - 1)),
                        cvals=[1] * (n - 1))]
            psi = np.zeros(1 << n, dtype=np.complex128)
            psi[-2] = psi[-1] = 1 / math.sqrt(2)
            st = O.apply_ops(psi, n, ops)
            jac = O.adjoint_jacobian(st, n, ops, [("N", "PauliY", n - 1)], [0])
            assert abs(jac[0, 0] - math.cos(p)) < 1e-14


def test_adjoint_decomposed_rot_table():
    # core/algorithms/tests/Test_AdjointJacobian.cpp:429-480 (margin 1e-4)
    expec = [[0, -9.90819496e-01, 0], [-8.18996553e-01, 1.62526544e-01, 0],
             [-0.203949, 0.48593716, 0], [0, 1, 0], [-2.03948985e-01, 4.85937177e-01, 0],
             [-8.18996598e-01, 1.62526487e-01, 0], [0, -9.90819511e-01, 0]]
    thetas = np.linspace(-2 * math.pi, 2 * math.pi, 7)
    for th, ex in zip(thetas, expec):
        lp = [th, th ** 3, math.sqrt(2) * th]
        ops = [O.Op("RZ", [0], [lp[0]]), O.Op("RY", [0], [lp[1]]), O.Op("RZ", [0], [lp[2]])]
        psi = np.array([1, -1], dtype=np.complex128) / math.sqrt(2)
        st = O.apply_ops(psi, 1, ops)
        jac = O.adjoint_jacobian(st, 1, ops, [("N", "PauliZ", 0)], [0, 1, 2])
        np.testing.assert_allclose(jac[0], ex, atol=1e-4)


# --------------------------------------------------------------------------
# Mathematics: unitarity, inverse, generator identity, brute force
# --------------------------------------------------------------------------
@pytest.mark.parametrize("name", ALL_GATES)
def test_gate_unitary_and_inverse(name):
    rng = np.random.default_rng(7)
    nw, params = gate_args(name, rng)
    u = O.gate_matrix(name, params, False, nw)
    ui = O.gate_matrix(name, params, True, nw)
    np.testing.assert_allclose(u @ u.conj().T, np.eye(u.shape[0]), atol=1e-14)
    np.testing.assert_allclose(ui @ u, np.eye(u.shape[0]), atol=1e-14)


GEN_GATES = ["RX", "RY", "RZ", "PhaseShift", "CRX", "CRY", "CRZ", "ControlledPhaseShift",
             "IsingXX", "IsingXY", "IsingYY", "IsingZZ", "SingleExcitation",
             "SingleExcitationMinus", "SingleExcitationPlus", "PSWAP", "DoubleExcitation",
             "DoubleExcitationMinus", "DoubleExcitationPlus", "MultiRZ", "GlobalPhase"]


@pytest.mark.parametrize("name", [g for g in GEN_GATES if g != "PSWAP"])
def test_generator_identity(name):
    # i s G U(theta) = dU/dtheta (Test_GateImplementations_Generator.cpp), at theta=0.37.
    # PSWAP is excluded: the reference's PSWAP generator (swap block, scale +1,
    # GateImplementationsLM.hpp:2666-2689) does not satisfy this identity; the
    # oracle reproduces it as is and test_generator_vs_reference pins it.
    nw = O.GATE_WIRES[name] if O.GATE_WIRES[name] > 0 else 3
    g, s = O.generator(name, nw)
    th, h = 0.37, 1e-6
    du = (O.gate_matrix(name, [th + h], False, nw) - O.gate_matrix(name, [th - h], False, nw)) / (2 * h)
    np.testing.assert_allclose(1j * s * g @ O.gate_matrix(name, [th], False, nw), du, atol=1e-9)
    np.testing.assert_allclose(g, g.conj().T, atol=0)


@pytest.mark.parametrize("trial", range(6))
def test_apply_matrix_brute_force(trial):
    rng = np.random.default_rng(100 + trial)
    n = 5
    k = int(rng.integers(1, 4))
    nc = int(rng.integers(0, 3))
    perm = list(rng.permutation(n))
    wires, ctrls = perm[:k], perm[k:k + nc]
    cvals = list(rng.integers(0, 2, size=nc))
    m = rng.standard_normal((1 << k, 1 << k)) + 1j * rng.standard_normal((1 << k, 1 << k))
    psi = rand_state(n, rng)
    got = O.apply_matrix(psi.copy(), n, m, wires, ctrls, cvals)
    exp = kron_full(n, m, wires, ctrls, cvals) @ psi
    np.testing.assert_allclose(got, exp, atol=1e-13)


def test_planted_wire_order_mistake_is_caught():
    # The brute-force check must fail for an oracle that reverses the wires.
    rng = np.random.default_rng(5)
    n, m = 4, rng.standard_normal((4, 4)) + 0j
    psi = rand_state(n, rng)
    wrong = O.apply_matrix(psi.copy(), n, m, [2, 0])  # planted: swapped order
    exp = kron_full(n, m, [0, 2]) @ psi
    assert np.abs(wrong - exp).max() > 1e-3


def test_probs_and_expval_consistency():
    rng = np.random.default_rng(3)
    n = 6
    psi = rand_state(n, rng)
    p = O.probs(psi, n)
    assert abs(p.sum() - 1) < 1e-14
    for w in range(n):
        sign = np.array([1 - 2 * ((i >> (n - 1 - w)) & 1) for i in range(1 << n)])
        assert abs(O.expval_named(psi, n, "PauliZ", [w]) - (p * sign).sum()) < 1e-14
    # marginal with a permuted wire order: brute force
    wires = [4, 1, 3]
    pm = O.probs(psi, n, wires)
    exp = np.zeros(8)
    for i in range(1 << n):
        o = 0
        for w in wires:
            o = (o << 1) | ((i >> (n - 1 - w)) & 1)
        exp[o] += p[i]
    np.testing.assert_allclose(pm, exp, atol=1e-15)


def test_adjoint_vs_finite_difference():
    rng = np.random.default_rng(11)
    n = 4
    ops = []
    for layer in range(2):
        for w in range(n):
            ops.append(O.Op(["RX", "RY", "RZ"][(w + layer) % 3], [w], [float(rng.uniform(0, 6))]))
        for w in range(n - 1):
            ops.append(O.Op("CNOT", [w, w + 1]))
    ops.append(O.Op("IsingXY", [0, 2], [0.3]))
    ops.append(O.Op("CRY", [3, 1], [1.1]))
    ops.append(O.Op("PhaseShift", [2], [0.5], ctrls=[0], cvals=[0]))
    ops.append(O.Op("RX", [1], [0.8], inverse=True))
    H = O.random_pauli_hamiltonian(n, 8, seed=3, probs_ixyz=(0.4, 0.2, 0.2, 0.2))
    obs = [H, ("N", "PauliZ", 1)]
    npar = sum(1 for op in ops if op.params)
    tp = list(range(npar))
    st = O.apply_ops(O.zero_state(n), n, ops)
    jac = O.adjoint_jacobian(st, n, ops, obs, tp)

    def energy(ps):
        it = iter(ps)
        ops2 = [O.Op(op.name, op.wires, [next(it)] if op.params else [], op.inverse,
                     op.matrix, op.ctrls, op.cvals) for op in ops]
        s = O.apply_ops(O.zero_state(n), n, ops2)
        return np.array([O.expval_obs(s, n, o) for o in obs])

    p0 = [op.params[0] for op in ops if op.params]
    h = 1e-6
    for t in tp:
        pp, pm = list(p0), list(p0)
        pp[t] += h
        pm[t] -= h
        fd = (energy(pp) - energy(pm)) / (2 * h)
        np.testing.assert_allclose(jac[:, t], fd, atol=1e-8)


# --------------------------------------------------------------------------
# Against the compiled reference (oracle/_ref)
# --------------------------------------------------------------------------
needs_ref = pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built")


@needs_ref
@pytest.mark.parametrize("name", ALL_GATES)
@pytest.mark.parametrize("inverse", [False, True])
def test_named_gate_vs_reference(name, inverse):
    rng = np.random.default_rng(hash((name, inverse)) % 2**32)
    n = 6
    nw, params = gate_args(name, rng)
    wires = list(rng.permutation(n)[:nw])
    psi = rand_state(n, rng)
    r = ref.RefStateVector(n)
    r.set_state(psi)
    r.apply(name, wires, inverse, params)
    got = O.apply_op(psi.copy(), n, name, wires, inverse, params)
    np.testing.assert_allclose(got, r.get_state(), atol=1e-13)


@needs_ref
@pytest.mark.parametrize("name", sorted(O.CONTROLLED_GATES))
def test_controlled_gate_vs_reference(name):
    rng = np.random.default_rng(hash(name) % 2**32)
    n = 7
    nw, params = gate_args(name, rng, k_multi=2)
    for nc in (1, 2):
        perm = list(rng.permutation(n))
        wires, ctrls = perm[:nw], perm[nw:nw + nc]
        cvals = list(rng.integers(0, 2, size=nc))
        inv = bool(rng.integers(0, 2))
        psi = rand_state(n, rng)
        r = ref.RefStateVector(n)
        r.set_state(psi)
        r.apply(name, wires, inv, params, ctrls, cvals)
        got = O.apply_op(psi.copy(), n, name, wires, inv, params, ctrls, cvals)
        np.testing.assert_allclose(got, r.get_state(), atol=1e-13, err_msg=f"nc={nc}")


@needs_ref
@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_matrix_vs_reference(k):
    rng = np.random.default_rng(k)
    n = 7
    for nc in (0, 1, 2):
        perm = list(rng.permutation(n))
        wires, ctrls = perm[:k], perm[k:k + nc]
        cvals = list(rng.integers(0, 2, size=nc))
        m = rng.standard_normal((1 << k, 1 << k)) + 1j * rng.standard_normal((1 << k, 1 << k))
        for inv in (False, True):
            psi = rand_state(n, rng)
            r = ref.RefStateVector(n)
            r.set_state(psi)
            r.apply_matrix(m, wires, inv, ctrls, cvals)
            got = O.apply_matrix(psi.copy(), n, m, wires, ctrls, cvals, inverse=inv)
            np.testing.assert_allclose(got, r.get_state(), atol=1e-12)


@needs_ref
@pytest.mark.parametrize("name", GEN_GATES)
def test_generator_vs_reference(name):
    rng = np.random.default_rng(hash(name) % 2**32)
    n = 6
    nw = O.GATE_WIRES[name] if O.GATE_WIRES[name] > 0 else 3
    for nc in (0, 1):
        if nc and name in ("CRX", "CRY", "CRZ", "ControlledPhaseShift"):
            continue
        perm = list(rng.permutation(n))
        wires, ctrls = perm[:nw], perm[nw:nw + nc]
        cvals = list(rng.integers(0, 2, size=nc))
        psi = rand_state(n, rng)
        r = ref.RefStateVector(n)
        r.set_state(psi)
        s_ref = r.apply_generator(name, wires, ctrls, cvals)
        mine = psi.copy()
        s = O.apply_generator(mine, n, name, wires, ctrls, cvals)
        assert s == s_ref
        np.testing.assert_allclose(mine, r.get_state(), atol=1e-13)


@needs_ref
def test_pauli_rot_vs_reference():
    rng = np.random.default_rng(9)
    n = 6
    for word in ("XYZ", "ZZ", "YIX", "X"):
        wires = list(rng.permutation(n)[:len(word)])
        th = float(rng.uniform(-3, 3))
        psi = rand_state(n, rng)
        r = ref.RefStateVector(n)
        r.set_state(psi)
        r.apply_pauli_rot(wires, th, word)
        got = O.apply_pauli_rot(psi.copy(), n, wires, th, word)
        np.testing.assert_allclose(got, r.get_state(), atol=1e-13)


@needs_ref
def test_measurements_vs_reference():
    rng = np.random.default_rng(21)
    n = 7
    psi = rand_state(n, rng)
    r = ref.RefStateVector(n)
    r.set_state(psi)
    np.testing.assert_allclose(O.probs(psi, n), r.probs(), atol=1e-15)
    for wires in ([0], [6, 2], [3, 1, 5], [0, 1, 2, 3, 4, 5, 6], [2, 0, 6, 4, 1]):
        np.testing.assert_allclose(O.probs(psi, n, wires), r.probs(wires), atol=1e-15)
    for name in ("PauliX", "PauliY", "PauliZ", "Hadamard", "Identity"):
        for w in range(n):
            assert abs(O.expval_named(psi, n, name, [w]) - r.expval_named(name, [w])) < 1e-14
    for k in (1, 2, 3):
        a = rng.standard_normal((1 << k, 1 << k)) + 1j * rng.standard_normal((1 << k, 1 << k))
        herm = a + a.conj().T
        wires = list(rng.permutation(n)[:k])
        assert abs(O.expval_matrix(psi, n, herm, wires) - r.expval_matrix(herm, wires)) < 1e-12
    H = O.random_pauli_hamiltonian(n, 20, seed=4, probs_ixyz=(0.4, 0.2, 0.2, 0.2))
    assert abs(O.expval_obs(psi, n, H) - r.expval_obs(H)) < 1e-12
    tp = ("T", [("PauliZ", 0), ("PauliX", 3)])
    assert abs(O.expval_obs(psi, n, tp) - r.expval_obs(tp)) < 1e-14


@needs_ref
def test_adjoint_vs_reference():
    rng = np.random.default_rng(31)
    n = 6
    ops = O.circuit_strongly_entangling(n, 2, seed=5)
    ops.append(O.Op("CRX", [0, 3], [0.4]))
    ops.append(O.Op("IsingZZ", [1, 2], [0.9], inverse=True))
    ops.append(O.Op("RY", [4], [0.2], ctrls=[5], cvals=[0]))
    ops.append(O.Op("QubitUnitary", [2, 5], [], matrix=O.gate_matrix("CNOT")))
    H = O.random_pauli_hamiltonian(n, 12, seed=6, probs_ixyz=(0.4, 0.2, 0.2, 0.2))
    obs = [H, ("N", "PauliX", 2), ("T", [("PauliZ", 0), ("PauliY", 4)])]
    npar = sum(1 for op in ops if op.params)
    tp = sorted(rng.choice(npar, size=npar // 2, replace=False).tolist())
    st = O.apply_ops(O.zero_state(n), n, ops)
    r = ref.RefStateVector(n)
    r.apply_ops(ops)
    np.testing.assert_allclose(st, r.get_state(), atol=1e-13)
    jref = r.adjoint_jacobian(ops, obs, tp)
    jmine = O.adjoint_jacobian(st, n, ops, obs, tp)
    np.testing.assert_allclose(jmine, jref, atol=1e-12)
    # single-observable path (AdjointJacobianLQubit.hpp:225-319)
    np.testing.assert_allclose(O.adjoint_jacobian(st, n, ops, [H], tp), r.adjoint_jacobian(ops, [H], tp),
                               atol=1e-12)


def _vjp_circuit():
    rng = np.random.default_rng(41)
    n = 5
    ops = O.circuit_strongly_entangling(n, 2, seed=7)
    ops.append(O.Op("CRZ", [0, 3], [0.4]))
    ops.append(O.Op("IsingXX", [1, 2], [0.9], inverse=True))
    ops.append(O.Op("RX", [4], [0.2], ctrls=[1], cvals=[0]))
    ops.append(O.Op("QubitUnitary", [2, 4], [], matrix=O.gate_matrix("CNOT")))
    ops.append(O.Op("PhaseShift", [3], [1.3]))
    npar = sum(1 for op in ops if op.params)
    tp = sorted(rng.choice(npar, size=npar - 3, replace=False).tolist())
    dy = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    return n, ops, tp, dy


def test_vjp_vs_finite_difference():
    """w_t = <dy | d psi / d theta_t> by central differences of the state."""
    n, ops, tp, dy = _vjp_circuit()
    st = O.apply_ops(O.zero_state(n), n, ops)
    w = O.vector_jacobian_product(st, n, ops, dy, tp)
    par = [op.params[0] for op in ops if op.params]

    def state(ps):
        it = iter(ps)
        ops2 = [O.Op(op.name, op.wires, [next(it)] if op.params else [], op.inverse,
                     op.matrix, op.ctrls, op.cvals) for op in ops]
        return O.apply_ops(O.zero_state(n), n, ops2)

    h = 1e-6
    for k, t in enumerate(tp):
        p1, p2 = list(par), list(par)
        p1[t] += h
        p2[t] -= h
        d = (state(p1) - state(p2)) / (2 * h)
        assert abs(np.vdot(dy, d) - w[k]) < 1e-8


def test_vjp_matches_adjoint_for_observable_cotangent():
    """With dy = O psi (O Hermitian), d<O>/d theta_t = 2 Re w_t."""
    n, ops, tp, _ = _vjp_circuit()
    st = O.apply_ops(O.zero_state(n), n, ops)
    H = O.random_pauli_hamiltonian(n, 10, seed=8, probs_ixyz=(0.4, 0.2, 0.2, 0.2))
    dy = O.obs_apply(H, st.copy(), n)
    w = O.vector_jacobian_product(st, n, ops, dy, tp)
    jac = O.adjoint_jacobian(st, n, ops, [H], tp)[0]
    np.testing.assert_allclose(2 * w.real, jac, atol=1e-12)


@needs_ref
def test_vjp_vs_reference():
    n, ops, tp, dy = _vjp_circuit()
    st = O.apply_ops(O.zero_state(n), n, ops)
    r = ref.RefStateVector(n)
    r.apply_ops(ops)
    np.testing.assert_allclose(O.vector_jacobian_product(st, n, ops, dy, tp),
                               r.vector_jacobian_product(ops, dy, tp), atol=1e-12)
