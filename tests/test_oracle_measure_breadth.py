"""Oracle pinning for variances, sparse-Hamiltonian expectations and samples.

Checks: closed forms (variances of Pauli observables on basis and product
states, a diagonal sparse operator against probabilities), invariants
(sample frequencies against the probabilities, sample column order against a
basis state), and the compiled reference (oracle/_ref) value for value and,
for samples, draw for draw.
"""
import math

import numpy as np
import pytest

from helpers import rand_state
from oracle import lq_oracle as O
from oracle import ref

needs_ref = pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built")

NONPARAM_GATES = [g for g in O.GATE_NPARAMS if O.GATE_NPARAMS[g] == 0 and O.GATE_WIRES[g] > 0]


def random_csr(n, nnz, rng, hermitian=True):
    d = 1 << n
    a = np.zeros((d, d), dtype=np.complex128)
    for _ in range(nnz):
        i, j = rng.integers(0, d, 2)
        v = rng.normal() + 1j * rng.normal()
        a[i, j] += v
        if hermitian:
            a[j, i] += np.conj(v)
    rows, cols = np.nonzero(a)
    indptr = np.searchsorted(rows, np.arange(d + 1))
    return indptr.astype(np.int64), cols.astype(np.int64), a[rows, cols], a


def test_var_closed_forms():
    n = 3
    z = O.zero_state(n)
    assert O.var_named(z, n, "PauliZ", [1]) == pytest.approx(0.0, abs=1e-15)
    assert O.var_named(z, n, "PauliX", [1]) == pytest.approx(1.0, abs=1e-15)
    # RY(t)|0>: <Z> = cos t, var Z = sin^2 t; <X> = sin t, var X = cos^2 t
    t = 0.7
    s = O.apply_op(z.copy(), n, "RY", [2], params=[t])
    assert O.var_named(s, n, "PauliZ", [2]) == pytest.approx(math.sin(t) ** 2, abs=1e-14)
    assert O.var_named(s, n, "PauliX", [2]) == pytest.approx(math.cos(t) ** 2, abs=1e-14)
    assert O.var_obs(s, n, ("N", "PauliZ", 2)) == pytest.approx(math.sin(t) ** 2, abs=1e-14)
    # Hamiltonian a Z2 + b X2 on RY(t)|0>: var = a^2 + b^2 - (a cos t + b sin t)^2
    a, b = 0.3, -1.1
    h = ("H", [a, b], [("N", "PauliZ", 2), ("N", "PauliX", 2)])
    assert O.var_obs(s, n, h) == pytest.approx(a * a + b * b - (a * math.cos(t) + b * math.sin(t)) ** 2, abs=1e-14)
    # unnormalised state, Identity: |psi|^2 - |psi|^4
    assert O.var_named(2 * s, n, "Identity", [0]) == pytest.approx(4 - 16, abs=1e-13)


def test_sparse_diagonal_matches_probs():
    n = 4
    rng = np.random.default_rng(1)
    st = rand_state(n, rng)
    d = rng.normal(size=1 << n)
    indptr = np.arange((1 << n) + 1)
    cols = np.arange(1 << n)
    p = np.abs(st) ** 2
    assert O.expval_sparse(st, indptr, cols, d) == pytest.approx(float(np.dot(d, p)), abs=1e-13)
    assert O.var_sparse(st, indptr, cols, d) == pytest.approx(float(np.dot(d * d, p) - np.dot(d, p) ** 2), abs=1e-13)
    # the sparse form of a dense Hermitian matches the dense expval
    indptr, cols, vals, a = random_csr(n, 20, rng)
    assert O.expval_sparse(st, indptr, cols, vals) == pytest.approx(float(np.real(np.vdot(st, a @ st))), abs=1e-12)


def test_samples_follow_probabilities_and_column_order():
    n = 3
    z = O.zero_state(n)
    basis = z.copy()
    basis[:] = 0
    basis[0b011] = 1.0  # wires 1 and 2 set
    s = O.generate_samples(basis, n, [2, 0, 1], 20, seed=3)
    assert (s == np.array([1, 0, 1], dtype=np.uint64)).all()
    rng = np.random.default_rng(2)
    st = rand_state(n, rng)
    shots = 20000
    smp = O.generate_samples(st, n, None, shots, seed=11)
    idx = (smp * (1 << np.arange(n - 1, -1, -1))).sum(axis=1)
    freq = np.bincount(idx.astype(np.int64), minlength=1 << n) / shots
    p = O.probs(st, n)
    # each frequency within 5 sigma of its probability
    assert (np.abs(freq - p) <= 5 * np.sqrt(p * (1 - p) / shots) + 1e-12).all()
    # a fixed seed repeats its draws
    assert np.array_equal(smp[:50], O.generate_samples(st, n, None, 50, seed=11))


@needs_ref
@pytest.mark.parametrize("prec", ["double", "float"])
def test_var_vs_reference(prec):
    n = 5
    rng = np.random.default_rng(7)
    st = rand_state(n, rng)
    r = ref.RefStateVector(n, prec)
    r.set_state(st)
    s0 = st.astype(np.complex64).astype(np.complex128) if prec == "float" else st
    tol = 1e-13 if prec == "double" else 2e-6
    for name in NONPARAM_GATES + ["Identity"]:
        k = O.GATE_WIRES.get(name, 1)
        wires = [int(w) for w in rng.permutation(n)[:k]]
        assert r.var_named(name, wires) == pytest.approx(O.var_named(s0, n, name, wires), abs=tol), name
    herm = rng.normal(size=(4, 4)) + 1j * rng.normal(size=(4, 4))
    herm = herm + herm.conj().T
    for obs in [("N", "PauliY", 3), ("T", [("PauliX", 0), ("PauliZ", 4)]), ("M", herm, [1, 3]),
                ("H", [0.4, -0.9, 1.3], [("N", "PauliZ", 1), ("T", [("PauliX", 0), ("PauliY", 2)]), ("M", herm, [4, 0])])]:
        assert r.var_obs(obs) == pytest.approx(O.var_obs(s0, n, obs), abs=tol * 10), obs[0]


@needs_ref
@pytest.mark.parametrize("prec", ["double", "float"])
def test_sparse_vs_reference(prec):
    n = 6
    rng = np.random.default_rng(8)
    st = rand_state(n, rng)
    r = ref.RefStateVector(n, prec)
    r.set_state(st)
    s0 = st.astype(np.complex64).astype(np.complex128) if prec == "float" else st
    tol = 1e-12 if prec == "double" else 1e-4
    for trial in range(3):
        indptr, cols, vals, _ = random_csr(n, 40 + 30 * trial, rng, hermitian=trial != 2)
        assert r.expval_sparse(indptr, cols, vals) == pytest.approx(O.expval_sparse(s0, indptr, cols, vals), abs=tol)
        assert r.var_sparse(indptr, cols, vals) == pytest.approx(O.var_sparse(s0, indptr, cols, vals), abs=tol)


@needs_ref
@pytest.mark.parametrize("prec", ["double", "float"])
@pytest.mark.parametrize("wires", [None, [4, 1], [0, 2, 3, 5, 1]])
@pytest.mark.parametrize("seed", [0, 42, 2 ** 32 + 5])
def test_samples_vs_reference_draw_for_draw(prec, wires, seed):
    n = 6
    rng = np.random.default_rng(9)
    st = rand_state(n, rng)
    r = ref.RefStateVector(n, prec)
    r.set_state(st)
    s0 = st.astype(np.complex64).astype(np.complex128) if prec == "float" else st
    a = r.generate_samples(wires, 300, seed)
    b = O.generate_samples(s0, n, wires, 300, seed, single=prec == "float")
    assert np.array_equal(a, b)


@needs_ref
@pytest.mark.parametrize("prec", ["double", "float"])
def test_apply_sparse_matrix_vs_reference(prec):
    n = 6
    rng = np.random.default_rng(12)
    for k, ctrls in ((1, []), (2, []), (3, [5]), (4, [0, 2])):
        st = rand_state(n, rng)
        wires = [w for w in rng.permutation(n) if w not in ctrls][:k]
        wires = [int(w) for w in wires]
        cvals = [int(x) for x in rng.integers(0, 2, len(ctrls))]
        indptr, cols, vals, _ = random_csr(k, 3 << k, rng, hermitian=False)
        r = ref.RefStateVector(n, prec)
        r.set_state(st)
        r.apply_sparse(indptr, cols, vals, wires, ctrls, cvals)
        s0 = st.astype(np.complex64).astype(np.complex128) if prec == "float" else st
        vq = vals.astype(np.complex64).astype(np.complex128) if prec == "float" else vals
        exp = O.apply_sparse_matrix(s0.copy(), n, indptr, cols, vq, wires, ctrls, cvals)
        np.testing.assert_allclose(r.get_state(), exp, atol=1e-12 if prec == "double" else 2e-5)
