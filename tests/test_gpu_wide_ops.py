"""Operators the round-one parity tests left out, engine vs the oracle.

* applyPauliRot (lq/StateVectorLQubit.hpp:492-501, GateImplementationsLM.hpp
  :575-629): words of X, Y, Z, I and other letters (which the reference treats
  as X), inverse on and off, both precisions, including a 12-letter word.
* expval(matrix, wires) on 4 to 8 wires (the reference pins 1 to 8 wires in
  lq/measurements/tests/Test_MeasurementsLQubit.cpp:205-447) and past 10.
* Dense matrices on 6 to 12 wires (tile ops up to 10 targets, the
  global-memory kernel beyond), with controls and inverse; var(matrix) and a
  Hermitian observable on 11 wires.

complex128 within 1e-12 (amplitudes) / 1e-10 (expectations), complex64
within 5e-5 of the FP64 oracle (1e-3 relative for the 2^k-term sums past 10
wires).
"""
import numpy as np
import pytest

from helpers import TOL_C128, TOL_C64, rand_state
from oracle import lq_oracle as O

pytestmark = pytest.mark.gpu

L = pytest.importorskip("lightning_amd.lightning_qubit_ops")

SV = {128: L.StateVectorC128, 64: L.StateVectorC64}
MEAS = {128: L.MeasurementsC128, 64: L.MeasurementsC64}
TOL = {128: TOL_C128, 64: TOL_C64}
CDT = {128: np.complex128, 64: np.complex64}


def make_sv(bits, n, psi):
    sv = SV[bits](n)
    sv.updateData(np.asarray(psi, dtype=CDT[bits]))
    return sv


def rand_unitary(d, rng):
    z = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def rand_hermitian(d, rng):
    a = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
    return (a + a.conj().T) / 2


@pytest.mark.parametrize("bits", [128, 64])
@pytest.mark.parametrize("n,word", [(7, "X"), (7, "Y"), (7, "ZZ"), (7, "XYZ"), (7, "IZX"), (12, "YXIZ"),
                                    (12, "ZAYX"), (12, "YYYYY"), (14, "XYZIXYZIXYZY")])
def test_pauli_rot(bits, n, word):
    rng = np.random.default_rng(abs(hash((bits, n, word))) % 2**32)
    wires = [int(w) for w in rng.permutation(n)[:len(word)]]
    for inverse in (False, True):
        theta = float(rng.uniform(-np.pi, np.pi))
        psi = rand_state(n, rng)
        sv = make_sv(bits, n, psi)
        sv.applyPauliRot(wires, inverse, [theta], word)
        exp = O.apply_pauli_rot(psi.copy(), n, wires, theta, word, inverse)
        np.testing.assert_allclose(sv.state(), exp, atol=TOL[bits], err_msg=f"{word} inv={inverse}")


@pytest.mark.parametrize("bits", [128, 64])
@pytest.mark.parametrize("k", [4, 5, 6, 7, 8, 11])
def test_expval_matrix_wide(bits, k):
    n = 12 if k <= 8 else 13
    rng = np.random.default_rng(100 + k)
    psi = rand_state(n, rng)
    sv = make_sv(bits, n, psi)
    m = MEAS[bits](sv)
    for _ in range(2):
        wires = [int(w) for w in rng.permutation(n)[:k]]
        h = rand_hermitian(1 << k, rng)
        exp = O.expval_matrix(psi, n, h, wires)
        got = m.expval(np.asarray(h, dtype=CDT[bits]), wires)
        tol = 1e-10 if bits == 128 else 1e-3 * max(1.0, np.abs(h).max())
        assert abs(got - exp) <= tol, (k, got, exp)


@pytest.mark.parametrize("k", [6, 8, 10, 11, 12])
def test_matrices_wide(k):
    n = 14
    rng = np.random.default_rng(200 + k)
    for nc, inverse in ((0, False), (1, True), (2, False)):
        perm = [int(w) for w in rng.permutation(n)]
        wires, ctrls = perm[:k], perm[k:k + nc]
        cvals = [bool(v) for v in rng.integers(0, 2, size=nc)]
        u = rand_unitary(1 << k, rng)
        psi = rand_state(n, rng)
        sv = make_sv(128, n, psi)
        sv.RY([perm[-1]], False, [0.3])  # a queued gate before the operator
        if nc:
            sv.applyControlledMatrix(u, ctrls, cvals, wires, inverse)
        else:
            sv.applyMatrix(u, wires, inverse)
        sv.RX([perm[-1]], False, [0.2])  # and one after it
        exp = O.apply_op(psi.copy(), n, "RY", [perm[-1]], False, [0.3])
        exp = O.apply_matrix(exp, n, u, wires, ctrls, cvals, inverse)
        exp = O.apply_op(exp, n, "RX", [perm[-1]], False, [0.2])
        np.testing.assert_allclose(sv.state(), exp, atol=TOL_C128, err_msg=f"k={k} nc={nc}")


def test_var_and_hermitian_observable_wide():
    n, k = 13, 11
    rng = np.random.default_rng(7)
    psi = rand_state(n, rng)
    sv = make_sv(128, n, psi)
    m = L.MeasurementsC128(sv)
    wires = [int(w) for w in rng.permutation(n)[:k]]
    h = rand_hermitian(1 << k, rng)
    assert abs(m.var(h, wires) - O.var_matrix(psi, n, h, wires)) < 1e-9
    rest = [w for w in range(n) if w not in wires]
    obs = L.TensorProdObsC128([L.HermitianObsC128(h, wires), L.NamedObsC128("PauliZ", [rest[0]])])
    phi = O.apply_matrix(psi.copy(), n, h, wires)
    phi = O.apply_matrix(phi, n, np.diag([1.0, -1.0]), [rest[0]])
    assert abs(m.expval(obs) - float(np.real(np.vdot(psi, phi)))) < 1e-10
    assert abs(m.expval(L.HermitianObsC128(h, wires)) - O.expval_matrix(psi, n, h, wires)) < 1e-10
