"""Matrix-core path for complex64 dense operators on the five lowest index
bits (lq_set_mfma_dense5): applyMatrix of a Haar-random 32x32 unitary on
wires n-5 .. n-1 (in order, reversed and shuffled, so the operator is
permuted into index-bit order on the host), inverse on and off, against the
FP64 oracle (lq/gates/cpu_kernels/GateImplementationsLM.hpp:454-498,
applyNQubitOp, restated in oracle/lq_oracle.py apply_matrix) and against
the default fused-tile path. Operators the matrix-core kernel does not take
(other wires, controls, complex128) must still give the oracle's result.

complex64 within 5e-5 of the FP64 oracle per amplitude (the 32-term fp32
sums of the GEMM carry ~1e-7 relative error).
"""
import numpy as np
import pytest

from helpers import TOL_C64, rand_state
from oracle import lq_oracle as O

pytestmark = pytest.mark.gpu

L = pytest.importorskip("lightning_amd.lightning_qubit_ops")


def rand_unitary(d, rng):
    z = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


@pytest.fixture
def mfma_on():
    prev = L.set_mfma_dense5(True)
    yield
    L.set_mfma_dense5(prev)


@pytest.mark.parametrize("n", [12, 13, 16])
@pytest.mark.parametrize("order", ["msb_first", "reversed", "shuffled"])
def test_dense5_mfma_matches_oracle(mfma_on, n, order):
    rng = np.random.default_rng(100 * n + len(order))
    wires = list(range(n - 5, n))
    if order == "reversed":
        wires = wires[::-1]
    elif order == "shuffled":
        wires = [int(w) for w in rng.permutation(wires)]
    for inverse in (False, True):
        u = rand_unitary(32, rng)
        psi = rand_state(n, rng)
        sv = L.StateVectorC64(n)
        sv.updateData(psi.astype(np.complex64))
        sv.applyMatrix(u.astype(np.complex64), wires, inverse)
        got = sv.state()
        exp = O.apply_matrix(psi.astype(np.complex64).astype(np.complex128), n, u.astype(np.complex64), wires,
                             inverse=inverse)
        np.testing.assert_allclose(got, exp, atol=TOL_C64)
        # the same operator through the fused tile path
        L.set_mfma_dense5(False)
        ref = L.StateVectorC64(n)
        ref.updateData(psi.astype(np.complex64))
        ref.applyMatrix(u.astype(np.complex64), wires, inverse)
        L.set_mfma_dense5(True)
        np.testing.assert_allclose(got, ref.state(), atol=TOL_C64)


def test_dense5_mfma_in_a_circuit(mfma_on):
    """Gates before and after the operator (queued, flushed around it), right
    after a reset (the deferred reset is written first)."""
    n = 14
    rng = np.random.default_rng(7)
    u = rand_unitary(32, rng)
    sv = L.StateVectorC64(n)
    sv.resetStateVector()
    ops = []
    for w in range(n):
        sv.RY([w], False, [0.3 + 0.1 * w])
        ops.append(O.Op("RY", [w], [0.3 + 0.1 * w]))
    sv.CNOT([0, n - 1], False, [])
    ops.append(O.Op("CNOT", [0, n - 1], []))
    sv.applyMatrix(u.astype(np.complex64), list(range(n - 5, n)), False)
    sv.RX([n - 2], False, [0.7])
    got = sv.state()
    st = O.apply_ops(O.zero_state(n), n, ops)
    st = O.apply_matrix(st, n, u.astype(np.complex64), list(range(n - 5, n)))
    st = O.apply_ops(st, n, [O.Op("RX", [n - 2], [0.7])])
    np.testing.assert_allclose(got, st, atol=TOL_C64)


@pytest.mark.parametrize("case", ["other_wires", "controlled", "c128"])
def test_dense5_operators_the_kernel_does_not_take(mfma_on, case):
    n = 12
    rng = np.random.default_rng(11)
    u = rand_unitary(32, rng)
    psi = rand_state(n, rng)
    if case == "c128":
        sv = L.StateVectorC128(n)
        sv.updateData(psi)
        sv.applyMatrix(u, list(range(n - 5, n)), False)
        np.testing.assert_allclose(sv.state(), O.apply_matrix(psi.copy(), n, u, list(range(n - 5, n))), atol=1e-12)
        return
    sv = L.StateVectorC64(n)
    sv.updateData(psi.astype(np.complex64))
    p64 = psi.astype(np.complex64).astype(np.complex128)
    u64 = u.astype(np.complex64)
    if case == "other_wires":
        wires = [0, 3, n - 3, n - 2, n - 1]
        sv.applyMatrix(u64, wires, False)
        exp = O.apply_matrix(p64, n, u64, wires)
    else:
        wires = list(range(n - 5, n))
        sv.applyControlledMatrix(u64, [2], [True], wires, False)
        exp = O.apply_matrix(p64, n, u64, wires, ctrls=[2], cvals=[1])
    np.testing.assert_allclose(sv.state(), exp, atol=TOL_C64)
