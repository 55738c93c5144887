"""Deferred product states, copies and batched gate submission on the GPU.

After a reset (or a basis state) the engine folds the leading one-qubit gates
of the next program, and global phases, into per-qubit factors of a product
state; the first specialised pass forms that state from the factors, every
other use writes it with a fill kernel. Each case compares with the oracle
(complex128 within 1e-12, complex64 within its stated tolerance). Also: a
copy out of a state is complete before the source's next write, and
``applyOperations`` matches gate-by-gate application.
"""
import math

import numpy as np
import pytest

from helpers import TOL_C128, TOL_C64, oracle_run, rand_state
from oracle import lq_oracle as O

pytestmark = pytest.mark.gpu

L = pytest.importorskip("lightning_amd.lightning_qubit_ops")
W = pytest.importorskip("lightning_amd.workloads")

SV = {128: L.StateVectorC128, 64: L.StateVectorC64}
TOL = {128: TOL_C128, 64: TOL_C64}


@pytest.fixture(autouse=True, scope="module")
def product_fold_on():
    """Folding is an engine option (off by default): these tests turn it on."""
    prev = L.set_engine_option("product_fold", 1)
    yield
    L.set_engine_option("product_fold", prev)


def one_qubit_layer(n, seed):
    rng = np.random.default_rng(seed)
    names = ["RX", "RY", "RZ", "Rot", "Hadamard", "S", "T", "SX", "PauliY", "PhaseShift"]
    ops = []
    for w in range(n):
        for _ in range(2):
            nm = names[int(rng.integers(len(names)))]
            ops.append(W.CircuitOp(nm, [w], list(rng.uniform(-math.pi, math.pi, O.GATE_NPARAMS[nm])),
                                   bool(rng.integers(0, 2))))
    ops.append(W.CircuitOp("GlobalPhase", [0], [0.37]))
    return ops


@pytest.mark.parametrize("bits", [128, 64])
@pytest.mark.parametrize("n", [12, 17])
def test_only_one_qubit_gates_after_reset(bits, n):
    """Every gate is folded: getState writes the product state itself."""
    ops = one_qubit_layer(n, seed=n)
    sv = SV[bits](n)
    sv.updateData(rand_state(n, np.random.default_rng(1)))
    sv.resetStateVector()
    W.apply_circuit(sv, ops)
    np.testing.assert_allclose(sv.state(), oracle_run(n, ops), atol=TOL[bits])


@pytest.mark.parametrize("bits", [128, 64])
def test_basis_state_then_layers_and_expval(bits):
    """Folded layer on a basis state, then entangling gates and a second
    layer; an expval right after, and once the specialised kernels are ready."""
    n = 17
    bitsv = [int(x) for x in np.random.default_rng(3).integers(0, 2, n)]
    idx = sum(b << (n - 1 - w) for w, b in enumerate(bitsv))
    psi0 = np.zeros(1 << n, dtype=np.complex128)
    psi0[idx] = 1.0
    ops = one_qubit_layer(n, seed=5) + W.hea_rot_cz(n, 2, seed=6)
    st = oracle_run(n, ops, psi0)
    exp = [O.expval_named(st, n, "PauliX", [4]), O.expval_named(st, n, "PauliZ", [16])]
    meas = {128: L.MeasurementsC128, 64: L.MeasurementsC64}[bits]
    for attempt in range(2):
        sv = SV[bits](n)
        sv.updateData(rand_state(n, np.random.default_rng(10 + attempt)))
        sv.setBasisState(bitsv, list(range(n)))
        W.apply_circuit(sv, ops)
        m = meas(sv)
        got = [m.expval("PauliX", [4]), m.expval("PauliZ", [16])]
        np.testing.assert_allclose(got, exp, atol=TOL[bits] * 10)
        np.testing.assert_allclose(sv.state(), st, atol=TOL[bits])
        if attempt == 0:
            L.jit_wait()


def test_copy_completes_before_source_is_rewritten():
    n = 22
    rng = np.random.default_rng(7)
    a, b = rand_state(n, rng), rand_state(n, rng)
    src, dst = L.StateVectorC128(n), L.StateVectorC128(n)
    src.updateData(a)
    dst.copy_from(src)
    src.updateData(b)  # enqueued right behind the copy on the source's stream
    np.testing.assert_array_equal(dst.state(), a)
    np.testing.assert_array_equal(src.state(), b)
    dst.copy_from(src)
    ops = W.hea_rot_cz(n, 1, seed=2)
    W.apply_circuit(src, ops)
    src.synchronize()
    np.testing.assert_array_equal(dst.state(), b)
    np.testing.assert_allclose(src.state(), oracle_run(n, ops, b), atol=TOL_C128)
    # a copy out of a state with a deferred reset and folded gates
    src.resetStateVector()
    W.apply_circuit(src, one_qubit_layer(n, seed=9))
    dst.copy_from(src)
    np.testing.assert_allclose(dst.state(), oracle_run(n, one_qubit_layer(n, seed=9)), atol=TOL_C128)


@pytest.mark.parametrize("bits", [128, 64])
def test_apply_operations_matches_gate_methods(bits):
    n = 9
    ops = W.hea_rot_cz(n, 2, seed=3) + [W.CircuitOp("CRX", [1, 4], [0.4], True), W.CircuitOp("Toffoli", [0, 2, 5]),
                                        W.CircuitOp("MultiRZ", [0, 3, 8], [0.9]), W.CircuitOp("SWAP", [2, 7])]
    a, b = SV[bits](n), SV[bits](n)
    a.applyOperations([o.name for o in ops], [o.wires for o in ops], [o.inverse for o in ops],
                      [o.params for o in ops])
    for o in ops:
        getattr(b, o.name)(o.wires, o.inverse, o.params)
    np.testing.assert_allclose(a.state(), b.state(), atol=TOL[bits])
    np.testing.assert_allclose(a.state(), oracle_run(n, ops), atol=TOL[bits])
    with pytest.raises(RuntimeError, match="number of operations"):
        a.applyOperations(["RX"], [[0], [1]], [False], [[0.1]])
