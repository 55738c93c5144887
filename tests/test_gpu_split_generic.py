"""Split plans in the generic register kernel.

Specialised passes run a plan whose 2x2 gates are split into phase / real
rotation / phase, with phases gathered into register tables (TOP_RPHASE) and
per-thread products (TOP_TPHASE) and rotations stored as [[1, -t], [t, 1]]
with their cosines in a table. While some of a program's kernels are still
compiling, that plan's missing passes run in the generic kernel. Here the
split plan is forced (set_split_fallback(2)) on programs whose kernels are
not compiled yet, so every pass runs generically, and the results are
compared with the oracle; then again with the specialised kernels.
"""
import numpy as np
import pytest

from helpers import TOL_C128, TOL_C64, ham_to_oracle, oracle_run, rand_state, to_oracle_ops
from oracle import lq_oracle as O

pytestmark = pytest.mark.gpu

L = pytest.importorskip("lightning_amd.lightning_qubit_ops")
W = pytest.importorskip("lightning_amd.workloads")

SV = {128: L.StateVectorC128, 64: L.StateVectorC64}
TOL = {128: TOL_C128, 64: TOL_C64 * 10}


@pytest.fixture
def split_generic():
    L.jit_wait()
    prev = L.set_split_fallback(2)
    yield
    L.set_split_fallback(prev)


@pytest.mark.parametrize("bits", [128, 64])
def test_split_plan_generic_then_specialised(split_generic, bits):
    n = 17
    # a structure no other test compiles: HEA layers, then RZ/PhaseShift
    # phases on every wire (thread-bit phase products) and a CZ ring
    ops = W.hea_rot_cz(n, 4, seed=31)
    ops += [W.CircuitOp("RZ" if w % 2 else "PhaseShift", [w], [0.17 * (w + 1)]) for w in range(n)]
    ops += [W.CircuitOp("CZ", [w, (w + 3) % n]) for w in range(n)]
    ops += [W.CircuitOp("RY", [w], [0.05 * (w + 2)]) for w in range(n)]
    psi0 = rand_state(n, np.random.default_rng(32))
    exp = oracle_run(n, ops, psi0)
    counts = []
    for attempt in range(2):
        sv = SV[bits](n)
        sv.updateData(psi0.astype(np.complex128 if bits == 128 else np.complex64))
        sv.pass_counts(reset=True)
        W.apply_circuit(sv, ops)
        np.testing.assert_allclose(sv.state(), exp, atol=TOL[bits])
        counts.append(sv.pass_counts(reset=True))
        if attempt == 0:
            L.jit_wait()
    # passes shared with programs compiled earlier may already be specialised
    assert counts[0]["generic"] > 0, counts
    assert counts[1]["specialised"] > 0, counts


def test_split_plan_generic_adjoint(split_generic):
    # two-state passes with generator dots, in the generic kernel
    n = 16
    ops = W.strongly_entangling(n, 2, seed=41)
    coeffs, words = W.random_pauli_hamiltonian(n, 8, seed=42, law=(0.5, 0.2, 0.15, 0.15))
    tp = list(range(0, sum(1 for op in ops if op.params), 2))
    st = oracle_run(n, ops)
    jexp = np.asarray(O.adjoint_jacobian(st, n, to_oracle_ops(ops), [ham_to_oracle(coeffs, words)], tp)).reshape(-1)
    sv = L.StateVectorC128(n)
    W.apply_circuit(sv, ops)
    H = W.hamiltonian_obs(coeffs, words, 128)
    jac = np.asarray(L.algorithms.AdjointJacobianC128()(sv, [H], W.to_ops(ops, 128), tp)).reshape(-1)
    np.testing.assert_allclose(jac, jexp, atol=1e-10)
