"""Variances, sparse-Hamiltonian expectations and samples: engine (HIP,
through the C-ABI) vs the oracle on seeded states.

complex128 values within 1e-12 (x the operator's scale); complex64 within
5e-4; complex128 samples draw for draw; complex64 samples for at least 99% of
the shots (the float alias table is built from FP64 probabilities in both, so
rows differ only where a draw falls on a rounding boundary).
"""
import numpy as np
import pytest

from helpers import TOL_C128, TOL_C64, ham_to_oracle, oracle_run, rand_state, to_oracle_ops
from oracle import lq_oracle as O

pytestmark = pytest.mark.gpu

L = pytest.importorskip("lightning_amd.lightning_qubit_ops")
W = pytest.importorskip("lightning_amd.workloads")

SV = {128: L.StateVectorC128, 64: L.StateVectorC64}
MEAS = {128: L.MeasurementsC128, 64: L.MeasurementsC64}
CDT = {128: np.complex128, 64: np.complex64}
TOL = {128: TOL_C128, 64: TOL_C64 * 10}
NONPARAM_GATES = [g for g in O.GATE_NPARAMS if O.GATE_NPARAMS[g] == 0 and O.GATE_WIRES[g] > 0]


def state(bits, n, seed):
    rng = np.random.default_rng(seed)
    psi = rand_state(n, rng)
    sv = SV[bits](n)
    sv.updateData(psi.astype(CDT[bits]))
    # the oracle sees exactly the amplitudes the engine holds
    return sv, psi.astype(CDT[bits]).astype(np.complex128), rng


def random_csr(n, nnz, rng):
    d = 1 << n
    rows = rng.integers(0, d, nnz)
    cols = rng.integers(0, d, nnz)
    vals = rng.normal(size=nnz) + 1j * rng.normal(size=nnz)
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    indptr = np.searchsorted(rows, np.arange(d + 1)).astype(np.int64)
    return indptr, cols.astype(np.int64), vals


@pytest.mark.parametrize("bits", [128, 64])
@pytest.mark.parametrize("n", [6, 13])
def test_var_named_and_matrix(bits, n):
    sv, psi, rng = state(bits, n, 21 + n)
    M = MEAS[bits](sv)
    tol = TOL[bits] * 10
    for name in NONPARAM_GATES + ["Identity"]:
        k = O.GATE_WIRES.get(name, 1)
        wires = [int(w) for w in rng.permutation(n)[:k]]
        assert abs(M.var(name, wires) - O.var_named(psi, n, name, wires)) < tol, name
    for k in (1, 2, 3):
        a = rng.standard_normal((1 << k, 1 << k)) + 1j * rng.standard_normal((1 << k, 1 << k))
        h = a + a.conj().T
        wires = [int(w) for w in rng.permutation(n)[:k]]
        assert abs(M.var(h, wires) - O.var_matrix(psi, n, h, wires)) < tol * 100


@pytest.mark.parametrize("bits", [128, 64])
def test_var_observables_after_gates(bits):
    n = 14
    ops = W.hea_rot_cz(n, 2, seed=4)
    sv = SV[bits](n)
    W.apply_circuit(sv, ops)
    psi = oracle_run(n, ops)
    M = MEAS[bits](sv)
    sfx = f"C{bits}"
    named = getattr(L, "NamedObs" + sfx)
    tensor = getattr(L, "TensorProdObs" + sfx)
    herm = getattr(L, "HermitianObs" + sfx)
    ham = getattr(L, "Hamiltonian" + sfx)
    tol = TOL[bits] * 100
    coeffs, words = W.random_pauli_hamiltonian(n, 12, seed=3, law=(0.5, 0.2, 0.15, 0.15))
    assert abs(M.var(W.hamiltonian_obs(coeffs, words, bits)) - O.var_obs(psi, n, ham_to_oracle(coeffs, words))) < tol
    t = tensor([named("PauliZ", [0]), named("PauliY", [13])])
    assert abs(M.var(t) - O.var_obs(psi, n, ("T", [("PauliZ", 0), ("PauliY", 13)]))) < tol
    rng = np.random.default_rng(2)
    a = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
    hm = a + a.conj().T
    assert abs(M.var(ham([0.7, 0.2], [herm(hm, [3, 8]), t])) -
               O.var_obs(psi, n, ("H", [0.7, 0.2], [("M", hm, [3, 8]), ("T", [("PauliZ", 0), ("PauliY", 13)])]))) < tol


@pytest.mark.parametrize("bits", [128, 64])
@pytest.mark.parametrize("n", [5, 12])
def test_sparse_expval_var_and_observable(bits, n):
    sv, psi, rng = state(bits, n, 31 + n)
    M = MEAS[bits](sv)
    indptr, cols, vals = random_csr(n, 8 << n, rng)
    vq = vals.astype(CDT[bits]).astype(np.complex128)
    scale = float(np.abs(vals).sum())
    tol = TOL[bits] * scale
    assert abs(M.expval(indptr, cols, vals) - O.expval_sparse(psi, indptr, cols, vq)) < tol
    assert abs(M.var(indptr, cols, vals) - O.var_sparse(psi, indptr, cols, vq)) < tol * scale
    sh = getattr(L, f"SparseHamiltonianC{bits}")(vals, cols, indptr, list(range(n)))
    obs = ("S", vq, cols, indptr, list(range(n)))
    assert abs(M.expval(sh) - O.expval_obs(psi, n, obs)) < tol
    assert abs(M.var(sh) - O.var_obs(psi, n, obs)) < tol * scale
    # a sparse Hamiltonian in a Hamiltonian sum goes through the apply path
    named = getattr(L, f"NamedObsC{bits}")
    hsum = getattr(L, f"HamiltonianC{bits}")([0.5, -2.0], [sh, named("PauliX", [1])])
    osum = ("H", [0.5, -2.0], [obs, ("N", "PauliX", 1)])
    assert abs(M.expval(hsum) - O.expval_obs(psi, n, osum)) < tol * 4


def test_sparse_rejects_bad_input():
    n = 4
    sv, psi, rng = state(128, n, 3)
    M = MEAS[128](sv)
    indptr, cols, vals = random_csr(n, 20, rng)
    bad = cols.copy()
    bad[3] = 1 << n  # column outside the state: rejected on the host
    with pytest.raises(L.LightningException, match="column index out of range"):
        M.expval(indptr, bad, vals)
    with pytest.raises(L.LightningException, match="incompatible sizes"):
        M.expval(indptr[:-1], cols[: indptr[-2]], vals[: indptr[-2]])
    sh = L.SparseHamiltonianC128(vals, cols, indptr, [0, 1])
    with pytest.raises(L.LightningException, match="wire count"):
        M.expval(sh)


@pytest.mark.parametrize("bits", [128, 64])
def test_samples_match_oracle_draw_for_draw(bits):
    n = 10
    sv, psi, rng = state(bits, n, 41)
    M = MEAS[bits](sv)
    M.set_random_seed(1234)
    for wires, shots in ((None, 500), ([7, 0, 3], 800), ([9], 300)):
        got = M.generate_samples(n, shots) if wires is None else M.generate_samples(wires, shots)
        exp = O.generate_samples(psi, n, wires, shots, 1234, single=bits == 64)
        assert got.shape == exp.shape
        same = (got == exp).all(axis=1).mean()
        if bits == 128:
            assert same == 1.0, wires
        else:
            assert same >= 0.99, (wires, same)
    # the seed is applied again on every call
    a = M.generate_samples([1, 2], 64)
    b = M.generate_samples([1, 2], 64)
    assert np.array_equal(a, b)


def test_samples_statistics_at_scale():
    n = 20
    ops = W.hea_rot_cz(n, 2, seed=8)
    sv = SV[128](n)
    W.apply_circuit(sv, ops)
    M = MEAS[128](sv)
    M.set_random_seed(5)
    shots = 40000
    smp = M.generate_samples([0, 5, 19], shots)
    idx = (smp * np.array([4, 2, 1], dtype=np.uint64)).sum(axis=1).astype(np.int64)
    freq = np.bincount(idx, minlength=8) / shots
    p = M.probs([0, 5, 19])
    assert (np.abs(freq - p) <= 5 * np.sqrt(p * (1 - p) / shots) + 1e-12).all()


@pytest.mark.parametrize("bits", [128, 64])
def test_apply_sparse_matrix(bits):
    n = 12
    rng = np.random.default_rng(51)
    for k, ctrls in ((1, []), (3, []), (2, [7]), (5, [0, 11])):
        sv, psi, _ = state(bits, n, 60 + k)
        wires = [int(w) for w in rng.permutation(n) if w not in ctrls][:k]
        cvals = [bool(x) for x in rng.integers(0, 2, len(ctrls))]
        indptr, cols, vals = random_csr(k, 3 << k, rng)
        vq = vals.astype(CDT[bits]).astype(np.complex128)
        if ctrls:
            sv.applyControlledSparseMatrix(indptr, cols, vals, ctrls, cvals, wires)
        else:
            sv.applySparseMatrix(indptr, cols, vals, wires)
        exp = O.apply_sparse_matrix(psi.copy(), n, indptr, cols, vq, wires, ctrls, [int(c) for c in cvals])
        got = np.empty(1 << n, dtype=CDT[bits])
        sv.getState(got)
        np.testing.assert_allclose(got, exp, atol=TOL[bits] * 10 * float(np.abs(vals).max()) * 4)
    with pytest.raises(L.LightningException, match="Inverse not implemented"):
        sv.applySparseMatrix(indptr, cols, vals, wires, inverse=True)

