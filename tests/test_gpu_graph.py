"""Recorded circuits (lq_sv_record / lq_rec_replay): a circuit captured once as
a HIP graph and replayed with new parameters. Every replay is checked against
the oracle's state for its parameters (complex128 1e-12, complex64 5e-6) and
against the eager engine run of the same circuit; replays are interleaved
with eager work on other states and with measurements on the recorded one,
and run before and after the specialised kernels of the circuit are compiled
(the second plan is captured again)."""
import numpy as np
import pytest

from helpers import oracle_run
from oracle import lq_oracle as O

pytestmark = pytest.mark.gpu

L = pytest.importorskip("lightning_amd.lightning_qubit_ops")
W = pytest.importorskip("lightning_amd.workloads")


def circuit(n, params):
    it = iter(params)
    op = W.CircuitOp
    ops = []
    for layer in range(3):
        for w in range(n):
            ops += [op("RZ", [w], [next(it)]), op("RY", [w], [next(it)]), op("RZ", [w], [next(it)])]
        for s in (0, 1):
            ops += [op("CZ", [w, w + 1]) for w in range(s, n - 1, 2)]
    ops += [op("CRX", [1, n - 2], [next(it)]), op("IsingXX", [0, n - 1], [next(it)]),
            op("PhaseShift", [3], [next(it)], inverse=True)]
    rng = np.random.default_rng(9)
    ops.append(op("QubitUnitary", [2, 5], matrix=W.haar_unitary(4, rng)))
    return ops


def nparams(n):
    return 9 * n + 3


def flat(ops):
    return [p for o in ops for p in o.params]


@pytest.mark.parametrize("bits", [128, 64])
@pytest.mark.parametrize("n", [12, 17])
def test_replays_match_oracle_and_eager(bits, n):
    rng = np.random.default_rng(n + bits)
    tol = 1e-12 if bits == 128 else 5e-6
    make = L.StateVectorC128 if bits == 128 else L.StateVectorC64
    sv = make(n)
    p0 = rng.uniform(0, 2 * np.pi, nparams(n))
    ops0 = circuit(n, p0)
    sv.PauliX([4], False, [])                     # queued gates are discarded by the recording
    rec = sv.record(W.to_ops(ops0, bits))
    np.testing.assert_allclose(sv.state(), oracle_run(n, ops0), atol=tol)
    other = make(n)
    M = (L.MeasurementsC128 if bits == 128 else L.MeasurementsC64)(sv)
    for k in range(4):
        if k == 2:
            L.jit_wait()                              # specialised kernels of the plan are ready now
        pk = rng.uniform(0, 2 * np.pi, nparams(n))
        ops = circuit(n, pk)
        assert len(flat(ops)) == rec.nparams
        rec.replay(flat(ops))
        W.apply_circuit(other, ops)                   # eager work between replays
        other.synchronize()
        want = oracle_run(n, ops)
        got = sv.state()
        np.testing.assert_allclose(got, want, atol=tol)
        np.testing.assert_allclose(got, other.state(), atol=tol)
        assert abs(M.expval("PauliZ", [3]) - O.expval_named(want, n, "PauliZ", [3])) < (1e-12 if bits == 128 else 1e-5)
        other.resetStateVector()
    st = rec.stats()
    assert st["replays"] + st["captures"] == 5 and st["captures"] >= 1


def test_two_recordings_on_one_state_and_errors():
    n = 10
    sv = L.StateVectorC128(n)
    rng = np.random.default_rng(1)
    pa, pb = rng.uniform(0, 6, nparams(n)), rng.uniform(0, 6, nparams(n))
    a = sv.record(W.to_ops(circuit(n, pa), 128))
    b = sv.record(W.to_ops(W.hea_rot_cz(n, 2, seed=4), 128))
    for _ in range(2):
        a.replay(pa)
        np.testing.assert_allclose(sv.state(), oracle_run(n, circuit(n, pa)), atol=1e-12)
        b.replay(flat(W.hea_rot_cz(n, 2, seed=4)))
        np.testing.assert_allclose(sv.state(), oracle_run(n, W.hea_rot_cz(n, 2, seed=4)), atol=1e-12)
    with pytest.raises(L.LightningException, match="wrong number of parameters"):
        a.replay(pa[:-1])
    del sv                                            # the recordings keep the state alive
    a.replay(pb)
    del a, b
