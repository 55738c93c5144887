"""complex64 runs with the FP64 error estimate and the automatic complex128
re-run (lightning_amd.precision): a bound the complex64 run meets keeps its
values (within the bound of the complex128 ones), a bound it cannot meet
forces the re-run, whose values equal the oracle's to 1e-10."""
import numpy as np
import pytest

from helpers import oracle_run
from oracle import lq_oracle as O

pytestmark = pytest.mark.gpu

P = pytest.importorskip("lightning_amd.precision")
W = pytest.importorskip("lightning_amd.workloads")

OBS = [("N", "PauliZ", 0), ("T", [("PauliX", 3), ("PauliY", 7)]), ("N", "PauliX", 11)]


def oracle_vals(n, ops):
    st = oracle_run(n, ops)
    return [O.expval_obs(st, n, o) for o in OBS]


def test_loose_bound_keeps_complex64():
    n = 16
    ops = W.hea_rot_cz(n, 6, seed=3)
    r = P.expvals_guarded(n, ops, OBS, bound=1e-4, spot_check=True)
    assert r.precision == "complex64" and not r.fallback
    assert r.estimate <= 1e-4 and r.details["norm_drift"] > 0.0
    np.testing.assert_allclose(r.values, oracle_vals(n, ops), atol=1e-4)


def test_tight_bound_forces_complex128():
    n = 16
    ops = W.hea_rot_cz(n, 6, seed=3)
    r = P.expvals_guarded(n, ops, OBS, bound=1e-13)
    assert r.fallback and r.precision == "complex128"
    assert r.estimate > 1e-13
    np.testing.assert_allclose(r.values, oracle_vals(n, ops), atol=1e-10)


def test_stated_bounds_hold_on_the_hea():
    """The stated bound of the 30-qubit configuration holds at 20 qubits of
    the same circuit: the estimate stays below it and so do the errors."""
    n = 20
    ops = W.hea_rot_cz(n, 20, seed=1)
    r = P.expvals_guarded(n, ops, OBS, bound=P.STATED_BOUNDS["hea_30"], spot_check=True)
    assert not r.fallback
    np.testing.assert_allclose(r.values, oracle_vals(n, ops), atol=P.STATED_BOUNDS["hea_30"])
