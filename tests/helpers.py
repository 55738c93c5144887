"""Shared test helpers: conversions between the engine's op lists and the oracle's."""
import numpy as np

from oracle import lq_oracle as O

TOL_C128 = 1e-12   # amplitudes / expvals / Jacobians in complex128 (within 1e-10 required)
TOL_C64 = 5e-5     # complex64 engine vs FP64 oracle, per value, for the small test circuits


def to_oracle_ops(ops):
    return [O.Op(op.name, list(op.wires), list(op.params), bool(op.inverse),
                 None if op.matrix is None else np.asarray(op.matrix, dtype=np.complex128),
                 list(op.ctrls), list(op.cvals) if op.cvals else [1] * len(op.ctrls)) for op in ops]


def oracle_run(n, ops, psi0=None):
    s = O.zero_state(n) if psi0 is None else np.asarray(psi0, dtype=np.complex128).copy()
    return O.apply_ops(s, n, to_oracle_ops(ops))


def ham_to_oracle(coeffs, words):
    return ("H", list(coeffs), [("T", list(w)) for w in words])


def rand_state(n, rng):
    s = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    return s / np.linalg.norm(s)
