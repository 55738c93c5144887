"""Op lists of the golden fixtures, rebuilt from the inputs stored in each file.

Ops are oracle ``Op`` records (name, wires, params, inverse, matrix, ctrls,
cvals), the form of the reference's serialized ops lists
(pennylane_lightning/lightning_base/_serialize.py:428-571); ``to_circuit``
turns them into the engine's ``CircuitOp``. Observables use the oracle's
tuple form: ("N", name, wire), ("T", [(name, wire), ...]), ("H", coeffs, [obs]).
"""
import math

import numpy as np

from oracle import lq_oracle as O

PAULI = ("Identity", "PauliX", "PauliY", "PauliZ")


def rxyz_cnot_ops(ang):
    layers, n, _ = ang.shape
    ops = []
    for layer in range(layers):
        for w in range(n):
            for g, name in enumerate(("RX", "RY", "RZ")):
                ops.append(O.Op(name, [w], [float(ang[layer, w, g])]))
        for w in range(n - 1):
            ops.append(O.Op("CNOT", [w, w + 1]))
    return ops


def hea_ops(ang):
    layers, n, _ = ang.shape
    ops = []
    for layer in range(layers):
        for w in range(n):
            ops.append(O.Op("Rot", [w], [float(x) for x in ang[layer, w]]))
        for w in range(0, n - 1, 2):
            ops.append(O.Op("CZ", [w, w + 1]))
        for w in range(1, n - 1, 2):
            ops.append(O.Op("CZ", [w, w + 1]))
    return ops


def strongly_entangling_ops(ang):
    layers, n, _ = ang.shape
    ops = []
    for layer in range(layers):
        for w in range(n):
            phi, theta, omega = (float(x) for x in ang[layer, w])
            ops += [O.Op("RZ", [w], [phi]), O.Op("RY", [w], [theta]), O.Op("RZ", [w], [omega])]
        r = layer % (n - 1) + 1
        for w in range(n):
            ops.append(O.Op("CNOT", [w, (w + r) % n]))
    return ops


def random_words(n, terms, seed, law=(0.7, 0.1, 0.1, 0.1)):
    """Pauli letters (terms x n, 0..3 = I, X, Y, Z) and N(0, 1) coefficients."""
    rng = np.random.default_rng(seed)
    coeffs = rng.standard_normal(terms)
    letters = rng.choice(4, size=(terms, n), p=list(law)).astype(np.int8)
    return letters, coeffs


def word_factors(row):
    f = [(PAULI[int(c)], w) for w, c in enumerate(row) if c != 0]
    return f if f else [("Identity", 0)]


def ham_spec(coeffs, letters):
    return ("H", [float(c) for c in coeffs], [("T", word_factors(row)) for row in letters])


def brickwork_inputs(n, layers, rng):
    """Haar 4x4 unitaries (layers x pairs x 4 x 4, unused slots zero) and the
    wires of one 3-controlled X per layer (layers x 4: controls, target)."""
    pairs = n // 2
    mats = np.zeros((layers, pairs, 4, 4), dtype=np.complex128)
    mcx = np.zeros((layers, 4), dtype=np.int64)
    for layer in range(layers):
        start = layer % 2
        for k, a in enumerate(range(start, n - 1, 2)):
            z = (rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))) / math.sqrt(2.0)
            q, r = np.linalg.qr(z)
            d = np.diag(r)
            mats[layer, k] = q * (d / np.abs(d))
        mcx[layer] = rng.choice(n, size=4, replace=False)
    return mats, mcx


def qft_brickwork_ops(n, mats, mcx):
    ops = []
    for i in range(n):
        ops.append(O.Op("Hadamard", [i]))
        for j in range(i + 1, n):
            ops.append(O.Op("ControlledPhaseShift", [j, i], [math.pi / (2 ** (j - i))]))
    for i in range(n // 2):
        ops.append(O.Op("SWAP", [i, n - 1 - i]))
    for layer in range(mats.shape[0]):
        start = layer % 2
        for k, a in enumerate(range(start, n - 1, 2)):
            ops.append(O.Op("QubitUnitary", [a, a + 1], [], False, mats[layer, k]))
        w = [int(x) for x in mcx[layer]]
        ops.append(O.Op("PauliX", [w[3]], [], False, None, w[:3], [1, 1, 1]))
    return ops


def to_circuit(ops):
    """Oracle ops -> engine CircuitOps."""
    from lightning_amd.workloads import CircuitOp
    return [CircuitOp(o.name, list(o.wires), list(o.params), bool(o.inverse), o.matrix, list(o.ctrls),
                      [bool(v) for v in o.cvals]) for o in ops]


def ham_words(letters):
    """Pauli words as lists of (name, wire) for the engine's Hamiltonian builder."""
    return [word_factors(row) for row in letters]
