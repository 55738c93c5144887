"""Seeded circuits of the measured workloads, as plain op lists.

The same lists feed the engine, the oracle in tests and the reference build,
so every engine consumes identical inputs. An op is a ``CircuitOp``; ``to_ops``
turns a list into an ``OpsStructC*`` for the adjoint Jacobian.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


@dataclass
class CircuitOp:
    name: str
    wires: list
    params: list = field(default_factory=list)
    inverse: bool = False
    matrix: np.ndarray | None = None
    ctrls: list = field(default_factory=list)
    cvals: list = field(default_factory=list)


def layers_rxyz_cnot(n: int = 20, layers: int = 10, seed: int = 0) -> list[CircuitOp]:
    """|0..0>, then per layer RX, RY, RZ on every wire (angles U[0, 2pi) from
    ``seed``) and a CNOT ladder (i, i+1), i = 0..n-2."""
    rng = np.random.default_rng(seed)
    ang = rng.uniform(0.0, 2 * math.pi, size=(layers, n, 3))
    ops = []
    for layer in range(layers):
        for w in range(n):
            for g, name in enumerate(("RX", "RY", "RZ")):
                ops.append(CircuitOp(name, [w], [float(ang[layer, w, g])]))
        for w in range(n - 1):
            ops.append(CircuitOp("CNOT", [w, w + 1]))
    return ops


def hea_rot_cz(n: int = 30, layers: int = 20, seed: int = 1) -> list[CircuitOp]:
    """Hardware-efficient ansatz: Rot(phi, theta, omega) on every wire (angles
    U[0, 2pi) from ``seed``), then CZ(i, i+1) for even i, then for odd i."""
    rng = np.random.default_rng(seed)
    ang = rng.uniform(0.0, 2 * math.pi, size=(layers, n, 3))
    ops = []
    for layer in range(layers):
        for w in range(n):
            ops.append(CircuitOp("Rot", [w], [float(x) for x in ang[layer, w]]))
        for w in range(0, n - 1, 2):
            ops.append(CircuitOp("CZ", [w, w + 1]))
        for w in range(1, n - 1, 2):
            ops.append(CircuitOp("CZ", [w, w + 1]))
    return ops


def hea_rot_cz_adjoint_form(n: int, layers: int, seed: int) -> list[CircuitOp]:
    """``hea_rot_cz`` with each Rot written as RZ(phi) RY(theta) RZ(omega), the
    form the adjoint method differentiates
    (pennylane_lightning/lightning_base/_serialize.py:526-529)."""
    out = []
    for op in hea_rot_cz(n, layers, seed):
        if op.name == "Rot":
            phi, theta, omega = op.params
            out += [CircuitOp("RZ", op.wires, [phi]), CircuitOp("RY", op.wires, [theta]),
                    CircuitOp("RZ", op.wires, [omega])]
        else:
            out.append(op)
    return out


def strongly_entangling(n: int = 24, layers: int = 6, seed: int = 2) -> list[CircuitOp]:
    """Strongly entangling layers in adjoint form: per layer and wire
    RZ(phi) RY(theta) RZ(omega) (angles U[0, 2pi) from ``seed``), then
    CNOT(i, (i + r) mod n) with r = layer mod (n - 1) + 1."""
    rng = np.random.default_rng(seed)
    ang = rng.uniform(0.0, 2 * math.pi, size=(layers, n, 3))
    ops = []
    for layer in range(layers):
        for w in range(n):
            phi, theta, omega = (float(x) for x in ang[layer, w])
            ops += [CircuitOp("RZ", [w], [phi]), CircuitOp("RY", [w], [theta]), CircuitOp("RZ", [w], [omega])]
        r = layer % (n - 1) + 1
        for w in range(n):
            ops.append(CircuitOp("CNOT", [w, (w + r) % n]))
    return ops


def random_pauli_hamiltonian(n: int, terms: int, seed: int, law=(0.7, 0.1, 0.1, 0.1)):
    """Random Pauli sum: each wire I/X/Y/Z with probabilities ``law``,
    coefficients N(0, 1). Returns (coeffs, words) with each word a list of
    (name, wire) factors; an all-identity word is [("Identity", 0)]."""
    rng = np.random.default_rng(seed)
    names = ("Identity", "PauliX", "PauliY", "PauliZ")
    coeffs = rng.standard_normal(terms)
    words = []
    for _ in range(terms):
        letters = rng.choice(4, size=n, p=list(law))
        f = [(names[c], w) for w, c in enumerate(letters) if c != 0]
        words.append(f if f else [("Identity", 0)])
    return [float(c) for c in coeffs], words


def haar_unitary(dim: int, rng: np.random.Generator) -> np.ndarray:
    """Haar-random unitary: QR of a complex Gaussian with the phase fix."""
    z = (rng.standard_normal((dim, dim)) + 1j * rng.standard_normal((dim, dim))) / math.sqrt(2.0)
    q, r = np.linalg.qr(z)
    d = np.diag(r)
    return q * (d / np.abs(d))


def qft_brickwork_mcx(n: int, layers: int = 8, mcx_per_layer: int = 1, seed: int = 5) -> list[CircuitOp]:
    """QFT (H and ControlledPhaseShift(pi/2^k) on all pairs, final swaps),
    then ``layers`` layers of Haar 4x4 unitaries on brickwork pairs (applied
    as matrices) each followed by ``mcx_per_layer`` PauliX gates with 3
    controls on random wires."""
    rng = np.random.default_rng(seed)
    ops = []
    for i in range(n):
        ops.append(CircuitOp("Hadamard", [i]))
        for j in range(i + 1, n):
            ops.append(CircuitOp("ControlledPhaseShift", [j, i], [math.pi / (2 ** (j - i))]))
    for i in range(n // 2):
        ops.append(CircuitOp("SWAP", [i, n - 1 - i]))
    for layer in range(layers):
        start = layer % 2
        for a in range(start, n - 1, 2):
            ops.append(CircuitOp("QubitUnitary", [a, a + 1], matrix=haar_unitary(4, rng)))
        for _ in range(mcx_per_layer):
            w = [int(x) for x in rng.choice(n, size=4, replace=False)]
            ops.append(CircuitOp("PauliX", [w[3]], ctrls=w[:3], cvals=[1, 1, 1]))
    return ops


def apply_circuit(sv, ops: list[CircuitOp]) -> None:
    """Apply an op list to a StateVectorC*: runs of plain named gates with
    one ``applyOperations`` call each, matrices and controlled gates through
    their own methods."""
    run = []

    def flush_run():
        if run:
            sv.applyOperations([o.name for o in run], [o.wires for o in run], [o.inverse for o in run],
                               [o.params for o in run])
            run.clear()

    for op in ops:
        if op.matrix is None and not op.ctrls:
            run.append(op)
            continue
        flush_run()
        if op.matrix is not None:
            if op.ctrls:
                sv.applyControlledMatrix(op.matrix, op.ctrls, op.cvals or [True] * len(op.ctrls), op.wires,
                                         op.inverse)
            else:
                sv.applyMatrix(op.matrix, op.wires, op.inverse)
        else:
            getattr(sv, op.name)(op.ctrls, op.cvals or [True] * len(op.ctrls), op.wires, op.inverse, op.params)
    flush_run()


def to_ops(ops: list[CircuitOp], bits: int = 128):
    """OpsStructC64/C128 of an op list (the adjoint Jacobian's input)."""
    from . import lightning_qubit_ops as L
    cls = L.OpsStructC128 if bits == 128 else L.OpsStructC64
    return cls([o.name for o in ops], [o.params for o in ops], [o.wires for o in ops], [o.inverse for o in ops],
               [o.matrix if o.matrix is not None else [] for o in ops], [o.ctrls for o in ops],
               [o.cvals or [True] * len(o.ctrls) for o in ops])


def hamiltonian_obs(coeffs, words, bits: int = 128):
    """HamiltonianC*(coeffs, [TensorProdObs(NamedObs...)]) as the serializer builds it
    (pennylane_lightning/lightning_base/_serialize.py:358-375)."""
    from . import lightning_qubit_ops as L
    sfx = f"C{bits}"
    named = getattr(L, "NamedObs" + sfx)
    tensor = getattr(L, "TensorProdObs" + sfx)
    ham = getattr(L, "Hamiltonian" + sfx)
    terms = [tensor([named(nm, [w]) for nm, w in word]) for word in words]
    return ham(coeffs, terms)
