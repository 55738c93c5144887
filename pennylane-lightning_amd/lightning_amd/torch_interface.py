"""Expectation values of a parametrised circuit as a differentiable torch function.

``circuit_expvals(params, n, ops, observables)`` runs the circuit on the
engine and returns ``<O_k>`` as a torch tensor; its backward pass is one
adjoint sweep on the device. The cotangent ``dy`` of the expvals folds into a
single Hamiltonian ``sum_k dy_k O_k`` whose adjoint Jacobian is the
vector-Jacobian product, as the reference does for expval tapes
(pennylane_lightning/lightning_base/_adjoint_jacobian.py:229-288,
``calculate_vjp``). The forward state stays on the device for the backward
pass (the role of the reference's ``_intermediate_states`` cache); the sweep
holds the final state, lambda and one observable state (the adjoint's
bra/ket/scratch), recomputing earlier states by applying inverse gates.

Parameters map to the parametric ops in order; a ``Rot`` consumes three and
is applied as RZ(phi) RY(theta) RZ(omega), the form the reference's
serializer gives the adjoint method
(pennylane_lightning/lightning_base/_serialize.py:526-529).

Observables are tuples: ``("N", name, wire)``, ``("T", [(name, wire), ...])``,
``("M", matrix, wires)`` (Hermitian) and ``("H", coeffs, [terms])``.
"""
from __future__ import annotations

import numpy as np
import torch

from . import lightning_qubit_ops as L
from .adjoint_jacobian import cotangent_observable, observable
from .workloads import CircuitOp, apply_circuit, to_ops


def adjoint_form(ops):
    """A plain Rot (not inverted, no controls) -> RZ(phi) RY(theta) RZ(omega),
    as the reference serializer decomposes ``isinstance(op, Rot)``; other ops
    unchanged (an inverted or controlled Rot keeps three parameters, which the
    adjoint method rejects as the reference does)."""
    out = []
    for op in ops:
        if op.name == "Rot" and op.matrix is None and not op.inverse and not op.ctrls:
            phi, theta, omega = op.params
            out += [CircuitOp("RZ", op.wires, [phi]), CircuitOp("RY", op.wires, [theta]),
                    CircuitOp("RZ", op.wires, [omega])]
        else:
            out.append(op)
    return out


def bind(ops, values):
    """Copy of ``ops`` with the parameters of the parametric ops replaced, in order."""
    it = iter(np.asarray(values, dtype=np.float64).tolist())
    out = []
    for op in ops:
        if op.params:
            out.append(CircuitOp(op.name, op.wires, [next(it) for _ in op.params], op.inverse, op.matrix,
                                 op.ctrls, op.cvals))
        else:
            out.append(op)
    return out


def num_params(ops) -> int:
    return sum(len(op.params) for op in ops)


class _CircuitExpvals(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, n, ops, observables, bits):
        vals = params.detach().to("cpu", torch.float64).numpy()
        bound = bind(ops, vals)
        sv = (L.StateVectorC128 if bits == 128 else L.StateVectorC64)(n)
        apply_circuit(sv, bound)
        m = (L.MeasurementsC128 if bits == 128 else L.MeasurementsC64)(sv)
        ev = np.array([m.expval(observable(o, bits)) for o in observables], dtype=np.float64)
        ctx.sv, ctx.bound, ctx.observables, ctx.bits = sv, bound, observables, bits
        ctx.param_dtype, ctx.param_device = params.dtype, params.device
        return torch.as_tensor(ev, dtype=params.dtype, device=params.device)

    @staticmethod
    def backward(ctx, grad_out):
        dy = grad_out.detach().to("cpu", torch.float64).numpy().reshape(-1)
        npar = num_params(ctx.bound)
        grad = np.zeros(npar, dtype=np.float64)
        ham = cotangent_observable(ctx.observables, dy)
        if ham is not None and npar:
            adj = (L.algorithms.AdjointJacobianC128 if ctx.bits == 128 else L.algorithms.AdjointJacobianC64)()
            grad = np.asarray(adj(ctx.sv, [observable(ham, ctx.bits)], to_ops(ctx.bound, ctx.bits),
                                  list(range(npar))), dtype=np.float64)
        ctx.sv.release_workspace()
        return (torch.as_tensor(grad, dtype=ctx.param_dtype, device=ctx.param_device), None, None, None, None)


def circuit_expvals(params: torch.Tensor, n: int, ops, observables, bits: int = 128) -> torch.Tensor:
    """<O_k> (k over ``observables``) after ``ops`` on |0...0> with the
    parametric ops' parameters taken from ``params`` in order; differentiable
    in ``params`` through the adjoint method."""
    ops = adjoint_form(ops)
    if params.numel() != num_params(ops):
        raise ValueError(f"expected {num_params(ops)} parameters, got {params.numel()}")
    return _CircuitExpvals.apply(params.reshape(-1), n, ops, list(observables), bits)

