"""ctypes access to oracle/_ref/liblq_ref.so -- TEST INFRASTRUCTURE ONLY.

The library is the reference lightning.qubit C++ core compiled from its own
sources (oracle/Makefile) plus oracle/ref_driver.cpp. Used by tests/ to pin
oracle/lq_oracle.py, by tests/golden/make_golden.py to write fixtures, and by
bench.py's cpu_baseline leg. ``available()`` is False where it was not built.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_ref", "liblq_ref.so")
_lib = None


def available() -> bool:
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise RuntimeError(f"{LIB_PATH} is not built (make -C oracle ref)")
        L = C.CDLL(LIB_PATH)
        L.ref_last_error.restype = C.c_char_p
        L.ref_sv_new.restype = C.c_void_p
        L.ref_sv_new.argtypes = [C.c_uint32, C.c_int]
        L.ref_sv_free.argtypes = [C.c_void_p, C.c_int]
        _lib = L
    return _lib


def _u32(xs: Sequence[int]):
    a = np.ascontiguousarray(np.asarray(list(xs), dtype=np.uint32))
    return a, a.ctypes.data_as(C.POINTER(C.c_uint32))


def _u8(xs: Sequence[int]):
    a = np.ascontiguousarray(np.asarray([int(bool(x)) for x in xs], dtype=np.uint8))
    return a, a.ctypes.data_as(C.POINTER(C.c_uint8))


def _check(rc: int):
    if rc != 0:
        raise RuntimeError(lib().ref_last_error().decode())


class RefStateVector:
    """The reference StateVectorLQubitManaged<double|float>."""

    def __init__(self, n: int, precision: str = "double"):
        self.n = n
        self.prec = 1 if precision == "double" else 0
        self.cdtype = np.complex128 if self.prec == 1 else np.complex64
        self.h = lib().ref_sv_new(n, self.prec)
        if not self.h:
            raise RuntimeError(lib().ref_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ref_sv_free(C.c_void_p(self.h), self.prec)
            self.h = None

    def _h(self):
        return C.c_void_p(self.h)

    def set_state(self, data: np.ndarray):
        a = np.ascontiguousarray(np.asarray(data, dtype=self.cdtype))
        _check(lib().ref_sv_set(self._h(), self.prec, a.ctypes.data_as(C.c_void_p), C.c_uint64(a.size)))

    def get_state(self) -> np.ndarray:
        out = np.empty(1 << self.n, dtype=self.cdtype)
        _check(lib().ref_sv_get(self._h(), self.prec, out.ctypes.data_as(C.c_void_p), C.c_uint64(out.size)))
        return out

    def apply(self, name: str, wires, inverse=False, params=(), ctrls=(), cvals=()):
        w, wp = _u32(wires)
        c, cp = _u32(ctrls)
        v, vp = _u8(cvals if len(cvals) else [1] * len(ctrls))
        p = np.ascontiguousarray(np.asarray(list(params), dtype=np.float64))
        _check(lib().ref_apply(self._h(), self.prec, name.encode(), wp, C.c_uint32(len(w)), cp, vp,
                               C.c_uint32(len(c)), int(bool(inverse)),
                               p.ctypes.data_as(C.POINTER(C.c_double)), C.c_uint32(len(p))))

    def apply_matrix(self, mat, wires, inverse=False, ctrls=(), cvals=()):
        m = np.ascontiguousarray(np.asarray(mat, dtype=self.cdtype).reshape(-1))
        w, wp = _u32(wires)
        c, cp = _u32(ctrls)
        v, vp = _u8(cvals if len(cvals) else [1] * len(ctrls))
        _check(lib().ref_apply_matrix(self._h(), self.prec, m.ctypes.data_as(C.c_void_p), wp,
                                      C.c_uint32(len(w)), cp, vp, C.c_uint32(len(c)), int(bool(inverse))))

    def apply_pauli_rot(self, wires, theta, word, inverse=False):
        w, wp = _u32(wires)
        _check(lib().ref_apply_pauli_rot(self._h(), self.prec, wp, C.c_uint32(len(w)), int(bool(inverse)),
                                         C.c_double(theta), word.encode()))

    def apply_generator(self, name, wires, ctrls=(), cvals=(), adj=False) -> float:
        w, wp = _u32(wires)
        c, cp = _u32(ctrls)
        v, vp = _u8(cvals if len(cvals) else [1] * len(ctrls))
        s = C.c_double()
        _check(lib().ref_apply_generator(self._h(), self.prec, name.encode(), wp, C.c_uint32(len(w)), cp, vp,
                                         C.c_uint32(len(c)), int(bool(adj)), C.byref(s)))
        return s.value

    def normalize(self):
        _check(lib().ref_normalize(self._h(), self.prec))

    def collapse(self, wire, branch):
        _check(lib().ref_collapse(self._h(), self.prec, C.c_uint32(wire), int(bool(branch))))

    def set_basis_state(self, bits, wires):
        b, bp = _u32(bits)
        w, wp = _u32(wires)
        _check(lib().ref_set_basis_state(self._h(), self.prec, bp, wp, C.c_uint32(len(w))))

    def set_state_vector(self, state, wires):
        s = np.ascontiguousarray(np.asarray(state, dtype=self.cdtype))
        w, wp = _u32(wires)
        _check(lib().ref_set_state_vector(self._h(), self.prec, s.ctypes.data_as(C.c_void_p), wp,
                                          C.c_uint32(len(w))))

    def probs(self, wires=None) -> np.ndarray:
        allw = wires is None
        wl = [] if allw else list(wires)
        w, wp = _u32(wl)
        out = np.empty(1 << (self.n if allw else len(wl)), dtype=np.float64)
        _check(lib().ref_probs(self._h(), self.prec, wp, C.c_uint32(len(wl)), int(allw),
                               out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def expval_named(self, name, wires) -> float:
        w, wp = _u32(wires)
        out = C.c_double()
        _check(lib().ref_expval_named(self._h(), self.prec, name.encode(), wp, C.c_uint32(len(w)), C.byref(out)))
        return out.value

    def expval_matrix(self, mat, wires) -> float:
        m = np.ascontiguousarray(np.asarray(mat, dtype=self.cdtype).reshape(-1))
        w, wp = _u32(wires)
        out = C.c_double()
        _check(lib().ref_expval_matrix(self._h(), self.prec, m.ctypes.data_as(C.c_void_p), wp,
                                       C.c_uint32(len(w)), C.byref(out)))
        return out.value

    def expval_obs(self, obs) -> float:
        from .lq_oracle import obs_to_spec
        out = C.c_double()
        _check(lib().ref_expval_obs(self._h(), self.prec, obs_to_spec(obs).encode(), C.byref(out)))
        return out.value

    def apply_sparse(self, indptr, indices, data, wires, ctrls=(), cvals=()):
        ip = np.ascontiguousarray(np.asarray(indptr, dtype=np.int64))
        ix = np.ascontiguousarray(np.asarray(indices, dtype=np.int64))
        va = np.ascontiguousarray(np.asarray(data, dtype=np.complex128))
        w, wp = _u32(wires)
        c, cp = _u32(ctrls)
        v, vp = _u8(cvals if len(cvals) else [1] * len(ctrls))
        i64 = C.POINTER(C.c_int64)
        _check(lib().ref_apply_sparse(self._h(), self.prec, ip.ctypes.data_as(i64), C.c_uint64(ip.size),
                                      ix.ctypes.data_as(i64), va.ctypes.data_as(C.c_void_p), C.c_uint64(va.size),
                                      wp, C.c_uint32(len(w)), cp, vp, C.c_uint32(len(c))))

    def var_named(self, name, wires) -> float:
        w, wp = _u32(wires)
        out = C.c_double()
        _check(lib().ref_var_named(self._h(), self.prec, name.encode(), wp, C.c_uint32(len(w)), C.byref(out)))
        return out.value

    def var_obs(self, obs) -> float:
        from .lq_oracle import obs_to_spec
        out = C.c_double()
        _check(lib().ref_var_obs(self._h(), self.prec, obs_to_spec(obs).encode(), C.byref(out)))
        return out.value

    def _sparse(self, variance, indptr, indices, data) -> float:
        ip = np.ascontiguousarray(np.asarray(indptr, dtype=np.int64))
        ix = np.ascontiguousarray(np.asarray(indices, dtype=np.int64))
        va = np.ascontiguousarray(np.asarray(data, dtype=np.complex128))
        out = C.c_double()
        i64 = C.POINTER(C.c_int64)
        _check(lib().ref_sparse(self._h(), self.prec, int(variance), ip.ctypes.data_as(i64), C.c_uint64(ip.size),
                                ix.ctypes.data_as(i64), va.ctypes.data_as(C.c_void_p), C.c_uint64(va.size),
                                C.byref(out)))
        return out.value

    def expval_sparse(self, indptr, indices, data) -> float:
        return self._sparse(False, indptr, indices, data)

    def var_sparse(self, indptr, indices, data) -> float:
        return self._sparse(True, indptr, indices, data)

    def generate_samples(self, wires, shots: int, seed: int) -> np.ndarray:
        """shots x len(wires) (all wires when wires is None) after setSeed(seed)."""
        allw = wires is None
        wl = [] if allw else list(wires)
        w, wp = _u32(wl)
        k = self.n if allw else len(wl)
        out = np.empty((shots, k), dtype=np.uint64)
        _check(lib().ref_generate_samples(self._h(), self.prec, wp, C.c_uint32(len(wl)), int(allw),
                                          C.c_uint64(shots), C.c_uint64(seed),
                                          out.ctypes.data_as(C.POINTER(C.c_uint64))))
        return out

    def apply_ops(self, ops):
        for op in ops:
            if op.matrix is not None and op.name not in _known():
                self.apply_matrix(op.matrix, op.wires, op.inverse, op.ctrls, op.cvals)
            else:
                self.apply(op.name, op.wires, op.inverse, op.params, op.ctrls, op.cvals)

    def adjoint_jacobian(self, ops, observables, trainable) -> np.ndarray:
        """AdjointJacobian on this (final) state; returns (nobs, ntp)."""
        from .lq_oracle import obs_to_spec
        nops = len(ops)
        names = (C.c_char_p * max(nops, 1))(*[op.name.encode() for op in ops])
        poff, woff, moff, coff = [0], [0], [0], [0]
        params, wires, mats, ctrls, cvals, invs = [], [], [], [], [], []
        for op in ops:
            params += [float(x) for x in op.params]
            poff.append(len(params))
            wires += list(op.wires)
            woff.append(len(wires))
            if op.matrix is not None:
                m = np.asarray(op.matrix, dtype=np.complex128).reshape(-1)
                mats += [x for z in m for x in (z.real, z.imag)]
            moff.append(len(mats) // 2)
            ctrls += list(op.ctrls)
            cvals += [int(bool(v)) for v in (op.cvals if len(op.cvals) else [1] * len(op.ctrls))]
            coff.append(len(ctrls))
            invs.append(int(bool(op.inverse)))
        arrs = dict(
            poff=np.asarray(poff, np.uint32), params=np.asarray(params + [0.0], np.float64),
            woff=np.asarray(woff, np.uint32), wires=np.asarray(wires + [0], np.uint32),
            invs=np.asarray(invs + [0], np.uint8), moff=np.asarray(moff, np.uint32),
            mats=np.asarray(mats + [0.0], np.float64), coff=np.asarray(coff, np.uint32),
            ctrls=np.asarray(ctrls + [0], np.uint32), cvals=np.asarray(cvals + [0], np.uint8))
        P = lambda a, t: a.ctypes.data_as(C.POINTER(t))
        specs = [obs_to_spec(o).encode() for o in observables]
        spec_arr = (C.c_char_p * max(len(specs), 1))(*specs)
        tp = np.asarray(list(trainable) + [0], dtype=np.uint64)
        jac = np.zeros(max(len(observables) * len(trainable), 1), dtype=np.float64)
        _check(lib().ref_adjoint(
            self._h(), self.prec, C.c_uint32(nops), names,
            P(arrs["poff"], C.c_uint32), P(arrs["params"], C.c_double),
            P(arrs["woff"], C.c_uint32), P(arrs["wires"], C.c_uint32),
            P(arrs["invs"], C.c_uint8), P(arrs["moff"], C.c_uint32),
            P(arrs["mats"], C.c_double), P(arrs["coff"], C.c_uint32),
            P(arrs["ctrls"], C.c_uint32), P(arrs["cvals"], C.c_uint8),
            C.c_uint32(len(specs)), spec_arr, P(tp, C.c_uint64), C.c_uint32(len(trainable)),
            P(jac, C.c_double)))
        return jac[: len(observables) * len(trainable)].reshape(len(observables), len(trainable))

    def vector_jacobian_product(self, ops, dy, trainable) -> np.ndarray:
        """VectorJacobianProduct on this (final) state; returns len(trainable) complex."""
        a = _pack_ops(ops)
        names = (C.c_char_p * max(len(ops), 1))(*[op.name.encode() for op in ops])
        P = lambda arr, t: arr.ctypes.data_as(C.POINTER(t))
        d = np.ascontiguousarray(np.asarray(dy).reshape(-1),
                                 dtype=np.complex128 if self.prec == 1 else np.complex64)
        tp = np.asarray(list(trainable) + [0], dtype=np.uint64)
        out = np.zeros(2 * max(len(trainable), 1), dtype=np.float64)
        _check(lib().ref_vjp(
            self._h(), self.prec, C.c_uint32(len(ops)), names,
            P(a["poff"], C.c_uint32), P(a["params"], C.c_double), P(a["woff"], C.c_uint32),
            P(a["wires"], C.c_uint32), P(a["invs"], C.c_uint8), P(a["moff"], C.c_uint32),
            P(a["mats"], C.c_double), P(a["coff"], C.c_uint32), P(a["ctrls"], C.c_uint32),
            P(a["cvals"], C.c_uint8), d.ctypes.data_as(C.c_void_p), C.c_uint64(d.size),
            P(tp, C.c_uint64), C.c_uint32(len(trainable)), P(out, C.c_double)))
        return out[: 2 * len(trainable)].view(np.complex128).copy()


def _pack_ops(ops):
    poff, woff, moff, coff = [0], [0], [0], [0]
    params, wires, mats, ctrls, cvals, invs = [], [], [], [], [], []
    for op in ops:
        params += [float(x) for x in op.params]
        poff.append(len(params))
        wires += list(op.wires)
        woff.append(len(wires))
        if op.matrix is not None:
            m = np.asarray(op.matrix, dtype=np.complex128).reshape(-1)
            mats += [x for z in m for x in (z.real, z.imag)]
        moff.append(len(mats) // 2)
        ctrls += list(op.ctrls)
        cvals += [int(bool(v)) for v in (op.cvals if len(op.cvals) else [1] * len(op.ctrls))]
        coff.append(len(ctrls))
        invs.append(int(bool(op.inverse)))
    return dict(
        poff=np.asarray(poff, np.uint32), params=np.asarray(params + [0.0], np.float64),
        woff=np.asarray(woff, np.uint32), wires=np.asarray(wires + [0], np.uint32),
        invs=np.asarray(invs + [0], np.uint8), moff=np.asarray(moff, np.uint32),
        mats=np.asarray(mats + [0.0], np.float64), coff=np.asarray(coff, np.uint32),
        ctrls=np.asarray(ctrls + [0], np.uint32), cvals=np.asarray(cvals + [0], np.uint8))


def _known():
    from .lq_oracle import GATE_NPARAMS
    return GATE_NPARAMS
