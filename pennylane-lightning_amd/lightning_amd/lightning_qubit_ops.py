"""Drop-in mirror of ``pennylane_lightning.lightning_qubit_ops`` on MI355X.

Class and method names, argument meaning and error behaviour follow the
reference bindings (pennylane_lightning/core/bindings/Bindings.hpp,
pennylane_lightning/core/simulators/lightning_qubit/bindings/LQubitBindings.hpp):

* ``StateVectorC64`` / ``StateVectorC128`` (Bindings.hpp:870-922,
  LQubitBindings.hpp:142-243): one method per gate name with the signatures
  ``Gate(wires, inverse, params)`` and
  ``Gate(controlled_wires, controlled_values, wires, inverse=False, params=[])``,
  ``applyMatrix``, ``applyControlledMatrix``, ``applyPauliRot``,
  ``setBasisState``, ``setStateVector``, ``resetStateVector``,
  ``updateData``, ``getState``, ``normalize``, ``collapse``, ``kernel_map``.
* ``MeasurementsC64`` / ``MeasurementsC128`` (Bindings.hpp:650-688,
  LQubitBindings.hpp:268-350): ``probs``, ``expval`` (observable, name or
  matrix form).
* ``observables.*`` (Bindings.hpp:467-591) and ``algorithms.*``
  (Bindings.hpp:698-851): ``OpsStructC*``, ``create_ops_listC*``,
  ``AdjointJacobianC*``.

Failures raise ``LightningException`` (a ``RuntimeError``) with the
reference's message format. The state lives in HBM; gate calls are queued
and run as fused passes when a result is read.
"""
from __future__ import annotations

import ctypes as C
import sys
import types
from typing import Sequence

import numpy as np

from . import _capi
from ._capi import LightningException, check, f64, lib, u8, u32

GATE_NAMES = [
    "Identity", "PauliX", "PauliY", "PauliZ", "Hadamard", "S", "SX", "T", "PhaseShift", "RX", "RY", "RZ",
    "Rot", "CNOT", "CY", "CZ", "SWAP", "IsingXX", "IsingXY", "IsingYY", "IsingZZ", "ControlledPhaseShift",
    "CRX", "CRY", "CRZ", "CRot", "SingleExcitation", "SingleExcitationMinus", "SingleExcitationPlus",
    "Toffoli", "CSWAP", "DoubleExcitation", "DoubleExcitationMinus", "DoubleExcitationPlus", "MultiRZ",
    "GlobalPhase", "PSWAP", "PCPhase",
]

_DTYPES = {64: (0, np.complex64, np.float32), 128: (1, np.complex128, np.float64)}


def backend_info():
    return {"NAME": "lightning.qubit (MI355X)", "ENGINE": lib.lq_backend_info().decode()}


def compile_info():
    return {"cpu.arch": "x86_64", "device.arch": "gfx950", "compiler.name": "hipcc"}


def runtime_info():
    n = C.c_int()
    check(lib.lq_device_count(C.byref(n)))
    return {"devices": n.value}


def jit_wait():
    """Block until every queued pass compile has finished."""
    check(lib.lq_jit_wait())


def set_split_fallback(mode: int) -> int:
    """What runs while some specialised kernels of a program are missing:
    0 the unsplit plan, 1 the cheaper by an op-count estimate (default),
    2 the split plan with its missing passes in the generic kernel. Returns
    the previous mode."""
    prev = C.c_int()
    check(lib.lq_set_split_fallback(int(mode), C.byref(prev)))
    return prev.value


def set_mfma_dense5(on: bool) -> bool:
    """Run complex64 dense operators on exactly the five lowest index bits
    (wires n-5 .. n-1, no controls) on the matrix cores. Returns the previous
    setting."""
    prev = C.c_int()
    check(lib.lq_set_mfma_dense5(1 if on else 0, C.byref(prev)))
    return bool(prev.value)


def set_engine_option(name: str, value: int) -> int:
    """Set a process-wide engine option ("split_fallback", "mfma_dense5",
    "product_fold", "plan_search", "adjoint_group"); returns the previous value."""
    prev = C.c_int()
    check(lib.lq_set_engine_option(name.encode(), int(value), C.byref(prev)))
    return prev.value


def adjoint_sweep_counts(reset: bool = True) -> dict:
    """Adjoint backward sweeps run in this process: shared by a group of
    observables ("grouped") and for one observable ("single")."""
    g, s = C.c_uint64(), C.c_uint64()
    check(lib.lq_adjoint_sweep_counts(int(bool(reset)), C.byref(g), C.byref(s)))
    return {"grouped": g.value, "single": s.value}


def jit_stats() -> dict:
    """Specialised kernels loaded, compiles failed, compiles pending."""
    r, f, p = C.c_int(), C.c_int(), C.c_int()
    check(lib.lq_jit_stats(C.byref(r), C.byref(f), C.byref(p)))
    return {"ready": r.value, "failed": f.value, "pending": p.value}


def _is_bool(x) -> bool:
    return isinstance(x, (bool, np.bool_))


class _DLDevice(C.Structure):
    _fields_ = [("device_type", C.c_int32), ("device_id", C.c_int32)]


class _DLDataType(C.Structure):
    _fields_ = [("code", C.c_uint8), ("bits", C.c_uint8), ("lanes", C.c_uint16)]


class _DLTensor(C.Structure):
    _fields_ = [("data", C.c_void_p), ("device", _DLDevice), ("ndim", C.c_int32), ("dtype", _DLDataType),
                ("shape", C.POINTER(C.c_int64)), ("strides", C.POINTER(C.c_int64)), ("byte_offset", C.c_uint64)]


class _DLManagedTensor(C.Structure):
    _fields_ = [("dl_tensor", _DLTensor), ("manager_ctx", C.c_void_p), ("deleter", C.c_void_p)]


_DLTENSOR = b"dltensor"
_capsule_new = C.PYFUNCTYPE(C.py_object, C.c_void_p, C.c_char_p, C.c_void_p)(("PyCapsule_New", C.pythonapi))
_capsule_is_valid = C.PYFUNCTYPE(C.c_int, C.c_void_p, C.c_char_p)(("PyCapsule_IsValid", C.pythonapi))
_capsule_pointer = C.PYFUNCTYPE(C.c_void_p, C.c_void_p, C.c_char_p)(("PyCapsule_GetPointer", C.pythonapi))


@C.CFUNCTYPE(None, C.c_void_p)
def _capsule_destructor(capsule):
    # a capsule still named "dltensor" was never consumed: release the export
    # (a consumer renames it and calls the tensor's deleter itself)
    if _capsule_is_valid(capsule, _DLTENSOR):
        lib.lq_dlpack_release(C.c_void_p(_capsule_pointer(capsule, _DLTENSOR)))


def _dlpack_capsule(managed: int):
    return _capsule_new(managed, _DLTENSOR, C.cast(_capsule_destructor, C.c_void_p))


def _is_device_tensor(x) -> bool:
    dev = getattr(x, "__dlpack_device__", None)
    return dev is not None and hasattr(x, "data_ptr") and dev()[0] in (2, 10) and not isinstance(x, _StateVectorBase)


def _torch_dtype_ok(t, cdtype) -> bool:
    import torch
    return t.dtype == (torch.complex64 if cdtype == np.complex64 else torch.complex128)


def _device_source(t, cdtype, size):
    """A torch GPU tensor as the source of a device copy: 1-D, contiguous, the
    state's precision and size; its producer stream is waited for."""
    import torch
    if t.dim() != 1:
        raise ValueError("Array must be 1-dimensional")
    if not _torch_dtype_ok(t, cdtype) or not t.is_contiguous() or t.numel() != size:
        raise LightningException("The size of provided data must be equal to the statevector size.")
    torch.cuda.current_stream(t.device).synchronize()
    return t


def _check_device_target(t, cdtype, size):
    import torch
    if t.dim() != 1 or not _torch_dtype_ok(t, cdtype) or not t.is_contiguous() or t.numel() != size:
        raise ValueError("output must be a contiguous 1-D tensor of the state's size and precision")
    torch.cuda.current_stream(t.device).synchronize()


class _StateVectorBase:
    _bits = 128

    def __init__(self, num_qubits: int):
        code, self._cdtype, self._rdtype = _DTYPES[self._bits]
        h = C.c_void_p()
        check(lib.lq_sv_create(int(num_qubits), code, C.byref(h)))
        self._h = h
        self._n = int(num_qubits)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib.lq_sv_destroy(h)
            self._h = None

    # -- size / state preparation -------------------------------------------
    def __len__(self):
        return 1 << self._n

    def size(self):
        return 1 << self._n

    @property
    def num_qubits(self):
        return self._n

    def resetStateVector(self):
        check(lib.lq_sv_reset(self._h))

    def setBasisState(self, state=None, wires=None, async_=False, **kw):
        state = [] if state is None else list(state)
        wires = list(range(len(state))) if wires is None else list(wires)
        b, bp, _ = u32(state)
        w, wp, nw = u32(wires)
        check(lib.lq_sv_set_basis_state(self._h, bp, wp, nw))

    def setStateVector(self, state, wires, async_=False, **kw):
        s = np.ascontiguousarray(np.asarray(state, dtype=self._cdtype).reshape(-1))
        w, wp, nw = u32(wires)
        if s.size != (1 << nw):
            raise LightningException("Inconsistent state and wires dimensions.")
        check(lib.lq_sv_set_state_vector(self._h, s.ctypes.data_as(C.c_void_p), wp, nw))

    def updateData(self, data):
        if _is_device_tensor(data):
            t = _device_source(data, self._cdtype, 1 << self._n)
            check(lib.lq_sv_update_data_device(self._h, C.c_void_p(t.data_ptr()), C.c_uint64(t.numel())))
            return
        a = np.asarray(data)
        if a.ndim != 1:
            raise ValueError("Array must be 1-dimensional")
        a = np.ascontiguousarray(a, dtype=self._cdtype)
        check(lib.lq_sv_update_data(self._h, a.ctypes.data_as(C.c_void_p), C.c_uint64(a.size)))

    def getState(self, state):
        if _is_device_tensor(state):
            _check_device_target(state, self._cdtype, 1 << self._n)
            check(lib.lq_sv_get_data_device(self._h, C.c_void_p(state.data_ptr()), C.c_uint64(1 << self._n)))
            return
        if state.shape[0] < (1 << self._n):
            raise ValueError("Output array is too small")
        if state.dtype != self._cdtype or not state.flags.c_contiguous:
            tmp = np.empty(1 << self._n, dtype=self._cdtype)
            check(lib.lq_sv_get_data(self._h, tmp.ctypes.data_as(C.c_void_p), C.c_uint64(tmp.size)))
            state[: tmp.size] = tmp
        else:
            check(lib.lq_sv_get_data(self._h, state.ctypes.data_as(C.c_void_p), C.c_uint64(1 << self._n)))

    def state(self) -> np.ndarray:
        out = np.empty(1 << self._n, dtype=self._cdtype)
        self.getState(out)
        return out

    def synchronize(self):
        check(lib.lq_sv_synchronize(self._h))

    # -- zero-copy exchange (DLPack) ------------------------------------------
    def __dlpack__(self, stream=None, max_version=None, dl_device=None, copy=None):
        """The amplitudes as a DLPack capsule sharing this state's device
        memory (``torch.from_dlpack(sv)``). Queued gates run and are waited
        for first, so the data is complete on every stream; later gates show
        in the consumer's tensor, which keeps the memory alive after this
        object is gone."""
        if copy:
            raise BufferError("the state is exported without a copy (copy=True is not offered)")
        p = C.c_void_p()
        check(lib.lq_sv_dlpack(self._h, C.byref(p)))
        return _dlpack_capsule(p.value)

    def __dlpack_device__(self):
        p = C.c_void_p()
        check(lib.lq_sv_dlpack(self._h, C.byref(p)))
        dev = _DLManagedTensor.from_address(p.value).dl_tensor.device
        out = (int(dev.device_type), int(dev.device_id))
        check(lib.lq_dlpack_release(p))
        return out

    # -- recorded circuits (HIP graphs) --------------------------------------
    def record(self, operations: "OpsStruct") -> "Recording":
        """Prepare |0...0>, apply ``operations`` and keep the program as a HIP
        graph; ``Recording.replay(params)`` re-runs it with new parameters."""
        return Recording(self, operations)

    def to_torch(self):
        """A torch tensor on the GPU viewing this state's amplitudes (no copy)."""
        import torch
        return torch.from_dlpack(self)

    def device_ptr(self) -> int:
        p = C.c_void_p()
        check(lib.lq_sv_device_ptr(self._h, C.byref(p)))
        return p.value

    def profile(self, enable: bool = True):
        """Time every tile-pass kernel with HIP events on this state's stream."""
        check(lib.lq_sv_profile(self._h, int(bool(enable))))

    def profile_read(self, reset: bool = True) -> dict:
        p, ms, b, mx = C.c_uint64(), C.c_double(), C.c_double(), C.c_double()
        check(lib.lq_sv_profile_read(self._h, int(bool(reset)), C.byref(p), C.byref(ms), C.byref(b), C.byref(mx)))
        return {"passes": p.value, "ms": ms.value, "bytes": b.value, "max_ms": mx.value}

    def profile_passes(self) -> list:
        """Per-pass records of the profiled passes since the last reset:
        kernel ms, algorithmic bytes, ops, register stages, LDS transposes,
        specialised (bool), tile bits."""
        cnt = C.c_uint64()
        check(lib.lq_sv_profile_passes(self._h, None, 0, C.byref(cnt)))
        buf = (C.c_double * (7 * max(cnt.value, 1)))()
        check(lib.lq_sv_profile_passes(self._h, buf, cnt.value, C.byref(cnt)))
        out = []
        for i in range(cnt.value):
            r = buf[7 * i:7 * i + 7]
            tb = int(r[6])
            out.append({"ms": r[0], "bytes": r[1], "ops": int(r[2]), "stages": int(r[3]), "transposes": int(r[4]),
                        "specialised": bool(r[5]), "tile_bits": [b for b in range(64) if (tb >> b) & 1]})
        return out

    def stream_ptr(self) -> int:
        s = C.c_void_p()
        check(lib.lq_sv_stream(self._h, C.byref(s)))
        return s.value or 0

    def release_workspace(self):
        check(lib.lq_sv_release_workspace(self._h))

    def overlap(self, other) -> complex:
        """<self|other> on the device, FP64 accumulation; the two states may
        differ in precision."""
        re, im = C.c_double(), C.c_double()
        check(lib.lq_sv_overlap(self._h, other._h, C.byref(re), C.byref(im)))
        return complex(re.value, im.value)

    def pass_counts(self, reset: bool = True) -> dict:
        """Tile passes run as specialised kernels and as generic tile kernels."""
        a, b = C.c_uint64(), C.c_uint64()
        check(lib.lq_sv_pass_counts(self._h, int(bool(reset)), C.byref(a), C.byref(b)))
        return {"specialised": a.value, "generic": b.value}

    def copy_from(self, other):
        check(lib.lq_sv_copy(self._h, other._h))

    # -- gates ----------------------------------------------------------------
    def _apply(self, name, wires, inverse=False, params=(), ctrls=(), cvals=()):
        # small ctypes arrays (cheaper per call than numpy conversions)
        nw, nc, np_ = len(wires), len(ctrls), len(params)
        wa = (C.c_uint32 * nw)(*wires) if nw else None
        ca = (C.c_uint32 * nc)(*ctrls) if nc else None
        va = (C.c_uint8 * nc)(*[1 if bool(x) else 0 for x in cvals]) if nc and len(cvals) else None
        pa = (C.c_double * np_)(*params) if np_ else None
        rc = lib.lq_sv_apply_op(self._h, _name_bytes(name), wa, nw, ca, va, nc, int(bool(inverse)), pa, np_)
        if rc:
            check(rc)

    def applyOperations(self, ops, wires, inverses, params=None):
        """Named gates in order with one C call
        (core/simulators/base/StateVectorBase.hpp:127-172)."""
        n = len(ops)
        if len(wires) != n or len(inverses) != n or (params is not None and len(params) != n):
            raise LightningException("Invalid arguments: number of operations, wires, inverses, and parameters "
                                     "must all be equal")
        if n == 0:
            return
        blob = b"\0".join(_name_bytes(o) for o in ops) + b"\0"
        woff = np.zeros(n + 1, dtype=np.uint32)
        np.cumsum([len(w) for w in wires], out=woff[1:])
        wflat = np.fromiter((x for w in wires for x in w), dtype=np.uint32, count=int(woff[-1]))
        inv = np.fromiter((1 if x else 0 for x in inverses), dtype=np.uint8, count=n)
        pp = pf = None
        if params is not None:
            poff = np.zeros(n + 1, dtype=np.uint32)
            np.cumsum([len(p) for p in params], out=poff[1:])
            pfl = np.fromiter((x for p in params for x in p), dtype=np.float64, count=int(poff[-1]))
            pp = poff.ctypes.data_as(C.POINTER(C.c_uint32))
            pf = pfl.ctypes.data_as(C.POINTER(C.c_double))
        check(lib.lq_sv_apply_operations(self._h, n, blob, woff.ctypes.data_as(C.POINTER(C.c_uint32)),
                                         wflat.ctypes.data_as(C.POINTER(C.c_uint32)),
                                         inv.ctypes.data_as(C.POINTER(C.c_uint8)), pp, pf))

    def applyMatrix(self, matrix, wires, inverse=False):
        m = np.ascontiguousarray(np.asarray(matrix, dtype=self._cdtype).reshape(-1))
        w, wp, nw = u32(wires)
        if m.size != (1 << (2 * nw)):
            raise LightningException("The size of matrix does not match with the given number of wires")
        check(lib.lq_sv_apply_matrix(self._h, m.ctypes.data_as(C.c_void_p), wp, nw, None, None, 0,
                                     int(bool(inverse))))

    def applyControlledMatrix(self, matrix, controlled_wires, controlled_values, wires, inverse=False):
        m = np.ascontiguousarray(np.asarray(matrix, dtype=self._cdtype).reshape(-1))
        w, wp, nw = u32(wires)
        c, cp, nc = u32(controlled_wires)
        if len(controlled_values) != nc:
            raise LightningException("`controlled_wires` must have the same size as `controlled_values`.")
        v, vp = u8(controlled_values)
        if m.size != (1 << (2 * nw)):
            raise LightningException("The size of matrix does not match with the given number of wires")
        check(lib.lq_sv_apply_matrix(self._h, m.ctypes.data_as(C.c_void_p), wp, nw, cp, vp, nc,
                                     int(bool(inverse))))

    def applySparseMatrix(self, row_map, col_idx, values, wires, inverse=False):
        """CSR operator on `wires` (LQubitBindings.hpp:150-161)."""
        self.applyControlledSparseMatrix(row_map, col_idx, values, [], [], wires, inverse)

    def applyControlledSparseMatrix(self, row_map, col_idx, values, controlled_wires, controlled_values, wires,
                                    inverse=False):
        """CSR operator on `wires` where the controls hold their values
        (LQubitBindings.hpp:163-176)."""
        ip, ix, va = _csr(row_map, col_idx, values, self._cdtype)
        w, wp, nw = u32(wires)
        c, cp, nc = u32(controlled_wires)
        if len(controlled_values) != nc:
            raise LightningException("`controlled_wires` must have the same size as `controlled_values`.")
        v, vp = u8(controlled_values)
        i64 = C.POINTER(C.c_int64)
        check(lib.lq_sv_apply_sparse_matrix(self._h, ip.ctypes.data_as(i64), ip.size, ix.ctypes.data_as(i64),
                                            va.ctypes.data_as(C.c_void_p), va.size, wp, nw, cp, vp, nc,
                                            int(bool(inverse))))

    def applyPauliRot(self, wires, inverse, params, word):
        w, wp, nw = u32(wires)
        check(lib.lq_sv_apply_pauli_rot(self._h, wp, nw, int(bool(inverse)), float(params[0]), word.encode()))

    def applyGenerator(self, name, wires, adj=False, controlled_wires=(), controlled_values=()):
        w, wp, nw = u32(wires)
        c, cp, nc = u32(controlled_wires)
        v, vp = u8(controlled_values if len(controlled_values) else [1] * nc)
        s = C.c_double()
        check(lib.lq_sv_apply_generator(self._h, name.encode(), wp, nw, cp, vp, nc, int(bool(adj)), C.byref(s)))
        return s.value

    def normalize(self):
        check(lib.lq_sv_normalize(self._h))

    def norm2(self) -> float:
        """<psi|psi> in FP64 on the device (queued gates run first)."""
        out = C.c_double()
        check(lib.lq_sv_norm2(self._h, C.byref(out)))
        return out.value

    def collapse(self, wire, branch):
        check(lib.lq_sv_collapse(self._h, int(wire), int(bool(branch))))

    def kernel_map(self):
        return {name: "TileLDS" for name in GATE_NAMES}

    def apply_ops(self, ops: "OpsStruct", adjoint: bool = False):
        check(lib.lq_sv_apply_ops(self._h, ops._h, int(bool(adjoint))))


_NAME_BYTES = {}


def _name_bytes(name: str) -> bytes:
    b = _NAME_BYTES.get(name)
    if b is None:
        b = _NAME_BYTES[name] = name.encode()
    return b


def _make_gate(name):
    def gate(self, *args, **kwargs):
        # Gate(wires, inverse, params) or
        # Gate(controlled_wires, controlled_values, wires, inverse=False, params=[])
        if len(args) == 3 and _is_bool(args[1]) and not kwargs:
            wires, inverse, params = args
            self._apply(name, wires, inverse, params)
            return
        if len(args) >= 3:
            ctrls, cvals, wires = args[0], args[1], args[2]
            inverse = args[3] if len(args) > 3 else kwargs.get("inverse", False)
            params = args[4] if len(args) > 4 else kwargs.get("params", [])
            if len(ctrls) != len(cvals):
                raise LightningException("`controlled_wires` must have the same size as `controlled_values`.")
            self._apply(name, wires, inverse, params, ctrls, cvals)
            return
        wires = args[0] if args else kwargs["wires"]
        inverse = args[1] if len(args) > 1 else kwargs.get("inverse", False)
        params = args[2] if len(args) > 2 else kwargs.get("params", [])
        self._apply(name, wires, inverse, params)

    gate.__name__ = name
    gate.__doc__ = f"Apply the {name} gate."
    return gate


for _g in GATE_NAMES:
    setattr(_StateVectorBase, _g, _make_gate(_g))


class StateVectorC64(_StateVectorBase):
    _bits = 64


class StateVectorC128(_StateVectorBase):
    _bits = 128



# ---------------------------------------------------------------------------
# Observables (core/observables/Observables.hpp:128-584)
# ---------------------------------------------------------------------------
class _Obs:
    _h = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib.lq_obs_destroy(h)
            self._h = None

    def __eq__(self, other):
        return type(self) is type(other) and self._key() == other._key()

    def __hash__(self):
        return hash(repr(self))


class NamedObs(_Obs):
    def __init__(self, name: str, wires: Sequence[int]):
        self._name, self._wires = name, [int(w) for w in wires]
        w, wp, nw = u32(self._wires)
        h = C.c_void_p()
        check(lib.lq_obs_named(name.encode(), wp, nw, C.byref(h)))
        self._h = h

    def get_wires(self):
        return list(self._wires)

    def _key(self):
        return (self._name, tuple(self._wires))

    def __repr__(self):
        return f"{self._name}{self._wires}"


class HermitianObs(_Obs):
    def __init__(self, matrix, wires: Sequence[int]):
        self._matrix = np.ascontiguousarray(np.asarray(matrix, dtype=np.complex128).reshape(-1))
        self._wires = [int(w) for w in wires]
        w, wp, nw = u32(self._wires)
        h = C.c_void_p()
        check(lib.lq_obs_hermitian(self._matrix.ctypes.data_as(C.c_void_p), wp, nw, C.byref(h)))
        self._h = h

    def get_wires(self):
        return list(self._wires)

    def get_matrix(self):
        return self._matrix.copy()

    def _key(self):
        return (tuple(self._matrix.tolist()), tuple(self._wires))

    def __repr__(self):
        return f"Hermitian{self._wires}"


class TensorProdObs(_Obs):
    def __init__(self, obs: Sequence[_Obs]):
        self._obs = list(obs)
        arr = (C.c_void_p * max(len(self._obs), 1))(*[o._h.value for o in self._obs])
        h = C.c_void_p()
        check(lib.lq_obs_tensor(arr, len(self._obs), C.byref(h)))
        self._h = h

    def get_wires(self):
        return sorted({w for o in self._obs for w in o.get_wires()})

    def get_ops(self):
        return list(self._obs)

    def _key(self):
        return tuple(o._key() for o in self._obs)

    def __repr__(self):
        return " @ ".join(repr(o) for o in self._obs)


class Hamiltonian(_Obs):
    def __init__(self, coeffs, obs: Sequence[_Obs]):
        self._coeffs = [float(c) for c in np.asarray(coeffs).reshape(-1)]
        self._obs = list(obs)
        if len(self._coeffs) != len(self._obs):
            raise LightningException("coefficients and observables differ in length")
        c, cp, _ = f64(self._coeffs)
        arr = (C.c_void_p * max(len(self._obs), 1))(*[o._h.value for o in self._obs])
        h = C.c_void_p()
        check(lib.lq_obs_hamiltonian(cp, arr, len(self._obs), C.byref(h)))
        self._h = h

    def get_wires(self):
        return sorted({w for o in self._obs for w in o.get_wires()})

    def get_coeffs(self):
        return list(self._coeffs)

    def get_ops(self):
        return list(self._obs)

    def _key(self):
        return (tuple(self._coeffs), tuple(o._key() for o in self._obs))

    def __repr__(self):
        return "Hamiltonian: { 'coeffs' : " + repr(self._coeffs) + ", 'observables' : " + repr(self._obs) + "}"


def _csr(indptr, indices, values, cdtype):
    """CSR arrays as the C-ABI takes them (int64 offsets and columns)."""
    ip = np.ascontiguousarray(np.asarray(indptr, dtype=np.int64).reshape(-1))
    ix = np.ascontiguousarray(np.asarray(indices, dtype=np.int64).reshape(-1))
    va = np.ascontiguousarray(np.asarray(values, dtype=cdtype).reshape(-1))
    if ix.size != va.size:
        raise LightningException("sparse matrix indices and values differ in length")
    return ip, ix, va


class SparseHamiltonian(_Obs):
    """SparseHamiltonian(data, indices, offsets, wires): CSR over the full
    state index (lq/observables/ObservablesLQubit.hpp:372-436)."""

    def __init__(self, data, indices, offsets, wires: Sequence[int]):
        self._offsets, self._indices, self._data = _csr(offsets, indices, data, np.complex128)
        self._wires = [int(w) for w in wires]
        w, wp, nw = u32(self._wires)
        h = C.c_void_p()
        i64 = C.POINTER(C.c_int64)
        check(lib.lq_obs_sparse(self._data.ctypes.data_as(C.c_void_p), self._indices.ctypes.data_as(i64),
                                self._data.size, self._offsets.ctypes.data_as(i64), self._offsets.size, wp, nw,
                                C.byref(h)))
        self._h = h

    def get_wires(self):
        return list(self._wires)

    def _key(self):
        return (tuple(self._data.tolist()), tuple(self._indices.tolist()), tuple(self._offsets.tolist()),
                tuple(self._wires))

    def __repr__(self):
        return "SparseHamiltonian: {\n'data' : " + repr(self._data.tolist()) + ",\n'indices' : " + \
            repr(self._indices.tolist()) + ",\n'offsets' : " + repr(self._offsets.tolist()) + "\n}"


# ---------------------------------------------------------------------------
# Measurements (lq/measurements/MeasurementsLQubit.hpp)
# ---------------------------------------------------------------------------
class _MeasurementsBase:
    def __init__(self, sv: _StateVectorBase):
        self._sv = sv  # the state vector must outlive the measurements, as in the reference
        self._seed = None  # MeasurementsBase::_deviceseed (MeasurementsBase.hpp:77-84)

    def set_random_seed(self, seed: int):
        """Seed for generate_samples; every later call draws from a generator
        seeded with it again (MeasurementsBase.hpp:77-84)."""
        self._seed = int(seed)

    def generate_samples(self, wires_or_num_wires, num_shots: int) -> np.ndarray:
        """generate_samples(num_wires, shots) for all wires (Bindings.hpp:634-640)
        or generate_samples(wires, shots) (LQubitBindings.hpp:330-336):
        shots x len(wires) array of 0/1, column j = wires[j]."""
        shots = int(num_shots)
        if isinstance(wires_or_num_wires, (int, np.integer)):
            nw_out = int(wires_or_num_wires)
            w, wp, nw = u32([])
            out = np.empty((shots, self._sv._n), dtype=np.uint64)
            check(lib.lq_generate_samples(self._sv._h, wp, 0, 1, shots, int(self._seed is not None),
                                          self._seed or 0, out.ctypes.data_as(C.POINTER(C.c_uint64))))
            return out.reshape(shots, nw_out) if nw_out == self._sv._n else out
        w, wp, nw = u32(wires_or_num_wires)
        out = np.empty((shots, nw), dtype=np.uint64)
        check(lib.lq_generate_samples(self._sv._h, wp, nw, 0, shots, int(self._seed is not None), self._seed or 0,
                                      out.ctypes.data_as(C.POINTER(C.c_uint64))))
        return out

    def _sparse_call(self, fn, indptr, indices, values):
        ip, ix, va = _csr(indptr, indices, values, self._sv._cdtype)
        out = C.c_double()
        i64 = C.POINTER(C.c_int64)
        check(fn(self._sv._h, ip.ctypes.data_as(i64), ip.size, ix.ctypes.data_as(i64), va.ctypes.data_as(C.c_void_p),
                 va.size, C.byref(out)))
        return out.value

    def var(self, *args):
        """var(observable) | var(name, wires) | var(matrix, wires) |
        var(row_map, col_idx, values) (MeasurementsLQubit.hpp:420-636)."""
        out = C.c_double()
        if len(args) == 1 and isinstance(args[0], _Obs):
            check(lib.lq_var_obs(self._sv._h, args[0]._h, C.byref(out)))
        elif len(args) == 2 and isinstance(args[0], str):
            w, wp, nw = u32(args[1])
            check(lib.lq_var_named(self._sv._h, args[0].encode(), wp, nw, C.byref(out)))
        elif len(args) == 2:
            m = np.ascontiguousarray(np.asarray(args[0], dtype=self._sv._cdtype).reshape(-1))
            w, wp, nw = u32(args[1])
            if m.size != (1 << (2 * nw)):
                raise LightningException("The size of matrix does not match with the given number of wires")
            check(lib.lq_var_matrix(self._sv._h, m.ctypes.data_as(C.c_void_p), wp, nw, C.byref(out)))
        elif len(args) == 3:
            return self._sparse_call(lib.lq_var_sparse, *args)
        else:
            raise TypeError("var(observable), var(name, wires), var(matrix, wires) or var(row_map, col_idx, values)")
        return out.value

    def probs(self, wires=None):
        if wires is None:
            out = np.empty(1 << self._sv._n, dtype=np.float64)
            w, wp, nw = u32([])
            check(lib.lq_probs(self._sv._h, wp, 0, 1, out.ctypes.data_as(C.POINTER(C.c_double))))
        else:
            w, wp, nw = u32(wires)
            out = np.empty(1 << nw, dtype=np.float64)
            check(lib.lq_probs(self._sv._h, wp, nw, 0, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out.astype(self._sv._rdtype) if self._sv._bits == 64 else out

    def expval(self, *args):
        out = C.c_double()
        if len(args) == 1 and isinstance(args[0], _Obs):
            check(lib.lq_expval_obs(self._sv._h, args[0]._h, C.byref(out)))
        elif len(args) == 2 and isinstance(args[0], str):
            w, wp, nw = u32(args[1])
            check(lib.lq_expval_named(self._sv._h, args[0].encode(), wp, nw, C.byref(out)))
        elif len(args) == 2:
            m = np.ascontiguousarray(np.asarray(args[0], dtype=self._sv._cdtype).reshape(-1))
            w, wp, nw = u32(args[1])
            if m.size != (1 << (2 * nw)):
                raise LightningException("The size of matrix does not match with the given number of wires")
            check(lib.lq_expval_matrix(self._sv._h, m.ctypes.data_as(C.c_void_p), wp, nw, C.byref(out)))
        elif len(args) == 3:
            return self._sparse_call(lib.lq_expval_sparse, *args)
        else:
            raise TypeError("expval(observable), expval(name, wires), expval(matrix, wires) or "
                            "expval(row_map, col_idx, values)")
        return out.value

    def expval_named_batch(self, names: Sequence[str], wires: Sequence[int]) -> np.ndarray:
        """Several one-wire named expvals from one read of the state."""
        arr = (C.c_char_p * max(len(names), 1))(*[n.encode() for n in names])
        w, wp, _ = u32(wires)
        out = np.empty(len(names), dtype=np.float64)
        check(lib.lq_expval_named_batch(self._sv._h, len(names), arr, wp, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out


class MeasurementsC64(_MeasurementsBase):
    pass


class MeasurementsC128(_MeasurementsBase):
    pass


# ---------------------------------------------------------------------------
# Algorithms (core/algorithms/JacobianData.hpp, lq/algorithms/AdjointJacobianLQubit.hpp)
# ---------------------------------------------------------------------------
class OpsStruct:
    """OpsData: names, params, wires, inverses, matrices[, ctrl wires, ctrl values]."""

    def __init__(self, ops_name, ops_params, ops_wires, ops_inverses, ops_matrices,
                 ops_controlled_wires=None, ops_controlled_values=None):
        n = len(ops_name)
        ops_controlled_wires = ops_controlled_wires or [[] for _ in range(n)]
        ops_controlled_values = ops_controlled_values or [[] for _ in range(n)]
        self.names = list(ops_name)
        self.params = [list(map(float, p)) for p in ops_params]
        self.wires = [list(map(int, w)) for w in ops_wires]
        self.inverses = [bool(i) for i in ops_inverses]
        self.matrices = [np.asarray(m, dtype=np.complex128).reshape(-1) if m is not None and len(m) else
                         np.zeros(0, np.complex128) for m in ops_matrices]
        self.ctrl_wires = [list(map(int, c)) for c in ops_controlled_wires]
        self.ctrl_values = [list(map(bool, v)) for v in ops_controlled_values]
        poff, woff, moff, coff = [0], [0], [0], [0]
        params, wires, mats, ctrls, cvals = [], [], [], [], []
        for i in range(n):
            params += self.params[i]
            poff.append(len(params))
            wires += self.wires[i]
            woff.append(len(wires))
            m = self.matrices[i]
            mats += [x for z in m for x in (z.real, z.imag)]
            moff.append(len(mats) // 2)
            ctrls += self.ctrl_wires[i]
            cvals += self.ctrl_values[i]
            coff.append(len(ctrls))
        keep = [np.asarray(poff, np.uint32), np.asarray(params + [0.0], np.float64),
                np.asarray(woff, np.uint32), np.asarray(wires + [0], np.uint32),
                np.asarray([int(b) for b in self.inverses] + [0], np.uint8), np.asarray(moff, np.uint32),
                np.asarray(mats + [0.0], np.float64), np.asarray(coff, np.uint32),
                np.asarray(ctrls + [0], np.uint32), np.asarray([int(b) for b in cvals] + [0], np.uint8)]
        names = (C.c_char_p * max(n, 1))(*[s.encode() for s in self.names])
        ptr = lambda a, t: a.ctypes.data_as(C.POINTER(t))
        h = C.c_void_p()
        check(lib.lq_ops_create(n, names, ptr(keep[0], C.c_uint32), ptr(keep[1], C.c_double),
                                ptr(keep[2], C.c_uint32), ptr(keep[3], C.c_uint32), ptr(keep[4], C.c_uint8),
                                ptr(keep[5], C.c_uint32), ptr(keep[6], C.c_double), ptr(keep[7], C.c_uint32),
                                ptr(keep[8], C.c_uint32), ptr(keep[9], C.c_uint8), C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib.lq_ops_destroy(h)
            self._h = None

    def getNumParOps(self):
        v = C.c_uint64()
        check(lib.lq_ops_num_par_ops(self._h, C.byref(v)))
        return v.value

    def __repr__(self):
        return f"Operations: [{', '.join(self.names)}]"


def _create_ops_list(ops_name, ops_params, ops_wires, ops_inverses, ops_matrices,
                     ops_controlled_wires=None, ops_controlled_values=None):
    return OpsStruct(ops_name, ops_params, ops_wires, ops_inverses, ops_matrices,
                     ops_controlled_wires, ops_controlled_values)


class Recording:
    """A circuit captured as a HIP graph on one state (lq_sv_record). ``replay``
    takes every op's parameters in op order (as many as the recorded ops
    carry, matrix ops' 0.0 placeholders included) and leaves
    U(params)|0...0> in the state."""

    def __init__(self, sv: "_StateVectorBase", operations: "OpsStruct"):
        h = C.c_void_p()
        check(lib.lq_sv_record(sv._h, operations._h, C.byref(h)))
        self._h = h
        self._sv = sv
        self.nparams = sum(len(p) for p in operations.params)

    def replay(self, params):
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float64).reshape(-1))
        check(lib.lq_rec_replay(self._h, p.ctypes.data_as(C.POINTER(C.c_double)) if p.size else None,
                                C.c_uint64(p.size)))

    def stats(self) -> dict:
        a, b = C.c_uint64(), C.c_uint64()
        check(lib.lq_rec_stats(self._h, C.byref(a), C.byref(b)))
        return {"captures": a.value, "replays": b.value}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib.lq_rec_destroy(h)
            self._h = None


class AdjointJacobian:
    """AdjointJacobianC*()(sv, observables, ops, trainable) -> [obs][param] array."""

    def __call__(self, sv: _StateVectorBase, observables, operations: OpsStruct, trainableParams,
                 apply_operations: bool = False, return_expvals: bool = False):
        obs = list(observables)
        tp = np.ascontiguousarray(np.asarray(sorted(int(t) for t in trainableParams), dtype=np.uint64))
        jac = np.zeros(max(len(obs) * tp.size, 1), dtype=np.float64)
        ev = np.zeros(max(len(obs), 1), dtype=np.float64)
        arr = (C.c_void_p * max(len(obs), 1))(*[o._h.value for o in obs])
        tpp = tp.ctypes.data_as(C.POINTER(C.c_uint64)) if tp.size else None
        check(lib.lq_adjoint_jacobian(sv._h, operations._h, arr, len(obs), tpp, tp.size, int(bool(apply_operations)),
                                      jac.ctypes.data_as(C.POINTER(C.c_double)),
                                      ev.ctypes.data_as(C.POINTER(C.c_double)) if return_expvals else None))
        out = jac[: len(obs) * tp.size]
        if sv._bits == 64:
            out = out.astype(np.float32)
        if return_expvals:
            return out, ev[: len(obs)]
        return out


class VectorJacobianProduct:
    """VectorJacobianProductC*()(sv, ops, dy, trainable) -> complex array of
    w_t = <dy | d psi / d theta_t> (LQubitBindings.hpp:412-457,
    VectorJacobianProduct.hpp:43-160)."""

    def __call__(self, sv: _StateVectorBase, operations: OpsStruct, dy, trainableParams,
                 apply_operations: bool = False):
        cdt = np.complex64 if sv._bits == 64 else np.complex128
        d = np.ascontiguousarray(np.asarray(dy).reshape(-1), dtype=cdt)
        tp = np.ascontiguousarray(np.asarray(sorted(int(t) for t in trainableParams), dtype=np.uint64))
        out = np.zeros(max(2 * tp.size, 2), dtype=np.float64)
        tpp = tp.ctypes.data_as(C.POINTER(C.c_uint64)) if tp.size else None
        check(lib.lq_vector_jacobian_product(sv._h, operations._h, d.ctypes.data_as(C.c_void_p), d.size, tpp,
                                             tp.size, int(bool(apply_operations)),
                                             out.ctypes.data_as(C.POINTER(C.c_double))))
        w = out[: 2 * tp.size].view(np.complex128)
        return w.astype(cdt)


def allocate_aligned_array(size: int, dtype, zero_init: bool = False) -> np.ndarray:
    """1-D host array of ``size`` elements whose data starts on a 64-byte
    boundary (Bindings.hpp:381-456)."""
    dt = np.dtype(dtype)
    raw = (np.zeros if zero_init else np.empty)(size * dt.itemsize + 64, dtype=np.uint8)
    off = (-raw.ctypes.data) % 64
    return raw[off: off + size * dt.itemsize].view(dt)


# Precision-suffixed names, as in the reference module; the two namespaces are
# also importable as submodules (``from ...lightning_qubit_ops.algorithms import``)
observables = types.ModuleType(__name__ + ".observables")
algorithms = types.ModuleType(__name__ + ".algorithms")
sys.modules[observables.__name__] = observables
sys.modules[algorithms.__name__] = algorithms
for _bits in (64, 128):
    for _cls in (NamedObs, HermitianObs, TensorProdObs, Hamiltonian, SparseHamiltonian):
        _sub = type(f"{_cls.__name__}C{_bits}", (_cls,), {})
        setattr(observables, _sub.__name__, _sub)
        globals()[_sub.__name__] = _sub
    _ops = type(f"OpsStructC{_bits}", (OpsStruct,), {})
    _adj = type(f"AdjointJacobianC{_bits}", (AdjointJacobian,), {})
    _vjp = type(f"VectorJacobianProductC{_bits}", (VectorJacobianProduct,), {})
    setattr(algorithms, _vjp.__name__, _vjp)
    globals()[_vjp.__name__] = _vjp
    setattr(algorithms, _ops.__name__, _ops)
    setattr(algorithms, _adj.__name__, _adj)
    setattr(algorithms, f"create_ops_listC{_bits}",
            (lambda cls: (lambda *a, **k: cls(*a, **k)))(_ops))
    globals()[_ops.__name__] = _ops
    globals()[_adj.__name__] = _adj
