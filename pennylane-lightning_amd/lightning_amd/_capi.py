"""ctypes binding of liblq_amd.so (include/lq_amd.h).

The library must be built (``__graft_entry__.build()`` or ``make -C
pennylane-lightning_amd``); importing this module without it raises
ImportError. There is no CPU fallback.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "liblq_amd.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built: run `make -C pennylane-lightning_amd` "
                      "(or __graft_entry__.build()) first")

# Hardware queues per process: HIP maps streams onto GPU_MAX_HW_QUEUES
# hardware queues (default 4), and each queue runs its kernels in order, so
# a stream sharing a queue waits behind every kernel another stream put
# there. A server runs its engine workers and packed launches on about ten
# streams; with 4 queues a packed launch waited behind a large request's
# passes and the batched server's throughput collapsed at random (p50 of
# packed requests 100-1000 ms); with 16 each stream has its own queue.
# Set before the HIP runtime starts (the caller's value is kept).
os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")

# torch (when installed) is imported first: the engine's HIP runtime
# libraries have the same sonames as the ones torch bundles, so the loader
# binds the engine to torch's copies and the process has one HIP runtime.
# Device memory then passes between the two without copies (DLPack), and
# either may initialise the GPU first.
try:
    import torch  # noqa: F401
except ImportError:  # pragma: no cover - the engine works without torch
    pass

lib = C.CDLL(LIB_PATH)

_P = C.c_void_p
_U32P = C.POINTER(C.c_uint32)
_U8P = C.POINTER(C.c_uint8)
_DP = C.POINTER(C.c_double)
_U64P = C.POINTER(C.c_uint64)
_I64P = C.POINTER(C.c_int64)

_SIGS = {
    "lq_last_error": (C.c_char_p, []),
    "lq_backend_info": (C.c_char_p, []),
    "lq_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "lq_sv_create": (C.c_int, [C.c_uint32, C.c_int, C.POINTER(_P)]),
    "lq_sv_destroy": (C.c_int, [_P]),
    "lq_sv_num_qubits": (C.c_int, [_P, _U32P]),
    "lq_sv_dtype": (C.c_int, [_P, C.POINTER(C.c_int)]),
    "lq_sv_reset": (C.c_int, [_P]),
    "lq_sv_set_basis_state": (C.c_int, [_P, _U32P, _U32P, C.c_uint32]),
    "lq_sv_set_state_vector": (C.c_int, [_P, _P, _U32P, C.c_uint32]),
    "lq_sv_update_data": (C.c_int, [_P, _P, C.c_uint64]),
    "lq_sv_get_data": (C.c_int, [_P, _P, C.c_uint64]),
    "lq_sv_copy": (C.c_int, [_P, _P]),
    "lq_sv_synchronize": (C.c_int, [_P]),
    "lq_sv_device_ptr": (C.c_int, [_P, C.POINTER(_P)]),
    "lq_sv_update_data_device": (C.c_int, [_P, _P, C.c_uint64]),
    "lq_sv_get_data_device": (C.c_int, [_P, _P, C.c_uint64]),
    "lq_sv_dlpack": (C.c_int, [_P, C.POINTER(_P)]),
    "lq_sv_record": (C.c_int, [_P, _P, C.POINTER(_P)]),
    "lq_adjoint_sweep_counts": (C.c_int, [C.c_int, _U64P, _U64P]),
    "lq_rec_replay": (C.c_int, [_P, _DP, C.c_uint64]),
    "lq_rec_stats": (C.c_int, [_P, _U64P, _U64P]),
    "lq_rec_destroy": (C.c_int, [_P]),
    "lq_dlpack_release": (C.c_int, [_P]),
    "lq_sv_apply_op": (C.c_int, [_P, C.c_char_p, _U32P, C.c_uint32, _U32P, _U8P, C.c_uint32, C.c_int, _DP,
                                 C.c_uint32]),
    "lq_sv_apply_operations": (C.c_int, [_P, C.c_uint32, C.c_char_p, _U32P, _U32P, _U8P, _U32P, _DP]),
    "lq_sv_apply_matrix": (C.c_int, [_P, _P, _U32P, C.c_uint32, _U32P, _U8P, C.c_uint32, C.c_int]),
    "lq_sv_apply_sparse_matrix": (C.c_int, [_P, _I64P, C.c_uint64, _I64P, _P, C.c_uint64, _U32P, C.c_uint32, _U32P,
                                            _U8P, C.c_uint32, C.c_int]),
    "lq_sv_apply_pauli_rot": (C.c_int, [_P, _U32P, C.c_uint32, C.c_int, C.c_double, C.c_char_p]),
    "lq_sv_apply_generator": (C.c_int, [_P, C.c_char_p, _U32P, C.c_uint32, _U32P, _U8P, C.c_uint32, C.c_int,
                                        _DP]),
    "lq_sv_normalize": (C.c_int, [_P]),
    "lq_sv_norm2": (C.c_int, [_P, _DP]),
    "lq_sv_collapse": (C.c_int, [_P, C.c_uint32, C.c_int]),
    "lq_probs": (C.c_int, [_P, _U32P, C.c_uint32, C.c_int, _DP]),
    "lq_expval_named": (C.c_int, [_P, C.c_char_p, _U32P, C.c_uint32, _DP]),
    "lq_expval_named_batch": (C.c_int, [_P, C.c_uint32, C.POINTER(C.c_char_p), _U32P, _DP]),
    "lq_expval_matrix": (C.c_int, [_P, _P, _U32P, C.c_uint32, _DP]),
    "lq_expval_obs": (C.c_int, [_P, _P, _DP]),
    "lq_var_named": (C.c_int, [_P, C.c_char_p, _U32P, C.c_uint32, _DP]),
    "lq_var_matrix": (C.c_int, [_P, _P, _U32P, C.c_uint32, _DP]),
    "lq_var_obs": (C.c_int, [_P, _P, _DP]),
    "lq_expval_sparse": (C.c_int, [_P, _I64P, C.c_uint64, _I64P, _P, C.c_uint64, _DP]),
    "lq_var_sparse": (C.c_int, [_P, _I64P, C.c_uint64, _I64P, _P, C.c_uint64, _DP]),
    "lq_generate_samples": (C.c_int, [_P, _U32P, C.c_uint32, C.c_int, C.c_uint64, C.c_int, C.c_uint64, _U64P]),
    "lq_obs_sparse": (C.c_int, [_P, _I64P, C.c_uint64, _I64P, C.c_uint64, _U32P, C.c_uint32, C.POINTER(_P)]),
    "lq_obs_named": (C.c_int, [C.c_char_p, _U32P, C.c_uint32, C.POINTER(_P)]),
    "lq_obs_hermitian": (C.c_int, [_P, _U32P, C.c_uint32, C.POINTER(_P)]),
    "lq_obs_tensor": (C.c_int, [C.POINTER(_P), C.c_uint32, C.POINTER(_P)]),
    "lq_obs_hamiltonian": (C.c_int, [_DP, C.POINTER(_P), C.c_uint32, C.POINTER(_P)]),
    "lq_obs_destroy": (C.c_int, [_P]),
    "lq_ops_create": (C.c_int, [C.c_uint32, C.POINTER(C.c_char_p), _U32P, _DP, _U32P, _U32P, _U8P, _U32P, _DP,
                                _U32P, _U32P, _U8P, C.POINTER(_P)]),
    "lq_ops_destroy": (C.c_int, [_P]),
    "lq_ops_num_par_ops": (C.c_int, [_P, _U64P]),
    "lq_sv_apply_ops": (C.c_int, [_P, _P, C.c_int]),
    "lq_adjoint_jacobian": (C.c_int, [_P, _P, C.POINTER(_P), C.c_uint32, _U64P, C.c_uint32, C.c_int, _DP, _DP]),
    "lq_vector_jacobian_product": (C.c_int, [_P, _P, _P, C.c_uint64, _U64P, C.c_uint32, C.c_int, _DP]),
    "lq_sv_release_workspace": (C.c_int, [_P]),
    "lq_sv_overlap": (C.c_int, [_P, _P, _DP, _DP]),
    "lq_sv_profile": (C.c_int, [_P, C.c_int]),
    "lq_sv_profile_read": (C.c_int, [_P, C.c_int, _U64P, _DP, _DP, _DP]),
    "lq_sv_profile_passes": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.c_uint64, C.POINTER(C.c_uint64)]),
    "lq_sv_stream": (C.c_int, [_P, C.POINTER(_P)]),
    "lq_jit_wait": (C.c_int, []),
    "lq_jit_shutdown": (C.c_int, []),
    "lq_jit_stats": (C.c_int, [C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "lq_set_split_fallback": (C.c_int, [C.c_int, C.POINTER(C.c_int)]),
    "lq_set_mfma_dense5": (C.c_int, [C.c_int, C.POINTER(C.c_int)]),
    "lq_set_engine_option": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_int)]),
    "lq_probs_noalloc_init": (C.c_int, []),
    "lq_probs_noalloc": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int, C.POINTER(C.c_uint32), C.c_uint32,
                                   C.c_void_p, C.c_void_p]),
    "lq_sv_pass_counts": (C.c_int, [_P, C.c_int, _U64P, _U64P]),
}

for _name, (_res, _args) in _SIGS.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args

# background pass compiles must not be in flight while the process tears down
atexit.register(lib.lq_jit_shutdown)


class LightningException(RuntimeError):
    """Raised for every failure reported by the engine."""


def check(rc: int) -> None:
    if rc != 0:
        raise LightningException(lib.lq_last_error().decode())


def u32(xs):
    a = np.ascontiguousarray(np.asarray(list(xs), dtype=np.uint32).reshape(-1))
    if a.size == 0:
        a = np.zeros(1, dtype=np.uint32)
        return a, a.ctypes.data_as(_U32P), 0
    return a, a.ctypes.data_as(_U32P), a.size


def u8(xs):
    a = np.ascontiguousarray(np.asarray([1 if bool(x) else 0 for x in xs], dtype=np.uint8))
    if a.size == 0:
        a = np.zeros(1, dtype=np.uint8)
    return a, a.ctypes.data_as(_U8P)


def f64(xs):
    a = np.ascontiguousarray(np.asarray(list(xs), dtype=np.float64).reshape(-1))
    if a.size == 0:
        a = np.zeros(1, dtype=np.float64)
        return a, a.ctypes.data_as(_DP), 0
    return a, a.ctypes.data_as(_DP), a.size
