"""The allocation-free probabilities entry point (lq_probs_noalloc) against
lq_probs on the same states, through the C-ABI.

lq_probs_noalloc reads a state the caller holds in device memory and writes
2^nw doubles to the caller's device memory, with its partial histograms in a
fixed device array of the library. Here the input is an engine state's
device memory (lq_sv_device_ptr) and the output buffer is another state's
device memory, so the test itself allocates nothing through torch. Two calls
in a row, the second on a smaller state with fewer wires, both precisions;
the results must equal lq_probs (the same kernels and block count: within
1e-15 absolute, the compensated FP64 sums of identical partials). Argument
errors come back as the reference's error format.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import oracle_run
from oracle import lq_oracle as O

pytestmark = pytest.mark.gpu

L = pytest.importorskip("lightning_amd.lightning_qubit_ops")
W = pytest.importorskip("lightning_amd.workloads")
from lightning_amd._capi import lib  # noqa: E402

SV = {128: L.StateVectorC128, 64: L.StateVectorC64}
DTYPE = {128: 1, 64: 0}


def prepared(bits, n, seed):
    ops = W.layers_rxyz_cnot(n, 2, seed=seed)
    sv = SV[bits](n)
    W.apply_circuit(sv, ops)
    sv.synchronize()
    return sv, ops


def lq_probs(sv, wires):
    """lq_probs(wires) through the C-ABI, in FP64 (the Python mirror casts a
    complex64 state's probabilities to float32, as the reference binding does)."""
    w = (C.c_uint32 * len(wires))(*wires)
    out = np.empty(1 << len(wires), dtype=np.float64)
    assert lib.lq_probs(sv._h, w, len(wires), 0, out.ctypes.data_as(C.POINTER(C.c_double))) == 0
    return out


def noalloc_probs(sv, bits, n, wires, out_sv):
    """probs through lq_probs_noalloc on sv's stream into out_sv's memory."""
    src = sv.device_ptr()
    dst = out_sv.device_ptr()
    out_sv.synchronize()  # any write of out_sv's own is done before the probs land there
    w = (C.c_uint32 * len(wires))(*wires)
    rc = lib.lq_probs_noalloc(C.c_void_p(src), n, DTYPE[bits], w, len(wires), C.c_void_p(dst),
                              C.c_void_p(sv.stream_ptr()))
    assert rc == 0, lib.lq_last_error().decode()
    sv.synchronize()
    raw = np.asarray(out_sv.state())
    return raw.view(np.float64 if raw.dtype == np.complex128 else np.float32)


@pytest.mark.parametrize("bits", [128, 64])
def test_noalloc_probs_match_lq_probs(bits):
    assert lib.lq_probs_noalloc_init() == 0, lib.lq_last_error().decode()
    # output buffer: device memory of a complex128 state of 12 qubits (64 KiB >= 2^10 doubles)
    out_sv = L.StateVectorC128(12)
    cases = [(16, [3, 0, 11, 7, 15, 9]), (12, [5, 2, 10])]  # the second call is smaller
    for n, wires in cases:
        sv, ops = prepared(bits, n, seed=n)
        got = noalloc_probs(sv, bits, n, wires, out_sv)[: 1 << len(wires)]
        ref = lq_probs(sv, wires)
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-15)
        exp = O.probs(oracle_run(n, ops), n, wires)
        np.testing.assert_allclose(got, exp, atol=1e-12 if bits == 128 else 5e-6)


def test_noalloc_probs_rejects_sizes_beyond_its_array():
    out_sv = L.StateVectorC128(12)
    sv, _ = prepared(128, 12, seed=1)
    w = (C.c_uint32 * 11)(*range(11))
    rc = lib.lq_probs_noalloc(C.c_void_p(sv.device_ptr()), 12, 1, w, 11, C.c_void_p(out_sv.device_ptr()), None)
    assert rc != 0
    msg = lib.lq_last_error().decode()
    assert "Error in PennyLane Lightning" in msg and "1..10 wires" in msg
    w2 = (C.c_uint32 * 2)(3, 3)
    assert lib.lq_probs_noalloc(C.c_void_p(sv.device_ptr()), 12, 1, w2, 2, C.c_void_p(out_sv.device_ptr()), None) != 0
    assert "unique" in lib.lq_last_error().decode()
