"""Energy + adjoint-gradient request serving with continuous batching.

``Server`` wraps the engine's request server (``lq_server_*`` in
include/lq_amd.h): requests whose state fits the packed kernel run many per
launch with their states in LDS; larger ones run through the fused-pass engine
on concurrent workers, admitted by HBM bytes (3 x 2^n amplitudes). With
``batched=False`` the same server runs every request alone, in arrival order,
through the engine -- the plain one-request-at-a-time path.

``serving_requests`` builds the serving workload: n in [10, 28] with
P(n) ~ 2^-((n-10)/2), depth uniform in 4..16, the hardware-efficient layers
(Rot on every wire, CZ on even then odd pairs) with the Rot written as
RZ RY RZ so every angle is trainable, fresh angles per request, and a
50-term random Pauli Hamiltonian per request (I/X/Y/Z with probabilities
0.7/0.1/0.1/0.1, N(0, 1) coefficients).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import lightning_qubit_ops as L
from ._capi import check, lib
from .workloads import CircuitOp, hamiltonian_obs, random_pauli_hamiltonian, to_ops

_P = C.c_void_p
_U64P = C.POINTER(C.c_uint64)
_U32P = C.POINTER(C.c_uint32)
_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int)


class _Config(C.Structure):
    _fields_ = [("dtype", C.c_int), ("batched", C.c_int), ("hbm_budget", C.c_uint64),
                ("small_max_qubits", C.c_int), ("max_batch", C.c_int), ("batch_streams", C.c_int),
                ("large_workers", C.c_int), ("prep_threads", C.c_int), ("lane_workers", C.c_int),
                ("lane_max_qubits", C.c_int)]


lib.lq_server_create.argtypes = [C.POINTER(_Config), C.POINTER(_P)]
lib.lq_server_create.restype = C.c_int
lib.lq_server_destroy.argtypes = [_P]
lib.lq_server_destroy.restype = C.c_int
lib.lq_server_submit.argtypes = [_P, C.c_uint32, _P, _P, _U64P, C.c_uint32, _U64P]
lib.lq_server_submit.restype = C.c_int
lib.lq_server_wait.argtypes = [_P, C.c_uint64, C.c_double, _IP, _DP, _DP, C.c_uint32, _DP, _IP]
lib.lq_server_wait.restype = C.c_int
lib.lq_server_replay.argtypes = [_P, C.c_uint32, _U32P, C.POINTER(_P), C.POINTER(_P), _U64P, _U64P, _DP, _DP, _DP,
                                 _DP, _IP, _DP]
lib.lq_server_replay.restype = C.c_int
lib.lq_server_stats.argtypes = [_P, _U64P]
lib.lq_server_stats.restype = C.c_int
lib.lq_server_pool_stats.argtypes = [_P, _U64P, _U64P]
lib.lq_server_pool_stats.restype = C.c_int
lib.lq_server_timing.argtypes = [_P, _DP, C.c_int]
lib.lq_server_timing.restype = C.c_int


@dataclass
class Request:
    """One request: its circuit (adjoint form), Hamiltonian and handles."""
    n: int
    depth: int
    ops: list
    coeffs: list
    words: list
    bits: int = 128
    _ops_h: object = None
    _ham_h: object = None

    @property
    def num_params(self) -> int:
        return sum(1 for op in self.ops if op.params)

    def handles(self):
        if self._ops_h is None:
            self._ops_h = to_ops(self.ops, self.bits)
            self._ham_h = hamiltonian_obs(self.coeffs, self.words, self.bits)
        return self._ops_h, self._ham_h


def hea_adjoint_ops(n: int, depth: int, rng: np.random.Generator) -> list:
    ang = rng.uniform(0.0, 2 * math.pi, size=(depth, n, 3))
    ops = []
    for layer in range(depth):
        for w in range(n):
            phi, theta, omega = (float(x) for x in ang[layer, w])
            ops += [CircuitOp("RZ", [w], [phi]), CircuitOp("RY", [w], [theta]), CircuitOp("RZ", [w], [omega])]
        for w in range(0, n - 1, 2):
            ops.append(CircuitOp("CZ", [w, w + 1]))
        for w in range(1, n - 1, 2):
            ops.append(CircuitOp("CZ", [w, w + 1]))
    return ops


def serving_requests(count: int, seed: int = 0, n_min: int = 10, n_max: int = 28, depth_min: int = 4,
                     depth_max: int = 16, terms: int = 50, bits: int = 128) -> list:
    rng = np.random.default_rng(seed)
    ns = np.arange(n_min, n_max + 1)
    p = 2.0 ** (-(ns - n_min) / 2.0)
    p /= p.sum()
    out = []
    for i in range(count):
        n = int(rng.choice(ns, p=p))
        depth = int(rng.integers(depth_min, depth_max + 1))
        r = np.random.default_rng([seed, i])
        ops = hea_adjoint_ops(n, depth, r)
        coeffs, words = random_pauli_hamiltonian(n, terms, seed=int(r.integers(1 << 31)))
        out.append(Request(n, depth, ops, coeffs, words, bits))
    return out


def poisson_arrivals(count: int, qps: float, seed: int = 1) -> np.ndarray:
    gaps = np.random.default_rng(seed).exponential(1.0 / qps, size=count)
    return np.cumsum(gaps) - gaps[0]


class Server:
    def __init__(self, bits: int = 128, batched: bool = True, hbm_budget: int = 0, small_max_qubits: int = -1,
                 max_batch: int = 0, batch_streams: int = 0, large_workers: int = 0, prep_threads: int = 0,
                 lane_workers: int = 0, lane_max_qubits: int = 0):
        cfg = _Config(1 if bits == 128 else 0, int(bool(batched)), int(hbm_budget), int(small_max_qubits),
                      int(max_batch), int(batch_streams), int(large_workers), int(prep_threads),
                      int(lane_workers), int(lane_max_qubits))
        self._h = C.c_void_p()
        check(lib.lq_server_create(C.byref(cfg), C.byref(self._h)))
        self.bits = bits

    def close(self):
        if self._h:
            check(lib.lq_server_destroy(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, req: Request) -> int:
        ops_h, ham_h = req.handles()
        tp = np.arange(req.num_params, dtype=np.uint64)
        rid = C.c_uint64()
        check(lib.lq_server_submit(self._h, req.n, ops_h._h, ham_h._h, tp.ctypes.data_as(_U64P), tp.size,
                                   C.byref(rid)))
        return rid.value

    def wait(self, rid: int, num_params: int, timeout: float = 600.0):
        done, e, lat, path = C.c_int(), C.c_double(), C.c_double(), C.c_int()
        g = np.zeros(max(num_params, 1))
        check(lib.lq_server_wait(self._h, rid, timeout, C.byref(done), C.byref(e), g.ctypes.data_as(_DP), num_params,
                                 C.byref(lat), C.byref(path)))
        if not done.value:
            raise TimeoutError(f"request {rid} not done after {timeout} s")
        return e.value, g[:num_params], lat.value, path.value

    def run(self, req: Request):
        """One request alone: (energy, gradient)."""
        e, g, _, _ = self.wait(self.submit(req), req.num_params)
        return e, g

    def replay(self, reqs: list, arrivals) -> dict:
        """Submit reqs[i] at arrivals[i] seconds and collect every result."""
        count = len(reqs)
        hs = [r.handles() for r in reqs]
        ns = np.asarray([r.n for r in reqs], dtype=np.uint32)
        ops_arr = (C.c_void_p * count)(*[h[0]._h.value for h in hs])
        ham_arr = (C.c_void_p * count)(*[h[1]._h.value for h in hs])
        off = np.zeros(count + 1, dtype=np.uint64)
        off[1:] = np.cumsum([r.num_params for r in reqs])
        tp = np.concatenate([np.arange(r.num_params, dtype=np.uint64) for r in reqs] + [np.zeros(1, np.uint64)])
        arr = np.ascontiguousarray(np.asarray(arrivals, dtype=np.float64))
        en = np.zeros(count)
        gr = np.zeros(int(off[-1]) + 1)
        lat = np.zeros(count)
        path = np.zeros(count, dtype=np.int32)
        wall = C.c_double()
        check(lib.lq_server_replay(self._h, count, ns.ctypes.data_as(_U32P), ops_arr, ham_arr,
                                   tp.ctypes.data_as(_U64P), off.ctypes.data_as(_U64P), arr.ctypes.data_as(_DP),
                                   en.ctypes.data_as(_DP), gr.ctypes.data_as(_DP), lat.ctypes.data_as(_DP),
                                   path.ctypes.data_as(_IP), C.byref(wall)))
        grads = [gr[int(off[i]):int(off[i + 1])] for i in range(count)]
        return {"energies": en, "grads": grads, "latency_ms": lat, "path": path, "wall_s": wall.value,
                "arrivals_s": arr.copy()}

    def timing(self, reset: bool = True) -> dict:
        """Packed-path timing maxima (ms) since the start or the last reset:
        a request's wait from ready to launch, the host time of a launch call,
        a launch's time until the batcher saw it done, and that launch's
        upload / kernel / download time on the GPU."""
        out = np.zeros(6)
        check(lib.lq_server_timing(self._h, out.ctypes.data_as(_DP), int(bool(reset))))
        return {"ready_wait_max_ms": float(out[0]), "launch_call_max_ms": float(out[1]),
                "launch_max_ms": float(out[2]), "slow_launch_upload_ms": float(out[3]),
                "slow_launch_kernel_ms": float(out[4]), "slow_launch_download_ms": float(out[5])}

    def stats(self) -> dict:
        out = np.zeros(6, dtype=np.uint64)
        check(lib.lq_server_stats(self._h, out.ctypes.data_as(_U64P)))
        keys = ("submitted", "completed", "packed", "engine", "launches", "hbm_in_use")
        st = {k: int(v) for k, v in zip(keys, out)}
        made, freed = C.c_uint64(), C.c_uint64()
        check(lib.lq_server_pool_stats(self._h, C.byref(made), C.byref(freed)))
        st.update({"states_made": made.value, "states_freed": freed.value})
        return st


def summarize(res: dict, count: int) -> dict:
    """Latency percentiles over every request; ``req_per_s`` is all requests
    over the whole replay (first arrival to last completion), and
    ``steady_req_per_s`` the completions inside the arrival window (first to
    last arrival) per second of that window -- the departure rate while
    requests arrive, which equals the offered rate when the server keeps up
    and is not diluted by the service time of the last few requests."""
    lat = np.asarray(res["latency_ms"])
    out = {"requests": count, "wall_s": res["wall_s"], "req_per_s": count / res["wall_s"] if res["wall_s"] else None,
           "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
           "mean_ms": float(lat.mean()), "packed": int((np.asarray(res["path"]) == 0).sum())}
    paths = np.asarray(res["path"])
    for code, name in ((0, "packed"), (1, "engine")):
        sel = lat[paths == code]
        if sel.size:
            out[f"{name}_p50_ms"] = float(np.percentile(sel, 50))
            out[f"{name}_p99_ms"] = float(np.percentile(sel, 99))
    arr = res.get("arrivals_s")
    if arr is not None and len(arr) > 1:
        arr = np.asarray(arr, dtype=np.float64)
        window = float(arr[-1] - arr[0])
        done = arr + lat * 1e-3
        if window > 0:
            out["steady_req_per_s"] = float((done <= arr[-1]).sum()) / window
    return out
