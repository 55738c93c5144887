"""Benchmark: 30-qubit complex128 gate application on one MI355X.

One step = the whole hardware-efficient ansatz of depth 20 on a 30-qubit
complex128 state (16 GiB, resident in HBM): reset to |0...0>, 600 Rot + 580
CZ gates, then <Z0 Z1> and <X29>. The engine fuses the gates into LDS tile
passes; every pass reads and writes the whole state once, except the first,
which forms the reset state itself and only writes it.

value = HBM % of peak over the whole job: the tile passes' algorithmic bytes
(amplitudes read + written: 2 x 16 GiB per pass, 16 GiB for the first) over
the wall time of the timed steps, against 8 TB/s. roofline.achieved is the
same bytes over the passes' own kernel time (HIP events on the engine's
stream). roofline.copy_kernel_GBps is the engine's own streaming copy kernel
(nontemporal 16-byte loads and stores, lq_sv_copy) over the same 16 GiB.

serving (the second half of the metric): energy + adjoint-gradient requests
through the batched server and through the plain one-request-at-a-time
path, at a few Poisson arrival rates each; the ratio of the best batched
throughput whose p99 is within the plain path's p99 to the plain path's
best throughput that keeps up with its arrivals. A point keeps up when the
requests completed while requests still arrive come to >= 90% of the
offered rate (steady_req_per_s); req_per_s (all requests over first arrival
to last completion) is reported beside it. serving_c64: the same sweep with
complex64 states (FP64 reductions).

cpu_baseline: the reference lightning.qubit C++ (oracle/_ref/liblq_ref.so,
built from the reference sources) on the host cores, on the first gates of
the same circuit at the same size, as effective GB/s with the per-gate byte
model below; the GPU's effective GB/s on the same model is reported beside it.

Per-gate (unfused) byte model: a gate reads and writes the amplitudes it can
change -- Rot: all (2 x 16 GiB), CZ: the |11> quarter (2 x 4 GiB).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

# before anything starts the HIP runtime: one hardware queue per stream for
# the serving workloads (the engine binding sets the same default)
os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "pennylane-lightning_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "30q gate-apply HBM % of peak; adjoint-grad req/s + p50/p99 vs plain GPU path"
HBM_PEAK = 8.0e12

# HBM bytes per specialised pass launch measured with rocprofv3 (FETCH_SIZE and
# WRITE_SIZE in separate runs of this script, each scaled by the same counter
# over the streaming copy kernel of a known size): (precision, qubits) ->
# (bytes, profile summary holding the runs)
MEASURED_TRAFFIC = {("c128", 30): (32453964496, "profiles/lqpass_hea30_c128_traffic.md")}


def gate_bytes(op, n, elem):
    s = (1 << n) * elem
    if op.name == "CZ":
        return 2 * s / 4
    if op.name == "CNOT":
        return 2 * s / 2
    return 2 * s


def cpu_baseline(n, ops, budget_s, elem):
    try:
        from oracle import ref
    except Exception:
        return None
    if not ref.available():
        return None
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    r = ref.RefStateVector(n, "double" if elem == 16 else "float")
    done, nbytes = 0, 0.0
    t0 = time.perf_counter()
    for op in ops:
        r.apply(op.name, op.wires, op.inverse, op.params)
        done += 1
        nbytes += gate_bytes(op, n, elem)
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    del r
    return {"value": round(nbytes / dt / 1e9, 3), "unit": "GB/s effective (per-gate byte model)",
            "cores": threads, "kind": "reference", "host": host_info(),
            "sample": f"first {done} of {len(ops)} gates of the same {n}-qubit circuit, "
                      f"{'complex128' if elem == 16 else 'complex64'}, {dt:.1f} s, OpenMP threads={threads}",
            "gates_per_s": round(done / dt, 4)}


def copy_rate(sv, reps=5):
    """HBM rate of the engine's streaming copy kernel (lq_sv_copy: reads and
    writes the whole state) between two states of sv's size, in bytes/s."""
    other = type(sv)(sv.num_qubits)
    other.copy_from(sv)
    other.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        other.copy_from(sv)
    other.synchronize()
    dt = (time.perf_counter() - t0) / reps
    nbytes = len(sv) * (16 if type(sv).__name__.endswith("128") else 8)
    del other
    return 2 * nbytes / dt


def host_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"nproc": usable, "cpu_count": os.cpu_count(), "cpu_model": model}


def cpu_adjoint_baseline(n, layers, terms):
    """Seconds for the reference (oracle/_ref) to run the forward circuit and
    the AdjointJacobian on the host cores."""
    try:
        from oracle import ref
        from oracle import lq_oracle as O
    except Exception:
        return None
    if not ref.available():
        return None
    from lightning_amd import workloads as W
    ops = W.strongly_entangling(n, layers, seed=2)
    coeffs, words = W.random_pauli_hamiltonian(n, terms, seed=3)
    oops = [O.Op(o.name, o.wires, o.params) for o in ops]
    H = ("H", coeffs, [("T", w) for w in words])
    r = ref.RefStateVector(n, "double")
    t0 = time.perf_counter()
    r.apply_ops(oops)
    r.adjoint_jacobian(oops, [H], list(range(sum(1 for o in ops if o.params))))
    dt = time.perf_counter() - t0
    del r
    return dt


def run_adjoint(args):
    """24-qubit complex128 strongly-entangling circuit, 6 layers (432 trainable
    parameters, 576 ops), Hamiltonian of 200 random Pauli words (I/X/Y/Z with
    probabilities 0.7/0.1/0.1/0.1, N(0,1) coefficients). One step = reset +
    forward + energy + the full 432-parameter adjoint Jacobian."""
    import torch
    from lightning_amd import lightning_qubit_ops as L
    from lightning_amd import workloads as W

    n, layers, terms = 24, 6, 200
    ops = W.strongly_entangling(n, layers, seed=2)
    coeffs, words = W.random_pauli_hamiltonian(n, terms, seed=3)
    sv = L.StateVectorC128(n)
    H = W.hamiltonian_obs(coeffs, words, 128)
    ops_s = W.to_ops(ops, 128)
    tp = list(range(sum(1 for o in ops if o.params)))
    adj = L.algorithms.AdjointJacobianC128()

    def step():
        sv.resetStateVector()
        W.apply_circuit(sv, ops)
        return adj(sv, [H], ops_s, tp, return_expvals=True)

    for _ in range(args.warmup):
        step()
    L.jit_wait()
    step()
    sv.synchronize()
    sv.pass_counts(reset=True)
    sv.profile(True)
    sv.profile_read(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        jac, ev = step()
    sv.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    st = sv.profile_read(reset=True)
    sv.profile(False)
    pc = sv.pass_counts(reset=True)
    ms = (t1 - t0) * 1e3 / args.steps
    S = (1 << n) * 16
    cpu_n = 20
    cpu_s = None if args.no_cpu else cpu_adjoint_baseline(cpu_n, layers, terms)
    out = {
        "metric": "24-qubit adjoint gradient (432 parameters, 200-term Hamiltonian)",
        "value": round(1e3 / ms, 3), "unit": "gradients/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "complex128", "data": "synthetic",
        "config": {"workload": "strongly-entangling 24 qubits x 6 layers, RZ RY RZ + CNOT(i, i+r), 200-term "
                               "Pauli Hamiltonian, energy + 432-parameter adjoint Jacobian",
                   "n_qubits": n, "layers": layers, "terms": terms, "params": len(tp)},
        "passes_per_step": st["passes"] / args.steps, "pass_ms_per_step": round(st["ms"] / args.steps, 3),
        "pass_GBps": round(st["bytes"] / (st["ms"] * 1e-3) / 1e9, 1) if st["ms"] else None,
        "bytes_per_step_over_S": round(st["bytes"] / args.steps / S, 1),
        "specialised_passes": pc["specialised"], "generic_passes": pc["generic"],
        "energy": float(ev[0]), "grad_norm": float(np.linalg.norm(jac)),
        "cpu_baseline": None if cpu_s is None else {
            "value": round(cpu_s, 3), "unit": f"s per gradient at {cpu_n} qubits (same structure)",
            "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), "kind": "reference",
            "sample": f"reference AdjointJacobian, {cpu_n} qubits x {layers} layers, {terms} terms; the "
                      f"24-qubit run is ~16x the state"},
    }
    print(json.dumps(out))


def run_adjoint_multi(args):
    """Several observables in one adjoint Jacobian: the 24-qubit, 6-layer
    strongly entangling circuit (432 parameters) with one named observable
    and with three (<Z_0>, <X_11>, <Y_23>), each call = reset + forward + the
    full Jacobian (one row per observable). Three observables share one
    backward sweep once that program's kernels are compiled."""
    from lightning_amd import lightning_qubit_ops as L
    from lightning_amd import workloads as W

    n, layers = 24, 6
    ops = W.strongly_entangling(n, layers, seed=2)
    sv = L.StateVectorC128(n)
    ops_s = W.to_ops(ops, 128)
    tp = list(range(sum(1 for o in ops if o.params)))
    adj = L.algorithms.AdjointJacobianC128()
    named = [("PauliZ", 0), ("PauliX", 11), ("PauliY", 23)]
    res = {}
    for m in (1, 3):
        obs = [L.NamedObsC128(nm, [w]) for nm, w in named[:m]]

        def step():
            sv.resetStateVector()
            W.apply_circuit(sv, ops)
            return adj(sv, obs, ops_s, tp)

        for _ in range(max(args.warmup, 1)):
            step()
        L.jit_wait()
        step()
        sv.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        sv.synchronize()
        res[m] = (time.perf_counter() - t0) * 1e3 / args.steps
    print(json.dumps({"metric": "adjoint Jacobian with 1 and 3 observables (24 qubits, 432 parameters)",
                      "ms_one_observable": round(res[1], 3), "ms_three_observables": round(res[3], 3),
                      "ratio": round(res[3] / res[1], 3), "n_gpus": 1, "dtype": "complex128", "data": "synthetic",
                      "steps": args.steps}))


def run_graph(args):
    """Circuit replays from a HIP graph against eager application: the
    hardware-efficient ansatz in adjoint form (per layer and wire RZ RY RZ,
    then CZ on even and odd pairs), 10 layers, complex128, at 12, 16 and 20
    qubits. Eager: reset + applyOperations + synchronize; replay: new
    parameters into the recorded graph + synchronize. Each figure is the mean
    over args.steps circuits with fresh angles."""
    from lightning_amd import lightning_qubit_ops as L
    from lightning_amd import workloads as W

    rng = np.random.default_rng(0)
    rows = []
    for n in (12, 16, 20):
        ops = W.hea_rot_cz_adjoint_form(n, 10, seed=1)
        npar = sum(len(o.params) for o in ops)
        sv = L.StateVectorC128(n)
        rec = sv.record(W.to_ops(ops, 128))
        sets = [rng.uniform(0, 2 * math.pi, npar) for _ in range(args.steps)]

        def bind(vals):
            it = iter(vals)
            return [W.CircuitOp(o.name, o.wires, [next(it) for _ in o.params], o.inverse, o.matrix, o.ctrls, o.cvals)
                    for o in ops]

        bound = [bind(v) for v in sets]
        for _ in range(max(args.warmup, 1)):
            sv.resetStateVector()
            W.apply_circuit(sv, bound[0])
            sv.synchronize()
            rec.replay(sets[0])
            sv.synchronize()
        L.jit_wait()
        rec.replay(sets[0])
        sv.resetStateVector()
        W.apply_circuit(sv, bound[0])
        sv.synchronize()
        t0 = time.perf_counter()
        for b in bound:
            sv.resetStateVector()
            W.apply_circuit(sv, b)
            sv.synchronize()
        eager = (time.perf_counter() - t0) * 1e3 / len(bound)
        t0 = time.perf_counter()
        for v in sets:
            rec.replay(v)
            sv.synchronize()
        replay = (time.perf_counter() - t0) * 1e3 / len(sets)
        rows.append({"n_qubits": n, "gates": len(ops), "eager_ms": round(eager, 4), "replay_ms": round(replay, 4),
                     "speedup": round(eager / replay, 3), **rec.stats()})
    print(json.dumps({"metric": "circuit time, HIP-graph replay vs eager (HEA 10 layers, complex128)", "rows": rows,
                      "n_gpus": 1, "dtype": "complex128", "data": "synthetic", "steps": args.steps}))


def serving_sweep(count, qps_batched, qps_plain, bits=128, seed=0, server_kw=None):
    """Replay `count` requests at each arrival rate through the batched
    server, then through the plain path; returns the summary object."""
    from lightning_amd import serving as S

    reqs = S.serving_requests(count, seed=seed, bits=bits)
    t0 = time.perf_counter()
    for r in reqs:
        r.handles()
    setup_s = time.perf_counter() - t0
    warm = [r for r in reqs if r.n <= 14][:64]
    warm_big = [r for r in S.serving_requests(count, seed=seed + 1, bits=bits) if r.n > 12]
    for r in warm_big:
        r.handles()
    rows = []
    # the batched sweep runs first: specialised kernels compiled in the
    # background during a sweep stay cached, so any such gain goes to the
    # plain path
    for batched, sweep in ((True, qps_batched), (False, qps_plain)):
        if not sweep:
            continue
        srv = S.Server(bits=bits, batched=batched, **(server_kw or {}))
        srv.replay(warm, np.zeros(len(warm)))
        # untimed warm-up of the server itself: the engine-path requests of
        # another stream of the same distribution (other seed), so the pool
        # holds states of the sizes that occur and the structures that recur
        # are planned before the first timed point (without it the first
        # point of a sweep ran cold, up to 2x slower)
        srv.replay(warm_big, np.zeros(len(warm_big)))
        for q in sweep:
            st0 = srv.stats()
            srv.timing(reset=True)
            res = srv.replay(reqs, S.poisson_arrivals(len(reqs), q, seed=7))
            st1 = srv.stats()
            tm = srv.timing(reset=True)
            summ = S.summarize(res, len(reqs))
            if batched:
                # packed path: longest ready -> launch wait and launch (call to done)
                summ.update({"packed_ready_wait_max_ms": round(tm["ready_wait_max_ms"], 2),
                             "packed_launch_max_ms": round(tm["launch_max_ms"], 2)})
            summ.update({"mode": "batched" if batched else "plain", "offered_qps": q,
                         "states_made": st1["states_made"] - st0["states_made"],
                         "states_freed": st1["states_freed"] - st0["states_freed"],
                         "hbm_in_use_GiB": round(st1["hbm_in_use"] / 2 ** 30, 1)})
            rows.append(summ)
            print(json.dumps({"serving_point": summ}), file=sys.stderr, flush=True)
        srv.close()

    # a point keeps up with its arrivals when, while requests arrive, it
    # completes them at >= 90% of the offered rate (steady_req_per_s); the
    # plain reference point is its fastest such point, and the batched point
    # its fastest such point whose p99 is within the plain point's p99
    def keeps_up(r):
        return r.get("steady_req_per_s", r["req_per_s"]) >= 0.9 * r["offered_qps"]

    def rate(r):
        return r.get("steady_req_per_s", r["req_per_s"])

    plain = [r for r in rows if r["mode"] == "plain"]
    if not plain or not qps_batched:  # one mode only (profiling runs): no ratio
        return {"ratio_at_equal_p99": None, "requests": count, "sweep": rows, "setup_s": round(setup_s, 2)}
    plain_ok = [r for r in plain if keeps_up(r)] or [min(plain, key=lambda r: r["p99_ms"])]
    best_plain = max(plain_ok, key=rate)
    ok = [r for r in rows if r["mode"] == "batched" and keeps_up(r) and r["p99_ms"] <= best_plain["p99_ms"]]
    best_b = max(ok, key=rate) if ok else None
    ratio = rate(best_b) / rate(best_plain) if best_b else None
    return {"ratio_at_equal_p99": round(ratio, 3) if ratio else None, "requests": count,
            "dtype": "complex128" if bits == 128 else "complex64",
            "offered_qps": {"batched": list(qps_batched), "plain": list(qps_plain)},
            "plain_best": best_plain, "batched_at_equal_p99": best_b, "sweep": rows, "setup_s": round(setup_s, 2)}


def run_serving(args):
    """Serving workload: requests with n in [10, 28] (P(n) ~ 2^-((n-10)/2)),
    depth 4..16 hardware-efficient layers in adjoint form, 50-term Pauli
    Hamiltonian; each asks for the energy and its full adjoint gradient
    (complex128). Poisson arrivals at each rate of a sweep, through the
    batched server and through the plain one-request-at-a-time path."""
    def rates(spec, default):
        spec = args.qps or spec or default
        return [] if spec == "none" else [float(x) for x in spec.split(",")]

    qps_plain = rates(args.sweep_plain, "40,60,80,100,120,140")
    qps_batched = rates(args.sweep_batched, "80,120,160,200,240,300")
    kw = {}
    if args.serve_small_max is not None:
        kw["small_max_qubits"] = args.serve_small_max
    if args.serve_workers is not None:
        kw["large_workers"] = args.serve_workers
    if args.serve_prep is not None:
        kw["prep_threads"] = args.serve_prep
    if args.serve_streams is not None:
        kw["batch_streams"] = args.serve_streams
    if args.serve_lane is not None:
        kw["lane_workers"] = args.serve_lane
    sv = serving_sweep(args.requests, qps_batched, qps_plain, bits=args.serve_bits, server_kw=kw)
    out = {
        "metric": "adjoint-grad req/s at equal p99, batched vs plain GPU path",
        "value": sv["ratio_at_equal_p99"], "unit": "x plain throughput (both keeping up with arrivals, batched p99 <= plain p99)",
        "n_gpus": 1, "steps": len(sv["sweep"]), "warmup": 64, "ms_per_step": None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "complex128" if args.serve_bits == 128 else "complex64",
        "data": "synthetic",
        "config": {"workload": f"{args.requests} requests, n 10..28 with P(n) ~ 2^-((n-10)/2), depth 4..16 HEA "
                               f"(RZ RY RZ + CZ), 50-term Pauli Hamiltonian, energy + full adjoint gradient",
                   "requests": args.requests, "jit": os.environ.get("LQ_JIT", "1")},
        "serving": sv,
    }
    print(json.dumps(out))


def run_qft33(args):
    """33-qubit QFT (33 H, 528 ControlledPhaseShift(pi/2^k), 16 SWAP), 8
    brickwork layers of Haar-random 4x4 unitaries (applyMatrix) each followed
    by one 3-controlled X on random wires; outputs <Z_i> for all wires and the
    marginal probabilities of wires 0-9. complex128 (128 GiB) timed, then the
    complex64 run (64 GiB) and its fidelity and output errors against it."""
    import torch
    from lightning_amd import lightning_qubit_ops as L
    from lightning_amd import workloads as W

    n = args.qubits if args.qubits != 30 else 33
    ops = W.qft_brickwork_mcx(n, layers=8, mcx_per_layer=1, seed=5)

    def outputs(sv, M):
        z = M.expval_named_batch(["PauliZ"] * n, list(range(n)))
        return np.asarray(z), np.asarray(M.probs(list(range(10))))

    sv = L.StateVectorC128(n)
    M = L.MeasurementsC128(sv)

    def step():
        sv.resetStateVector()
        W.apply_circuit(sv, ops)
        return outputs(sv, M)

    for _ in range(args.warmup):
        step()
    L.jit_wait()
    sv.synchronize()
    sv.pass_counts(reset=True)
    sv.profile(True)
    sv.profile_read(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        z128, p128 = step()
    sv.synchronize()
    t1 = time.perf_counter()
    st = sv.profile_read(reset=True)
    sv.profile(False)
    pc = sv.pass_counts(reset=True)
    ms = (t1 - t0) * 1e3 / args.steps
    sv64 = L.StateVectorC64(n)
    W.apply_circuit(sv64, ops)
    z64, p64 = outputs(sv64, L.MeasurementsC64(sv64))
    ov = sv.overlap(sv64)
    n64 = sv64.overlap(sv64).real
    fid = abs(ov) ** 2 / n64
    out = {
        "metric": f"{n}-qubit QFT + Haar brickwork + MCX", "value": round(st["bytes"] / (st["ms"] * 1e-3) / 1e9, 1)
        if st["ms"] else None, "unit": "GB/s (pass bytes / pass time)", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 1), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "complex128", "data": "synthetic",
        "config": {"workload": f"QFT({n}) + 8 Haar 4x4 brickwork layers + 8 3-controlled X, <Z_i> + probs(0..9)",
                   "n_qubits": n, "ops": len(ops)},
        "passes_per_step": st["passes"] / args.steps, "hbm_frac": round(st["bytes"] / (st["ms"] * 1e-3) / HBM_PEAK, 4)
        if st["ms"] else None, "specialised_passes": pc["specialised"], "generic_passes": pc["generic"],
        "complex64": {"fidelity": fid, "max_abs_dZ": float(np.abs(z64 - z128).max()),
                      "max_abs_dprobs": float(np.abs(p64 - p128).max()), "norm64": n64},
        "z_mean_abs": float(np.abs(z128).mean()),
    }
    print(json.dumps(out))


def run_mfma5(args):
    """Matrix cores for fused dense blocks, complex64: one Haar-random 32x32
    unitary on wires 25..29 of a 30-qubit complex64 state (the five lowest
    index bits), applied (a) as a 64x64 real GEMM per 32-amplitude group on
    the MFMA units, (b) as a dense tile op in the engine's default path, and
    (c) the gates such a block fuses in the hardware-efficient ansatz (one
    layer on those wires: 5 Rot + 4 CZ) through the engine's specialised
    passes. Times per application (stream-synchronised), HBM rate over
    8 GiB read + 8 GiB written, and the GEMM's fp32 rate (256 flop per
    amplitude)."""
    from lightning_amd import lightning_qubit_ops as L

    n = args.qubits
    rng = np.random.default_rng(3)
    z = rng.standard_normal((32, 32)) + 1j * rng.standard_normal((32, 32))
    q, r = np.linalg.qr(z)
    u = (q * (np.diag(r) / np.abs(np.diag(r)))).astype(np.complex64)
    wires = list(range(n - 5, n))
    sv = L.StateVectorC64(n)
    sv.resetStateVector()
    sv.synchronize()
    S = (1 << n) * 8
    reps = max(args.steps, 3)

    def timed(fn):
        for _ in range(2):
            fn()
        sv.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        sv.synchronize()
        return (time.perf_counter() - t0) * 1e3 / reps

    def dense():
        sv.applyMatrix(u, wires, False)
        sv.synchronize()

    L.set_mfma_dense5(True)
    ms_mfma = timed(dense)
    L.set_mfma_dense5(False)
    ms_tile = timed(dense)
    ang = rng.uniform(0, 2 * np.pi, size=(5, 3))

    def layer():
        for j, w in enumerate(wires):
            sv.Rot([w], False, [float(x) for x in ang[j]])
        for a in (n - 5, n - 3, n - 4, n - 2):
            sv.CZ([a, a + 1], False, [])
        sv.synchronize()

    layer()
    L.jit_wait()
    ms_gates = timed(layer)
    out = {"metric": "complex64 fused 2^5 x 2^5 block: matrix cores vs engine paths", "n_qubits": n,
           "dtype": "complex64", "n_gpus": 1, "data": "synthetic", "reps": reps,
           "mfma_gemm": {"ms": round(ms_mfma, 3), "GBps": round(2 * S / ms_mfma / 1e6, 1),
                         "fp32_TFLOPs": round(256 * (1 << n) / ms_mfma / 1e9, 1)},
           "dense_tile_op": {"ms": round(ms_tile, 3), "GBps": round(2 * S / ms_tile / 1e6, 1)},
           "factored_gates_5rot_4cz": {"ms": round(ms_gates, 3), "GBps": round(2 * S / ms_gates / 1e6, 1)}}
    print(json.dumps(out))


def run_precision(args):
    """complex64 against complex128 for each workload: state fidelity and
    output errors (expvals, probabilities, energies, gradients)."""
    from lightning_amd import lightning_qubit_ops as L
    from lightning_amd import workloads as W
    from lightning_amd import serving as S

    rows = {}

    def pair(n, ops):
        a = L.StateVectorC128(n)
        W.apply_circuit(a, ops)
        b = L.StateVectorC64(n)
        W.apply_circuit(b, ops)
        return a, b

    # 20-qubit RX/RY/RZ + CNOT, 10 layers: <Z_i> and probs
    n = 20
    a, b = pair(n, W.layers_rxyz_cnot(n, 10, seed=0))
    ma, mb = L.MeasurementsC128(a), L.MeasurementsC64(b)
    za = np.array([ma.expval("PauliZ", [w]) for w in range(n)])
    zb = np.array([mb.expval("PauliZ", [w]) for w in range(n)])
    rows["rxyz_cnot_20"] = {"fidelity": abs(a.overlap(b)) ** 2, "max_abs_dZ": float(np.abs(za - zb).max()),
                            "max_abs_dprobs": float(np.abs(ma.probs() - mb.probs()).max())}
    del a, b, ma, mb
    # 30-qubit HEA depth 20: <Z0 Z1>, <X29>
    n = 30
    a, b = pair(n, W.hea_rot_cz(n, 20, seed=1))
    zz = [L.TensorProdObsC128([L.NamedObsC128("PauliZ", [0]), L.NamedObsC128("PauliZ", [1])]),
          L.TensorProdObsC64([L.NamedObsC64("PauliZ", [0]), L.NamedObsC64("PauliZ", [1])])]
    ea = [L.MeasurementsC128(a).expval(zz[0]), L.MeasurementsC128(a).expval("PauliX", [29])]
    eb = [L.MeasurementsC64(b).expval(zz[1]), L.MeasurementsC64(b).expval("PauliX", [29])]
    rows["hea_30"] = {"fidelity": abs(a.overlap(b)) ** 2, "abs_dZ0Z1": abs(ea[0] - eb[0]), "abs_dX29": abs(ea[1] - eb[1])}
    del a, b
    # 24-qubit adjoint: energy and the 432 gradients
    n = 24
    ops = W.strongly_entangling(n, 6, seed=2)
    coeffs, words = W.random_pauli_hamiltonian(n, 200, seed=3)
    tp = list(range(sum(1 for o in ops if o.params)))
    res = []
    for bits in (128, 64):
        sv = (L.StateVectorC128 if bits == 128 else L.StateVectorC64)(n)
        W.apply_circuit(sv, ops)
        adj = (L.algorithms.AdjointJacobianC128 if bits == 128 else L.algorithms.AdjointJacobianC64)()
        j, e = adj(sv, [W.hamiltonian_obs(coeffs, words, bits)], W.to_ops(ops, bits), tp, return_expvals=True)
        res.append((np.asarray(j, dtype=np.float64), float(e[0])))
        del sv
    rows["adjoint_24"] = {"abs_denergy": abs(res[0][1] - res[1][1]),
                          "max_abs_dgrad": float(np.abs(res[0][0] - res[1][0]).max()),
                          "rel_dgrad_norm": float(np.linalg.norm(res[0][0] - res[1][0]) / np.linalg.norm(res[0][0]))}
    # serving sample (n <= 16, packed kernel): energies and gradients
    reqs = S.serving_requests(64, seed=11, n_max=16)
    outs = []
    for bits in (128, 64):
        for r in reqs:
            r.bits, r._ops_h, r._ham_h = bits, None, None
        srv = S.Server(bits=bits, batched=True)
        outs.append(srv.replay(reqs, np.zeros(len(reqs))))
        srv.close()
    dg = max(float(np.abs(x - y).max()) for x, y in zip(outs[0]["grads"], outs[1]["grads"]))
    rows["serving_sample_64"] = {"max_abs_denergy": float(np.abs(outs[0]["energies"] - outs[1]["energies"]).max()),
                                 "max_abs_dgrad": dg}
    print(json.dumps({"metric": "complex64 error against complex128", "rows": rows, "dtype": "complex64",
                      "n_gpus": 1, "data": "synthetic"}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--layers", type=int, default=20)
    ap.add_argument("--precision", choices=["c128", "c64"], default="c128")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-only", action="store_true",
                    help="only run the plain streaming kernel (4 GiB read + 4 GiB written per launch)")
    ap.add_argument("--workload", choices=["hea30", "adjoint24", "serving", "qft33", "precision", "mfma5",
                                           "adjoint_multi", "graph"],
                    default="hea30",
                    help="hea30: the gate-application line (default); adjoint24: the 24-qubit adjoint gradient; "
                         "serving: energy + gradient requests, batched vs plain; qft33: 33-qubit QFT + Haar "
                         "brickwork + 3-controlled X with complex64 fidelity; precision: complex64 error per "
                         "workload against the complex128 run")
    ap.add_argument("--requests", type=int, default=1024)
    ap.add_argument("--no-serving", action="store_true", help="default workload: skip the serving measurement")
    ap.add_argument("--serving-requests", type=int, default=1024)
    ap.add_argument("--no-serving-c64", action="store_true", help="skip the complex64 serving sweep")
    ap.add_argument("--serving-qps-batched", default="550,750,950,1150,1350")
    ap.add_argument("--serving-qps-plain", default="170,220,260")
    ap.add_argument("--qps", default="", help="comma-separated arrival rates for the serving sweep")
    ap.add_argument("--serve-bits", type=int, choices=[64, 128], default=128,
                    help="serving workload: state precision (complex64 keeps FP64 reductions)")
    ap.add_argument("--sweep-batched", default="",
                    help="serving workload: arrival rates of the batched server ('none': skip that mode)")
    ap.add_argument("--sweep-plain", default="", help="serving workload: arrival rates of the plain path ('none': skip)")
    ap.add_argument("--serve-small-max", type=int, default=None,
                    help="serving: largest qubit count run in the packed kernel (default: the server's)")
    ap.add_argument("--serve-prep", type=int, default=None, help="serving workload: packed-program host threads")
    ap.add_argument("--serve-streams", type=int, default=None, help="serving workload: packed launches in flight")
    ap.add_argument("--serve-lane", type=int, default=None,
                    help="serving workload: engine workers reserved for requests of <= 16 qubits (negative: none)")
    ap.add_argument("--serve-workers", type=int, default=None,
                    help="serving: concurrent engine workers for larger requests (default: the server's)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU sharding is not built "
                                   "(run with --gpus 1)", "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    torch.cuda.set_device(0)
    if args.copy_only:
        from lightning_amd import lightning_qubit_ops as L
        sv = L.StateVectorC128(28)
        sv.synchronize()
        r = copy_rate(sv)
        print(json.dumps({"kernel": "stream_copy_kernel (lq_sv_copy) complex128", "bytes_read": 4 << 30,
                          "bytes_written": 4 << 30, "GBps": round(r / 1e9, 1)}))
        return
    if args.workload == "adjoint24":
        return run_adjoint(args)
    if args.workload == "serving":
        return run_serving(args)
    if args.workload == "qft33":
        return run_qft33(args)
    if args.workload == "precision":
        return run_precision(args)
    if args.workload == "mfma5":
        return run_mfma5(args)
    if args.workload == "adjoint_multi":
        return run_adjoint_multi(args)
    if args.workload == "graph":
        return run_graph(args)
    from lightning_amd import lightning_qubit_ops as L
    from lightning_amd import workloads as W

    n, elem = args.qubits, (16 if args.precision == "c128" else 8)
    ops = W.hea_rot_cz(n, args.layers, seed=1)
    SV = L.StateVectorC128 if elem == 16 else L.StateVectorC64
    sv = SV(n)
    sfx = "C128" if elem == 16 else "C64"
    zz = getattr(L, "TensorProdObs" + sfx)([getattr(L, "NamedObs" + sfx)("PauliZ", [0]),
                                             getattr(L, "NamedObs" + sfx)("PauliZ", [1])])
    x_last = getattr(L, "NamedObs" + sfx)("PauliX", [n - 1])
    M = (L.MeasurementsC128 if elem == 16 else L.MeasurementsC64)(sv)

    def step():
        sv.resetStateVector()
        W.apply_circuit(sv, ops)
        return M.expval(zz), M.expval(x_last)

    for _ in range(args.warmup):
        step()
    # the warmup queued the compiles of the specialised pass kernels; wait for
    # them and run once more so the kernels are loaded before the timed steps
    L.jit_wait()
    step()
    sv.synchronize()
    sv.pass_counts(reset=True)
    sv.profile(True)
    sv.profile_read(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ev = step()
    sv.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    recs = sv.profile_passes()
    st = sv.profile_read(reset=True)
    sv.profile(False)
    pc = sv.pass_counts(reset=True)
    per_step = int(round(st["passes"] / args.steps)) if args.steps else 0
    # the last step's passes, in launch order
    pass_table = [[round(r["ms"], 3), round(r["bytes"] / 1e9, 2), r["ops"], r["stages"], r["transposes"],
                   int(r["specialised"]), r["tile_bits"]] for r in (recs[-per_step:] if per_step else [])]
    real = "double" if elem == 16 else "float"
    if pc["generic"] == 0:
        kernel = "lqpass (specialised register tile pass, hiprtc)"
    elif pc["specialised"] == 0:
        kernel = f"tile_reg_kernel<{real},1,4>"
    else:
        kernel = f"lqpass + tile_reg_kernel<{real},1,4> (mixed)"

    ms_step = (t1 - t0) * 1e3 / args.steps
    passes_per_step = st["passes"] / args.steps
    pass_ms = st["ms"] / max(st["passes"], 1)
    bytes_per_pass = st["bytes"] / max(st["passes"], 1)
    achieved = bytes_per_pass / (pass_ms * 1e-3) if pass_ms > 0 else 0.0
    # whole job: the passes' algorithmic bytes over the wall time of the
    # timed steps (fill, reductions and launch gaps included)
    value = 100.0 * (st["bytes"] / args.steps) / (ms_step * 1e-3) / HBM_PEAK if ms_step > 0 else 0.0
    copy_bps = copy_rate(sv)
    cpu = None if args.no_cpu else cpu_baseline(n, ops, args.cpu_seconds, elem)
    traffic = MEASURED_TRAFFIC.get((args.precision, n))
    serving = serving_c64 = None
    if not args.no_serving:
        del sv, M
        serving = serving_sweep(args.serving_requests, [float(x) for x in args.serving_qps_batched.split(",")],
                                [float(x) for x in args.serving_qps_plain.split(",")])
        if not args.no_serving_c64:
            # the same stream with complex64 states (FP64 reductions)
            serving_c64 = serving_sweep(args.serving_requests,
                                        [float(x) for x in args.serving_qps_batched.split(",")],
                                        [float(x) for x in args.serving_qps_plain.split(",")], bits=64)
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "% of 8 TB/s HBM peak (30-qubit complex128 gate passes)" if elem == 16 else
                "% of 8 TB/s HBM peak (complex64 gate passes)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex128" if elem == 16 else "complex64",
        "data": "synthetic",
        "config": {"workload": f"{n}-qubit hardware-efficient ansatz, depth {args.layers}: Rot on every wire "
                               f"then CZ(i,i+1) even i then odd i; <Z0 Z1> and <X{n-1}>",
                   "n_qubits": n, "layers": args.layers, "gates": len(ops), "seed": 1},
        "roofline": {"bound": "hbm", "achieved": round(achieved / 1e9, 1), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK, 4),
                     "traffic": traffic[0] if traffic else None, "traffic_unit": "bytes per launch",
                     "traffic_source": traffic[1] if traffic else None,
                     "kernel": kernel, "launches_per_step": passes_per_step, "avg_launch_ms": round(pass_ms, 4),
                     "bytes_per_launch": bytes_per_pass,
                     "specialised_passes": pc["specialised"], "generic_passes": pc["generic"],
                     "copy_kernel_GBps": round(copy_bps / 1e9, 1),
                     "frac_of_copy": round(achieved / copy_bps, 4) if copy_bps else None},
        "pass_table": {"columns": ["kernel_ms", "GB", "ops", "stages", "lds_transposes", "specialised", "tile_bits"],
                       "rows": pass_table},
        "gates_per_s": round(len(ops) / (ms_step * 1e-3), 1),
        "expvals": {"Z0Z1": ev[0], f"X{n-1}": ev[1]},
        "cpu_baseline": cpu,
        "serving": serving,
        "serving_c64": serving_c64,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
