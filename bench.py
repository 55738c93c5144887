"""Benchmark: 30-qubit complex128 gate application on one MI355X.

One step = the whole hardware-efficient ansatz of depth 20 on a 30-qubit
complex128 state (16 GiB, resident in HBM): reset to |0...0>, 600 Rot + 580
CZ gates, then <Z0 Z1> and <X29>. The engine fuses the gates into LDS tile
passes; every pass reads and writes the whole state once.

value = HBM % of peak achieved by the tile passes: their algorithmic bytes
(amplitudes read + written, 2 x 16 GiB per pass) over their kernel time,
measured with HIP events on the engine's stream, against 8 TB/s.

cpu_baseline: the reference lightning.qubit C++ (oracle/_ref/liblq_ref.so,
built from the reference sources) on the host cores, on the first gates of
the same circuit at the same size, as effective GB/s with the per-gate byte
model below; the GPU's effective GB/s on the same model is reported beside it.

Per-gate (unfused) byte model: a gate reads and writes the amplitudes it can
change -- Rot: all (2 x 16 GiB), CZ: the |11> quarter (2 x 4 GiB).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "pennylane-lightning_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "30q gate-apply HBM % of peak; adjoint-grad req/s + p50/p99 vs plain GPU path"
HBM_PEAK = 8.0e12

# HBM bytes per specialised pass launch measured with rocprofv3 (FETCH_SIZE and
# WRITE_SIZE in separate runs of this script, each scaled by the same counter
# over torch.neg of a known 2 GiB in / 2 GiB out): (precision, qubits) -> bytes
MEASURED_TRAFFIC = {("c128", 30): 34358717135}


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
            "cores": threads, "kind": "reference",
            "sample": f"first {done} of {len(ops)} gates of the same {n}-qubit circuit, "
                      f"{'complex128' if elem == 16 else 'complex64'}, {dt:.1f} s, OpenMP threads={threads}",
            "gates_per_s": round(done / dt, 4)}


def copy_rate(nbytes, reps=5):
    """HBM rate of a plain streaming kernel (torch.neg: reads and writes
    nbytes each) on torch's current stream, in bytes/s."""
    import torch

    src = torch.ones(nbytes // 16, dtype=torch.complex128, device="cuda")
    dst = torch.empty_like(src)
    torch.neg(src, out=dst)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        torch.neg(src, out=dst)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    torch.cuda.empty_cache()
    return 2 * nbytes / (ms * 1e-3)


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
    ap.add_argument("--workload", choices=["hea30", "adjoint24"], default="hea30",
                    help="hea30: the gate-application line (default); adjoint24: the 24-qubit adjoint gradient")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU sharding is not built "
                                   "(run with --gpus 1)", "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    torch.cuda.set_device(0)
    if args.copy_only:
        r = copy_rate(4 << 30)
        print(json.dumps({"kernel": "torch.neg complex128", "bytes_read": 4 << 30, "bytes_written": 4 << 30,
                          "GBps": round(r / 1e9, 1)}))
        return
    if args.workload == "adjoint24":
        return run_adjoint(args)
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
    st = sv.profile_read(reset=True)
    sv.profile(False)
    pc = sv.pass_counts(reset=True)
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
    value = 100.0 * (st["bytes"] / (st["ms"] * 1e-3)) / HBM_PEAK if st["ms"] > 0 else 0.0
    unfused = sum(gate_bytes(op, n, elem) for op in ops)

    copy_bps = copy_rate((1 << n) * elem)
    cpu = None if args.no_cpu else cpu_baseline(n, ops, args.cpu_seconds, elem)
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
                     "traffic": MEASURED_TRAFFIC.get((args.precision, n)), "traffic_unit": "bytes per launch",
                     "kernel": kernel, "launches_per_step": passes_per_step, "avg_launch_ms": round(pass_ms, 4),
                     "bytes_per_launch": bytes_per_pass,
                     "specialised_passes": pc["specialised"], "generic_passes": pc["generic"],
                     "copy_kernel_GBps": round(copy_bps / 1e9, 1),
                     "frac_of_copy": round(achieved / copy_bps, 4) if copy_bps else None},
        "gates_per_s": round(len(ops) / (ms_step * 1e-3), 1),
        "effective_unfused_GBps": round(unfused / (ms_step * 1e-3) / 1e9, 1),
        "expvals": {"Z0Z1": ev[0], f"X{n-1}": ev[1]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
