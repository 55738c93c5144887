"""Write the golden fixtures in this folder from the reference lightning.qubit.

The reference is the C++ core of pennylane_lightning compiled from its own
sources by ``make -C oracle ref`` (oracle/_ref/liblq_ref.so, driven through
oracle/ref.py). Every fixture stores its inputs (angles, matrices, Pauli words,
coefficients) next to the reference outputs, so the tests rebuild the exact
op lists from the file and never depend on a random generator.

    python tests/golden/make_golden.py            # all fixtures
    python tests/golden/make_golden.py --only NAME

Sizes: the n=30 and n=24 fixtures need ~34 GiB and ~1 GiB of host memory and
a few minutes on 8 cores.
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "pennylane-lightning_amd")]

from oracle import lq_oracle as O  # noqa: E402
from oracle import ref as R  # noqa: E402

import golden_circuits as G  # noqa: E402


def run_ref(n, ops):
    r = R.RefStateVector(n, "double")
    r.apply_ops(ops)
    return r


def fx_rxyz(name):
    n, layers = 20, 10
    ang = np.random.default_rng(0).uniform(0.0, 2 * np.pi, size=(layers, n, 3))
    ops = G.rxyz_cnot_ops(ang)
    r = run_ref(n, ops)
    z = np.array([r.expval_named("PauliZ", [w]) for w in range(n)])
    st = r.get_state()
    p = np.abs(st) ** 2
    pr = r.probs()
    idx = np.sort(np.random.default_rng(100).choice(1 << n, size=256, replace=False))
    np.savez_compressed(os.path.join(HERE, name), angles=ang, expval_z=z, probs_block_sums=pr.reshape(1024, -1).sum(1),
                        sample_idx=idx, sample_probs=pr[idx], sample_amps=st[idx],
                        probs_check=np.abs(p[idx] - pr[idx]).max())


def fx_hea_small(name):
    n, layers = 12, 4
    ang = np.random.default_rng(1).uniform(0.0, 2 * np.pi, size=(layers, n, 3))
    ops = G.hea_ops(ang)
    r = run_ref(n, ops)
    zz = r.expval_obs(("T", [("PauliZ", 0), ("PauliZ", 1)]))
    x = r.expval_named("PauliX", [n - 1])
    np.savez_compressed(os.path.join(HERE, name), angles=ang, state=r.get_state(), expval_z0z1=zz,
                        expval_x_last=x)


def fx_hea30(name):
    n, layers = 30, 20
    ang = np.random.default_rng(1).uniform(0.0, 2 * np.pi, size=(layers, n, 3))
    ops = G.hea_ops(ang)
    t = time.time()
    r = run_ref(n, ops)
    zz = r.expval_obs(("T", [("PauliZ", 0), ("PauliZ", 1)]))
    x = r.expval_named("PauliX", [n - 1])
    z = np.array([r.expval_named("PauliZ", [w]) for w in (0, 7, 15, 29)])
    print(f"  n=30 reference run {time.time() - t:.0f} s", flush=True)
    np.savez_compressed(os.path.join(HERE, name), angles=ang, expval_z0z1=zz, expval_x_last=x,
                        z_wires=np.array([0, 7, 15, 29]), expval_z=z)


def fx_adjoint(name, n, layers, terms, seed_ang, seed_h):
    ang = np.random.default_rng(seed_ang).uniform(0.0, 2 * np.pi, size=(layers, n, 3))
    letters, coeffs = G.random_words(n, terms, seed_h)
    ops = G.strongly_entangling_ops(ang)
    obs = [G.ham_spec(coeffs, letters), ("N", "PauliZ", n // 2)]
    npar = sum(1 for op in ops if op.params)
    tp = list(range(npar))
    t = time.time()
    r = run_ref(n, ops)
    ev = np.array([r.expval_obs(o) for o in obs])
    jac = r.adjoint_jacobian(ops, obs, tp)
    print(f"  adjoint n={n} reference run {time.time() - t:.0f} s", flush=True)
    np.savez_compressed(os.path.join(HERE, name), angles=ang, letters=letters, coeffs=coeffs, jacobian=jac,
                        expvals=ev)


def fx_qft(name):
    n, layers = 12, 8
    rng = np.random.default_rng(5)
    mats, mcx = G.brickwork_inputs(n, layers, rng)
    ops = G.qft_brickwork_ops(n, mats, mcx)
    r = run_ref(n, ops)
    z = np.array([r.expval_named("PauliZ", [w]) for w in range(n)])
    p10 = r.probs(list(range(10)))
    np.savez_compressed(os.path.join(HERE, name), mats=mats, mcx=mcx, expval_z=z, probs_w0_9=p10)


FIXTURES = {
    "rxyz_cnot_n20_l10.npz": fx_rxyz,
    "hea_rot_cz_n12_l4.npz": fx_hea_small,
    "qft_haar_mcx_n12_l8.npz": fx_qft,
    "strongly_entangling_n10_l2_h30.npz": lambda nm: fx_adjoint(nm, 10, 2, 30, 2, 4),
    "strongly_entangling_n24_l6_h200.npz": lambda nm: fx_adjoint(nm, 24, 6, 200, 2, 3),
    "hea_rot_cz_n30_l20.npz": fx_hea30,
}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", default=None)
    args = ap.parse_args()
    if not R.available():
        sys.exit("oracle/_ref/liblq_ref.so is not built: run `make -C oracle ref` first")
    for name, fn in FIXTURES.items():
        if args.only and args.only != name:
            continue
        print(name, flush=True)
        fn(name)


if __name__ == "__main__":
    main()
