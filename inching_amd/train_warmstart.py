"""Train the warm-start network on modes computed by the plain solver.

    python -m inching_amd.train_warmstart --structures 2000 --epochs 20 --out inching_amd/weights/warmstart_gnn.npz

Training set: seeded globules (0.058 atoms/A^3) with atom counts log-uniform in
[--min-atoms, --max-atoms]; their lowest K non-trivial modes come from the batched
matrix-free float64 TRLM (``batch.solve_batch``), solved in batches; the network is
trained with Adam on the subspace loss (``warmstart.subspace_loss``). Every forward and
backward pass of the network runs in the library's HIP kernels.
"""
import argparse
import json
import math
import sys
import time

import numpy as np
import torch


def build_samples(count, lo, hi, K, seed, batch=128, log=print):
    from . import synth
    from .batch import solve_batch
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(math.log(lo), math.log(hi), size=count)).astype(np.int64)
    samples = []
    for b0 in range(0, count, batch):
        t = time.time()
        structs = [synth.globule(int(n), seed=seed * 100003 + b0 + i) for i, n in enumerate(sizes[b0:b0 + batch])]
        r = solve_batch(structs, k=K, precision="mf64", tol_rel=1e-7, maxiter=300)
        A = r.A
        rp = A.rowptr
        for s in range(len(structs)):
            a0, a1 = int(r.seg_atoms[s]), int(r.seg_atoms[s + 1])
            rows = A.rowptr[a0:a1 + 1]
            col = (A.col[int(rows[0]):int(rows[-1])] - a0).contiguous()
            rows = (rows - rows[0]).contiguous()
            U = r.eigvec[:, 3 * a0:3 * a1].T.contiguous()          # [3n, K], orthonormal per structure
            xyz = A.xyz[a0:a1].contiguous()
            from .hessian import rigid_body_basis
            samples.append({"xyz": xyz, "rowptr": rows, "col": col, "U": U, "Q": rigid_body_basis(xyz)})
        del rp
        log("batch %d-%d solved (%d converged) in %.1fs" % (b0, b0 + len(structs), r.converged, time.time() - t))
    return samples, sizes


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--structures", type=int, default=2000)
    ap.add_argument("--min-atoms", type=int, default=1000)
    ap.add_argument("--max-atoms", type=int, default=20000)
    ap.add_argument("--modes", type=int, default=32)
    ap.add_argument("--channels", type=int, default=32)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    from . import warmstart as W

    def log(msg):
        print(msg, flush=True)

    t0 = time.time()
    samples, sizes = build_samples(a.structures, a.min_atoms, a.max_atoms, a.modes, a.seed, log=log)
    t_data = time.time() - t0
    net = W.WarmStartNet(channels=a.channels, seed=a.seed)
    t1 = time.time()
    hist = W.train(net, samples, epochs=a.epochs, lr=a.lr, log=log)
    t_train = time.time() - t1
    meta = {"structures": a.structures, "atoms_min": a.min_atoms, "atoms_max": a.max_atoms, "modes": a.modes,
            "channels": a.channels, "epochs": a.epochs, "lr": a.lr, "seed": a.seed, "loss_history": hist,
            "seconds_data": round(t_data, 1), "seconds_train": round(t_train, 1), "total_atoms": int(sizes.sum())}
    net.save(a.out or W.WEIGHTS, meta=meta)
    log(json.dumps({k: v for k, v in meta.items() if k != "loss_history"} | {"final_loss": hist[-1]}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
