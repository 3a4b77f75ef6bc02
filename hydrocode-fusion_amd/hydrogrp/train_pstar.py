"""Train the learned p* guess on the GPU and write hydrogrp/pstar_mlp.npz.

    python -m hydrogrp.train_pstar [--samples 10000000] [--epochs 12]

Data: fresh samples of the training distribution (hgrp_mlp_data), labels from
the exact solver at Newton tolerance 1e-12; the forward/backward/Adam kernels
are the library's own (hgrp_mlp_train).  Prints the loss per pass and the
Newton iteration counts with and without the guess on held-out samples.
"""
import argparse
import json
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from hydrogrp import mlp  # noqa: E402


def evaluate(params, seed=777, count=1 << 16, tol=1e-12):
    _, _, states, ok = mlp.data(seed, count, tol)
    st = states[ok]
    plain = mlp.riemann_guess_batch(st, tol=tol)
    learned = mlp.riemann_guess_batch(st, tol=tol, params=params)
    nl = learned[:, 4]
    return {"held_out": int(len(st)), "iters_plain": float(np.mean(plain[:, 4])),
            "iters_learned": float(np.mean(nl)), "fallbacks": int(np.sum(nl > 3)),
            "max_rel_dp": float(np.max(np.abs(learned[:, 1] - plain[:, 1]) / plain[:, 1]))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=10_000_000)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--out", default=mlp.WEIGHTS)
    a = ap.parse_args()
    p0 = mlp.init_params(a.seed)
    params, losses = mlp.train(p0, seed=a.seed, n_samples=a.samples, epochs=a.epochs, batch=a.batch, lr=a.lr)
    for e, l in enumerate(losses):
        print(f"pass {e}: mse {l:.3e}", flush=True)
    ev = evaluate(params)
    print(json.dumps(ev), flush=True)
    np.savez(a.out, params=params, losses=losses, samples=a.samples, epochs=a.epochs, batch=a.batch, lr=a.lr,
             seed=a.seed, eval=json.dumps(ev))
    print("wrote", a.out, flush=True)


if __name__ == "__main__":
    main()
