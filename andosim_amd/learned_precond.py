"""Training and evaluation of the learned block preconditioner.

The network, its loss and its kernels are in csrc/learned.hip; this module
drives them through the C ABI (native.Trainer):

  systems = training_systems()            # Newton systems of the batched sheets
  w, losses = train(systems, steps=8000)  # Adam on the GPU
  table = compare(w, eval_systems)        # PCG iterations, plain vs learned

`python -m andosim_amd.learned_precond --out weights.npy` trains with the
defaults and writes the weights (ando_set_precond_weights loads them).
"""
from __future__ import annotations

import argparse

import numpy as np

from . import native, precond_data, scenes


def training_systems(count: int = 12, seed: int = 3):
    """Later substeps (20..100, every 6th) of the batched workload's first
    `count` sheets, plus sheets draped over a sphere (sides 17-41)."""
    data = precond_data.collect(precond_data.training_scenes(count, seed), steps=100, every=6, start=20)
    drapes = [scenes.small_sheet(s, seed=k, pins=0, wall=False, rigid=True) for k, s in enumerate([17, 25, 33, 41])]
    data += precond_data.collect(drapes, steps=60, every=4, start=8)
    return data


def train(systems, steps: int = 8000, lr: float = 3e-4, seed: int = 100, init_seed: int = 5, device: int = 0):
    tr = native.Trainer(device)
    for s in systems:
        tr.add_system(*s)
    tr.init_weights(init_seed)
    losses = tr.train(steps, lr=lr, seed=seed)
    return tr.get_weights(), losses


def compare(weights, systems, tol: float = 1e-6, maxit: int = 1000, device: int = 0):
    """Per system: (plain iterations, converged, learned preconditioner +
    warm start iterations, converged, block-Jacobi from the learned warm start
    iterations, converged); no fallbacks."""
    tr = native.Trainer(device)
    ids = [tr.add_system(*s) for s in systems]
    tr.set_weights(weights)
    out = []
    for i in ids:
        ip, cp = tr.solve(i, 0, tol, maxit)
        il, cl = tr.solve(i, 1, tol, maxit)
        iw, cw = tr.solve(i, 2, tol, maxit)
        out.append((ip, cp, il, cl, iw, cw))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--steps", type=int, default=8000)
    ap.add_argument("--lr", type=float, default=3e-4)
    a = ap.parse_args(argv)
    systems = training_systems()
    w, losses = train(systems, a.steps, a.lr)
    np.save(a.out, w)
    print(f"{len(systems)} systems, loss {np.mean(losses[:200]):.4g} -> {np.mean(losses[-200:]):.4g}", flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
