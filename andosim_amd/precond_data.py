"""Newton systems for training and evaluating the learned preconditioner.

Each system is the first Newton linearisation of a substep, taken from a run
of the plain GPU solver (corrected physics by default): H and b = -g at x with
inertia target x_hat = x + dt v (Integrator::compute_gradient /
assemble_system_matrix, integrator.cpp:246-544).
Training data comes from the batched workload's sheets (side U{17..65}, 1-4
pins, wind, ground wall).
"""
from __future__ import annotations

import numpy as np

from . import native, scenes


def collect(scene_list, steps: int, every: int = 1, start: int = 0, device: int = 0, corrected_physics: bool = True):
    """Returns a list of (row_ptr, cols, H (nb,3,3), b (3n,)) float32/int32
    systems, from runs with the corrected physics unless told otherwise (the
    matrix is the mesh part; self-contact coupling blocks are not in it)."""
    out = []
    for sc in scene_list:
        sim = native.GpuSim(sc, device=device)
        if corrected_physics:
            sim.set_mode(fp64_near=True, corrected_physics=True)
        rp, ci = sim.pattern()
        dt = np.float32(sc.params["dt"])
        for s in range(steps):
            sim.substep()
            x, v = sim.get("x"), sim.get("v")
            if not (np.isfinite(x).all() and np.isfinite(v).all()):
                break
            if s < start or (s - start) % every:
                continue
            g, H = sim.linearize((x + dt * v).astype(np.float32))
            if not (np.isfinite(g).all() and np.isfinite(H).all()) or not np.any(g):
                continue
            out.append((rp.copy(), ci.copy(), H.copy(), (-g).astype(np.float32)))
    return out


def training_scenes(count: int, seed: int = 3):
    """Sheets of the batched workload (its first `count` instances)."""
    return scenes.batch256(seed=seed, count=count)
