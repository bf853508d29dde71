"""Benchmark: AndoSim per-step cloth solve on one MI355X.

One "step" is one add-on substep (operators.py:900-905): State.apply_gravity
followed by Integrator.step, on the 257x257 cloth (66,049 vertices, 131,072
triangles) falling onto two crossed cylinders (r = 0.1 m, 4,800 triangles),
faithful mode, PCG tolerance 1e-6.  The state is resident in HBM for the whole
run; no host transfer happens inside the timed region except the scalars that
steer Newton, line search and PCG.

Prints ONE JSON line (see the driver contract) with a roofline entry for the
dominant kernel (the BSR SpMV of PCG, timed with HIP events on the library's
stream) and a CPU baseline (the oracle, one thread, on a bounded sample of the
same steps).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "sim steps/s + PCG iters/step, 66k-vert cloth: learned-precond vs plain PCG, 1e-6 tol"
HBM_PEAK_GBS = 8000.0


def spmv_bytes(n: int, nb: int) -> int:
    """Algorithmic HBM bytes of one k_pcg_spmv launch: block values (36 B) and
    column (4 B) per block, row pointers, z and p_old read once, p_new and Ap
    written once (12 B per vertex each)."""
    return 40 * nb + 4 * (n + 1) + 4 * 12 * n


def update_bytes(n: int) -> int:
    """k_pcg_update: p, Ap, Pinv (36 B), x, r read; x, r, z written."""
    return 120 * n


def init_bytes(n: int) -> int:
    """PCG set-up: diagonal blocks (36 B) + index read, rhs read; P^-1 (36 B), r, x, z, p written."""
    return 40 * n + 12 * n + 36 * n + 4 * 12 * n


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="cyl257")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-timing", action="store_true", help="run without per-launch kernel events")
    ap.add_argument("--pcg", choices=["persistent", "two-kernel"], default="persistent")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the multi-GPU design is described "
              "in DESIGN.md and is not built.", file=sys.stderr)
        return 2

    from andosim_amd import native, scenes

    sc = scenes.WORKLOADS[a.workload]()
    two = a.pcg == "two-kernel"
    sim = native.GpuSim(sc, fp64_near=True, kernel_timing=False)
    sim.set_mode(fp64_near=True, kernel_timing=False, pcg_two_kernel=two)
    n = len(sc.verts)
    rp, ci = sim.pattern()
    nb = len(ci)

    for _ in range(a.warmup):
        sim.substep()
    x0 = sim.get("x").copy()
    v0 = sim.get("v").copy()
    rig0 = [sim.rigid_state(b) for b in range(sim.nbodies)]
    sim.set_mode(fp64_near=True, kernel_timing=not a.no_timing, pcg_two_kernel=two)

    pcg_iters = newton = timed_iters = 0
    t_launch = [0.0, 0.0]
    n_launch = [0, 0]
    kernel = None
    t0 = time.perf_counter()
    for _ in range(a.steps):
        sim.substep()
        st = sim.stats()
        pcg_iters += st["pcg_iterations"]
        newton += st["newton_iterations"]
        timed_iters += st["timed_iterations"]
        kernel = st["pcg_kernel"]
        for k in range(2):
            t_launch[k] += st["timed_ms"][k]
            n_launch[k] += st["timed_launches"][k]
    elapsed = time.perf_counter() - t0
    steps_per_s = a.steps / elapsed

    roof = None
    if n_launch[0] and kernel == 1:
        # One launch = one whole PCG solve: algorithmic bytes of its iterations + set-up.
        by_total = timed_iters * (spmv_bytes(n, nb) + update_bytes(n)) + n_launch[0] * init_bytes(n)
        secs = t_launch[0] * 1e-3
        ach = by_total / secs / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "k_pcg_persistent",
                "bytes_per_launch": int(by_total / n_launch[0]), "avg_launch_us": round(secs / n_launch[0] * 1e6, 2),
                "launches": n_launch[0], "iterations_per_launch": round(timed_iters / n_launch[0], 2),
                "us_per_iteration": round(secs / max(timed_iters, 1) * 1e6, 2),
                "bytes_per_iteration": spmv_bytes(n, nb) + update_bytes(n)}
    elif n_launch[0]:
        avg_s = t_launch[0] / n_launch[0] * 1e-3
        by = spmv_bytes(n, nb)
        ach = by / avg_s / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "k_pcg_spmv",
                "bytes_per_launch": by, "avg_launch_us": round(avg_s * 1e6, 2), "launches": n_launch[0]}
        if n_launch[1]:
            avg_u = t_launch[1] / n_launch[1] * 1e-3
            roof["update_kernel_gbs"] = round(update_bytes(n) / avg_u / 1e9, 1)
            roof["update_avg_launch_us"] = round(avg_u * 1e6, 2)

    cpu = None
    if not a.no_cpu:
        from oracle.oracle import OracleSim

        o = OracleSim(sc, 32)
        o.set("x", x0.astype(np.float64))
        o.set("v", v0.astype(np.float64))
        for b, rs in enumerate(rig0):
            o.set_rigid_pose(b, rs)
        k = 0
        c0 = time.perf_counter()
        while k < a.steps:
            o.substep()
            k += 1
            if time.perf_counter() - c0 > a.cpu_seconds:
                break
        cs = time.perf_counter() - c0
        cpu = {"value": round(k / cs, 4), "unit": "steps/s", "cores": 1, "kind": "port",
               "sample": f"{k} substeps of {a.workload} starting at step {a.warmup} (oracle, FP32, 1 thread, "
                         f"{cs:.1f} s)"}

    out = {
        "metric": METRIC, "value": round(steps_per_s, 3), "unit": "steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(1e3 * elapsed / a.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{a.workload}: 257x257 cloth (66,049 v, 131,072 tris) onto two crossed cylinders "
                               "r=0.1 m (4,800 rigid tris); dt=0.01, PCG tol 1e-6, faithful mode",
                   "vertices": n, "blocks": nb, "precond": "block-Jacobi (plain PCG)",
                   "pcg_kernel": {1: "persistent", 0: "two-kernel"}.get(kernel, "none")},
        "pcg_iters_per_step": round(pcg_iters / a.steps, 2),
        "newton_iters_per_step": round(newton / a.steps, 2),
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
