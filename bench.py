"""Benchmark: AndoSim per-step cloth solve on one MI355X.

Headline workload: cyl257, a 257x257 cloth (66,049 vertices, 131,072
triangles) falling onto two crossed cylinders (r = 0.1 m, 4,800 triangles);
material and parameters of the 21x21 drop (E=3e5, nu=0.3, t=1 mm, rho=200,
dt=0.01, gap 1 mm, 5% strain limit, PCG tol 1e-6), run with the corrected
physics (ANDO_MODE_CORRECTED_PHYSICS: membrane force from the current
positions, self and rigid contact candidates from swept boxes, FP64 CCD line
search).  One "step" is one add-on substep (operators.py:900-905):
State.apply_gravity followed by Integrator.step, with the state resident in
HBM.  --warmup substeps from the initial state run untimed (first contact is
at substep 9), then --steps contiguous substeps are timed; every timed state
is checked finite and the intersection validator runs on the last one.

The steps/s value comes from a pass without per-launch events, with the
two-level Schwarz preconditioner (ANDO_MODE_TWO_LEVEL; --precond block-jacobi
times the reference's block-Jacobi PCG instead).  A second pass from the same
start state records HIP events around every PCG launch on the library's
stream; the roofline entry for the dominant kernel comes from it.  The same
window is then stepped again with the other preconditioner and with the
learned preconditioner, for PCG iterations per step side by side.  The CPU baseline is the oracle (the same corrected physics,
FP32) on the substep half-way through the timed window, from the GPU's state
there, its element-wise loops on all host threads OpenMP is given; the FP64
oracle on that substep gives the accuracy figures.  The faithful mode (the
reference's step as written) is reported beside it on the window where it
stays finite.
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

_T0 = time.perf_counter()


def progress(msg: str) -> None:
    """One line on stderr per finished stage (and every few substeps of a long
    run), with the seconds since start, so a long run is never silent."""
    print(f"[bench {time.perf_counter() - _T0:7.1f} s] {msg}", file=sys.stderr, flush=True)


def leg(name: str, fn, *args, **kw):
    """Runs one of the legs reported beside the headline; a solver error in it
    (RuntimeError from the library) is recorded as {"error": ...} instead of
    ending the run, so the headline line is always printed."""
    try:
        return fn(*args, **kw)
    except RuntimeError as e:
        progress(f"{name}: error: {e}")
        return {"error": str(e)}


METRIC = "sim steps/s + PCG iters/step, 66k-vert cloth: learned-precond vs plain PCG, 1e-6 tol"
HBM_PEAK_GBS = 8000.0


def spmv_bytes(n: int, nb: int) -> int:
    """Algorithmic HBM bytes of one SpMV pass: block values (36 B) and column
    (4 B) per block, row pointers, z and p_old read once, p_new and Ap written
    once (12 B per vertex each)."""
    return 40 * nb + 4 * (n + 1) + 4 * 12 * n


def update_bytes(n: int) -> int:
    """Vector update: p, Ap, P^-1 (36 B), x, r read; x, r, z written."""
    return 120 * n


def tl_update_bytes(n: int, nb: int, npatch: int) -> int:
    """Two-level update (k_tl_update): the patch rows' blocks and columns
    (40 B per block), p, Ap, x, r and the block-Jacobi block read, x, r, z
    written (120 B per row), the coarse inverse (4 B per entry, 3 npatch
    squared entries) and the patch sums of r and Ap (16 B per coarse unknown)."""
    nc = 3 * npatch
    return 40 * nb + 120 * n + 4 * nc * nc + 16 * nc


def measured_traffic(workload: str, n: int, nb: int, kernel_prefix: str):
    """HBM bytes per launch of the kernel from the committed rocprofv3 PMC
    summary (FETCH_SIZE and WRITE_SIZE, each scaled by the copy-kernel
    calibration), or None when no summary matches this workload."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}.json")
    try:
        with open(path) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    if d.get("vertices") != n or d.get("blocks") != nb:
        return None
    for name, rec in d.get("kernels", {}).items():
        if name == kernel_prefix or name.endswith(" " + kernel_prefix):
            return int(round(rec["traffic_bytes"]))
    return None


def corrected_contact_leg(steps: int):
    """Dense self contact: 8 sheets of 65x65 stacked 0.08 mm apart (jittered
    +-0.03 mm, so they start interpenetrating in places) just above a ground
    wall, run for `steps` substeps in faithful mode (self broad phase empty, as
    the reference), in corrected mode (LBVH self PT/EE contacts and barriers,
    the reference's sampled CCD in the line search), in corrected mode with
    additive CCD and with the corrected physics; steps/s, contacts per step,
    PCG iterations and failures and the intersection validator at the end of
    each run."""
    from andosim_amd import native, scenes

    out = {"workload": "stack_resting: 8 sheets 65x65 (33,800 v, 65,536 tris, 98,304 edges) 0.08 mm apart, "
                       "jitter 0.03 mm, ground wall", "steps": steps}
    for mode in ("faithful", "corrected", "corrected_additive_ccd", "corrected_physics"):
        progress(f"stack_resting leg: {mode}")
        sc = scenes.stack_resting()
        sim = native.GpuSim(sc)
        sim.set_mode(fp64_near=True, corrected=(mode != "faithful"), additive_ccd=(mode == "corrected_additive_ccd"),
                     corrected_physics=(mode == "corrected_physics"))
        t = 0.0
        contacts, pcg = [], []
        ls_fail = pcg_fail = 0
        done = 0
        for _ in range(steps):
            t0 = time.perf_counter()
            sim.substep()
            t += time.perf_counter() - t0
            st = sim.stats()
            contacts.append(st["contacts"])
            pcg.append(st["pcg_iterations"])
            ls_fail += st["line_search_failures"]
            pcg_fail += st["pcg_failures"]
            done += 1
            if not finite_state(sim):
                break
        et, ins = sim.count_intersections()
        out[mode] = {"steps_per_s": round(done / t, 2), "steps_run": done, "finite": finite_state(sim),
                     "contacts_per_step": contacts, "pcg_iters_per_step": round(float(np.mean(pcg)), 1),
                     "line_search_failures": ls_fail, "pcg_failures": pcg_fail,
                     "edge_triangle_intersections_at_end": et}
    return out


def per_frame_leg(sim, sc, reps: int = 20, with_cpu: bool = True):
    """The add-on's per-frame consumers (operators.py:872-880, 909, 947,
    967-975) at the state the timed window ends in: host-clock ms per call of
    each device entry point (the call includes its small readback), and the
    oracle's CPU restatement of the same reductions for comparison."""
    import ctypes as C
    from andosim_amd import native

    L, h = sim.L, sim.h

    def avg_ms(fn):
        fn()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        return round(1e3 * (time.perf_counter() - t0) / reps, 4)

    cnt = C.c_int32(0)
    x = np.zeros((sim.n, 3), np.float32)
    v = np.zeros((sim.n, 3), np.float32)
    out = {"energy_ms": avg_ms(sim.energy),
           "detect_contacts_ms": avg_ms(lambda: native.check(L.ando_compute_contacts(
               h, sim.bodies, sim.nbodies, None, 0, C.byref(cnt)))),
           "cfl_inputs_ms": avg_ms(sim.cfl_inputs),
           "get_state_ms": avg_ms(lambda: native.check(L.ando_get_state(h, native._fp(x), native._fp(v), sim.n))),
           "get_positions_ms": avg_ms(lambda: native.check(L.ando_get_state(h, native._fp(x), None, sim.n)))}
    # The module's get_positions: positions only, by DMA into a pinned buffer.
    buf = C.c_void_p()
    native.check(L.ando_host_alloc(12 * sim.n, C.byref(buf)))
    pinned = C.cast(buf, C.POINTER(C.c_float))
    out["get_positions_pinned_ms"] = avg_ms(lambda: native.check(L.ando_get_state(h, pinned, None, sim.n)))
    native.check(L.ando_host_free(buf))
    cs, count = sim.contacts(cap=1 << 20)
    out["contacts"] = count
    out["metrics_ms"] = avg_ms(lambda: sim.contact_metrics(float(sc.params["contact_gap_max"]), True))
    out["list_contacts_ms"] = avg_ms(lambda: sim.contacts(cap=max(count, 1)))
    if with_cpu:
        from oracle import oracle as orc
        o = orc.OracleSim(sc, 32)
        xs, vs = sim.get("x"), sim.get("v")
        t0 = time.perf_counter()
        for _ in range(reps):
            o.energy(xs, vs, sim.masses)
        out["cpu_oracle_energy_ms"] = round(1e3 * (time.perf_counter() - t0) / reps, 4)
        arrs = ([c.type for c in cs], [list(c.idx) for c in cs], [c.gap for c in cs], [list(c.normal) for c in cs])
        t0 = time.perf_counter()
        for _ in range(reps):
            orc.contact_metrics(*arrs, vs, float(sc.params["contact_gap_max"]), True, 32)
        out["cpu_oracle_metrics_ms"] = round(1e3 * (time.perf_counter() - t0) / reps, 4)
    return out


def batched_leg(steps: int, count: int = 256, accuracy: bool = False):
    """The batched workload (`count` cloths of side U{17..65}, 1-4 pins, wind,
    ground wall; about 481k vertices for 256), corrected physics: cloth-steps/s stepping the
    contexts one after another (the reference's only way) and through
    ando_step_batch (64 host threads, a stream per cloth, the solves of
    threads that reach PCG together run in batched launches); and the
    faithful step batched the same way, for comparison."""
    from andosim_amd import native, scenes

    out = {"workload": f"batch{count}: {count} independent cloths, sides U{{17..65}}", "substeps": steps}
    scs = scenes.batch256(count=count)
    big = int(np.argmax([len(sc.verts) for sc in scs]))
    for mode in ("sequential", "batched", "batched_faithful"):
        progress(f"batched leg: {mode}")
        sims = [native.GpuSim(sc) for sc in scs]
        for g in sims:
            g.set_mode(fp64_near=True, corrected_physics=(mode != "batched_faithful"))
        for g in sims:  # warm-up substep (constants, allocations)
            g.substep()
        t = 0.0
        before = None
        for k in range(steps):
            if mode == "batched" and k == steps - 1:
                before = snapshot(sims[big])
            t0 = time.perf_counter()
            if mode == "sequential":
                for g in sims:
                    g.substep()
            else:
                for g in sims:
                    g.apply_gravity()
                native.step_batch(sims, threads=64)
            t += time.perf_counter() - t0
        out[mode] = {"cloth_steps_per_s": round(count * steps / t, 1),
                     "vertices": int(sum(g.n for g in sims)),
                     "pcg_failures_last_substep": int(sum(g.stats()["pcg_failures"] for g in sims)),
                     "finite": all(finite_state(g) for g in sims)}
        if mode == "batched_faithful":
            out[mode]["note"] = ("the reference's step as written, for comparison with earlier measurements: the "
                                 "explicit gravity pre-step lets the cloths pass the ground wall")
        if mode != "sequential":
            out[mode]["host_threads"] = 64
        if mode == "batched":
            if accuracy and before is not None:
                out[mode]["accuracy_vs_fp64"] = accuracy_vs_fp64(scs[big], before, sims[big].get("x").astype(np.float64))
                out[mode]["accuracy_vs_fp64"]["cloth"] = f"the largest ({len(scs[big].verts)} vertices), last substep"
            out[mode]["batched_solve_share"] = round(
                sum(g.stats()["batched_solves"] for g in sims) / max(1, sum(g.stats()["pcg_solves"] for g in sims)), 3)
    return out


def oracle_substep(sc, state, bits: int, physics: bool):
    """One oracle substep (FP32 or FP64) from a saved (x, v, rigid poses) state;
    returns the positions and the wall-clock seconds."""
    from oracle import oracle as orc
    o = orc.OracleSim(sc, bits)
    o.set_physics(physics)
    x0, v0, rig = state
    o.set("x", x0.astype(np.float64))
    o.set("v", v0.astype(np.float64))
    for b, rs in enumerate(rig):
        o.set_rigid_pose(b, rs)
    t0 = time.perf_counter()
    o.substep()
    return o.get("x").astype(np.float64), time.perf_counter() - t0, o.stats()


def host_cpu_model() -> str:
    """The host CPU's model name (from /proc/cpuinfo), or ''."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return ""


def rel_err(a, b) -> float:
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), 1e-30))


def accuracy_vs_fp64(sc, state, x_gpu, x_fp32=None, physics: bool = True, extra=None, envelope: bool = False):
    """max |x - x64| / max |x64| after one substep from `state` for the GPU and
    for the FP32 oracle, against the FP64 oracle (host threads); `extra` maps
    names to further GPU results from the same state.  With `envelope`, also
    the FP32 oracle from the state with one coordinate moved by one ulp: how
    far rounding alone moves an FP32 substep (contact-heavy substeps amplify
    it), the scale against which the GPU's distance is read."""
    x64, t64, _ = oracle_substep(sc, state, 64, physics)
    if x_fp32 is None:
        x_fp32, _, _ = oracle_substep(sc, state, 32, physics)
    out = {"reference": "oracle in FP64 on the host, same start state, one substep",
           "gpu_max_rel_err": float(f"{rel_err(x_gpu, x64):.3e}"),
           "fp32_oracle_max_rel_err": float(f"{rel_err(x_fp32, x64):.3e}"),
           "fp64_oracle_seconds": round(t64, 1)}
    for name, x in (extra or {}).items():
        out[name + "_max_rel_err"] = float(f"{rel_err(x, x64):.3e}")
    if envelope:
        x0, v0, rig = state
        xp = x0.copy()
        i = int(np.argmax(np.abs(x0[:, 2])))
        xp[i, 2] = np.nextafter(xp[i, 2], np.float32(np.inf), dtype=np.float32)
        x32p, _, _ = oracle_substep(sc, (xp, v0, rig), 32, physics)
        out["fp32_oracle_one_ulp_max_rel_err"] = float(f"{rel_err(x32p, x64):.3e}")
        out["fp32_oracle_one_ulp_vs_fp32_oracle"] = float(f"{rel_err(x32p, x_fp32):.3e}")
        out["gpu_vs_fp32_oracle"] = float(f"{rel_err(x_gpu, x_fp32):.3e}")
        out["one_ulp"] = f"z of vertex {i} moved up by one FP32 ulp"
    return out


def snapshot(sim):
    return (sim.get("x").copy(), sim.get("v").copy(), [sim.rigid_state(b) for b in range(sim.nbodies)])


def restore_state(sim, state):
    x0, v0, rig = state
    sim.set_state(x0, v0)
    for b, rs in enumerate(rig):
        st = sim.bodies[b]
        for k in range(3):
            st.position[k] = float(rs["position"][k])
            st.linear_velocity[k] = float(rs["linear_velocity"][k])
            st.angular_velocity[k] = float(rs["angular_velocity"][k])
        for k in range(9):
            st.rotation[k] = float(rs["rotation"].ravel()[k])


def run_steps(sim, count: int, check_every: int = 0, snap_at=None, snaps=None):
    """`count` substeps; per-step (seconds, stats); stops at a non-finite
    state.  With snap_at, the state before substep snap_at (counted from 0)
    is appended to snaps, outside the timed calls.  Returns (records, finite)."""
    recs = []
    for k in range(count):
        if snap_at is not None and k == snap_at:
            snaps.append(snapshot(sim))
        t0 = time.perf_counter()
        sim.substep()
        recs.append((time.perf_counter() - t0, sim.stats()))
        if k % 25 == 24:
            progress(f"  substep {k + 1} of {count}")
        if (check_every and k % check_every == check_every - 1) or k == count - 1:
            if not finite_state(sim):
                return recs, False
    return recs, True


def summarize(recs):
    el = sum(r[0] for r in recs)
    n = max(1, len(recs))
    st = [r[1] for r in recs]
    return {"steps_per_s": round(len(recs) / el, 3) if el > 0 else None,
            "ms_per_step": round(1e3 * el / n, 2),
            "pcg_iters_per_step": round(sum(t["pcg_iterations"] for t in st) / n, 1),
            "pcg_solves": sum(t["pcg_solves"] for t in st),
            "pcg_failures": sum(t["pcg_failures"] for t in st),
            "line_search_failures": sum(t["line_search_failures"] for t in st),
            "newton_iters_per_step": round(sum(t["newton_iterations"] for t in st) / n, 2),
            "max_candidates": max((t["max_candidates"] for t in st), default=0),
            "capped_trials": sum(t["capped_trials"] for t in st)}


def physics_leg(name: str, steps: int, accuracy: bool, two_level: bool = False):
    """A workload run from its initial state with the corrected physics for
    `steps` substeps: throughput, solver counters, finiteness and the
    intersection validator at the end; optionally the FP64-oracle error of
    the GPU's substep from the state half-way through."""
    from andosim_amd import native, scenes
    progress(f"{name} leg: {steps} substeps, corrected physics")
    sc = scenes.WORKLOADS[name]()
    sim = native.GpuSim(sc, fp64_near=True)
    sim.set_mode(fp64_near=True, corrected_physics=True, two_level=two_level)
    half = steps // 2
    recs, ok = run_steps(sim, half, 10)
    state = snapshot(sim)
    x_mid = None
    if ok:
        sim.substep()
        x_mid = sim.get("x").astype(np.float64)
        restore_state(sim, state)
        more, ok = run_steps(sim, steps - half, 10)
        recs += more
    out = {"workload": name, "vertices": int(len(sc.verts)), "steps_run": len(recs), "finite": ok,
           "precond": "two-level" if two_level else "block-Jacobi", **summarize(recs)}
    et, ins = sim.count_intersections()
    out["intersections_at_end"] = {"edge_triangle": et, "vertices_inside_bodies": ins}
    if accuracy and x_mid is not None:
        out["accuracy_vs_fp64"] = accuracy_vs_fp64(sc, state, x_mid)
        out["accuracy_vs_fp64"]["at_substep"] = half
    return out


def faithful_leg(name: str, max_explore: int, steps: int):
    """The reference's step as written (faithful mode) on the same workload:
    it diverges a few substeps after first contact (the strain barrier's
    m/gap^2 stiffness stalls PCG at its cap; the FP32 oracle does the same), so
    it is timed on the substeps from first contact to the last finite one."""
    from andosim_amd import native, scenes
    progress(f"{name} leg: faithful mode")
    sc = scenes.WORKLOADS[name]()
    sim = native.GpuSim(sc, fp64_near=True)
    first, saved, last_ok = None, None, -1
    for k in range(max_explore):
        snap = snapshot(sim) if first is None else None
        sim.substep()
        if not finite_state(sim):
            break
        last_ok = k
        if first is None and sim.stats()["contacts"] > 0:
            first, saved = k, snap
    if first is None:
        return {"workload": name, "error": "no contact within the explored substeps"}
    window = last_ok - first + 1
    recs = []
    while len(recs) < steps:
        restore_state(sim, saved)
        more, ok = run_steps(sim, min(window, steps - len(recs)))
        recs += more
        if not ok:
            break
    return {"workload": name, "mode": "faithful", "window": [first, last_ok],
            "first_nonfinite_step": last_ok + 1 if last_ok + 1 < max_explore else None,
            "steps_timed": len(recs), **summarize(recs)}


def learned_comparison(sim, sc, window, restore, replay, a):
    """Learned preconditioner vs plain block-Jacobi PCG on the same window.

    The network is trained here, on the GPU, on Newton systems of the batched
    workload's sheets (never on this workload).  Two comparisons:
      * the first Newton system of every window step (taken from the plain
        run), solved by plain PCG and by the learned preconditioner with its
        warm start (no fallback);
      * the window replayed from its start state with plain block-Jacobi PCG
        and with ANDO_MODE_LEARNED_PRECOND (the learned preconditioner for
        half the iteration cap, then block-Jacobi from its iterate; or the
        warm start only): PCG iterations, Newton iterations, failures per
        step, steps/s and the validator at the end.
    """
    from andosim_amd import learned_precond, native

    progress("learned preconditioner: training")
    t0 = time.perf_counter()
    systems = learned_precond.training_systems()
    w, losses = learned_precond.train(systems, steps=a.train_steps)
    t_train = time.perf_counter() - t0
    progress(f"learned preconditioner: trained in {t_train:.1f} s")
    # First Newton systems of the window steps, from the plain run.
    restore()
    dt = np.float32(sc.params["dt"])
    rp, ci = sim.pattern()
    firsts = []
    for _ in window:
        x, v = sim.get("x"), sim.get("v")
        g, H = sim.linearize((x + dt * v).astype(np.float32))
        firsts.append((rp, ci, H, (-g).astype(np.float32)))
        sim.substep()
    table = learned_precond.compare(w, firsts, tol=float(sc.params["pcg_tol"]),
                                    maxit=int(sc.params["pcg_max_iters"]))
    # Whole steps in each learned mode.
    sim.set_precond_weights(w)

    def steps(mode):
        """The window replayed from its start state with the given solver."""
        progress(f"learned comparison: {mode} solves")
        if mode == "plain":
            sim.set_mode(fp64_near=True, kernel_timing=False, corrected_physics=True)
        else:
            sim.set_mode(fp64_near=True, kernel_timing=False, learned=True, warm_start_only=(mode == "warm"),
                         corrected_physics=True)
        restore()
        recs, nonfinite = [], None
        for k in window:
            t1 = time.perf_counter()
            sim.substep()
            recs.append((time.perf_counter() - t1, sim.stats()))
            if not finite_state(sim):
                nonfinite = k
                break
        sim.set_mode(fp64_near=True, kernel_timing=False, learned=False, corrected_physics=True)
        el = sum(r[0] for r in recs)
        st = [r[1] for r in recs]
        return {"pcg_iters_per_step": round(sum(t["pcg_iterations"] for t in st) / len(recs), 2),
                "steps_per_s": round(len(recs) / el, 3), "nonfinite_at_step": nonfinite,
                "newton_iters_per_step": round(sum(t["newton_iterations"] for t in st) / len(recs), 2),
                "pcg_failures": sum(t["pcg_failures"] for t in st),
                "line_search_failures": sum(t["line_search_failures"] for t in st),
                "solves": sum(t["pcg_solves"] for t in st),
                "learned_solves": sum(t["learned_solves"] for t in st),
                "continued_in_block_jacobi": sum(t["learned_fallbacks"] for t in st),
                "block_jacobi_iters_per_step": round(sum(t["fallback_iterations"] for t in st) / len(recs), 2),
                "intersections_at_end": list(sim.count_intersections())}

    return {
        "window_substeps": [window[0], window[-1]],
        "plain": steps("plain"),
        "precond_and_warm_start": steps("learned"),
        "warm_start_only": steps("warm"),
        "first_newton_system": {"plain_iters": [t[0] for t in table], "plain_converged": [t[1] for t in table],
                                "learned_iters": [t[2] for t in table], "learned_converged": [t[3] for t in table],
                                "warm_start_only_iters": [t[4] for t in table],
                                "warm_start_only_converged": [t[5] for t in table]},
        "training": {"systems": len(systems), "steps": a.train_steps, "seconds": round(t_train, 1),
                     "loss_first": round(float(np.mean(losses[:200])), 4),
                     "loss_last": round(float(np.mean(losses[-200:])), 4),
                     "data": "batched workload sheets (side U{17..65}, 1-4 pins, wind), substeps 20..100, "
                             "and 4 sheets draped over a sphere, corrected physics"},
    }


def finite_state(sim) -> bool:
    x = sim.get("x")
    v = sim.get("v")
    return bool(np.isfinite(x).all() and np.isfinite(v).all() and np.abs(x).max() < 1e3)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="timed contiguous substeps")
    ap.add_argument("--warmup", type=int, default=10, help="untimed substeps from the initial state")
    ap.add_argument("--workload", default="cyl257")
    ap.add_argument("--event-steps", type=int, default=30, help="substeps of the kernel-event pass")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the FP64 accuracy")
    ap.add_argument("--no-timing", action="store_true", help="skip the kernel-event pass")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-preconditioner comparison")
    ap.add_argument("--learned-steps", type=int, default=30)
    ap.add_argument("--train-steps", type=int, default=8000)
    ap.add_argument("--no-legs", action="store_true", help="skip the other workloads")
    ap.add_argument("--no-leg-accuracy", action="store_true",
                    help="skip the FP64-oracle error of a mid-run substep of every workload leg (about 20 s of host time)")
    ap.add_argument("--corrected-steps", type=int, default=20)
    ap.add_argument("--batched-steps", type=int, default=5)
    ap.add_argument("--precond", choices=("two-level", "block-jacobi"), default="two-level",
                    help="preconditioner of the timed run (the other one is run on the same window beside it)")
    ap.add_argument("--autotune", action="store_true",
                    help="every context picks its PCG launch shape by timing (ANDO_PCG_AUTOTUNE=1)")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); sharding the solve over several GPUs "
              "is not built.", file=sys.stderr)
        return 2
    if a.autotune:
        os.environ.setdefault("ANDO_PCG_AUTOTUNE", "1")
    from andosim_amd import native, scenes

    sc = scenes.WORKLOADS[a.workload]()
    sim = native.GpuSim(sc, fp64_near=True)
    tl = a.precond == "two-level"

    def mode(two_level, timing=False):
        sim.set_mode(fp64_near=True, corrected_physics=True, two_level=two_level, kernel_timing=timing)

    mode(False)
    n = len(sc.verts)
    rp, ci = sim.pattern()
    nb = len(ci)

    progress(f"{a.workload}: {n} vertices, {nb} blocks; {a.warmup} warm-up substeps")
    warm, ok = run_steps(sim, a.warmup, 1)
    if not ok:
        print(json.dumps({"metric": METRIC, "error": "non-finite state during warm-up"}), flush=True)
        return 1
    saved = snapshot(sim)
    mid_snap = []
    mid = a.steps // 2
    mode(tl)
    progress(f"timed run: {a.steps} substeps ({a.precond})")
    recs, ok = run_steps(sim, a.steps, 1, snap_at=mid, snaps=mid_snap)
    if not ok:
        print(json.dumps({"metric": METRIC, "error": f"non-finite state at timed substep {len(recs)}"}), flush=True)
        return 1
    head = summarize(recs)
    progress(f"timed run: {head['steps_per_s']} steps/s, {head['pcg_iters_per_step']} PCG iterations per substep")
    et, ins = sim.count_intersections()
    last = recs[-1][1]
    kernel, lanes = last["pcg_kernel"], last.get("spmv_lanes", 0)
    wg_slices, vec_block = last.get("spmv_wg_slices", 4), last.get("pcg_vec_block", 256)
    tuned = last.get("pcg_autotuned", 0) == 1

    # Event pass: from the same start state, every PCG launch bracketed by events.
    roof = None
    if not a.no_timing:
        progress("event pass: PCG launches bracketed by HIP events")
        restore_state(sim, saved)
        mode(tl, timing=True)
        try:
            ev_recs, _ = run_steps(sim, min(a.event_steps, a.steps), 1)
            ev_err = None
        except RuntimeError as e:
            progress(f"event pass: error: {e}")
            ev_recs, ev_err = [], str(e)
        mode(tl)
        # The pass replays the timed substeps from the same state: the same
        # iteration counts show that the timed kernels are the ones measured.
        replay = [r[1]["pcg_iterations"] for r in ev_recs] == [r[1]["pcg_iterations"] for r in recs[:len(ev_recs)]]
        if ev_err:
            roof = {"error": ev_err}
            ev_recs = []
        t_launch = [sum(r[1]["timed_ms"][k] for r in ev_recs) for k in range(2)]
        n_launch = [sum(r[1]["timed_launches"][k] for r in ev_recs) for k in range(2)]
        cpl_bytes = sum(r[1]["timed_coupling_bytes"] for r in ev_recs)
        ev_elapsed = sum(r[0] for r in ev_recs)
        if ev_recs and n_launch[0] and n_launch[1]:
            avg_s = t_launch[0] / n_launch[0] * 1e-3
            avg_u = t_launch[1] / n_launch[1] * 1e-3
            # Self-contact coupling blocks multiplied in the SpMV launches.
            cby = cpl_bytes / n_launch[0]
            sb = spmv_bytes(n, nb) + cby
            if tl:
                npatch = last.get("tl_patches", 0) or -(-n // vec_block)
                ub = tl_update_bytes(n, nb, npatch)
                kname, by, avg = "ando::tl::k_tl_update", ub, avg_u
                if avg_s > avg_u:
                    kname, by, avg = "ando::tl::k_tl_spmv", sb, avg_s
            else:
                ub = update_bytes(n) + (12 * n if cpl_bytes > 0 else 0)  # + Ap2 with coupling rows
                kname = {3: "ando::k_pcg_spmv_sell"}.get(kernel, "ando::k_pcg_spmv<3>")
                if kernel == 3 and lanes in (2, 4):
                    kname = f"ando::k_pcg_spmv_sell_split<{lanes}, {wg_slices}>"
                by, avg = sb, avg_s
            ach = by / avg / 1e9
            roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(ach / HBM_PEAK_GBS, 4),
                    "traffic": measured_traffic(f"{a.workload}_{'two_level' if tl else 'corrected_physics'}", n, nb,
                                                kname),
                    "kernel": kname, "bytes_per_launch": int(round(by)), "avg_launch_us": round(avg * 1e6, 2),
                    "launches": n_launch[0],
                    "spmv_gbs": round(sb / avg_s / 1e9, 1), "spmv_avg_launch_us": round(avg_s * 1e6, 2),
                    "spmv_bytes_per_launch": int(round(sb)), "coupling_bytes_per_launch": int(round(cby)),
                    "update_gbs": round(ub / avg_u / 1e9, 1), "update_avg_launch_us": round(avg_u * 1e6, 2),
                    "update_bytes_per_launch": int(ub),
                    "pcg_kernels_share_of_step": round((t_launch[0] + t_launch[1]) * 1e-3 / ev_elapsed, 3),
                    "event_pass_ms_per_step": round(1e3 * ev_elapsed / len(ev_recs), 3),
                    "event_pass_replays_timed_iterations": replay}

    progress("per-frame consumers")
    frame = leg("per-frame consumers", per_frame_leg, sim, sc, with_cpu=not a.no_cpu)

    # The same window from the same start state with the other preconditioner.
    progress(f"same window with {'block-Jacobi' if tl else 'the two-level preconditioner'}")
    restore_state(sim, saved)
    mode(not tl)
    try:
        o_recs, o_ok = run_steps(sim, a.steps, 1)
        o_err = None
    except RuntimeError as e:
        progress(f"other preconditioner: error: {e}")
        o_recs, o_ok, o_err = [], False, str(e)
    o_et, o_ins = sim.count_intersections()
    mode(False)
    tl_recs = recs if tl else o_recs
    tl_desc = ("two-level: Morton patches of %d rows, Chebyshev degree %s, coarse 3 unknowns per patch, coarse "
               "inverse rebuilt every %s solves" % (tl_recs[-1][1].get("pcg_vec_block", 0) if tl_recs else 0,
                                                    os.environ.get("ANDO_TL_CHEB", "9"),
                                                    os.environ.get("ANDO_TL_COARSE_EVERY", "2")))
    timed = {**summarize(recs), "intersections_at_end": [et, ins]}
    other = ({"error": o_err} if o_err else
             {**summarize(o_recs), "finite": o_ok, "intersections_at_end": [o_et, o_ins]})
    preconditioners = {
        "window_substeps": [a.warmup, a.warmup + len(recs) - 1],
        "block_jacobi": other if tl else timed,
        "two_level_schwarz": {**(timed if tl else other), "pcg_kernel": tl_desc},
    }
    if not o_err:
        preconditioners["two_level_iteration_ratio"] = round(
            preconditioners["block_jacobi"]["pcg_iters_per_step"] /
            max(preconditioners["two_level_schwarz"]["pcg_iters_per_step"], 1e-9), 2)
        preconditioners["two_level_speedup"] = round(
            preconditioners["two_level_schwarz"]["steps_per_s"] /
            max(preconditioners["block_jacobi"]["steps_per_s"], 1e-9), 3)

    learned = None
    if not a.no_learned:
        window = list(range(a.warmup, a.warmup + min(a.learned_steps, a.steps)))
        learned = leg("learned preconditioner", learned_comparison, sim, sc, window,
                      lambda: restore_state(sim, saved), None, a)

    legs = None
    if not a.no_legs:
        acc = not (a.no_leg_accuracy or a.no_cpu)
        legs = {"sphere129_corrected_physics": leg("sphere129", physics_leg, "sphere129", 200, acc, tl),
                "stack65x8_corrected_physics": leg("stack65x8", physics_leg, "stack65x8", 300, acc, tl),
                "cyl257_faithful": leg("faithful", faithful_leg, a.workload, 60, 20),
                "batched": leg("batched", batched_leg, a.batched_steps, accuracy=acc),
                "stack_resting_modes": leg("resting modes", corrected_contact_leg, a.corrected_steps)}

    cpu = accuracy = None
    if not a.no_cpu:
        from oracle import oracle as orc
        progress("CPU baseline: one oracle substep (FP32), then the FP64 oracle for the accuracy")
        threads = orc.get_threads()
        orc.set_threads(threads)
        # One substep from the GPU's state half-way through the timed window
        # (the cloth is on the cylinders there), as a bounded sample of the
        # timed work.
        state = mid_snap[0] if mid_snap else saved
        at = a.warmup + (mid if mid_snap else 0)
        x32, cs, st32 = oracle_substep(sc, state, 32, True)
        gpu_mid = summarize(recs[mid:mid + 1]) if mid_snap else head
        cpu = {"value": round(1.0 / cs, 5), "unit": "steps/s", "cores": threads, "kind": "port",
               "host_cpu": host_cpu_model(),
               "sample": f"substep {at} of {a.workload} (half-way through the timed window) from the GPU's state, "
                         f"oracle FP32 with the corrected physics, OpenMP loops on {threads} threads (results "
                         f"identical to 1 thread), {cs:.1f} s, {int(st32['pcg_iters'])} PCG iterations (the GPU: "
                         f"{gpu_mid['ms_per_step']} ms, {gpu_mid['pcg_iters_per_step']} PCG iterations on that "
                         f"substep)"}
        def gpu_substep_accuracy():
            restore_state(sim, state)
            mode(tl)
            sim.substep()
            x_gpu = sim.get("x").astype(np.float64)
            # The same substep with the other preconditioner on the GPU.
            restore_state(sim, state)
            mode(not tl)
            sim.substep()
            x_other = sim.get("x").astype(np.float64)
            mode(False)
            acc = accuracy_vs_fp64(sc, state, x_gpu, x32,
                                   extra={"gpu_block_jacobi" if tl else "gpu_two_level": x_other}, envelope=True)
            acc["at_substep"] = at
            return acc

        accuracy = leg("accuracy", gpu_substep_accuracy)

    out = {
        "metric": METRIC, "value": head["steps_per_s"], "unit": "steps/s", "n_gpus": 1, "steps": len(recs),
        "warmup": a.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{a.workload}: 257x257 cloth (66,049 v, 131,072 tris) onto two crossed cylinders "
                               "r=0.1 m (4,800 rigid tris); dt=0.01, 5% strain limit, PCG tol 1e-6, "
                               "corrected physics with self contact",
                   "vertices": n, "blocks": nb,
                   "precond": tl_desc if tl else "block-Jacobi (the reference's PCG)",
                   "timed_substeps": [a.warmup, a.warmup + len(recs) - 1],
                   "pcg_kernel": {0: "two-kernel, BSR row SpMV",
                                  4: "two-kernel two-level: patch SpMV (two lanes per row), patch update with "
                                     "Chebyshev steps in LDS and the coarse correction",
                                  3: f"two-kernel, sliced-ELL SpMV ({lanes} lane{'s' if lanes > 1 else ''} per row, "
                                     f"{wg_slices} slice{'s' if wg_slices != 1 else ''} per workgroup, "
                                     f"{vec_block}-thread vector kernels{', autotuned' if tuned else ''})"
                                  }.get(kernel, "none")},
        "pcg_iters_per_step": head["pcg_iters_per_step"],
        "newton_iters_per_step": head["newton_iters_per_step"],
        "pcg_solves": head["pcg_solves"], "pcg_failures": head["pcg_failures"],
        "line_search_failures": head["line_search_failures"], "max_candidates": head["max_candidates"],
        "capped_trials": head["capped_trials"],
        "intersections_at_end": {"edge_triangle": et, "vertices_inside_bodies": ins},
        "roofline": roof, "cpu_baseline": cpu, "accuracy_vs_fp64": accuracy, "preconditioners": preconditioners,
        "learned_precond": learned, "legs": legs, "per_frame_consumers": frame,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
