"""Two-level Schwarz PCG (ANDO_MODE_TWO_LEVEL) on the GPU.

* Solves through ando_pcg: converged, the float64 residual of the returned
  solution within the tolerance's reach, and the iteration count next to the
  CPU model's (oracle/two_level.py, float64, same row order, patches, power
  estimate, Chebyshev degree and coarse level) -- irregular random patterns
  (rows longer than one register chunk, a last partial patch) in the given
  order, cloth systems in the mesh's Morton order.
* Far fewer iterations than block-Jacobi PCG on the 257 x 257 drape's systems.
* Substeps of the corrected physics from identical states: within three times
  the FP32 envelope of the FP64 oracle's substep (the oracle solves with
  block-Jacobi PCG; both solves stop somewhere inside the PCG tolerance).
"""
import numpy as np
import pytest

from andosim_amd import scenes
from oracle import two_level as tlm
from oracle.oracle import OracleSim

pytestmark = pytest.mark.gpu

native = pytest.importorskip("andosim_amd.native")

from test_gpu_sell import random_spd_bsr  # noqa: E402
from test_gpu_physics import fp32_envelope, rel_err  # noqa: E402


def dense_residual(rp, ci, vals, x, b):
    A = tlm.bsr_to_csr(rp, ci, vals, len(rp) - 1)
    return float(np.max(np.abs(A @ np.asarray(x, np.float64).ravel() - np.asarray(b, np.float64).ravel())))


@pytest.mark.parametrize("n,seed", [(300, 1), (1000, 2), (4097, 3)])
def test_two_level_random_patterns(n, seed):
    print(f"two-level random n={n}", flush=True)
    rp, ci, vals, b = random_spd_bsr(n, seed, extra_rows=(0, n // 2, n - 1))
    sim = native.GpuSim()
    sim.set_mode(two_level=True)
    x, it, ok = sim.pcg(rp, ci, vals, b, tol=1e-6, maxit=2000)
    assert ok
    assert dense_residual(rp, ci, vals, x, b) <= 1e-5 * np.max(np.abs(b))
    A = tlm.bsr_to_csr(rp, ci, vals, n)
    _, it_m, ok_m = tlm.solve(A, b, np.arange(n), maxit=2000)
    assert ok_m
    assert abs(it - it_m) <= max(3, 0.15 * it_m), (it, it_m)
    plain = native.GpuSim()
    _, it_p, _ = plain.pcg(rp, ci, vals, b, tol=1e-6, maxit=2000)
    print(f"  iterations two-level {it} (model {it_m}), block-Jacobi {it_p}", flush=True)


@pytest.mark.parametrize("env,degree,coarse", [({"ANDO_TL_CHEB": "3"}, 3, True), ({"ANDO_TL_CHEB": "7"}, 7, True),
                                              ({"ANDO_TL_COARSE": "0"}, 9, False)])
def test_two_level_options_match_model(monkeypatch, env, degree, coarse):
    """The solver's options (read when the context is created): another
    Chebyshev degree, or no coarse level; iteration counts next to the CPU
    model's with the same settings, residual within reach."""
    print(f"two-level options {env}", flush=True)
    n = 1000
    rp, ci, vals, b = random_spd_bsr(n, 7, extra_rows=(0, n // 2, n - 1))
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    sim = native.GpuSim()
    sim.set_mode(two_level=True)
    x, it, ok = sim.pcg(rp, ci, vals, b, tol=1e-6, maxit=2000)
    assert ok
    assert dense_residual(rp, ci, vals, x, b) <= 1e-5 * np.max(np.abs(b))
    A = tlm.bsr_to_csr(rp, ci, vals, n)
    _, it_m, ok_m = tlm.solve(A, b, np.arange(n), maxit=2000, degree=degree, coarse=coarse)
    print(f"  iterations {it} (model {it_m})", flush=True)
    assert ok_m
    assert abs(it - it_m) <= max(3, 0.15 * it_m), (it, it_m)


def test_two_level_coarse_every_solve_drape(monkeypatch):
    """ANDO_TL_COARSE_EVERY=1 (the coarse inverse rebuilt for every solve of
    the step, not every second one): a 65 x 65 cylinder drape stays finite
    and intersection-free over 30 substeps, with the two-level solver."""
    print("two-level drape, coarse inverse every solve", flush=True)
    monkeypatch.setenv("ANDO_TL_COARSE_EVERY", "1")
    sc = scenes.cyl257(s=65)
    g = native.GpuSim(sc)
    g.set_mode(fp64_near=True, corrected_physics=True, two_level=True)
    for k in range(30):
        g.substep()
        assert g.stats()["pcg_kernel"] == 4
    assert np.all(np.isfinite(g.get("x")))
    assert g.count_intersections() == (0, 0)


def test_two_level_cold_start_drape(monkeypatch):
    """ANDO_TL_WARM=0 (every Newton solve from x0 = 0, as the reference's
    PCG) and the warm start: the 65 x 65 cylinder drape stays finite and
    intersection-free over 30 substeps either way (the iteration counts of
    the two are printed; which is lower depends on the workload)."""
    print("two-level drape, cold and warm starts", flush=True)
    sc = scenes.cyl257(s=65)
    iters = {}
    for warm in ("0", "1"):
        monkeypatch.setenv("ANDO_TL_WARM", warm)
        g = native.GpuSim(sc)
        g.set_mode(fp64_near=True, corrected_physics=True, two_level=True)
        total = 0
        for _ in range(30):
            g.substep()
            total += g.stats()["pcg_iterations"]
        assert np.all(np.isfinite(g.get("x")))
        assert g.count_intersections() == (0, 0)
        iters[warm] = total
    print(f"  PCG iterations over 30 substeps: cold {iters['0']}, warm {iters['1']}", flush=True)


def cloth_system(sc, substeps):
    """The first Newton system after `substeps` corrected-physics substeps."""
    sim = native.GpuSim(sc)
    sim.set_mode(fp64_near=True, corrected_physics=True)
    for _ in range(substeps):
        sim.substep()
    dt = np.float32(sc.params["dt"])
    x, v = sim.get("x"), sim.get("v")
    g, H = sim.linearize((x + dt * v).astype(np.float32))
    rp, ci = sim.pattern()
    return sim, rp, ci, H.reshape(-1, 9).astype(np.float32), (-g).astype(np.float32).ravel()


@pytest.mark.parametrize("name,make,substeps", [
    ("sheet33", lambda: scenes.small_sheet(33, pins=2, wall=True), 6),
    ("sphere129", scenes.sphere129, 14),
])
def test_two_level_cloth_matches_model(name, make, substeps):
    print(f"two-level cloth {name}", flush=True)
    sc = make()
    sim, rp, ci, H, b = cloth_system(sc, substeps)
    sim.set_mode(fp64_near=True, corrected_physics=True, two_level=True)
    x, it, ok = sim.pcg(rp, ci, H, b, tol=1e-6, maxit=1000)
    assert ok
    n = len(rp) - 1
    assert dense_residual(rp, ci, H, x, b) <= 1e-5 * np.max(np.abs(b))
    A = tlm.bsr_to_csr(rp, ci, H, n)
    _, it_m, ok_m = tlm.solve(A, b, tlm.morton_order(sc.verts), maxit=1000, precision=32)
    _, it_64, _ = tlm.solve(A, b, tlm.morton_order(sc.verts), maxit=1000)
    print(f"  iterations two-level {it} (model with float32 vectors {it_m}, float64 {it_64})", flush=True)
    assert ok_m
    assert abs(it - it_m) <= max(3, 0.2 * it_m), (it, it_m)


def test_two_level_cuts_iterations_on_cyl257():
    """The 257 x 257 drape two substeps after first contact: block-Jacobi needs
    hundreds of iterations (or the cap), the two-level solver at least three
    times fewer, and its solution meets the residual test in float64."""
    print("two-level cyl257", flush=True)
    sc = scenes.cyl257()
    sim, rp, ci, H, b = cloth_system(sc, 11)
    x_p, it_p, ok_p = sim.pcg(rp, ci, H, b, tol=1e-6, maxit=1000)
    sim.set_mode(fp64_near=True, corrected_physics=True, two_level=True)
    x, it, ok = sim.pcg(rp, ci, H, b, tol=1e-6, maxit=1000)
    print(f"  block-Jacobi {it_p} ({ok_p}), two-level {it} ({ok})", flush=True)
    assert ok
    assert dense_residual(rp, ci, H, x, b) <= 1e-5 * np.max(np.abs(b))
    assert 3 * it <= it_p, (it, it_p)


def test_two_level_substeps_match_oracle():
    """sphere129 past first contact: substeps from identical states with the
    two-level solver against the FP64 oracle."""
    print("two-level substeps sphere129", flush=True)
    sc = scenes.sphere129()
    g = native.GpuSim(sc)
    g.set_mode(fp64_near=True, corrected_physics=True)
    for _ in range(12):
        g.substep()
    for k in range(2):
        x0, v0 = g.get("x").copy(), g.get("v").copy()
        t = native.GpuSim(sc)
        t.set_mode(fp64_near=True, corrected_physics=True, two_level=True)
        t.set_state(x0, v0)
        t.substep()
        st = t.stats()
        assert st["pcg_kernel"] == 4
        o64 = OracleSim(sc, 64)
        o64.set_physics(True)
        o64.set("x", x0)
        o64.set("v", v0)
        o64.substep()
        x64, v64 = o64.get("x"), o64.get("v")
        ex, ev = fp32_envelope(sc, x0, v0, x64, v64, tries=1)
        xt, vt = t.get("x"), t.get("v")
        print(f"  substep {k}: two-level err {rel_err(xt, x64):.3e}, FP32 envelope {ex:.3e}, "
              f"pcg {st['pcg_per_solve'][:st['pcg_solves']]}", flush=True)
        assert rel_err(xt, x64) <= 3.0 * ex + 1e-6, (k, rel_err(xt, x64), ex)
        assert float(np.max(np.abs(vt - v64))) <= 3.0 * ev + 1e-4
        g.substep()


def test_two_level_replay_from_a_saved_state_is_exact():
    """A substep depends only on the state it starts from: after 14 substeps
    of the 65 x 65 cylinder drape (in contact from substep 9), the state is
    saved and set again, 6 more substeps run, the state is set back and the
    same 6 run again.  The coarse-inverse schedule (rebuilt every second
    solve) restarts at ando_set_state, so the replay repeats every PCG
    iteration count and every position bit for bit, whatever the solve count
    of the run before."""
    print("two-level replay from a saved state", flush=True)
    sc = scenes.cyl257(s=65)
    g = native.GpuSim(sc)
    g.set_mode(fp64_near=True, corrected_physics=True, two_level=True)
    for _ in range(14):
        g.substep()
    x0, v0 = g.get("x").copy(), g.get("v").copy()
    rig0 = [g.rigid_state(b) for b in range(g.nbodies)]

    def restore():
        g.set_state(x0, v0)
        for b, rs in enumerate(rig0):
            st = g.bodies[b]
            for k in range(3):
                st.position[k] = float(rs["position"][k])
                st.linear_velocity[k] = float(rs["linear_velocity"][k])
                st.angular_velocity[k] = float(rs["angular_velocity"][k])
            for k in range(9):
                st.rotation[k] = float(rs["rotation"].ravel()[k])

    runs = []
    for rep in range(2):
        restore()
        iters, solves = [], 0
        for _ in range(6):
            g.substep()
            st = g.stats()
            iters.append(st["pcg_iterations"])
            solves += st["pcg_solves"]
        runs.append((iters, g.get("x").copy()))
        print(f"  run {rep}: PCG iterations {iters} ({solves} solves)", flush=True)
    assert runs[0][0] == runs[1][0]
    assert np.array_equal(runs[0][1], runs[1][1])


def test_two_level_drape_stays_intersection_free():
    """cyl65 drape with the two-level solver: finite and intersection-free."""
    print("two-level drape cyl65", flush=True)
    sc = scenes.cyl257(s=65)
    g = native.GpuSim(sc)
    g.set_mode(fp64_near=True, corrected_physics=True, two_level=True)
    for k in range(60):
        g.substep()
        if k % 10 == 9:
            assert np.all(np.isfinite(g.get("x"))), k
            assert g.count_intersections() == (0, 0), k
