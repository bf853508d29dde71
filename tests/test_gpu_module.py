"""The drop-in module on the GPU, and the optional step terms (friction,
restitution, damping) against the oracle.

ando_barrier_core.Integrator.step is driven the way the Blender add-on drives
it (operators.py:900-905): State.apply_gravity, then Integrator.step.
"""
import os
import sys

import numpy as np
import pytest

from andosim_amd import scenes
from oracle.oracle import OracleSim

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "andosim_amd"))
native = pytest.importorskip("andosim_amd.native")


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


def sync_state(o, g, nb=0):
    g.set_state(o.get("x").astype(np.float32), o.get("v").astype(np.float32))
    for b in range(nb):
        rs = o.rigid_state(b)
        s = g.bodies[b]
        for k in range(3):
            s.position[k] = float(np.float32(rs["position"][k]))
            s.linear_velocity[k] = float(np.float32(rs["linear_velocity"][k]))
            s.angular_velocity[k] = float(np.float32(rs["angular_velocity"][k]))
        for k in range(9):
            s.rotation[k] = float(np.float32(rs["rotation"].ravel()[k]))


def friction_scene():
    sc = scenes.small_sheet(9, pins=0, wall=False, rigid=True)
    sc.params.update(enable_friction=True, friction_mu=0.5)
    return sc


def restitution_scene(rigid):
    sc = scenes.small_sheet(9, pins=0, wall=not rigid, rigid=rigid, z=0.02 if not rigid else 0.3)
    sc.params.update(contact_restitution=0.5, velocity_damping=0.1)
    return sc


@pytest.mark.parametrize("advance", [2, 4])
def test_friction_linearize_bit_exact(advance):
    sc = friction_scene()
    o = OracleSim(sc, 32)
    for _ in range(advance):
        o.substep()
    g = native.GpuSim(sc, fp64_near=False)
    sync_state(o, g, 1)
    x = o.get("x").astype(np.float32)
    v = o.get("v").astype(np.float32)
    xhat = (x + np.float32(sc.params["dt"]) * v).astype(np.float32)
    go, Ho = o.linearize(xhat)
    gg, Hg = g.linearize(xhat)
    assert np.array_equal(gg, go.astype(np.float32).reshape(-1))
    assert np.array_equal(Hg, Ho.astype(np.float32))


@pytest.mark.parametrize("make,nb", [(lambda: restitution_scene(False), 0), (lambda: restitution_scene(True), 1)])
def test_optional_terms_step_parity(make, nb):
    sc = make()
    o = OracleSim(sc, 32)
    worst = 0.0
    for _ in range(8):
        g = native.GpuSim(sc)
        sync_state(o, g, nb)
        o.substep()
        g.substep()
        worst = max(worst, rel_err(g.get("x"), o.get("x")))
        if nb:
            assert np.allclose(g.rigid_state(0)["linear_velocity"], o.rigid_state(0)["linear_velocity"],
                               rtol=1e-4, atol=1e-12)
    assert worst < 1e-6, worst


def test_friction_step_against_fp64():
    """With friction the Newton system is stiff (k_f up to mu |k g| / 1e-5), so two
    FP32 PCG solves that both meet the 1e-6 relative residual can differ by far
    more than 1e-6 in x.  The check is therefore against the FP64 oracle from the
    same state: the GPU step must be at least as close to it as the FP32 oracle."""
    sc = friction_scene()
    o32 = OracleSim(sc, 32)
    for k in range(8):
        x, v = o32.get("x"), o32.get("v")
        o64 = OracleSim(sc, 64)
        o64.set("x", x)
        o64.set("v", v)
        g = native.GpuSim(sc)
        sync_state(o32, g, 1)
        o64.set_rigid_pose(0, o32.rigid_state(0))
        o32.substep()
        o64.substep()
        g.substep()
        e_gpu = rel_err(g.get("x"), o64.get("x"))
        e_cpu = rel_err(o32.get("x"), o64.get("x"))
        # Both FP32 solves sit at the FP32 floor times the system's condition
        # number; their distances to the FP64 step are of the same order.
        assert e_gpu <= 5.0 * e_cpu + 1e-5, (k, e_gpu, e_cpu)


def build_module_scene(abc, sc):
    mesh = abc.Mesh()
    mesh.initialize(sc.verts, sc.tris, abc.Material(**sc.material))
    state = abc.State()
    state.initialize(mesh)
    cons = abc.Constraints()
    for vi, t in sc.pins:
        cons.add_pin(int(vi), t)
    for n, off, gap in sc.walls:
        cons.add_wall(n, off, gap)
    p = abc.SimParams()
    for k, v in sc.params.items():
        setattr(p, k, v)
    bodies = []
    for r in sc.rigids:
        b = abc.RigidBody()
        b.initialize(r.verts, r.tris, r.density)
        bodies.append(b)
    return mesh, state, cons, p, bodies


@pytest.mark.parametrize("make", [scenes.drop21, lambda: scenes.small_sheet(9, pins=0, wall=False, rigid=True)])
def test_module_step_matches_oracle(make):
    abc = pytest.importorskip("ando_barrier_core")
    sc = make()
    mesh, state, cons, p, bodies = build_module_scene(abc, sc)
    o = OracleSim(sc, 32)
    g = np.asarray(sc.gravity, np.float32)
    worst = 0.0
    for _ in range(20):
        state.apply_gravity(g, p.dt)
        if bodies:
            abc.Integrator.step(mesh, state, cons, p, bodies)
        else:
            abc.Integrator.step(mesh, state, cons, p)
        o.substep()
        worst = max(worst, rel_err(state.get_positions(), o.get("x")))
    assert worst < 1e-4, worst
    st = abc.Integrator.last_stats()
    assert st.pcg_solves >= 1 and st.newton_iterations >= 1
    assert st.pcg_kernel == 3 and st.spmv_lanes == 2 and st.learned_solves == 0
    if bodies:
        assert np.allclose(bodies[0].position, o.rigid_state(0)["position"], atol=1e-6)
    # The per-frame consumers of the add-on still work after GPU steps.
    contacts = abc.Integrator.compute_contacts(mesh, state, bodies) if bodies else abc.Integrator.compute_contacts(
        mesh, state)
    metrics = abc.CollisionValidator.compute_metrics(mesh, state, contacts, p.contact_gap_max, True)
    assert metrics.num_total_contacts == len(contacts)
    diag = abc.EnergyTracker.compute(mesh, state, cons, p)
    assert diag.max_velocity > 0.0


def test_module_host_edits_reach_device():
    """set_velocities after a GPU step is honoured by the next step."""
    abc = pytest.importorskip("ando_barrier_core")
    sc = scenes.small_sheet(7, pins=0, wall=False)
    mesh, state, cons, p, _ = build_module_scene(abc, sc)
    state.apply_gravity((0.0, 0.0, -9.81), p.dt)
    abc.Integrator.step(mesh, state, cons, p)
    state.set_velocities(np.zeros((mesh.num_vertices(), 3), np.float32))
    x0 = state.get_positions().copy()
    abc.Integrator.step(mesh, state, cons, p)
    # No gravity and zero velocity: the cloth stays (up to the strain-free rest gradient).
    assert np.max(np.abs(state.get_positions() - x0)) < 1e-6


def test_module_readback_arrays_are_independent():
    """get_positions / get_velocities while the device holds the newer state:
    each call returns a new (N, 3) float32 array over its own pinned buffer;
    arrays from earlier calls keep their values through later steps and later
    reads (buffers go back to the pool only when their array is released), and
    the values equal the pageable copy of the module's host path."""
    abc = pytest.importorskip("ando_barrier_core")
    sc = scenes.small_sheet(9, pins=0, wall=True)
    mesh, state, cons, p, _ = build_module_scene(abc, sc)
    n = mesh.num_vertices()
    kept = []
    for _ in range(6):
        state.apply_gravity((0.0, 0.0, -9.81), p.dt)
        abc.Integrator.step(mesh, state, cons, p)
        x, v = state.get_positions(), state.get_velocities()
        assert x.shape == (n, 3) and x.dtype == np.float32 and x.flags.c_contiguous and x.flags.writeable
        x2 = state.get_positions()
        assert np.array_equal(x, x2) and not np.shares_memory(x, x2)
        x2[:] = 0.0
        del x2
        kept.append((x.copy(), x, v.copy(), v))
    for xc, x, vc, v in kept:
        assert np.array_equal(xc, x) and np.array_equal(vc, v)
    assert not np.array_equal(kept[0][0], kept[-1][0])
    xs, vs = kept[-1][0], kept[-1][2]
    state.set_velocities(vs)  # pulls both arrays into the host copies first
    assert np.array_equal(state.get_positions(), xs) and np.array_equal(state.get_velocities(), vs)


def test_module_pinned_pool_is_bounded():
    """Reads of states of several sizes: released arrays return their buffers
    to the pool (at most four idle per size), nothing stays live once every
    array is gone, and live plus idle bytes never pass the 256 MB cap."""
    abc = pytest.importorskip("ando_barrier_core")
    import gc
    gc.collect()
    cap = 256 << 20
    live0, idle0 = abc._pinned_pool_bytes()  # earlier tests may leave idle buffers
    for s in (5, 9, 13, 17):
        sc = scenes.small_sheet(s, pins=0, wall=True)
        mesh, state, cons, p, _ = build_module_scene(abc, sc)
        state.apply_gravity((0.0, 0.0, -9.81), p.dt)
        abc.Integrator.step(mesh, state, cons, p)
        arrays = [state.get_positions() for _ in range(7)]
        live, idle = abc._pinned_pool_bytes()
        assert live >= 7 * 12 * s * s and live + idle <= cap
        del arrays
        live, idle = abc._pinned_pool_bytes()
        assert live + idle <= cap and idle - idle0 <= 4 * 12 * (5 * 5 + 9 * 9 + 13 * 13 + 17 * 17)
    gc.collect()
    assert abc._pinned_pool_bytes()[0] == live0


def light_sphere_scene(density=50.0):
    """The 9x9 sheet falling on a sphere light enough to be pushed around."""
    sc = scenes.small_sheet(9, pins=0, wall=False, rigid=True)
    r = sc.rigids[0]
    sc.rigids[0] = scenes.Rigid(r.verts, r.tris, density)
    return sc


def test_rigid_coupling_moves_a_light_body_like_the_oracle():
    """apply_rigid_coupling's per-body force and torque sums run on the device
    (double sums of the reference's float terms); the body's velocities after
    each step match the FP32 oracle's, and the coupling does move the body
    (rigid |v| > 1e-6, as test_hybrid.cpp:12-71 asks of a coupled step)."""
    sc = light_sphere_scene()
    o = OracleSim(sc, 32)
    moved = False
    for k in range(10):
        g = native.GpuSim(sc)
        sync_state(o, g, 1)
        # The torque nearly cancels over the symmetric drape, so its float sum
        # carries rounding noise far above 1e-5 of the result; the FP64 oracle
        # from the same state measures that noise.
        o64 = OracleSim(sc, 64)
        o64.set("x", o.get("x"))
        o64.set("v", o.get("v"))
        o64.set_rigid_pose(0, o.rigid_state(0))
        o.substep()
        o64.substep()
        g.substep()
        rg, ro, r64 = g.rigid_state(0), o.rigid_state(0), o64.rigid_state(0)
        for key in ("linear_velocity", "angular_velocity"):
            scale = max(float(np.max(np.abs(ro[key]))), 1e-9)
            noise = float(np.max(np.abs(ro[key] - r64[key])))
            assert np.max(np.abs(rg[key] - ro[key])) <= max(1e-5 * scale, 2.0 * noise), (k, key, rg[key], ro[key])
        assert np.allclose(rg["position"], ro["position"], atol=1e-7)
        assert rel_err(g.get("x"), o.get("x")) < 1e-6
        moved = moved or float(np.linalg.norm(ro["linear_velocity"])) > 1e-6
    assert moved


@pytest.mark.parametrize("corrected", [False, True])
def test_rigid_integrate_on_the_device_spins_like_the_oracle(corrected):
    """RigidBody::integrate (rigid_body.cpp:151-176) runs on the device after
    the coupling sums: a light sphere given a linear and an angular velocity
    under the falling sheet translates and rotates (axis-angle update) like
    the FP32 oracle's body.  Each step starts from the oracle's state; the
    FP64 oracle from the same state measures the float rounding of the
    nearly cancelling contact sums."""
    sc = light_sphere_scene()
    o = OracleSim(sc, 32)
    o.set_physics(corrected)
    rs = o.rigid_state(0)
    rs["linear_velocity"] = np.array([0.05, -0.02, 0.1])
    rs["angular_velocity"] = np.array([0.7, -2.1, 3.4])
    o.set_rigid_pose(0, rs)
    g = native.GpuSim(sc)
    g.set_mode(corrected_physics=corrected)
    contacts = 0
    for k in range(6):
        sync_state(o, g, 1)
        o64 = OracleSim(sc, 64)
        o64.set_physics(corrected)
        o64.set("x", o.get("x"))
        o64.set("v", o.get("v"))
        o64.set_rigid_pose(0, o.rigid_state(0))
        o.substep()
        o64.substep()
        g.substep()
        contacts += g.stats()["rigid_contacts"]
        rg, ro, r64 = g.rigid_state(0), o.rigid_state(0), o64.rigid_state(0)
        for key in ("position", "rotation", "linear_velocity", "angular_velocity"):
            scale = max(float(np.max(np.abs(ro[key]))), 1e-6)
            noise = float(np.max(np.abs(ro[key] - r64[key])))
            assert np.max(np.abs(rg[key] - ro[key])) <= max(2e-6 * scale, 2.0 * noise), (k, key, rg[key], ro[key])
        assert rel_err(g.get("x"), o.get("x")) < 1e-6
    assert float(np.max(np.abs(rg["rotation"] - np.eye(3)))) > 1e-2
    assert contacts > 0


def test_static_rigids_leave_bodies_unchanged():
    sc = scenes.small_sheet(9, pins=0, wall=False, rigid=True)
    a, b = native.GpuSim(sc), native.GpuSim(sc)
    b.set_mode(static_rigids=True)
    pose0 = b.rigid_state(0)
    contacts = 0
    for _ in range(8):
        a.substep()
        b.substep()
        contacts += b.stats()["rigid_contacts"]
    pb = b.rigid_state(0)
    for key in ("position", "linear_velocity", "angular_velocity", "rotation"):
        assert np.array_equal(pb[key], pose0[key]), key
    # The default body (density 1e12) barely moves, so the cloth agrees.
    assert rel_err(b.get("x"), a.get("x")) < 1e-5
    assert contacts > 0
