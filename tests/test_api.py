"""ando_barrier_core API semantics on the host (no GPU needed).

Restates the reference's tests/test_state_constraints.py:21-229 against the
drop-in module, plus the binding's documented exceptions
(src/py/bindings.cpp:39-66,102-153,244-356).
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "andosim_amd"))

abc = pytest.importorskip("ando_barrier_core")


def test_required_names_exist():
    # operators.py:11-21 and __init__.py:15-24 of the add-on.
    for name in ["Integrator", "EnergyTracker", "CollisionValidator", "AdaptiveTimestep", "RigidBody", "Mesh",
                 "State", "Constraints", "Material", "SimParams", "Triangle", "ContactType", "Contact",
                 "Elasticity", "EnergyDiagnostics", "CollisionMetrics", "barrier_energy", "barrier_gradient",
                 "barrier_hessian", "create_mesh_from_blender", "version"]:
        assert hasattr(abc, name), name
    assert abc.version() == "ando_barrier_core v1.0.0"


def test_mesh_initialisation_and_counts():
    mesh = abc.Mesh()
    assert mesh.num_vertices() == 0
    assert mesh.num_triangles() == 0
    mesh.initialize([[0.0, 0.0, 0.0], [0.1, 0.0, 0.0], [0.0, 0.1, 0.0]], [0, 1, 2], abc.Material())
    assert mesh.num_vertices() == 3
    assert mesh.num_triangles() == 1
    assert mesh.triangles.shape == (1, 3)


def test_mesh_shape_errors():
    mesh = abc.Mesh()
    with pytest.raises(ValueError):
        mesh.initialize([[0.0, 0.0]], [0, 1, 2], abc.Material())
    with pytest.raises(ValueError):
        mesh.initialize([[0.0, 0.0, 0.0]], [0, 1], abc.Material())
    with pytest.raises(ValueError):
        mesh.initialize([[0.0, 0.0, 0.0]], [[0, 1]], abc.Material())


def test_state_requires_initialised_mesh():
    with pytest.raises(ValueError):
        abc.State().initialize(abc.Mesh())


def test_state_initialisation_copies_vertex_data():
    material = abc.Material(density=2.0, thickness=0.5)
    vertices = np.array([[0.0, 0.0, 0.0], [0.2, 0.0, 0.0]], dtype=np.float32)
    mesh = abc.Mesh()
    mesh.initialize(vertices, [[0, 1, 1]], material)
    state = abc.State()
    state.initialize(mesh)
    masses = state.get_masses()
    assert np.allclose(masses, masses[0])
    mesh.vertices[0, 0] = 42.0
    assert np.allclose(state.get_positions(), [[0.0, 0.0, 0.0], [0.2, 0.0, 0.0]])


def test_state_apply_gravity_updates_velocities():
    mesh = abc.Mesh()
    mesh.initialize([[0.0, 0.0, 0.0]], [0, 0, 0], abc.Material())
    state = abc.State()
    state.initialize(mesh)
    p0 = state.get_positions().copy()
    v0 = state.get_velocities().copy()
    g = (0.0, -9.81, 0.0)
    state.apply_gravity(g, 0.5)
    v = state.get_velocities()
    assert np.allclose(v, v0 + np.array(g) * 0.5)
    assert np.allclose(state.get_positions(), p0 + v * 0.5)


def test_state_apply_gravity_validation():
    state = abc.State()
    with pytest.raises(RuntimeError):
        state.apply_gravity((0.0, -9.81, 0.0), 0.1)
    mesh = abc.Mesh()
    mesh.initialize([[0.0, 0.0, 0.0]], [0, 0, 0], abc.Material())
    state.initialize(mesh)
    with pytest.raises(ValueError):
        state.apply_gravity((0.0, -9.81), 0.1)


def test_constraints_track_pins_and_walls():
    c = abc.Constraints()
    assert c.num_active_pins() == 0
    assert c.num_active_contacts() == 0
    c.add_pin(0, (0.0, 0.0, 0.0))
    c.add_pin(1, (1.0, 0.0, 0.0))
    assert c.num_active_pins() == 2
    assert np.allclose(c._pins[0], (0.0, 0.0, 0.0))
    c.add_wall((0.0, 1.0, 0.0), offset=0.1, gap=0.01)
    c.add_wall((0.0, 0.0, 1.0), offset=-0.2, gap=0.05)
    assert c.num_active_contacts() == 2
    assert c._walls[0][1:] == (0.1, 0.01)
    with pytest.raises(ValueError):
        c.add_pin(0, (1.0, 2.0))
    with pytest.raises(ValueError):
        c.add_wall((0.0, 1.0), 0.0, 0.1)


def test_wall_normal_is_normalized():
    c = abc.Constraints()
    c.add_wall((0.0, 0.0, 2.0), 0.0, 0.001)
    assert np.allclose(c._walls[0][0], (0.0, 0.0, 1.0))


def test_material_kwargs_and_unknown_key():
    m = abc.Material(youngs_modulus=3e5, poisson_ratio=0.3)
    assert m.youngs_modulus == pytest.approx(3e5)
    with pytest.raises(AttributeError):
        abc.Material(colour=1.0)


def test_simparams_defaults():
    p = abc.SimParams()
    assert p.dt == pytest.approx(0.002)
    assert p.beta_max == pytest.approx(0.25)
    assert p.max_newton_steps == 8
    assert p.pcg_tol == pytest.approx(1e-3)
    assert p.pcg_max_iters == 1000
    assert p.contact_gap_max == pytest.approx(0.001)
    assert not p.enable_friction and not p.enable_strain_limiting


def test_adaptive_timestep_handles_degenerate_inputs():
    mesh = abc.Mesh()
    mesh.vertices = np.zeros((0, 3), dtype=np.float32)
    mesh.triangles = np.zeros((0, 3), dtype=np.int32)
    assert abc.AdaptiveTimestep.compute_min_edge_length(mesh) == 0.0
    assert abc.AdaptiveTimestep.compute_cfl_timestep(max_velocity=1.0, min_edge_length=0.0, safety=0.5) == 0.0
    assert abc.AdaptiveTimestep.compute_max_velocity([]) == 0.0


def test_compute_next_dt_respects_bounds_when_increasing():
    mesh = abc.Mesh()
    mesh.initialize([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]], [0, 1, 2], abc.Material())
    slow = np.full(9, 1e-4, dtype=np.float32)
    dt1 = abc.AdaptiveTimestep.compute_next_dt(slow, mesh, 0.02, 0.001, 0.05, 0.5)
    assert dt1 == pytest.approx(min(0.05, 0.02 * 1.5))
    dt2 = abc.AdaptiveTimestep.compute_next_dt(slow, mesh, dt1, 0.001, 0.05, 0.5)
    assert 0.001 <= dt2 <= 0.05
    assert dt2 == pytest.approx(min(0.05, dt1 * 1.5))


def test_compute_next_dt_clamps_to_dt_min():
    mesh = abc.Mesh()
    mesh.initialize([[0.0, 0.0, 0.0], [0.01, 0.0, 0.0], [0.0, 0.01, 0.0]], [0, 1, 2], abc.Material())
    fast = np.full(9, 1e6, dtype=np.float32)
    dt = abc.AdaptiveTimestep.compute_next_dt(fast, mesh, current_dt=0.1, dt_min=1e-5, dt_max=1.0, safety=0.5)
    assert 1e-5 <= dt <= 1e-5 * 1.01


def test_barrier_known_answers():
    # tests/test_barrier_derivatives.cpp:21-58: g=0.005, gmax=0.01, k=1000.
    assert abc.barrier_energy(0.005, 0.01, 1000.0) == pytest.approx(6.25e-3, rel=1e-5)
    assert abc.barrier_gradient(0.005, 0.01, 1000.0) == pytest.approx(-3.75, rel=1e-5)
    assert abc.barrier_hessian(0.005, 0.01, 1000.0) == pytest.approx(1500.0, rel=1e-5)
    assert abc.barrier_energy(0.015, 0.01, 1000.0) == 0.0
    assert abc.barrier_energy(0.01, 0.01, 1000.0) == 0.0
    assert abc.barrier_energy(-0.001, 0.01, 1000.0) > 0.0


def test_rigid_body_mass_and_world_vertices():
    b = abc.RigidBody()
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 2, 0], [0, 0, 3]], np.float32)
    b.initialize(V, np.array([[0, 1, 2], [0, 1, 3]], np.int32), 10.0)
    assert b.mass == pytest.approx(60.0)
    assert np.allclose(b.world_vertices(), V, atol=1e-6)
    assert np.allclose(b.position, V.mean(axis=0), atol=1e-6)
    b.apply_impulse((0.0, 0.0, 0.0), (60.0, 0.0, 0.0))
    assert np.allclose(b.linear_velocity, (1.0, 0.0, 0.0))
    b.integrate(0.5)
    assert np.allclose(b.position, V.mean(axis=0) + np.array([0.5, 0, 0]), atol=1e-6)


def test_energy_tracker_and_validator_on_host():
    mesh = abc.Mesh()
    mesh.initialize([[0.0, 0.0, 0.0], [0.1, 0.0, 0.0], [0.0, 0.1, 0.0]], [0, 1, 2], abc.Material())
    state = abc.State()
    state.initialize(mesh)
    state.apply_gravity((0.0, 0.0, -9.81), 0.01)
    d = abc.EnergyTracker.compute(mesh, state, abc.Constraints(), abc.SimParams())
    m = state.get_masses()
    v = state.get_velocities()
    assert d.kinetic_energy == pytest.approx(0.5 * float(np.sum(m * np.sum(v * v, axis=1))), rel=1e-5)
    assert d.max_velocity == pytest.approx(0.0981, rel=1e-5)
    met = abc.CollisionValidator.compute_metrics(mesh, state, [], 0.001, True)
    assert met.num_total_contacts == 0 and met.quality_level() == 0
    assert met.quality_description() == "Excellent - No penetrations"


def test_adaptive_timestep_accepts_a_host_state():
    """compute_next_dt / compute_max_velocity also take the State itself (an
    additive form); on a state that never reached the GPU they read its host
    velocities."""
    mesh = abc.Mesh()
    mesh.initialize([[0.0, 0.0, 0.0], [0.1, 0.0, 0.0], [0.0, 0.1, 0.0]], [0, 1, 2], abc.Material())
    state = abc.State()
    state.initialize(mesh)
    state.apply_gravity((0.0, 0.0, -9.81), 0.01)
    a = abc.AdaptiveTimestep.compute_next_dt(state, mesh, 0.01, 1e-4, 0.02, 0.5)
    b = abc.AdaptiveTimestep.compute_next_dt(state.get_velocities(), mesh, 0.01, 1e-4, 0.02, 0.5)
    assert a == b
    assert abc.AdaptiveTimestep.compute_max_velocity(state) == pytest.approx(0.0981, rel=1e-6)
