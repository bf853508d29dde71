"""The add-on's per-frame consumers on the GPU: EnergyTracker::compute and
CollisionValidator::compute_metrics, against the oracle.

Per-vertex and per-face terms are the reference's float expressions on both
sides, so the device's double sums must equal the oracle's double sums of the
same float terms up to double rounding (relative 1e-9 here); counts, minima
and maxima are exact.
"""
import os
import sys

import numpy as np
import pytest

from andosim_amd import scenes
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "andosim_amd"))

native = pytest.importorskip("andosim_amd.native")

REL = 1e-9


def _close(a, b, scale):
    return abs(a - b) <= REL * scale


def _run(sc, steps, corrected=False):
    g = native.GpuSim(sc)
    if corrected:
        g.set_mode(corrected=True)
    for _ in range(steps):
        g.substep()
    return g


def _check_energy(g, sc, deform_mesh):
    x, v = g.get("x"), g.get("v")
    o = orc.OracleSim(sc, 32)
    if deform_mesh:
        # Elastic energy reads the mesh vertices (F from mesh.vertices); give the
        # mesh the stepped positions so the faces are strained.
        native.check(g.L.ando_set_mesh_vertices(g.h, native._fp(x), g.n))
        o.set("mesh", x)
    e = g.energy()
    _, ref = o.energy(x, v, g.masses)
    m = g.masses.astype(np.float64)
    ke_scale = 0.5 * float(np.sum(m * np.sum(v.astype(np.float64) ** 2, axis=1)))
    assert _close(e["kinetic_energy"], ref["kinetic_energy"], ke_scale)
    p_scale = float(np.sum(m * np.abs(v).sum(axis=1)))
    L_scale = p_scale * float(np.abs(x).max())
    assert np.all(np.abs(e["linear_momentum"] - ref["linear_momentum"]) <= REL * p_scale)
    assert np.all(np.abs(e["angular_momentum"] - ref["angular_momentum"]) <= REL * L_scale)
    assert e["max_velocity"] == ref["max_velocity"]
    assert _close(e["elastic_energy"], ref["elastic_energy"], max(ref["elastic_energy"], 1e-30))
    if deform_mesh:
        assert ref["elastic_energy"] > 0.0
    return e


@pytest.mark.parametrize("name,make,steps", [
    ("sheet9_rigid", lambda: scenes.small_sheet(9, pins=0, wall=False, rigid=True), 6),
    ("sphere129", lambda: scenes.sphere129(), 4),
])
def test_energy_matches_oracle(name, make, steps):
    print(f"energy {name}", flush=True)
    sc = make()
    g = _run(sc, steps)
    _check_energy(g, sc, deform_mesh=False)
    _check_energy(g, sc, deform_mesh=True)


def _listing_arrays(cs):
    types = np.array([c.type for c in cs], np.int32)
    idx = np.array([list(c.idx) for c in cs], np.int32).reshape(-1, 4)
    gaps = np.array([c.gap for c in cs], np.float64)
    normals = np.array([list(c.normal) for c in cs], np.float64).reshape(-1, 3)
    return types, idx, gaps, normals


def _check_metrics(gm, om, count):
    for k in orc.METRIC_KEYS[:10]:
        assert gm[k] == om[k], k
    for k in ("min_gap", "max_gap", "max_penetration", "max_relative_velocity", "ccd_effectiveness"):
        assert gm[k] == np.float32(om[k]), k
    for k, n in (("sum_gap", "num_total_contacts"), ("sum_penetration", "num_penetrations"),
                 ("sum_relative_velocity", "num_relative_velocities")):
        scale = max(abs(om[k]), 1e-30)
        assert abs(gm[k] - om[k]) <= REL * scale, k
        avg = {"sum_gap": "avg_gap", "sum_penetration": "avg_penetration",
               "sum_relative_velocity": "avg_relative_velocity"}[k]
        expect = np.float32(np.float32(gm[k]) / np.float32(gm[n])) if gm[n] else np.float32(0.0)
        assert gm[avg] == expect, avg
        # The reference's float accumulation lies within the recursive-sum bound.
        eps = float(np.finfo(np.float32).eps)
        if gm[n]:
            assert abs(gm[avg] - om[avg]) <= 2 * count * eps * abs(om[avg]) + 1e-30, avg


@pytest.mark.parametrize("name,make,steps,corrected", [
    ("sheet9_rigid", lambda: scenes.small_sheet(9, pins=0, wall=False, rigid=True), 6, False),
    ("stack_resting33", lambda: scenes.stack_resting(s=33, sheets=4), 2, True),
])
def test_contact_metrics_table_matches_oracle(name, make, steps, corrected):
    print(f"metrics {name}", flush=True)
    sc = make()
    g = _run(sc, steps, corrected)
    cs, count = g.contacts(cap=1 << 20)
    assert count == len(cs) and count > 0
    v = g.get("v")
    types, idx, gaps, normals = _listing_arrays(cs)
    for ccd in (True, False):
        gap_max = float(sc.params["contact_gap_max"])
        gm = g.contact_metrics(gap_max, ccd)
        om = orc.contact_metrics(types, idx, gaps, normals, v, gap_max, ccd, 32)
        _check_metrics(gm, om, count)
        # The same contacts handed over as a list give the same result.
        gl = g.contact_metrics(gap_max, ccd, contacts=cs)
        assert gl == gm
    if corrected:
        assert gm["num_point_triangle"] + gm["num_edge_edge"] > 0 and gm["num_relative_velocities"] > 0
    else:
        assert gm["num_relative_velocities"] == 0  # rigid contacts only


def test_contact_metrics_reference_case_on_device():
    """The contact list of test_collision_validator.cpp:54-95 evaluated with
    device-resident velocities."""
    sc = scenes.small_sheet(5, pins=0, wall=False)
    g = native.GpuSim(sc)
    vel = np.zeros((g.n, 3), np.float32)
    vel[:4] = [[0.0, -1.0, 0.0], [0.0, 0.25, 0.0], [0.0, -0.4, 0.0], [0.0, 0.1, 0.0]]
    g.set_state(v=vel)
    specs = [(0, [0, 1, 2, 3], -5.0e-4), (1, [0, 1, 2, 3], 2.0e-4), (2, [2, -1, -1, -1], -1.5e-3)]
    cs = []
    for t, ix, gap in specs:
        c = native.Contact()
        c.type = t
        c.idx[:] = ix
        c.gap = gap
        c.normal[:] = [0.0, 1.0, 0.0]
        cs.append(c)
    for ccd in (True, False):
        gm = g.contact_metrics(0.002, ccd, contacts=cs)
        om = orc.contact_metrics(*_listing_arrays(cs), vel, 0.002, ccd, 32)
        _check_metrics(gm, om, 3)
    gm = g.contact_metrics(0.002, True, contacts=cs)
    assert gm["has_major_penetration"] == 1 and gm["has_tunneling"] == 1
    assert gm["max_relative_velocity"] == pytest.approx(1.0 - 0.05 / 3.0, rel=1e-6)
    empty = g.contact_metrics(0.002, True, contacts=[])
    assert empty["num_total_contacts"] == 0 and empty["ccd_effectiveness"] == 0.0
    bad = native.Contact()
    bad.type = 0
    bad.idx[:] = [0, 1, 2, g.n]
    with pytest.raises(RuntimeError, match="out of range"):
        g.contact_metrics(0.002, True, contacts=[bad])


def test_contact_metrics_refuse_a_replaced_table():
    sc = scenes.small_sheet(9, pins=0, wall=False, rigid=True)
    g = _run(sc, 2)
    g.contacts()
    g.contact_metrics(1e-3, True)
    g.substep()  # the step detects again
    with pytest.raises(RuntimeError):
        g.contact_metrics(1e-3, True)


def test_module_frame_consumers_use_the_device():
    """EnergyTracker.compute and CollisionValidator.compute_metrics through
    ando_barrier_core after GPU steps: the device results equal the host
    restatement's up to float summation order."""
    abc = pytest.importorskip("ando_barrier_core")
    sc = scenes.small_sheet(9, pins=0, wall=False, rigid=True)
    mesh = abc.Mesh()
    mesh.initialize(sc.verts, sc.tris, abc.Material(**sc.material))
    state = abc.State()
    state.initialize(mesh)
    cons = abc.Constraints()
    p = abc.SimParams()
    for k, val in sc.params.items():
        setattr(p, k, val)
    bodies = []
    for r in sc.rigids:
        b = abc.RigidBody()
        b.initialize(r.verts, r.tris, r.density)
        bodies.append(b)
    for _ in range(6):
        state.apply_gravity(np.asarray(sc.gravity, np.float32), p.dt)
        abc.Integrator.step(mesh, state, cons, p, bodies)
    contacts = abc.Integrator.compute_contacts(mesh, state, bodies)
    assert len(contacts) > 0
    dev = abc.CollisionValidator.compute_metrics(mesh, state, contacts, p.contact_gap_max, True)
    host = abc.CollisionValidator.compute_metrics(mesh, state, contacts[::-1], p.contact_gap_max, True)
    for k in ("num_total_contacts", "num_point_triangle", "num_edge_edge", "num_wall", "num_penetrations",
              "num_ccd_contacts", "num_broad_phase_contacts", "min_gap", "max_gap", "max_penetration",
              "max_relative_velocity", "is_stable", "has_tunneling", "has_major_penetration"):
        assert getattr(dev, k) == getattr(host, k), k
    assert dev.avg_gap == pytest.approx(host.avg_gap, rel=1e-5)
    d = abc.EnergyTracker.compute(mesh, state, cons, p)
    x, v, m = state.get_positions(), state.get_velocities(), state.get_masses()
    o = orc.OracleSim(sc, 32)
    _, ref = o.energy(x, v, m)
    assert d.kinetic_energy == pytest.approx(ref["kinetic_energy"], rel=1e-6)
    assert d.max_velocity == np.float32(ref["max_velocity"])
    assert np.allclose(d.linear_momentum, ref["linear_momentum"], rtol=1e-5, atol=1e-12)
    assert abc.EnergyTracker.compute_max_velocity(state) == d.max_velocity
    assert abc.EnergyTracker.compute_kinetic_energy(state) == d.kinetic_energy


def _f32_sqn(d):
    d = d.astype(np.float32)
    return d[:, 0] * d[:, 0] + (d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])


def test_cfl_inputs_exact():
    """max |v| and the minimum edge length are a maximum and a minimum of float
    expressions: the device values equal the float32 evaluation exactly."""
    sc = scenes.sphere129()
    g = _run(sc, 3)
    vmax, emin = g.cfl_inputs()
    v = g.get("v")
    assert vmax == np.sqrt(np.max(_f32_sqn(v)), dtype=np.float32)
    T = np.asarray(sc.tris)
    E = np.unique(np.sort(np.concatenate([T[:, [0, 1]], T[:, [1, 2]], T[:, [2, 0]]]), axis=1), axis=0)
    V = np.asarray(sc.verts, np.float32)
    d = V[E[:, 1]] - V[E[:, 0]]
    assert emin == np.sqrt(np.min(_f32_sqn(d)), dtype=np.float32)


def test_module_adaptive_timestep_accepts_the_state():
    abc = pytest.importorskip("ando_barrier_core")
    sc = scenes.small_sheet(9, pins=0, wall=False)
    mesh = abc.Mesh()
    mesh.initialize(sc.verts, sc.tris, abc.Material(**sc.material))
    state = abc.State()
    state.initialize(mesh)
    cons = abc.Constraints()
    p = abc.SimParams()
    for k, val in sc.params.items():
        setattr(p, k, val)
    for _ in range(3):
        state.apply_gravity(np.asarray(sc.gravity, np.float32), p.dt)
        abc.Integrator.step(mesh, state, cons, p)
    on_device = abc.AdaptiveTimestep.compute_next_dt(state, mesh, 0.01, 1e-4, 0.02, 0.5)
    on_host = abc.AdaptiveTimestep.compute_next_dt(state.get_velocities(), mesh, 0.01, 1e-4, 0.02, 0.5)
    assert on_device == on_host
    assert abc.AdaptiveTimestep.compute_max_velocity(state) == abc.AdaptiveTimestep.compute_max_velocity(
        state.get_velocities())
