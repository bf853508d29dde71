"""Corrected contact mode on the GPU: the LBVH candidate lists against the
oracle's brute-force enumeration of the same predicate.

The candidate/contact lists must be identical after sorting, bit for bit in
FP32 (same narrow-phase arithmetic on the same positions); the GPU list is
already in the brute-force order (PT by vertex then triangle, EE by e1 then e2).
"""
import numpy as np
import pytest

from andosim_amd import scenes
from oracle.oracle import OracleSim

pytestmark = pytest.mark.gpu

native = pytest.importorskip("andosim_amd.native")


def jittered_stack(s=9, gap=5e-5, seed=3, sheets=2, shift=0.013, jitter=1e-4):
    rng = np.random.default_rng(seed)
    Vs, Ts = [], []
    for k in range(sheets):
        V, T = scenes.grid(s, 0.2, 0.2 + k * gap)
        V[:, 0] += k * shift
        V[:, 2] += rng.uniform(-jitter, jitter, len(V))
        Vs.append(V)
        Ts.append(T + k * s * s)
    V = np.concatenate(Vs).astype(np.float32)
    T = np.concatenate(Ts).astype(np.int32)
    return scenes.Scene("stack", V, T, dict(scenes.BASE_MATERIAL), scenes.base_params(), [], [], [])


def crumpled(s=17, seed=5):
    """A sheet folded over itself: a smooth sheet mapped onto a tight S curve."""
    V, T = scenes.grid(s, 0.3, 0.2)
    u = (V[:, 0] - V[:, 0].min()) / (V[:, 0].max() - V[:, 0].min())
    ang = 3.0 * np.pi * u
    r = 0.02 + 0.001 * np.random.default_rng(seed).random(len(V))
    V = np.stack([r * np.sin(ang) + 0.004 * u, V[:, 1], 0.2 + r * np.cos(ang) * 0.1 + 0.0005 * ang], axis=1)
    return scenes.Scene("crumpled", V.astype(np.float32), T, dict(scenes.BASE_MATERIAL), scenes.base_params(),
                        [], [], [])


def contact_table_gpu(cg):
    rows = [[c.type, *c.idx, c.gap, *c.normal, *c.witness_q, *c.weights] for c in cg]
    return np.array(rows, np.float32).reshape(-1, 16)


def contact_table_oracle(co):
    A = np.concatenate([co[:, 0:5], co[:, 6:7], co[:, 7:10], co[:, 10:13], co[:, 13:17]], axis=1)
    return A.astype(np.float32).reshape(-1, 16)


@pytest.mark.parametrize("make", [lambda: jittered_stack(), lambda: jittered_stack(15, 2e-5, 7, 3),
                                  lambda: crumpled()])
def test_lbvh_contacts_match_brute_force(make):
    sc = make()
    o = OracleSim(sc, 32)
    o.set_contact_mode(True)
    co = o.contacts()
    g = native.GpuSim(sc)
    g.set_mode(corrected=True)
    cg, cnt = g.contacts()
    assert cnt == len(co)
    A, B = contact_table_oracle(co), contact_table_gpu(cg)
    # Same order as the brute-force enumeration, element for element.
    assert np.array_equal(A[:, :5], B[:, :5])
    assert np.array_equal(A, B)
    assert (A[:, 0] == 0).any()


def test_faithful_mode_has_no_self_contacts():
    g = native.GpuSim(jittered_stack())
    cg, cnt = g.contacts()
    assert cnt == 0


def sync_state(o, g):
    g.set_state(o.get("x").astype(np.float32), o.get("v").astype(np.float32))


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


def resting_stack(gap=8e-5, z0=6e-4):
    """Two jittered sheets, the lower one 0.57-0.63 mm above the ground wall,
    the upper one `gap` above it: PT and EE contacts that change step to step."""
    sc = jittered_stack(9, gap, 11, 2, 0.013, 3e-5)
    sc.verts[:, 2] += np.float32(z0 - 0.2)
    sc.walls = [(np.array([0, 0, 1], np.float32), 0.0, 1e-3)]
    return sc


@pytest.mark.parametrize("gap", [8e-5, 3e-5])
def test_corrected_linearize_bit_exact(gap):
    """Gradient and system matrix with self-contact barriers (FP64 island off):
    bit-identical to the FP32 oracle in corrected mode."""
    sc = resting_stack(gap)
    o = OracleSim(sc, 32)
    o.set_contact_mode(True)
    g = native.GpuSim(sc, fp64_near=False)
    g.set_mode(fp64_near=False, corrected=True)
    sync_state(o, g)
    x = o.get("x").astype(np.float32)
    v = o.get("v").astype(np.float32)
    xhat = (x + np.float32(sc.params["dt"]) * v).astype(np.float32)
    go, Ho = o.linearize(xhat)
    gg, Hg = g.linearize(xhat)
    assert len(o.contacts()) > 0
    assert np.array_equal(gg.reshape(-1), go.astype(np.float32).reshape(-1))
    assert np.array_equal(Hg, Ho.astype(np.float32))


@pytest.mark.parametrize("additive", [False, True])
def test_corrected_step_parity(additive):
    """Substeps from identical states with self contacts active: positions
    within 1e-6 relative of the FP32 oracle (same contacts, same line-search
    CCD — the reference's samples or additive CCD —, PCG to the same tolerance
    including the coupling blocks)."""
    print(f"corrected step parity additive={additive}", flush=True)
    sc = resting_stack(8e-5)
    o = OracleSim(sc, 32)
    o.set_contact_mode(True)
    o.set_ccd_mode(additive)
    if additive:
        # The upper sheet moves down 1 mm per substep onto the lower one (0.08 mm
        # below), so contact pairs close within a step.
        v = o.get("v")
        v[len(v) // 2:, 2] = -0.1
        o.set("v", v)
    errs, alphas = [], []
    for k in range(8):
        g = native.GpuSim(sc)
        g.set_mode(corrected=True, additive_ccd=additive)
        sync_state(o, g)
        o.substep()
        g.substep()
        errs.append(rel_err(g.get("x"), o.get("x")))
        assert g.stats()["contacts"] == o.stats()["contacts"]
        alphas.append(g.stats()["last_alpha"])
    assert max(errs) < 1e-6, errs


def crossing_sheets(s=12):
    """A horizontal sheet and a sheet tilted 30 degrees through it."""
    V1, T1 = scenes.grid(s, 0.2, 0.2)
    V2, T2 = scenes.grid(s, 0.2, 0.0)
    th = np.radians(30.0)
    y, z = V2[:, 1].copy(), V2[:, 2].copy()
    V2[:, 1] = y * np.cos(th) - z * np.sin(th) + 0.003
    V2[:, 2] = 0.2 + y * np.sin(th) + z * np.cos(th)
    V = np.concatenate([V1, V2]).astype(np.float32)
    T = np.concatenate([T1, T2 + s * s]).astype(np.int32)
    return scenes.Scene("cross", V, T, dict(scenes.BASE_MATERIAL), scenes.base_params(), [], [], [])


def sheet_through_sphere():
    sc = scenes.small_sheet(15, pins=0, wall=False, rigid=True, z=0.3)
    sc.rigids[0].verts = (sc.rigids[0].verts + np.array([0, 0, 0.105], np.float32)).astype(np.float32)
    return sc


@pytest.mark.parametrize("make", [crossing_sheets, crumpled, sheet_through_sphere, lambda: jittered_stack()])
def test_intersection_counts_match_brute_force(make):
    sc = make()
    o = OracleSim(sc, 32)
    g = native.GpuSim(sc)
    assert g.count_intersections() == o.count_intersections()
    if make is crossing_sheets:
        assert o.count_intersections()[0] > 0
    if make is sheet_through_sphere:
        assert o.count_intersections()[1] > 0


def test_one_triangle_faithful_contacts_match_oracle():
    """Faithful mode on a one-triangle mesh lists the reference's PT contacts
    (its BVH has one primitive-bearing leaf): identical to the oracle's list."""
    sc = scenes.one_triangle()
    o = OracleSim(sc, 32)
    co = o.contacts()
    g = native.GpuSim(sc)
    cg, cnt = g.contacts()
    assert cnt == len(co) == 2
    assert np.array_equal(contact_table_oracle(co), contact_table_gpu(cg))


@pytest.mark.parametrize("friction", [False, True])
def test_one_triangle_faithful_linearize_bit_exact(friction):
    """Gradient and matrix with the one-triangle PT barrier and (faithful mode)
    its friction on the point, the points having slid since the initial
    positions: bit-identical to the FP32 oracle."""
    sc = scenes.one_triangle(enable_friction=friction, friction_mu=0.5)
    o = OracleSim(sc, 32)
    x = o.get("x")
    x[3, 0] += 2e-3
    x[4, 1] -= 1e-3
    o.set("x", x)
    v = o.get("v")
    v[3] = [0.01, 0.0, -0.002]
    o.set("v", v)
    g = native.GpuSim(sc, fp64_near=False)
    g.set_mode(fp64_near=False)
    sync_state(o, g)
    xhat = (x + np.float32(sc.params["dt"]) * v).astype(np.float32)
    go, Ho = o.linearize(xhat)
    gg, Hg = g.linearize(xhat)
    assert len(o.contacts()) == 2
    assert np.array_equal(gg.reshape(-1), go.astype(np.float32).reshape(-1))
    assert np.array_equal(Hg, Ho.astype(np.float32))
    if friction:
        o0 = OracleSim(scenes.one_triangle(), 32)
        o0.set("x", x)
        o0.set("v", v)
        g0, _ = o0.linearize(xhat)
        assert not np.array_equal(g0.reshape(-1, 3)[3], go.reshape(-1, 3)[3])   # friction acts on the point


def test_one_triangle_faithful_steps_match_oracle():
    """Substeps of the one-triangle mesh with friction and restitution from
    identical states: positions within 1e-6 and velocities within 1e-5 of the
    FP32 oracle, same contact counts."""
    sc = scenes.one_triangle(enable_friction=True, friction_mu=0.5, contact_restitution=0.5)
    o = OracleSim(sc, 32)
    v = o.get("v")
    v[3] = [0.01, 0.0, -0.002]
    v[4] = [0.0, -0.01, 0.002]
    o.set("v", v)
    for k in range(4):
        g = native.GpuSim(sc)
        sync_state(o, g)
        o.substep()
        g.substep()
        assert g.stats()["contacts"] == o.stats()["contacts"]
        assert rel_err(g.get("x"), o.get("x")) < 1e-6, k
        dv = np.max(np.abs(g.get("v").astype(np.float64) - o.get("v")))
        assert dv <= 1e-5 * max(1.0, float(np.max(np.abs(o.get("v"))))), (k, dv)


@pytest.mark.parametrize("name,make,corrected", [
    ("stack", lambda: jittered_stack(), True),
    ("crumpled", lambda: crumpled(), True),
    ("crossing", lambda: crossing_sheets(), True),
    ("one_triangle", lambda: scenes.one_triangle(), False),
])
def test_self_restitution_bit_exact(name, make, corrected):
    """Integrator::apply_contact_restitution with self contacts (PT points, EE
    ends with the second edge's normal negated; sequential per vertex in list
    order): velocities bit-identical to the FP32 oracle."""
    print(f"self restitution {name}", flush=True)
    sc = make()
    sc.params["contact_restitution"] = 0.5
    o = OracleSim(sc, 32)
    o.set_contact_mode(corrected)
    rng = np.random.default_rng(11)
    v0 = rng.normal(scale=0.1, size=(len(sc.verts), 3)).astype(np.float32)
    o.set("v", v0)
    g = native.GpuSim(sc)
    g.set_mode(corrected=corrected)
    sync_state(o, g)
    assert len(o.contacts()) > 0
    o.restitution()
    g.apply_restitution()
    vo = o.get("v").astype(np.float32)
    assert not np.array_equal(vo, v0)
    assert np.array_equal(g.get("v"), vo)
