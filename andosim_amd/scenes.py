"""Deterministic synthetic scenes for the cloth-step workloads.

Every input is generated from a project-written SplitMix64 stream, so the CPU
oracle and the GPU path receive identical float32 arrays.  The cloth grid uses
the reference's triangulation of a quad (v0,v1,v2), (v0,v2,v3)
(demos/demo_utils.h:110-147), laid in the XY plane with +z up because the
workloads use a ground wall with normal (0,0,1).

Workloads:
  drop21   : 21x21 sheet (441 v), two top corners pinned, ground wall, 100 steps.
  sphere129: 129x129 sheet draped over an icosphere obstacle (r=0.25 m), 200 steps.
  stack65x8: 8 sheets of 65x65 in one mesh, stacked 5 mm apart, ground wall, 300 steps.
  batch256 : 256 sheets with side U{17..65}, 1-4 pins, wind N(0,2) N/m^2, 100 steps.
  cyl257   : 257x257 sheet (66,049 v) onto two crossed cylinders (r=0.1 m), 300 steps.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

MASK64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014) with 53-bit uniforms."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, a: float = 0.0, b: float = 1.0) -> float:
        return a + (b - a) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def uniforms(self, count: int, a: float, b: float) -> np.ndarray:
        return np.array([self.uniform(a, b) for _ in range(count)], dtype=np.float64)

    def normal(self) -> float:
        # Box-Muller on two 53-bit uniforms.
        u1 = max(self.uniform(), 1e-300)
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)

    def randint(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1))


def grid(s: int, size: float = 1.0, z: float = 0.3, center=(0.0, 0.0)):
    """s x s vertex sheet of side `size` at height z, quads split (v0,v1,v2),(v0,v2,v3)."""
    i = np.arange(s, dtype=np.float64)
    u = i / (s - 1)
    xs = center[0] + (u - 0.5) * size
    ys = center[1] + (u - 0.5) * size
    X, Y = np.meshgrid(xs, ys, indexing="xy")  # row j = y index, column i = x index
    verts = np.stack([X.ravel(), Y.ravel(), np.full(s * s, z)], axis=1)
    jj, ii = np.meshgrid(np.arange(s - 1), np.arange(s - 1), indexing="ij")
    v0 = (jj * s + ii).ravel()
    v1 = (jj * s + ii + 1).ravel()
    v2 = ((jj + 1) * s + ii + 1).ravel()
    v3 = ((jj + 1) * s + ii).ravel()
    tris = np.empty((2 * v0.size, 3), dtype=np.int32)
    tris[0::2] = np.stack([v0, v1, v2], axis=1)
    tris[1::2] = np.stack([v0, v2, v3], axis=1)
    return verts, tris


def icosphere(level: int, radius: float, center=(0.0, 0.0, 0.0)):
    """Subdivided icosahedron: level 4 gives 2,562 vertices and 5,120 triangles."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.array(verts) * radius + np.array(center)
    return V, np.array(faces, dtype=np.int32)


def cylinder(radius: float, length: float, n_around: int, n_along: int, axis: int, center):
    """Closed cylinder along `axis` (0=x, 1=y), outward-facing triangles, fan caps."""
    ang = 2.0 * math.pi * np.arange(n_around) / n_around
    ts = np.linspace(-0.5 * length, 0.5 * length, n_along + 1)
    rings = []
    for tcoord in ts:
        a = np.cos(ang) * radius
        b = np.sin(ang) * radius
        if axis == 0:
            pts = np.stack([np.full(n_around, tcoord), a, b], axis=1)
        else:
            pts = np.stack([a, np.full(n_around, tcoord), b], axis=1)
        rings.append(pts)
    V = np.concatenate(rings, axis=0)
    tris = []
    for k in range(n_along):
        for j in range(n_around):
            a = k * n_around + j
            b = k * n_around + (j + 1) % n_around
            c = (k + 1) * n_around + (j + 1) % n_around
            d = (k + 1) * n_around + j
            tris += [(a, b, c), (a, c, d)]
    c0 = len(V)
    c1 = c0 + 1
    caps = np.zeros((2, 3))
    caps[0, axis] = ts[0]
    caps[1, axis] = ts[-1]
    V = np.concatenate([V, caps], axis=0)
    last = n_along * n_around
    for j in range(n_around):
        tris.append((c0, (j + 1) % n_around, j))
        tris.append((c1, last + j, last + (j + 1) % n_around))
    V = V + np.array(center, dtype=np.float64)
    return V, np.array(tris, dtype=np.int32)


@dataclass
class Rigid:
    verts: np.ndarray
    tris: np.ndarray
    density: float


@dataclass
class Scene:
    name: str
    verts: np.ndarray          # (n,3) float32
    tris: np.ndarray           # (m,3) int32
    material: dict
    params: dict
    pins: list = field(default_factory=list)    # (vertex, target(3,))
    walls: list = field(default_factory=list)   # (normal(3,), offset, gap)
    rigids: list = field(default_factory=list)  # Rigid
    gravity: np.ndarray = field(default_factory=lambda: np.array([0.0, 0.0, -9.81], dtype=np.float32))
    steps: int = 100


BASE_MATERIAL = dict(youngs_modulus=3e5, poisson_ratio=0.3, density=200.0, thickness=1e-3, bending_stiffness=0.0)


def base_params(**kw) -> dict:
    p = dict(dt=0.01, beta_max=0.25, min_newton_steps=2, max_newton_steps=8, pcg_tol=1e-6, pcg_max_iters=1000,
             contact_gap_max=1e-3, wall_gap=1e-3, enable_ccd=True, enable_friction=False, friction_mu=0.1,
             friction_epsilon=1e-5, velocity_damping=0.0, contact_restitution=0.0, enable_strain_limiting=True,
             strain_limit=0.05, strain_tau=0.05)
    p.update(kw)
    return p


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def drop21(seed: int = 0, s: int = 21) -> Scene:
    rng = SplitMix64(0xA0D0 + seed)
    V, T = grid(s, 1.0, 0.3)
    V[:, 2] += rng.uniforms(len(V), -1e-4, 1e-4)
    V = _f32(V)
    top = [(s - 1) * s, s * s - 1]
    pins = [(i, V[i].copy()) for i in top]
    walls = [(np.array([0, 0, 1], np.float32), 0.0, 1e-3)]
    return Scene("drop21", V, T, dict(BASE_MATERIAL), base_params(), pins, walls, [], steps=100)


def _tilt(V, rng, max_deg):
    th = math.radians(rng.uniform(0.0, max_deg))
    c, s_ = math.cos(th), math.sin(th)
    cen = V.mean(axis=0)
    d = V - cen
    R = np.array([[1, 0, 0], [0, c, -s_], [0, s_, c]])
    return d @ R.T + cen


def sphere129(seed: int = 1, s: int = 129) -> Scene:
    rng = SplitMix64(0xB129 + seed)
    V, T = grid(s, 1.0, 0.30)
    V = _f32(_tilt(V, rng, 15.0))
    SV, ST = icosphere(4, 0.25)
    rig = [Rigid(_f32(SV), ST, 1e12)]
    return Scene("sphere129", V, T, dict(BASE_MATERIAL), base_params(), [], [], rig, steps=200)


def stack65x8(seed: int = 2, s: int = 65, sheets: int = 8) -> Scene:
    rng = SplitMix64(0xC065 + seed)
    Vs, Ts = [], []
    for k in range(sheets):
        V, T = grid(s, 1.0, 0.05 + 0.005 * k)
        th = math.radians(rng.uniform(0.0, 90.0))
        c, s_ = math.cos(th), math.sin(th)
        off = np.array([rng.uniform(-0.1, 0.1), rng.uniform(-0.1, 0.1), 0.0])
        R = np.array([[c, -s_, 0], [s_, c, 0], [0, 0, 1]])
        V = V @ R.T + off
        Ts.append(T + k * s * s)
        Vs.append(V)
    V = _f32(np.concatenate(Vs))
    T = np.ascontiguousarray(np.concatenate(Ts).astype(np.int32))
    walls = [(np.array([0, 0, 1], np.float32), 0.0, 1e-3)]
    return Scene("stack65x8", V, T, dict(BASE_MATERIAL), base_params(), [], walls, [], steps=300)


def batch256(seed: int = 3, count: int = 256) -> list:
    rng = SplitMix64(0xD256 + seed)
    out = []
    for b in range(count):
        s = rng.randint(17, 65)
        V, T = grid(s, 1.0, 0.3)
        V = _f32(V)
        boundary = [i for i in range(s)] + [(s - 1) * s + i for i in range(s)]
        npins = rng.randint(1, 4)
        pins = []
        for _ in range(npins):
            vi = boundary[rng.randint(0, len(boundary) - 1)]
            pins.append((vi, V[vi].copy()))
        wind = np.array([rng.normal() * 2.0, rng.normal() * 2.0, 0.0])
        # Wind pressure w (N/m^2) enters as acceleration w / (rho * thickness).
        acc = wind / (BASE_MATERIAL["density"] * BASE_MATERIAL["thickness"])
        g = _f32(np.array([0.0, 0.0, -9.81]) + acc)
        walls = [(np.array([0, 0, 1], np.float32), 0.0, 1e-3)]
        out.append(Scene(f"batch256[{b}]", V, T, dict(BASE_MATERIAL), base_params(), pins, walls, [], g, 100))
    return out


def cyl257(seed: int = 4, s: int = 257, n_around: int = 48, n_along: int = 24) -> Scene:
    rng = SplitMix64(0xE257 + seed)
    V, T = grid(s, 1.0, 0.35)
    V = _f32(_tilt(V, rng, 5.0))
    A = cylinder(0.1, 1.2, n_around, n_along, 0, (0.0, 0.0, 0.0))
    B = cylinder(0.1, 1.2, n_around, n_along, 1, (0.0, 0.0, 0.21))
    rig = [Rigid(_f32(A[0]), A[1], 1e12), Rigid(_f32(B[0]), B[1], 1e12)]
    return Scene("cyl257", V, T, dict(BASE_MATERIAL), base_params(), [], [], rig, steps=300)


def stack_resting(seed: int = 5, s: int = 65, sheets: int = 8, gap: float = 8e-5, jitter: float = 3e-5,
                  z0: float = 6e-4) -> Scene:
    """Dense self-contact: `sheets` s x s sheets (1 m) stacked `gap` apart in one
    mesh, each shifted U(-1,1) cm and jittered U(-jitter, jitter) in z, the
    lowest z0 above a ground wall.  Starts in contact, so the corrected mode's
    PT/EE barriers act from the first substep (the falling stack's sheets all
    fall at the same speed and, with the explicit gravity pre-step, pass the
    ground together before they touch)."""
    rng = SplitMix64(0x5E57 + seed)
    Vs, Ts = [], []
    for k in range(sheets):
        V, T = grid(s, 1.0, z0 + k * gap)
        V[:, 0] += rng.uniform(-0.01, 0.01)
        V[:, 1] += rng.uniform(-0.01, 0.01)
        V[:, 2] += rng.uniforms(len(V), -jitter, jitter)
        Ts.append(T + k * s * s)
        Vs.append(V)
    V = _f32(np.concatenate(Vs))
    T = np.ascontiguousarray(np.concatenate(Ts).astype(np.int32))
    walls = [(np.array([0, 0, 1], np.float32), 0.0, 1e-3)]
    return Scene("stack_resting", V, T, dict(BASE_MATERIAL), base_params(), [], walls, [], steps=60)


WORKLOADS = {"drop21": drop21, "sphere129": sphere129, "stack65x8": stack65x8, "cyl257": cyl257,
             "stack_resting": stack_resting}


def one_triangle(**params) -> Scene:
    """One triangle in the plane z = 0.3 plus isolated vertices: the mesh shape
    for which the reference's self broad phase is not empty (a single
    primitive-bearing leaf).  Vertices 3 and 4 lie over the interior within the
    1e-4 point box (above and below), vertex 5 over the interior 2 mm away,
    vertex 6 far off."""
    V = np.array([[0, 0, 0.3], [0.1, 0, 0.3], [0, 0.1, 0.3],
                  [0.02, 0.03, 0.3 + 5e-5], [0.03, 0.02, 0.3 - 4e-5], [0.025, 0.025, 0.302], [0.5, 0.5, 0.3]],
                 np.float32)
    T = np.array([[0, 1, 2]], np.int32)
    return Scene("one_triangle", V, T, dict(BASE_MATERIAL), base_params(**params), [], [], [])


def small_sheet(s: int = 9, seed: int = 7, pins: int = 2, wall: bool = True, rigid: bool = False,
                strain: bool = True, z: float = 0.3) -> Scene:
    """Tiny scenes for fast parity tests (same material and parameters as drop21)."""
    rng = SplitMix64(0xF000 + seed)
    V, T = grid(s, 0.4, z)
    V[:, 2] += rng.uniforms(len(V), -1e-4, 1e-4)
    V = _f32(V)
    pl = [(i, V[i].copy()) for i in ([(s - 1) * s, s * s - 1][:pins])]
    wl = [(np.array([0, 0, 1], np.float32), 0.0, 1e-3)] if wall else []
    rg = []
    if rigid:
        SV, ST = icosphere(2, 0.1, (0.0, 0.0, z - 0.105))
        rg = [Rigid(_f32(SV), ST, 1e12)]
    return Scene(f"small{s}", V, T, dict(BASE_MATERIAL), base_params(enable_strain_limiting=strain), pl, wl, rg)
