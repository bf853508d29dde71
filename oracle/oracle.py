"""ctypes wrapper over the CPU oracle (oracle/liborc.so) — TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, and only as the checker or the CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")

PARAM_ORDER = ["dt", "beta_max", "min_newton_steps", "max_newton_steps", "pcg_tol", "pcg_max_iters",
               "contact_gap_max", "wall_gap", "enable_ccd", "enable_friction", "friction_mu", "friction_epsilon",
               "velocity_damping", "contact_restitution", "enable_strain_limiting", "strain_limit", "strain_tau"]
MATERIAL_ORDER = ["youngs_modulus", "poisson_ratio", "density", "thickness", "bending_stiffness"]


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", HERE, "liborc.so"], check=True,
                       stdout=subprocess.DEVNULL)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_int]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_last_error.restype = C.c_char_p
        L.orc_set_mesh.argtypes = [C.c_void_p, C.c_int, _dp, C.c_int, _ip, _dp]
        L.orc_init_state.argtypes = [C.c_void_p]
        L.orc_set_params.argtypes = [C.c_void_p, _dp]
        L.orc_add_pin.argtypes = [C.c_void_p, C.c_int, _dp]
        L.orc_add_wall.argtypes = [C.c_void_p, _dp, C.c_double, C.c_double]
        L.orc_add_rigid.argtypes = [C.c_void_p, C.c_int, _dp, C.c_int, _ip, C.c_double]
        L.orc_apply_gravity.argtypes = [C.c_void_p, _dp, C.c_double]
        L.orc_step.argtypes = [C.c_void_p]
        L.orc_num_vertices.argtypes = [C.c_void_p]
        L.orc_num_blocks.argtypes = [C.c_void_p]
        L.orc_get.argtypes = [C.c_void_p, C.c_int, _dp]
        L.orc_set.argtypes = [C.c_void_p, C.c_int, _dp]
        L.orc_stats.argtypes = [C.c_void_p, _dp]
        L.orc_pcg_per_solve.argtypes = [C.c_void_p, _ip, C.c_int]
        L.orc_rigid_state.argtypes = [C.c_void_p, C.c_int, _dp]
        L.orc_set_rigid_state.argtypes = [C.c_void_p, C.c_int, _dp]
        L.orc_rigid_integrate.argtypes = [C.c_void_p, C.c_int, _dp, _dp, C.c_double]
        L.orc_pattern.argtypes = [C.c_void_p, _ip, _ip]
        L.orc_linearize.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.orc_pcg.argtypes = [C.c_int, C.c_int, _ip, _ip, _dp, _dp, _dp, C.c_double, C.c_int, C.POINTER(C.c_int)]
        L.orc_contacts.argtypes = [C.c_void_p, _dp, C.c_int]
        L.orc_barrier.restype = C.c_double
        L.orc_barrier.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double]
        L.orc_narrow_pt.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.orc_narrow_ee.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.orc_enforce_spd.argtypes = [C.c_int, _dp]
        L.orc_pin_stiffness.restype = C.c_double
        L.orc_pin_stiffness.argtypes = [C.c_int, C.c_double, _dp, _dp, C.c_double]
        L.orc_svd32.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
        L.orc_line_search.restype = C.c_double
        L.orc_line_search.argtypes = [C.c_void_p, _dp, C.c_double, C.c_void_p]
        L.orc_set_contacts_pt.argtypes = [C.c_void_p, C.c_int, _ip]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_set_contact_mode.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double]
        L.orc_count_intersections.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_get_threads.restype = C.c_int
        L.orc_energy.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_set_ccd_mode.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_physics.argtypes = [C.c_void_p, C.c_int]
        L.orc_restitution.argtypes = [C.c_void_p]
        L.orc_set_capture.argtypes = [C.c_void_p, C.c_int]
        L.orc_get_capture.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_additive_ccd.restype = C.c_double
        L.orc_additive_ccd.argtypes = [C.c_int, C.c_int, _dp, _dp]
        L.orc_accd_sep.restype = C.c_double
        L.orc_accd_sep.argtypes = [C.c_int, _dp, _dp]
        L.orc_contact_metrics.argtypes = [C.c_int, C.c_int, _ip, _ip, _dp, _dp, _dp, C.c_double, C.c_int, _dp]
        _lib = L
    return _lib


def set_threads(k: int) -> None:
    """Threads for the oracle's element-wise loops; results do not depend on it."""
    lib().orc_set_threads(int(k))


def get_threads() -> int:
    return int(lib().orc_get_threads())


def _d(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _chk(rc):
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())


ENERGY_KEYS = ["kinetic_energy", "elastic_energy", "linear_momentum", "angular_momentum", "max_velocity"]
METRIC_KEYS = ["num_point_triangle", "num_edge_edge", "num_wall", "num_total_contacts", "num_penetrations",
               "num_ccd_contacts", "num_broad_phase_contacts", "num_relative_velocities", "has_major_penetration",
               "has_tunneling", "min_gap", "max_gap", "avg_gap", "max_penetration", "avg_penetration",
               "ccd_effectiveness", "max_relative_velocity", "avg_relative_velocity", "sum_gap",
               "sum_penetration", "sum_relative_velocity"]


def _energy_dict(a):
    return {"kinetic_energy": a[0], "elastic_energy": a[1], "linear_momentum": np.array(a[2:5]),
            "angular_momentum": np.array(a[5:8]), "max_velocity": a[8]}


def additive_ccd(kind: str, x0, p, precision=32) -> float:
    """Additive CCD of one PT ("pt": point, then triangle) or EE ("ee") pair
    moving by p over [0, 1]: 1 if the motion is free, else a conservative
    time of impact."""
    return float(lib().orc_additive_ccd(int(precision), 0 if kind == "pt" else 1, _d(x0).reshape(-1),
                                        _d(p).reshape(-1)))


def accd_sep(kind: str, x0, p) -> float:
    """The corrected physics line search's time of impact (FP64, keeping 32
    FP32 ulps of the pair's largest coordinate, or half the initial gap when
    that is smaller) of one PT ("pt") or EE ("ee") pair moving by p."""
    return float(lib().orc_accd_sep(0 if kind == "pt" else 1, _d(x0).reshape(-1), _d(p).reshape(-1)))


def contact_metrics(types, idx, gaps, normals, velocities, gap_max, ccd, precision=32):
    """CollisionValidator::compute_metrics on a contact list (see orc_contact_metrics)."""
    t = np.ascontiguousarray(np.asarray(types, np.int32).reshape(-1))
    k = len(t)
    ix = np.ascontiguousarray(np.asarray(idx, np.int32).reshape(k, 4)) if k else np.zeros((1, 4), np.int32)
    g = _d(gaps).reshape(-1) if k else np.zeros(1)
    nr = _d(normals).reshape(-1) if k else np.zeros(3)
    vel = _d(velocities).reshape(-1)
    out = np.zeros(len(METRIC_KEYS))
    _chk(lib().orc_contact_metrics(int(precision), k, t if k else np.zeros(1, np.int32), ix.ravel(), g, nr, vel,
                                   float(gap_max), int(bool(ccd)), out))
    d = dict(zip(METRIC_KEYS, out.tolist()))
    for key in METRIC_KEYS[:10]:
        d[key] = int(d[key])
    return d


class OracleSim:
    """One oracle simulation: mesh + state + constraints + params (+ rigid bodies)."""

    def __init__(self, scene=None, precision: int = 32):
        self.L = lib()
        self.h = self.L.orc_create(precision)
        self.prec = precision
        if scene is not None:
            self.load(scene)

    def __del__(self):
        try:
            self.L.orc_destroy(self.h)
        except Exception:
            pass

    def load(self, sc):
        L = self.L
        mat = _d([sc.material[k] for k in MATERIAL_ORDER])
        _chk(L.orc_set_mesh(self.h, len(sc.verts), _d(sc.verts).ravel(), len(sc.tris),
                            np.ascontiguousarray(sc.tris, dtype=np.int32).ravel(), mat))
        _chk(L.orc_init_state(self.h))
        self.set_params(sc.params)
        for vi, t in sc.pins:
            _chk(L.orc_add_pin(self.h, int(vi), _d(t)))
        for nrm, off, gap in sc.walls:
            _chk(L.orc_add_wall(self.h, _d(nrm), float(off), float(gap)))
        for r in sc.rigids:
            _chk(L.orc_add_rigid(self.h, len(r.verts), _d(r.verts).ravel(), len(r.tris),
                                 np.ascontiguousarray(r.tris, dtype=np.int32).ravel(), float(r.density)))
        self.n = L.orc_num_vertices(self.h)
        self.gravity = np.asarray(sc.gravity, dtype=np.float64)
        self.dt = float(sc.params["dt"])

    def set_params(self, p: dict):
        _chk(self.L.orc_set_params(self.h, _d([float(p[k]) for k in PARAM_ORDER])))

    def apply_gravity(self, g=None, dt=None):
        _chk(self.L.orc_apply_gravity(self.h, _d(self.gravity if g is None else g), self.dt if dt is None else dt))

    def step(self):
        _chk(self.L.orc_step(self.h))

    def substep(self):
        """One add-on substep: apply_gravity then Integrator.step (operators.py:900-905)."""
        self.apply_gravity()
        self.step()

    def get(self, which: str) -> np.ndarray:
        sel = {"x": 0, "v": 1, "m": 2, "mesh": 3, "egrad": 4}[which]
        out = np.zeros(self.n if sel == 2 else 3 * self.n)
        _chk(self.L.orc_get(self.h, sel, out))
        return out if sel == 2 else out.reshape(-1, 3)

    def set(self, which: str, arr):
        _chk(self.L.orc_set(self.h, {"x": 0, "v": 1, "m": 2, "mesh": 3, "xprev": 4}[which], _d(arr).ravel()))

    def set_contact_mode(self, corrected: bool, pt_box: float = 1e-4, keep_gap: float = 0.01):
        """Faithful (self broad phase empty) or corrected (brute-force PT/EE with
        the intended predicate, both barriers in the energy)."""
        self.L.orc_set_contact_mode(self.h, 1 if corrected else 0, pt_box, keep_gap)

    def capture_solve(self, k: int):
        """Keep the system of the k-th PCG solve of the following steps."""
        _chk(self.L.orc_set_capture(self.h, int(k)))

    def captured(self):
        """(BSR values (nb, 3, 3), rhs) of the captured solve, or None."""
        nb = self.L.orc_num_blocks(self.h)
        H = np.zeros(9 * nb)
        b = np.zeros(3 * self.n)
        if not self.L.orc_get_capture(self.h, H, b):
            return None
        return H.reshape(nb, 3, 3), b

    def set_physics(self, corrected: bool):
        """Corrected physics (implicit gravity, membrane force from the current
        positions, attracting pins, per-iteration contact barriers over the
        step's swept-box candidates, CCD line search); implies the corrected
        contact mode."""
        _chk(self.L.orc_set_physics(self.h, 1 if corrected else 0))

    def set_ccd_mode(self, additive: bool):
        """Corrected mode line search: additive CCD (True) or the reference's
        10 distance samples (False)."""
        _chk(self.L.orc_set_ccd_mode(self.h, 1 if additive else 0))

    def energy(self, x, v, m):
        """EnergyTracker::compute on (x, v, m) with this mesh: (reference
        accumulation, the same per-item terms summed in double)."""
        out = np.zeros(18)
        _chk(self.L.orc_energy(self.h, _d(x).ravel(), _d(v).ravel(), _d(m).ravel(), out))
        return _energy_dict(out[:9]), _energy_dict(out[9:])

    def count_intersections(self):
        """Brute-force (edge-triangle crossings, vertices inside rigid bodies)."""
        et, ins = C.c_int(0), C.c_int(0)
        self.L.orc_count_intersections(self.h, C.byref(et), C.byref(ins))
        return et.value, ins.value

    def stats(self) -> dict:
        out = np.zeros(11)
        self.L.orc_stats(self.h, out)
        keys = ["newton_iters", "pcg_iters", "pcg_solves", "pcg_failures", "line_search_failures", "contacts",
                "rigid_contacts", "strain_active", "beta", "asym_count", "spd_general_calls"]
        d = dict(zip(keys, out.tolist()))
        buf = np.zeros(4096, dtype=np.int32)
        cnt = self.L.orc_pcg_per_solve(self.h, buf, len(buf))
        d["pcg_per_solve"] = buf[:min(cnt, len(buf))].tolist()
        return d

    def rigid_state(self, rb: int) -> dict:
        out = np.zeros(19)
        _chk(self.L.orc_rigid_state(self.h, rb, out))
        return dict(position=out[0:3], linear_velocity=out[3:6], angular_velocity=out[6:9],
                    rotation=out[9:18].reshape(3, 3), mass=out[18])

    def set_rigid_pose(self, rb: int, rs: dict):
        st = np.concatenate([np.asarray(rs["position"], np.float64), np.asarray(rs["linear_velocity"], np.float64),
                             np.asarray(rs["angular_velocity"], np.float64),
                             np.asarray(rs["rotation"], np.float64).ravel()])
        _chk(self.L.orc_set_rigid_state(self.h, rb, _d(st)))

    def rigid_integrate(self, rb: int, force, torque, dt: float):
        """RigidBody::integrate of body rb with the given force and torque."""
        _chk(self.L.orc_rigid_integrate(self.h, rb, _d(force), _d(torque), float(dt)))

    def pattern(self):
        nb = self.L.orc_num_blocks(self.h)
        rp = np.zeros(self.n + 1, dtype=np.int32)
        ci = np.zeros(nb, dtype=np.int32)
        self.L.orc_pattern(self.h, rp, ci)
        return rp, ci

    def linearize(self, xhat):
        nb = self.L.orc_num_blocks(self.h)
        g = np.zeros(3 * self.n)
        H = np.zeros(9 * nb)
        _chk(self.L.orc_linearize(self.h, _d(xhat).ravel(), g, H))
        return g, H.reshape(nb, 3, 3)

    def restitution(self):
        """Integrator::apply_contact_restitution on the current state."""
        _chk(self.L.orc_restitution(self.h))

    def contacts(self) -> np.ndarray:
        cap = 1 << 16
        buf = np.zeros(17 * cap)
        cnt = self.L.orc_contacts(self.h, buf, cap)
        if cnt > cap:
            buf = np.zeros(17 * cnt)
            cnt = self.L.orc_contacts(self.h, buf, cnt)
        return buf[:17 * cnt].reshape(cnt, 17)

    def line_search(self, d, ext=1.25) -> float:
        return self.L.orc_line_search(self.h, _d(d).ravel(), ext, None)

    def line_search_held(self, d):
        """Corrected physics: (alpha, the direction after the line search held
        the vertices of pairs within their separation that would not move apart)."""
        out = np.zeros(3 * self.n)
        a = self.L.orc_line_search(self.h, _d(d).ravel(), 1.25, out.ctypes.data_as(C.c_void_p))
        return a, out.reshape(-1, 3)

    def set_pt_contacts(self, idx4):
        idx = np.ascontiguousarray(np.asarray(idx4, dtype=np.int32).reshape(-1, 4))
        self.L.orc_set_contacts_pt(self.h, len(idx), idx.ravel())


def pcg(rp, ci, vals, b, x0=None, tol=1e-6, maxit=1000, precision=32):
    """Block-Jacobi PCG on a BSR matrix; returns (x, iterations, converged)."""
    n = len(rp) - 1
    x = np.zeros(3 * n) if x0 is None else _d(x0).ravel().copy()
    it = C.c_int(0)
    _chk(lib().orc_pcg(precision, n, np.ascontiguousarray(rp, dtype=np.int32),
                       np.ascontiguousarray(ci, dtype=np.int32), _d(vals).ravel(), _d(b).ravel(), x, tol, maxit,
                       C.byref(it)))
    iters = it.value
    return x, (iters if iters >= 0 else -iters - 1), iters >= 0


def barrier(which: str, g, gmax, k, precision=32) -> float:
    return lib().orc_barrier(precision, {"energy": 0, "gradient": 1, "hessian": 2}[which], g, gmax, k)


def narrow_pt(p, a, b, c, precision=32):
    out = np.zeros(7)
    lib().orc_narrow_pt(precision, _d(p), _d(a), _d(b), _d(c), out)
    return out[0], out[1:4], out[4:7]


def narrow_ee(p0, p1, q0, q1, precision=32):
    out = np.zeros(10)
    lib().orc_narrow_ee(precision, _d(p0), _d(p1), _d(q0), _d(q1), out)
    return out[0], out[1:4], out[4:7], out[7:10]


def enforce_spd(H, precision=32):
    h = _d(H).ravel().copy()
    lib().orc_enforce_spd(precision, h)
    return h.reshape(3, 3)


def pin_stiffness(mass, offset, H, min_gap, precision=32):
    return lib().orc_pin_stiffness(precision, mass, _d(offset), _d(H).ravel(), min_gap)


def svd32(F, precision=32):
    U = np.zeros(6)
    s = np.zeros(2)
    V = np.zeros(4)
    ok = lib().orc_svd32(precision, _d(F).ravel(), U, s, V)
    return U.reshape(3, 2), s, V.reshape(2, 2), bool(ok)
