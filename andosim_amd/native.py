"""ctypes binding of the C ABI in include/andosim_hip.h (libandosim_hip.so).

`GpuSim` drives one cloth through the ABI the way the Blender add-on drives
`ando_barrier_core` (operators.py:900-905): apply_gravity, then step.
Loading fails loudly when the HIP library is missing; nothing here falls back
to a CPU implementation.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libandosim_hip.so")

ABI_SYMBOLS = [
    "ando_abi_version", "ando_last_error", "ando_create", "ando_destroy", "ando_set_mesh", "ando_set_mesh_vertices",
    "ando_lumped_masses", "ando_set_previous_positions",
    "ando_set_state", "ando_get_state", "ando_set_pins", "ando_set_walls", "ando_set_mode", "ando_apply_gravity",
    "ando_step",
    "ando_get_stats", "ando_compute_contacts", "ando_count_intersections", "ando_get_pattern", "ando_linearize",
    "ando_pcg", "ando_apply_restitution",
    "ando_nn_num_weights", "ando_set_precond_weights", "ando_trainer_create", "ando_trainer_destroy",
    "ando_trainer_add_system", "ando_trainer_init_weights", "ando_trainer_set_weights", "ando_trainer_get_weights",
    "ando_trainer_forward", "ando_trainer_loss_grad", "ando_trainer_train", "ando_trainer_solve",
    "ando_energy_diagnostics", "ando_contact_metrics", "ando_cfl_inputs", "ando_step_batch",
    "ando_host_alloc", "ando_host_free", "ando_pcg_noalloc",
]


class Material(C.Structure):
    _fields_ = [("youngs_modulus", C.c_float), ("poisson_ratio", C.c_float), ("density", C.c_float),
                ("thickness", C.c_float), ("bending_stiffness", C.c_float)]


class Params(C.Structure):
    _fields_ = [("dt", C.c_float), ("beta_max", C.c_float), ("min_newton_steps", C.c_int32),
                ("max_newton_steps", C.c_int32), ("friction_min_newton_steps", C.c_int32), ("pcg_tol", C.c_float),
                ("pcg_max_iters", C.c_int32), ("contact_gap_max", C.c_float), ("wall_gap", C.c_float),
                ("enable_ccd", C.c_int32), ("contact_normal_epsilon", C.c_float), ("barrier_tolerance", C.c_float),
                ("enable_friction", C.c_int32), ("friction_mu", C.c_float), ("friction_epsilon", C.c_float),
                ("friction_tangent_threshold", C.c_float), ("velocity_damping", C.c_float),
                ("contact_restitution", C.c_float), ("enable_strain_limiting", C.c_int32),
                ("strain_limit", C.c_float), ("strain_tau", C.c_float), ("strain_epsilon", C.c_float),
                ("strain_svd_epsilon", C.c_float), ("hessian_epsilon", C.c_float), ("min_gap", C.c_float)]


class RigidBody(C.Structure):
    _fields_ = [("num_vertices", C.c_int32), ("num_triangles", C.c_int32),
                ("local_vertices", C.POINTER(C.c_float)), ("triangles", C.POINTER(C.c_int32)),
                ("mass", C.c_float), ("inertia_body_inv", C.c_float * 9), ("position", C.c_float * 3),
                ("rotation", C.c_float * 9), ("linear_velocity", C.c_float * 3),
                ("angular_velocity", C.c_float * 3)]


class Contact(C.Structure):
    _fields_ = [("type", C.c_int32), ("idx", C.c_int32 * 4), ("rigid_body_index", C.c_int32),
                ("vertex_count", C.c_int32), ("gap", C.c_float), ("normal", C.c_float * 3),
                ("witness_p", C.c_float * 3), ("witness_q", C.c_float * 3), ("weights", C.c_float * 4)]


class Energy(C.Structure):
    _fields_ = [("kinetic_energy", C.c_double), ("elastic_energy", C.c_double),
                ("linear_momentum", C.c_double * 3), ("angular_momentum", C.c_double * 3),
                ("max_velocity", C.c_double)]


class CollisionMetrics(C.Structure):
    _fields_ = [(k, C.c_int32) for k in (
        "num_point_triangle", "num_edge_edge", "num_wall", "num_total_contacts", "num_penetrations",
        "num_ccd_contacts", "num_broad_phase_contacts", "num_relative_velocities", "has_major_penetration",
        "has_tunneling")] + [(k, C.c_float) for k in (
            "min_gap", "max_gap", "avg_gap", "max_penetration", "avg_penetration", "ccd_effectiveness",
            "max_relative_velocity", "avg_relative_velocity")] + [(k, C.c_double) for k in (
                "sum_gap", "sum_penetration", "sum_relative_velocity")]


MAX_RECORDED = 64


class Stats(C.Structure):
    _fields_ = [("newton_iterations", C.c_int32), ("pcg_iterations", C.c_int32), ("pcg_solves", C.c_int32),
                ("pcg_failures", C.c_int32), ("line_search_failures", C.c_int32), ("contacts", C.c_int32),
                ("rigid_contacts", C.c_int32), ("strain_active", C.c_int32), ("beta", C.c_float),
                ("last_alpha", C.c_float), ("pcg_per_solve", C.c_int32 * MAX_RECORDED), ("step_ms", C.c_double),
                ("timed_launches", C.c_int32 * 2), ("timed_ms", C.c_double * 2), ("timed_iterations", C.c_int32),
                ("pcg_kernel", C.c_int32), ("learned_solves", C.c_int32), ("learned_fallbacks", C.c_int32),
                ("fallback_iterations", C.c_int32), ("spmv_lanes", C.c_int32),
                ("spmv_wg_slices", C.c_int32), ("batched_solves", C.c_int32), ("pcg_vec_block", C.c_int32),
                ("pcg_autotuned", C.c_int32), ("long_hit_sorts", C.c_int32), ("max_candidates", C.c_int32),
                ("held_iterations", C.c_int32), ("capped_trials", C.c_int32),
                ("timed_coupling_bytes", C.c_double)]


_lib = None
# Every live context and trainer, released by _close_all at interpreter exit
# (before module teardown), so no device object outlives the interpreter.
_live = weakref.WeakSet()


def _close_all():
    for obj in list(_live):
        obj.close()


def load():
    """Load libandosim_hip.so; raises if it is missing (no fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"HIP library not built: {LIB_PATH} (run python -m andosim_amd.build)")
    L = C.CDLL(LIB_PATH)
    fp, ip = C.POINTER(C.c_float), C.POINTER(C.c_int32)
    L.ando_last_error.restype = C.c_char_p
    L.ando_abi_version.restype = C.c_int32
    L.ando_create.argtypes = [C.c_int32, C.POINTER(C.c_void_p)]
    L.ando_destroy.argtypes = [C.c_void_p]
    L.ando_set_mesh.argtypes = [C.c_void_p, fp, C.c_int32, ip, C.c_int32, C.POINTER(Material)]
    L.ando_set_mesh_vertices.argtypes = [C.c_void_p, fp, C.c_int32]
    L.ando_lumped_masses.argtypes = [fp, C.c_int32, ip, C.c_int32, C.POINTER(Material), fp]
    L.ando_set_state.argtypes = [C.c_void_p, fp, fp, fp, C.c_int32]
    L.ando_get_state.argtypes = [C.c_void_p, fp, fp, C.c_int32]
    L.ando_host_alloc.argtypes = [C.c_size_t, C.POINTER(C.c_void_p)]
    L.ando_host_free.argtypes = [C.c_void_p]
    L.ando_set_previous_positions.argtypes = [C.c_void_p, fp, C.c_int32]
    L.ando_set_pins.argtypes = [C.c_void_p, ip, fp, C.c_int32]
    L.ando_set_walls.argtypes = [C.c_void_p, fp, fp, fp, C.c_int32]
    L.ando_apply_gravity.argtypes = [C.c_void_p, fp, C.c_float]
    L.ando_set_mode.argtypes = [C.c_void_p, C.c_int32]
    L.ando_step.argtypes = [C.c_void_p, C.POINTER(Params), C.POINTER(RigidBody), C.c_int32]
    L.ando_get_stats.argtypes = [C.c_void_p, C.POINTER(Stats)]
    L.ando_compute_contacts.argtypes = [C.c_void_p, C.POINTER(RigidBody), C.c_int32, C.POINTER(Contact), C.c_int32,
                                        ip]
    L.ando_get_pattern.argtypes = [C.c_void_p, ip, ip, ip]
    L.ando_count_intersections.argtypes = [C.c_void_p, C.POINTER(RigidBody), C.c_int32, ip, ip]
    L.ando_linearize.argtypes = [C.c_void_p, C.POINTER(Params), fp, C.POINTER(RigidBody), C.c_int32, fp, fp]
    L.ando_apply_restitution.argtypes = [C.c_void_p, C.POINTER(Params), C.POINTER(RigidBody), C.c_int32]
    L.ando_pcg.argtypes = [C.c_void_p, C.c_int32, ip, ip, fp, fp, fp, C.c_float, C.c_int32, ip, ip]
    L.ando_pcg_noalloc.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_float, C.c_int32, C.c_void_p, ip, ip]
    dp = C.POINTER(C.c_double)
    L.ando_nn_num_weights.restype = C.c_int32
    L.ando_set_precond_weights.argtypes = [C.c_void_p, fp, C.c_int32]
    L.ando_trainer_create.argtypes = [C.c_int32, C.POINTER(C.c_void_p)]
    L.ando_trainer_destroy.argtypes = [C.c_void_p]
    L.ando_trainer_add_system.argtypes = [C.c_void_p, C.c_int32, ip, ip, fp, fp, ip]
    L.ando_trainer_init_weights.argtypes = [C.c_void_p, C.c_uint64]
    L.ando_trainer_set_weights.argtypes = [C.c_void_p, fp, C.c_int32]
    L.ando_trainer_get_weights.argtypes = [C.c_void_p, fp, C.c_int32]
    L.ando_trainer_forward.argtypes = [C.c_void_p, C.c_int32, fp, fp]
    L.ando_trainer_loss_grad.argtypes = [C.c_void_p, C.c_int32, dp, fp]
    L.ando_trainer_train.argtypes = [C.c_void_p, C.c_int32, C.c_float, C.c_uint64, dp]
    L.ando_trainer_solve.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_float, C.c_int32, ip, ip]
    L.ando_energy_diagnostics.argtypes = [C.c_void_p, C.POINTER(Energy)]
    L.ando_contact_metrics.argtypes = [C.c_void_p, C.POINTER(Contact), C.c_int32, C.c_float, C.c_int32,
                                       C.POINTER(CollisionMetrics)]
    L.ando_cfl_inputs.argtypes = [C.c_void_p, fp, fp]
    L.ando_step_batch.argtypes = [C.POINTER(C.c_void_p), C.POINTER(Params), C.POINTER(C.POINTER(RigidBody)),
                                  C.POINTER(C.c_int32), C.c_int32, C.c_int32]
    _lib = L
    atexit.register(_close_all)
    return L


def _f(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def _i(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def check(rc):
    if rc != 0:
        raise RuntimeError(f"andosim_hip error {rc}: {load().ando_last_error().decode()}")


MATERIAL_KEYS = ["youngs_modulus", "poisson_ratio", "density", "thickness", "bending_stiffness"]


def material_struct(mat: dict) -> Material:
    return Material(**{k: float(mat[k]) for k in MATERIAL_KEYS})


def lumped_masses(verts, tris, mat: dict) -> np.ndarray:
    """State::compute_lumped_masses through the library's host helper."""
    V, T = _f(verts), _i(tris)
    out = np.zeros(len(V), np.float32)
    m = material_struct(mat)
    check(load().ando_lumped_masses(_fp(V), len(V), _ip(T), len(T), C.byref(m), _fp(out)))
    return out


def normalized3(v) -> np.ndarray:
    """Vec3::normalized in float32 with the reference's sum order x*x + (y*y + z*z)."""
    a = np.asarray(v, dtype=np.float32)
    sq = np.float32(a[0] * a[0] + np.float32(a[1] * a[1] + a[2] * a[2]))
    return (a / np.sqrt(sq)).astype(np.float32) if sq > 0 else a.copy()


def make_params(p: dict) -> Params:
    """SimParams with the reference defaults (types.h:60-100) overridden by `p`."""
    q = Params()
    q.dt = 0.002
    q.beta_max = 0.25
    q.min_newton_steps = 2
    q.max_newton_steps = 8
    q.friction_min_newton_steps = 32
    q.pcg_tol = 1e-3
    q.pcg_max_iters = 1000
    q.contact_gap_max = 0.001
    q.wall_gap = 0.001
    q.enable_ccd = 1
    q.contact_normal_epsilon = 1e-8
    q.barrier_tolerance = 1e-12
    q.enable_friction = 0
    q.friction_mu = 0.1
    q.friction_epsilon = 1e-5
    q.friction_tangent_threshold = 1e-6
    q.velocity_damping = 0.0
    q.contact_restitution = 0.0
    q.enable_strain_limiting = 0
    q.strain_limit = 0.05
    q.strain_tau = 0.05
    q.strain_epsilon = 0.0
    q.strain_svd_epsilon = 1e-6
    q.hessian_epsilon = 1e-8
    q.min_gap = 1e-8
    for k, v in p.items():
        setattr(q, k, int(v) if isinstance(getattr(q, k), int) else float(v))
    return q


class RigidHandle:
    """Host copy of one rigid body (rigid_body.cpp:20-92) plus its ABI struct."""

    def __init__(self, verts, tris, density):
        V = np.asarray(verts, dtype=np.float32)
        cen = np.zeros(3, np.float32)
        for v in V:  # centroid summed in order, then divided (rigid_body.cpp:30-35)
            cen = (cen + v).astype(np.float32)
        cen = (cen / np.float32(len(V))).astype(np.float32)
        self.local = _f(V - cen)
        self.tris = _i(tris)
        world = self.local + cen
        mn, mx = world.min(axis=0), world.max(axis=0)
        ext = (mx - mn).astype(np.float32)
        vol = np.float32(max(ext[0], np.float32(1e-6))) * np.float32(max(ext[1], np.float32(1e-6))) * \
            np.float32(max(ext[2], np.float32(1e-6)))
        mass = np.float32(max(np.float32(density) * np.float32(vol), np.float32(1e-3)))
        x2, y2, z2 = ext * ext
        I = np.array([(mass / np.float32(12)) * (y2 + z2), (mass / np.float32(12)) * (x2 + z2),
                      (mass / np.float32(12)) * (x2 + y2)], dtype=np.float32)
        I = np.maximum(I, np.float32(1e-6))
        s = RigidBody()
        s.num_vertices = len(self.local)
        s.num_triangles = len(self.tris)
        s.local_vertices = _fp(self.local)
        s.triangles = _ip(self.tris)
        s.mass = float(mass)
        for k in range(9):
            s.inertia_body_inv[k] = 0.0
            s.rotation[k] = 1.0 if k % 4 == 0 else 0.0
        for k in range(3):
            s.inertia_body_inv[4 * k] = float(np.float32(1.0) / I[k])
            s.position[k] = float(cen[k])
            s.linear_velocity[k] = 0.0
            s.angular_velocity[k] = 0.0
        self.struct = s


class GpuSim:
    """One cloth resident on the GPU, driven through the C ABI."""

    def __init__(self, scene=None, device: int = 0, fp64_near: bool = True, kernel_timing: bool = False):
        self.L = load()
        self.h = None
        h = C.c_void_p()
        check(self.L.ando_create(device, C.byref(h)))
        self.h = h
        _live.add(self)
        self.bodies = None
        self.nbodies = 0
        self.set_mode(fp64_near, kernel_timing)
        if scene is not None:
            self.load(scene)

    def close(self):
        """Release the device context (idempotent)."""
        h, self.h = getattr(self, "h", None), None
        if h is not None:
            self.L.ando_destroy(h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load(self, sc, masses=None):
        V = _f(sc.verts)
        T = _i(sc.tris)
        mat = material_struct(sc.material)
        check(self.L.ando_set_mesh(self.h, _fp(V), len(V), _ip(T), len(T), C.byref(mat)))
        self.n = len(V)
        self.masses = lumped_masses(V, T, sc.material) if masses is None else _f(masses)
        zeros = np.zeros_like(V)
        check(self.L.ando_set_state(self.h, _fp(V), _fp(zeros), _fp(self.masses), self.n))
        if sc.pins:
            pv = _i([p[0] for p in sc.pins])
            pt = _f(np.stack([p[1] for p in sc.pins]))
            check(self.L.ando_set_pins(self.h, _ip(pv), _fp(pt), len(pv)))
        if sc.walls:
            wn = _f(np.stack([normalized3(w[0]) for w in sc.walls]))
            wo = _f([w[1] for w in sc.walls])
            wg = _f([w[2] for w in sc.walls])
            check(self.L.ando_set_walls(self.h, _fp(wn), _fp(wo), _fp(wg), len(wo)))
        self.rigid = [RigidHandle(r.verts, r.tris, r.density) for r in sc.rigids]
        self.nbodies = len(self.rigid)
        if self.nbodies:
            self.bodies = (RigidBody * self.nbodies)(*[r.struct for r in self.rigid])
        self.params = make_params(sc.params)
        self.gravity = _f(sc.gravity)
        self.dt = float(sc.params["dt"])

    def set_mode(self, fp64_near: bool = True, kernel_timing: bool = False,
                 learned: bool = False, warm_start_only: bool = False, corrected: bool = False,
                 row_spmv: bool = False, static_rigids: bool = False, additive_ccd: bool = False,
                 corrected_physics: bool = False, two_level: bool = False):
        flags = ((1 if fp64_near else 0) | (2 if kernel_timing else 0) |
                 (8 if learned else 0) | (16 if learned and warm_start_only else 0) | (32 if corrected else 0) |
                 (64 if row_spmv else 0) | (128 if static_rigids else 0) | (256 if additive_ccd else 0) |
                 (512 if corrected_physics else 0) | (1024 if two_level else 0))
        check(self.L.ando_set_mode(self.h, flags))

    def apply_gravity(self, g=None, dt=None):
        g = self.gravity if g is None else _f(g)
        check(self.L.ando_apply_gravity(self.h, _fp(g), self.dt if dt is None else dt))

    def step(self):
        check(self.L.ando_step(self.h, C.byref(self.params), self.bodies, self.nbodies))

    def substep(self):
        self.apply_gravity()
        self.step()

    def get(self, which: str) -> np.ndarray:
        x = np.zeros((self.n, 3), np.float32)
        v = np.zeros((self.n, 3), np.float32)
        check(self.L.ando_get_state(self.h, _fp(x), _fp(v), self.n))
        return {"x": x, "v": v}[which]

    def set_state(self, x=None, v=None):
        check(self.L.ando_set_state(self.h, _fp(_f(x)) if x is not None else None,
                                    _fp(_f(v)) if v is not None else None, None, self.n))

    def stats(self) -> dict:
        s = Stats()
        check(self.L.ando_get_stats(self.h, C.byref(s)))
        d = {k: getattr(s, k) for k, _ in Stats._fields_ if k not in ("pcg_per_solve", "timed_launches", "timed_ms")}
        d["timed_launches"] = list(s.timed_launches)
        d["timed_ms"] = list(s.timed_ms)
        d["pcg_per_solve"] = list(s.pcg_per_solve)[:min(s.pcg_solves, MAX_RECORDED)]
        return d

    def rigid_state(self, rb: int) -> dict:
        s = self.bodies[rb]
        return dict(position=np.array(s.position[:]), linear_velocity=np.array(s.linear_velocity[:]),
                    angular_velocity=np.array(s.angular_velocity[:]),
                    rotation=np.array(s.rotation[:]).reshape(3, 3), mass=s.mass)

    def pattern(self):
        nb = C.c_int32(0)
        check(self.L.ando_get_pattern(self.h, None, None, C.byref(nb)))
        rp = np.zeros(self.n + 1, np.int32)
        ci = np.zeros(nb.value, np.int32)
        check(self.L.ando_get_pattern(self.h, _ip(rp), _ip(ci), C.byref(nb)))
        return rp, ci

    def linearize(self, xhat):
        rp, ci = self.pattern()
        g = np.zeros(3 * self.n, np.float32)
        H = np.zeros(9 * len(ci), np.float32)
        xh = _f(xhat).ravel()
        check(self.L.ando_linearize(self.h, C.byref(self.params), _fp(xh), self.bodies, self.nbodies, _fp(g),
                                    _fp(H)))
        return g, H.reshape(-1, 3, 3)

    def apply_restitution(self):
        """Integrator::apply_contact_restitution on the current state."""
        check(self.L.ando_apply_restitution(self.h, C.byref(self.params), self.bodies, self.nbodies))

    def contacts(self, cap=1 << 16):
        buf = (Contact * cap)()
        cnt = C.c_int32(0)
        check(self.L.ando_compute_contacts(self.h, self.bodies, self.nbodies, buf, cap, C.byref(cnt)))
        return [buf[i] for i in range(min(cnt.value, cap))], cnt.value

    def energy(self) -> dict:
        """EnergyTracker::compute on the device (sums in double)."""
        e = Energy()
        check(self.L.ando_energy_diagnostics(self.h, C.byref(e)))
        return {"kinetic_energy": e.kinetic_energy, "elastic_energy": e.elastic_energy,
                "linear_momentum": np.array(e.linear_momentum[:]),
                "angular_momentum": np.array(e.angular_momentum[:]), "max_velocity": e.max_velocity}

    def contact_metrics(self, gap_max: float, ccd: bool, contacts=None) -> dict:
        """CollisionValidator::compute_metrics on the device: over the contacts
        the last contacts() call listed (contacts=None), or over a given list
        of Contact structures."""
        m = CollisionMetrics()
        if contacts is None:
            check(self.L.ando_contact_metrics(self.h, None, 0, gap_max, int(bool(ccd)), C.byref(m)))
        else:
            arr = (Contact * max(len(contacts), 1))(*contacts)
            check(self.L.ando_contact_metrics(self.h, arr, len(contacts), gap_max, int(bool(ccd)), C.byref(m)))
        return {k: getattr(m, k) for k, _ in CollisionMetrics._fields_}

    def cfl_inputs(self):
        """(max |v|, min edge length) reduced on the device."""
        mv, me = C.c_float(0.0), C.c_float(0.0)
        check(self.L.ando_cfl_inputs(self.h, C.byref(mv), C.byref(me)))
        return mv.value, me.value

    def count_intersections(self):
        """(edge-triangle crossings, cloth vertices inside rigid bodies)."""
        et, ins = C.c_int32(0), C.c_int32(0)
        check(self.L.ando_count_intersections(self.h, self.bodies, self.nbodies, C.byref(et), C.byref(ins)))
        return et.value, ins.value

    def set_precond_weights(self, w):
        w = _f(w).ravel()
        check(self.L.ando_set_precond_weights(self.h, _fp(w), len(w)))

    def pcg(self, rp, ci, vals, b, tol=1e-6, maxit=1000):
        n = len(rp) - 1
        x = np.zeros(3 * n, np.float32)
        it = C.c_int32(0)
        conv = C.c_int32(0)
        rp_, ci_, v_, b_ = _i(rp), _i(ci), _f(vals).ravel(), _f(b).ravel()
        check(self.L.ando_pcg(self.h, n, _ip(rp_), _ip(ci_), _fp(v_), _fp(b_), _fp(x), tol, maxit, C.byref(it),
                              C.byref(conv)))
        return x, it.value, bool(conv.value)


def pcg_noalloc(device: int, num_rows: int, num_blocks: int, row_ptr, cols, values, b, x, tol: float = 1e-6,
                maxit: int = 1000, stream=None):
    """ando_pcg_noalloc on device buffers given as integer addresses (for
    example torch tensors' data_ptr()); returns (iterations, converged)."""
    it = C.c_int32(0)
    conv = C.c_int32(0)
    check(load().ando_pcg_noalloc(device, num_rows, num_blocks, C.c_void_p(row_ptr), C.c_void_p(cols),
                                  C.c_void_p(values), C.c_void_p(b), C.c_void_p(x), tol, maxit,
                                  C.c_void_p(stream) if stream else None, C.byref(it), C.byref(conv)))
    return it.value, bool(conv.value)


def step_batch(sims, threads: int = 0):
    """ando_step on every GpuSim of `sims` at once (each has its own context
    and stream); apply_gravity is each caller's, as in GpuSim.substep."""
    L = load()
    k = len(sims)
    ctxs = (C.c_void_p * max(k, 1))(*[s.h.value for s in sims])
    params = (Params * max(k, 1))(*[s.params for s in sims])
    bodies = (C.POINTER(RigidBody) * max(k, 1))(*[C.cast(s.bodies, C.POINTER(RigidBody)) if s.nbodies else None
                                                    for s in sims])
    nb = (C.c_int32 * max(k, 1))(*[s.nbodies for s in sims])
    check(L.ando_step_batch(ctxs, params, bodies, nb, k, threads))


def nn_num_weights() -> int:
    return int(load().ando_nn_num_weights())


class Trainer:
    """Learned-preconditioner trainer on one device (ando_trainer_*)."""

    def __init__(self, device: int = 0):
        self.L = load()
        self.h = None
        h = C.c_void_p()
        check(self.L.ando_trainer_create(device, C.byref(h)))
        self.h = h
        _live.add(self)
        self.nw = nn_num_weights()
        self.sizes = []

    def close(self):
        """Release the trainer's device memory (idempotent)."""
        h, self.h = getattr(self, "h", None), None
        if h is not None:
            self.L.ando_trainer_destroy(h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_system(self, rp, ci, vals, b) -> int:
        rp_, ci_, v_, b_ = _i(rp), _i(ci), _f(vals).ravel(), _f(b).ravel()
        idx = C.c_int32(0)
        check(self.L.ando_trainer_add_system(self.h, len(rp_) - 1, _ip(rp_), _ip(ci_), _fp(v_), _fp(b_),
                                             C.byref(idx)))
        self.sizes.append(len(rp_) - 1)
        return idx.value

    def init_weights(self, seed: int = 1):
        check(self.L.ando_trainer_init_weights(self.h, C.c_uint64(seed)))

    def set_weights(self, w):
        w = _f(w).ravel()
        check(self.L.ando_trainer_set_weights(self.h, _fp(w), len(w)))

    def get_weights(self) -> np.ndarray:
        w = np.zeros(self.nw, np.float32)
        check(self.L.ando_trainer_get_weights(self.h, _fp(w), self.nw))
        return w

    def forward(self, i: int):
        n = self.sizes[i]
        P = np.zeros(9 * n, np.float32)
        y = np.zeros(3 * n, np.float32)
        check(self.L.ando_trainer_forward(self.h, i, _fp(P), _fp(y)))
        return P.reshape(n, 3, 3), y.reshape(n, 3)

    def loss_grad(self, i: int):
        loss = C.c_double(0.0)
        g = np.zeros(self.nw, np.float32)
        check(self.L.ando_trainer_loss_grad(self.h, i, C.byref(loss), _fp(g)))
        return loss.value, g

    def train(self, steps: int, lr: float = 1e-3, seed: int = 7) -> np.ndarray:
        out = np.zeros(max(steps, 1), np.float64)
        check(self.L.ando_trainer_train(self.h, steps, C.c_float(lr), C.c_uint64(seed),
                                        out.ctypes.data_as(C.POINTER(C.c_double))))
        return out[:steps]

    def solve(self, i: int, learned, tol: float = 1e-6, maxit: int = 1000):
        """learned: False/0 plain, True/1 learned preconditioner + warm start, 2 warm start only."""
        it = C.c_int32(0)
        conv = C.c_int32(0)
        check(self.L.ando_trainer_solve(self.h, i, int(learned), C.c_float(tol), maxit, C.byref(it),
                                        C.byref(conv)))
        return it.value, bool(conv.value)
