// ando_barrier_core — drop-in Python module for AndoSim's Blender add-on.
//
// Mirrors the reference binding (src/py/bindings.cpp:27-634): same class and
// function names, argument names, defaults, dtypes and exceptions.  The data
// model (Mesh, State, Constraints, RigidBody) lives on the host as in the
// reference; Integrator.step and Integrator.compute_contacts run on the GPU
// through the C ABI of include/andosim_hip.h.  A State that has been stepped
// keeps its positions and velocities resident on the device; host copies are
// refreshed lazily when Python reads them.
//
// EnergyTracker and CollisionValidator.compute_metrics (the add-on's
// per-frame consumers) run on the device when the state is resident there
// (for the validator: when the list is the state's last contact listing);
// otherwise they, AdaptiveTimestep and the Elasticity/barrier statics are
// plain C++ restatements of
// src/core/{energy_tracker,collision_validator,adaptive_timestep,elasticity,barrier}.cpp.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "andosim_hip.h"
#include "model.h"

namespace py = pybind11;

namespace {

using ando::Material;
using FArr = py::array_t<float, py::array::c_style | py::array::forcecast>;
using IArr = py::array_t<int32_t, py::array::c_style | py::array::forcecast>;

struct V3 {
  float x, y, z;
};
inline float dot3(V3 a, V3 b) { return a.x * b.x + (a.y * b.y + a.z * b.z); }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline V3 normalized(V3 a) {
  float z = dot3(a, a);
  if (z > 0.0f) {
    float s = std::sqrt(z);
    return {a.x / s, a.y / s, a.z / s};
  }
  return a;
}

V3 to_v3(py::handle obj, const char* what) {
  auto arr = py::array_t<float, py::array::forcecast>::ensure(obj);
  if (!arr || arr.ndim() != 1 || arr.shape(0) != 3) throw py::value_error(what);
  auto a = arr.unchecked<1>();
  return {a(0), a(1), a(2)};
}
py::array_t<float> v3_array(V3 v) {
  py::array_t<float> r({py::ssize_t(3)});
  auto m = r.mutable_unchecked<1>();
  m(0) = v.x;
  m(1) = v.y;
  m(2) = v.z;
  return r;
}
py::array_t<float> nx3(const float* data, size_t n) {
  py::array_t<float> r({py::ssize_t(n), py::ssize_t(3)});
  if (n) std::memcpy(r.mutable_data(), data, 3 * n * sizeof(float));
  return r;
}

// types.h:60-100
struct SimParams {
  float dt = 0.002f;
  float beta_max = 0.25f;
  int min_newton_steps = 2;
  int max_newton_steps = 8;
  int friction_min_newton_steps = 32;
  float pcg_tol = 1e-3f;
  int pcg_max_iters = 1000;
  float contact_gap_max = 0.001f;
  float wall_gap = 0.001f;
  bool enable_ccd = true;
  float contact_normal_epsilon = 1e-8f;
  float barrier_tolerance = 1e-12f;
  bool enable_friction = false;
  float friction_mu = 0.1f;
  float friction_epsilon = 1e-5f;
  float friction_tangent_threshold = 1e-6f;
  float velocity_damping = 0.0f;
  float contact_restitution = 0.0f;
  bool enable_strain_limiting = false;
  float strain_limit = 0.05f;
  float strain_tau = 0.05f;
  float strain_epsilon = 0.0f;
  float strain_svd_epsilon = 1e-6f;
  float hessian_epsilon = 1e-8f;
  float min_gap = 1e-8f;

  ando_params abi() const {
    ando_params p{};
    p.dt = dt;
    p.beta_max = beta_max;
    p.min_newton_steps = min_newton_steps;
    p.max_newton_steps = max_newton_steps;
    p.friction_min_newton_steps = friction_min_newton_steps;
    p.pcg_tol = pcg_tol;
    p.pcg_max_iters = pcg_max_iters;
    p.contact_gap_max = contact_gap_max;
    p.wall_gap = wall_gap;
    p.enable_ccd = enable_ccd ? 1 : 0;
    p.contact_normal_epsilon = contact_normal_epsilon;
    p.barrier_tolerance = barrier_tolerance;
    p.enable_friction = enable_friction ? 1 : 0;
    p.friction_mu = friction_mu;
    p.friction_epsilon = friction_epsilon;
    p.friction_tangent_threshold = friction_tangent_threshold;
    p.velocity_damping = velocity_damping;
    p.contact_restitution = contact_restitution;
    p.enable_strain_limiting = enable_strain_limiting ? 1 : 0;
    p.strain_limit = strain_limit;
    p.strain_tau = strain_tau;
    p.strain_epsilon = strain_epsilon;
    p.strain_svd_epsilon = strain_svd_epsilon;
    p.hessian_epsilon = hessian_epsilon;
    p.min_gap = min_gap;
    return p;
  }
};

struct Triangle {
  int32_t v[3] = {-1, -1, -1};
};

struct Mesh {
  std::vector<float> vertices;     // 3n
  std::vector<int32_t> triangles;  // 3m
  Material material;
  ando::RestState rest;
  uint64_t version = 0;

  size_t num_vertices() const { return vertices.size() / 3; }
  size_t num_triangles() const { return triangles.size() / 3; }
  // Mesh::initialize / compute_rest_state (mesh.cpp:8-75).
  void compute_rest_state() {
    rest = ando::compute_rest_state(vertices.data(), (int32_t)num_vertices(), triangles.data(),
                                    (int32_t)num_triangles());
    ++version;
  }
  V3 vert(size_t i) const { return {vertices[3 * i], vertices[3 * i + 1], vertices[3 * i + 2]}; }
};

// One device context per State (created on the first GPU call).
struct Session {
  ando_ctx* ctx = nullptr;
  const Mesh* mesh = nullptr;
  uint64_t mesh_version = ~0ull;
  const void* constraints = nullptr;
  uint64_t constraints_version = ~0ull;
  uint64_t listing = 0;     // id of the last contact listing of this session
  int32_t listed = -1;      // its length
  ~Session() {
    if (ctx) ando_destroy(ctx);
  }
};

void abi_check(int rc) {
  if (rc != ANDO_OK) {
    std::string msg = ando_last_error();
    if (rc == ANDO_EINVAL) throw py::value_error(msg);
    throw std::runtime_error("andosim_hip: " + msg);
  }
}

// Page-locked buffers behind the arrays get_positions / get_velocities return
// while the device holds the newer state: the device copies into the buffer by
// DMA and the array owns it.  A released array hands its buffer back to a pool
// (at most four free buffers per size), so steady per-frame reads allocate
// nothing.  Live and idle buffers together are capped at kPinnedCap bytes: a
// request that would pass the cap first frees idle buffers, oldest first, and
// if the live buffers alone leave no room the read copies through the state's
// pageable host arrays instead.  The pool lives until the process ends.
constexpr size_t kPinnedCap = size_t(256) << 20;
struct PinnedPool {
  std::mutex mu;
  std::vector<std::pair<size_t, void*>> idle;  // (bytes, buffer), oldest first
  std::unordered_map<void*, size_t> live;      // buffer -> bytes
  size_t live_bytes = 0, idle_bytes = 0;
};
PinnedPool& pinned_pool() {
  static PinnedPool* pool = new PinnedPool();
  return *pool;
}
void* pinned_take(size_t bytes) {
  PinnedPool& pool = pinned_pool();
  std::vector<void*> evict;
  {
    std::lock_guard<std::mutex> g(pool.mu);
    for (size_t k = 0; k < pool.idle.size(); ++k)
      if (pool.idle[k].first == bytes) {
        void* p = pool.idle[k].second;
        pool.idle.erase(pool.idle.begin() + (std::ptrdiff_t)k);
        pool.idle_bytes -= bytes;
        pool.live[p] = bytes;
        pool.live_bytes += bytes;
        return p;
      }
    if (pool.live_bytes + bytes > kPinnedCap) return nullptr;
    while (!pool.idle.empty() && pool.live_bytes + pool.idle_bytes + bytes > kPinnedCap) {
      evict.push_back(pool.idle.front().second);
      pool.idle_bytes -= pool.idle.front().first;
      pool.idle.erase(pool.idle.begin());
    }
    // Reserve the bytes before allocating outside the lock.
    pool.live_bytes += bytes;
  }
  for (void* q : evict) (void)ando_host_free(q);
  void* p = nullptr;
  if (ando_host_alloc(bytes, &p) != ANDO_OK) {
    std::lock_guard<std::mutex> g(pool.mu);
    pool.live_bytes -= bytes;
    return nullptr;
  }
  std::lock_guard<std::mutex> g(pool.mu);
  pool.live[p] = bytes;
  return p;
}
void pinned_give(void* p) {
  PinnedPool& pool = pinned_pool();
  {
    std::lock_guard<std::mutex> g(pool.mu);
    auto it = pool.live.find(p);
    if (it == pool.live.end()) return;
    const size_t bytes = it->second;
    pool.live.erase(it);
    pool.live_bytes -= bytes;
    int same = 0;
    for (const auto& e : pool.idle) same += e.first == bytes;
    if (same < 4 && pool.live_bytes + pool.idle_bytes + bytes <= kPinnedCap) {
      pool.idle.push_back({bytes, p});
      pool.idle_bytes += bytes;
      return;
    }
  }
  (void)ando_host_free(p);
}
// Bytes the pool holds (live, idle); test hook.
std::pair<size_t, size_t> pinned_pool_bytes() {
  PinnedPool& pool = pinned_pool();
  std::lock_guard<std::mutex> g(pool.mu);
  return {pool.live_bytes, pool.idle_bytes};
}

struct State {
  std::vector<float> positions, velocities, masses, positions_prev;
  std::shared_ptr<Session> session;
  bool host_dirty = true;     // host copies newer than the device
  bool device_dirty = false;  // device newer than the host copies

  size_t num_vertices() const { return positions.size() / 3; }

  // State::initialize (state.cpp:7-17).
  void initialize(const Mesh& mesh) {
    if (mesh.num_vertices() == 0) throw std::invalid_argument("Mesh must be initialised before creating a state");
    positions = mesh.vertices;
    positions_prev = mesh.vertices;
    velocities.assign(mesh.vertices.size(), 0.0f);
    masses = ando::lumped_masses((int32_t)mesh.num_vertices(), mesh.triangles.data(), (int32_t)mesh.num_triangles(),
                                 mesh.rest.area, mesh.material);
    session.reset();
    host_dirty = true;
    device_dirty = false;
  }
  void pull() {
    if (device_dirty && session && session->ctx) {
      abi_check(ando_get_state(session->ctx, positions.data(), velocities.data(), (int32_t)num_vertices()));
      device_dirty = false;
    }
  }
  // get_positions (pos) / get_velocities: a new (N, 3) array, as the
  // reference's copy out of its Eigen vector (bindings.cpp:248-267).  While the
  // device holds the newer state, only the one array is copied, by DMA into a
  // pinned buffer that the returned array owns; the host copies stay stale.
  py::array_t<float> read(bool pos) {
    const size_t n = num_vertices();
    if (device_dirty && session && session->ctx && n > 0) {
      if (void* buf = pinned_take(3 * n * sizeof(float))) {
        float* f = static_cast<float*>(buf);
        const int rc = ando_get_state(session->ctx, pos ? f : nullptr, pos ? nullptr : f, (int32_t)n);
        if (rc != ANDO_OK) {
          pinned_give(buf);
          abi_check(rc);
        }
        py::capsule owner(buf, [](void* p) { pinned_give(p); });
        return py::array_t<float>({py::ssize_t(n), py::ssize_t(3)},
                                  {py::ssize_t(3 * sizeof(float)), py::ssize_t(sizeof(float))}, f, owner);
      }
    }
    pull();
    return nx3(pos ? positions.data() : velocities.data(), n);
  }
};

struct Pin {
  int32_t vertex_idx;
  V3 target;
  bool active = true;
};
struct Wall {
  V3 normal;
  float offset, gap;
  bool active = true;
};
struct Constraints {
  std::vector<Pin> pins;
  std::vector<Wall> walls;
  uint64_t version = 0;
  size_t num_active_pins() const {
    return (size_t)std::count_if(pins.begin(), pins.end(), [](const Pin& p) { return p.active; });
  }
  size_t num_active_walls() const {
    return (size_t)std::count_if(walls.begin(), walls.end(), [](const Wall& w) { return w.active; });
  }
};

enum class ContactType { POINT_TRIANGLE = 0, EDGE_EDGE = 1, WALL = 2, RIGID_POINT_TRIANGLE = 3 };

// ContactPair (collision.h:78-103).
struct ContactPair {
  ContactType type = ContactType::POINT_TRIANGLE;
  int32_t idx0 = -1, idx1 = -1, idx2 = -1, idx3 = -1;
  float gap = 0.0f;
  V3 normal{0, 0, 0}, witness_p{0, 0, 0}, witness_q{0, 0, 0};
  int vertex_count = 0;
  int rigid_body_index = -1;
  // Set when the contact comes from a device listing: which listing, and its
  // position in it.  A list that is exactly that listing is validated on the
  // device without being read back.
  uint64_t listing = 0;
  int32_t slot = -1;
};

// RigidBody (rigid_body.h:16-86, rigid_body.cpp:20-184) on the host.
struct RigidBody {
  std::vector<float> local;        // 3nv
  std::vector<int32_t> tris;       // 3nt
  float mass = 1.0f;
  float Ib[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  float Ibinv[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  V3 position{0, 0, 0};
  float rotation[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  V3 lin{0, 0, 0}, ang{0, 0, 0};
  V3 force{0, 0, 0}, torque{0, 0, 0};

  static void mul33(const float* A, const float* B, float* O) {
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) O[3 * r + c] = A[3 * r] * B[c] + (A[3 * r + 1] * B[3 + c] + A[3 * r + 2] * B[6 + c]);
  }
  static V3 mulv(const float* A, V3 v) {
    return {A[0] * v.x + (A[1] * v.y + A[2] * v.z), A[3] * v.x + (A[4] * v.y + A[5] * v.z),
            A[6] * v.x + (A[7] * v.y + A[8] * v.z)};
  }
  void initialize(const std::vector<float>& verts, const std::vector<int32_t>& t, float density) {
    tris = t;
    size_t nv = verts.size() / 3;
    if (nv == 0) {
      local.clear();
      mass = 0.0f;
      std::fill(Ib, Ib + 9, 0.0f);
      std::fill(Ibinv, Ibinv + 9, 0.0f);
      return;
    }
    V3 cen{0, 0, 0};
    for (size_t i = 0; i < nv; ++i) cen = {cen.x + verts[3 * i], cen.y + verts[3 * i + 1], cen.z + verts[3 * i + 2]};
    float fn = (float)nv;
    cen = {cen.x / fn, cen.y / fn, cen.z / fn};
    local.resize(verts.size());
    for (size_t i = 0; i < nv; ++i) {
      local[3 * i] = verts[3 * i] - cen.x;
      local[3 * i + 1] = verts[3 * i + 1] - cen.y;
      local[3 * i + 2] = verts[3 * i + 2] - cen.z;
    }
    position = cen;
    std::fill(rotation, rotation + 9, 0.0f);
    rotation[0] = rotation[4] = rotation[8] = 1.0f;
    lin = ang = {0, 0, 0};
    float mn[3] = {std::numeric_limits<float>::max(), std::numeric_limits<float>::max(),
                   std::numeric_limits<float>::max()};
    float mx[3] = {-mn[0], -mn[1], -mn[2]};
    for (size_t i = 0; i < nv; ++i) {
      float w[3] = {local[3 * i] + position.x, local[3 * i + 1] + position.y, local[3 * i + 2] + position.z};
      for (int k = 0; k < 3; ++k) {
        mn[k] = std::min(mn[k], w[k]);
        mx[k] = std::max(mx[k], w[k]);
      }
    }
    float ext[3] = {mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2]};
    float vol = std::max(ext[0], 1e-6f) * std::max(ext[1], 1e-6f) * std::max(ext[2], 1e-6f);
    mass = std::max(density * vol, 1e-3f);
    float x2 = ext[0] * ext[0], y2 = ext[1] * ext[1], z2 = ext[2] * ext[2];
    std::fill(Ib, Ib + 9, 0.0f);
    Ib[0] = (mass / 12.0f) * (y2 + z2);
    Ib[4] = (mass / 12.0f) * (x2 + z2);
    Ib[8] = (mass / 12.0f) * (x2 + y2);
    for (int i = 0; i < 3; ++i)
      if (Ib[4 * i] < 1e-6f) Ib[4 * i] = 1e-6f;
    // 3x3 cofactor inverse (as Eigen's fixed-size inverse).
    auto cof = [&](int i, int j) {
      int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
      return Ib[3 * i1 + j1] * Ib[3 * i2 + j2] - Ib[3 * i1 + j2] * Ib[3 * i2 + j1];
    };
    float det = cof(0, 0) * Ib[0] + (cof(1, 0) * Ib[3] + cof(2, 0) * Ib[6]);
    float inv = 1.0f / det;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) Ibinv[3 * i + j] = cof(j, i) * inv;
    force = torque = {0, 0, 0};
  }
  std::vector<float> world_vertices() const {
    std::vector<float> out(local.size());
    for (size_t i = 0; i < local.size() / 3; ++i) {
      V3 rv = mulv(rotation, {local[3 * i], local[3 * i + 1], local[3 * i + 2]});
      out[3 * i] = position.x + rv.x;
      out[3 * i + 1] = position.y + rv.y;
      out[3 * i + 2] = position.z + rv.z;
    }
    return out;
  }
  void inertia_world_inv(float* O) const {
    float RI[9], Rt[9];
    mul33(rotation, Ibinv, RI);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) Rt[3 * r + c] = rotation[3 * c + r];
    mul33(RI, Rt, O);
  }
  // rigid_body.cpp:113-124
  void apply_impulse(V3 p, V3 J) {
    if (mass <= 0.0f) return;
    lin = {lin.x + J.x / mass, lin.y + J.y / mass, lin.z + J.z / mass};
    V3 r = sub(p, position);
    float Iw[9];
    inertia_world_inv(Iw);
    V3 dw = mulv(Iw, cross(r, J));
    ang = {ang.x + dw.x, ang.y + dw.y, ang.z + dw.z};
  }
  // rigid_body.cpp:151-176
  void integrate(float dt) {
    if (mass <= 0.0f) return;
    ando_rigid_body b = abi();
    double F[3] = {force.x, force.y, force.z}, T[3] = {torque.x, torque.y, torque.z};
    (void)F;
    (void)T;
    V3 acc{force.x / mass, force.y / mass, force.z / mass};
    lin = {lin.x + acc.x * dt, lin.y + acc.y * dt, lin.z + acc.z * dt};
    float Iw[9];
    inertia_world_inv(Iw);
    V3 aa = mulv(Iw, torque);
    ang = {ang.x + aa.x * dt, ang.y + aa.y * dt, ang.z + aa.z * dt};
    position = {position.x + lin.x * dt, position.y + lin.y * dt, position.z + lin.z * dt};
    float om = std::sqrt(dot3(ang, ang));
    if ((double)om > 1e-8) {
      float th = om * dt;
      V3 ax{ang.x / om, ang.y / om, ang.z / om};
      float s = std::sin(th), c = std::cos(th);
      V3 sa{s * ax.x, s * ax.y, s * ax.z}, ca{(1.0f - c) * ax.x, (1.0f - c) * ax.y, (1.0f - c) * ax.z};
      float R[9];
      float tmp = ca.x * ax.y;
      R[1] = tmp - sa.z; R[3] = tmp + sa.z;
      tmp = ca.x * ax.z;
      R[2] = tmp + sa.y; R[6] = tmp - sa.y;
      tmp = ca.y * ax.z;
      R[5] = tmp - sa.x; R[7] = tmp + sa.x;
      R[0] = ca.x * ax.x + c; R[4] = ca.y * ax.y + c; R[8] = ca.z * ax.z + c;
      float Rn[9];
      mul33(R, rotation, Rn);
      std::memcpy(rotation, Rn, sizeof(Rn));
    }
    force = torque = {0, 0, 0};
    (void)b;
  }
  ando_rigid_body abi() const {
    ando_rigid_body b{};
    b.num_vertices = (int32_t)(local.size() / 3);
    b.num_triangles = (int32_t)(tris.size() / 3);
    b.local_vertices = local.data();
    b.triangles = tris.data();
    b.mass = mass;
    std::memcpy(b.inertia_body_inv, Ibinv, sizeof(Ibinv));
    b.position[0] = position.x; b.position[1] = position.y; b.position[2] = position.z;
    std::memcpy(b.rotation, rotation, sizeof(rotation));
    b.linear_velocity[0] = lin.x; b.linear_velocity[1] = lin.y; b.linear_velocity[2] = lin.z;
    b.angular_velocity[0] = ang.x; b.angular_velocity[1] = ang.y; b.angular_velocity[2] = ang.z;
    return b;
  }
  void take(const ando_rigid_body& b) {
    position = {b.position[0], b.position[1], b.position[2]};
    std::memcpy(rotation, b.rotation, sizeof(rotation));
    lin = {b.linear_velocity[0], b.linear_velocity[1], b.linear_velocity[2]};
    ang = {b.angular_velocity[0], b.angular_velocity[1], b.angular_velocity[2]};
    force = torque = {0, 0, 0};
  }
};

// ---------------------------------------------------------------------------
// Device session management.
// ---------------------------------------------------------------------------
ando_stats g_last_stats{};

Session& ensure_session(State& state) {
  if (!state.session) {
    auto s = std::make_shared<Session>();
    abi_check(ando_create(0, &s->ctx));
    state.session = s;
    state.host_dirty = true;
  }
  return *state.session;
}

void sync_to_device(const Mesh& mesh, State& state, const Constraints* cons) {
  if (state.num_vertices() != mesh.num_vertices())
    throw py::value_error("state and mesh have different vertex counts");
  Session& s = ensure_session(state);
  if (s.mesh != &mesh || s.mesh_version != mesh.version) {
    ando_material mat{mesh.material.youngs_modulus, mesh.material.poisson_ratio, mesh.material.density,
                      mesh.material.thickness, mesh.material.bending_stiffness};
    state.pull();
    abi_check(ando_set_mesh(s.ctx, mesh.vertices.data(), (int32_t)mesh.num_vertices(), mesh.triangles.data(),
                            (int32_t)mesh.num_triangles(), &mat));
    s.mesh = &mesh;
    s.mesh_version = mesh.version;
    s.constraints = nullptr;
    state.host_dirty = true;
  }
  if (state.host_dirty) {
    abi_check(ando_set_state(s.ctx, state.positions.data(), state.velocities.data(), state.masses.data(),
                             (int32_t)state.num_vertices()));
    abi_check(ando_set_previous_positions(s.ctx, state.positions_prev.data(), (int32_t)state.num_vertices()));
    state.host_dirty = false;
  }
  if (cons && (s.constraints != cons || s.constraints_version != cons->version)) {
    std::vector<int32_t> pv;
    std::vector<float> pt;
    for (const Pin& p : cons->pins)
      if (p.active) {
        pv.push_back(p.vertex_idx);
        pt.insert(pt.end(), {p.target.x, p.target.y, p.target.z});
      }
    abi_check(ando_set_pins(s.ctx, pv.data(), pt.data(), (int32_t)pv.size()));
    std::vector<float> wn, wo, wg;
    for (const Wall& w : cons->walls)
      if (w.active) {
        wn.insert(wn.end(), {w.normal.x, w.normal.y, w.normal.z});
        wo.push_back(w.offset);
        wg.push_back(w.gap);
      }
    abi_check(ando_set_walls(s.ctx, wn.data(), wo.data(), wg.data(), (int32_t)wo.size()));
    s.constraints = cons;
    s.constraints_version = cons->version;
  }
}

std::vector<RigidBody*> rigid_handles(py::object rigid_list) {
  std::vector<RigidBody*> out;
  if (rigid_list.is_none()) return out;
  for (auto item : rigid_list) out.push_back(&item.cast<RigidBody&>());
  return out;
}

// ---------------------------------------------------------------------------
// Host-only restatements (◇).
// ---------------------------------------------------------------------------
// Mesh::compute_F (mesh.cpp:126-150), with mesh.vertices.
void compute_F(const Mesh& mesh, size_t f, float F[2][2]) {
  const int32_t* t = &mesh.triangles[3 * f];
  V3 v0 = mesh.vert(t[0]), v1 = mesh.vert(t[1]), v2 = mesh.vert(t[2]);
  V3 e1 = sub(v1, v0), e2 = sub(v2, v0);
  V3 n = normalized(cross(e1, e2));
  V3 t1 = normalized(e1);
  V3 t2 = cross(n, t1);
  float D[2][2] = {{dot3(e1, t1), dot3(e2, t1)}, {dot3(e1, t2), dot3(e2, t2)}};
  const float* M = &mesh.rest.dm_inv[4 * f];
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 2; ++c) F[r][c] = D[r][0] * M[c] + D[r][1] * M[2 + c];
}

// Elasticity::compute_energy (elasticity.cpp:5-14, 115-126).
float elastic_energy(const Mesh& mesh) {
  float e = 0.0f;
  float mu = ando::shear_modulus(mesh.material);
  for (size_t f = 0; f < mesh.num_triangles(); ++f) {
    float F[2][2];
    compute_F(mesh, f, F);
    float k = mesh.rest.area[f] * mesh.material.thickness * mu;
    float d00 = F[0][0] - 1.0f, d01 = F[0][1], d10 = F[1][0], d11 = F[1][1] - 1.0f;
    float sq = (d00 * d00 + d10 * d10) + (d01 * d01 + d11 * d11);
    e += k * sq;
  }
  return e;
}

// Elasticity::compute_gradient (elasticity.cpp:18-61).
void elastic_gradient(const Mesh& mesh, float* g, size_t len) {
  std::fill(g, g + len, 0.0f);
  float mu = ando::shear_modulus(mesh.material);
  for (size_t f = 0; f < mesh.num_triangles(); ++f) {
    const int32_t* t = &mesh.triangles[3 * f];
    float F[2][2];
    compute_F(mesh, f, F);
    float k = mesh.rest.area[f] * mesh.material.thickness * mu;
    float twok = (float)(2.0 * (double)k);
    float P[2][2] = {{twok * (F[0][0] - 1.0f), twok * (F[0][1] - 0.0f)}, {twok * (F[1][0] - 0.0f), twok * (F[1][1] - 1.0f)}};
    const float* M = &mesh.rest.dm_inv[4 * f];
    float H[2][2];
    for (int r = 0; r < 2; ++r)
      for (int c = 0; c < 2; ++c) H[r][c] = P[r][0] * M[2 * c] + P[r][1] * M[2 * c + 1];
    V3 v0 = mesh.vert(t[0]), v1 = mesh.vert(t[1]), v2 = mesh.vert(t[2]);
    V3 e1 = sub(v1, v0), e2 = sub(v2, v0);
    V3 n = normalized(cross(e1, e2));
    V3 t1 = normalized(e1);
    V3 t2 = cross(n, t1);
    V3 f1{H[0][0] * t1.x + H[1][0] * t2.x, H[0][0] * t1.y + H[1][0] * t2.y, H[0][0] * t1.z + H[1][0] * t2.z};
    V3 f2{H[0][1] * t1.x + H[1][1] * t2.x, H[0][1] * t1.y + H[1][1] * t2.y, H[0][1] * t1.z + H[1][1] * t2.z};
    V3 f0{-(f1.x + f2.x), -(f1.y + f2.y), -(f1.z + f2.z)};
    V3 fs[3] = {f0, f1, f2};
    for (int a = 0; a < 3; ++a) {
      size_t b = 3 * (size_t)t[a];
      if (b + 2 >= len) continue;
      g[b] += fs[a].x;
      g[b + 1] += fs[a].y;
      g[b + 2] += fs[a].z;
    }
  }
}

// barrier.cpp:29-65
float barrier_energy(float g, float gmax, float k) {
  if (gmax <= 0.0f || g >= gmax) return 0.0f;
  float d = gmax - g;
  if (d <= 0.0f) return 0.0f;
  float inv = 1.0f / gmax;
  return 0.5f * k * inv * d * d * d;
}
float barrier_gradient(float g, float gmax, float k) {
  if (gmax <= 0.0f || g >= gmax) return 0.0f;
  float d = gmax - g;
  float inv = 1.0f / gmax;
  return -1.5f * k * inv * d * d;
}
float barrier_hessian(float g, float gmax, float k) {
  if (gmax <= 0.0f || g >= gmax) return 0.0f;
  float d = gmax - g;
  float inv = 1.0f / gmax;
  return 3.0f * k * inv * d;
}

// energy_tracker.h:11-30
struct EnergyDiagnostics {
  float kinetic_energy = 0, elastic_energy = 0, barrier_energy = 0, total_energy = 0;
  V3 linear_momentum{0, 0, 0}, angular_momentum{0, 0, 0};
  float max_velocity = 0, max_acceleration = 0;
  int num_contacts = 0, num_pins = 0;
  float energy_drift_percent = 0, energy_drift_absolute = 0;
  // energy_tracker.cpp:78-86
  void update_drift(float prev) {
    if ((double)prev > 1e-12) {
      energy_drift_absolute = total_energy - prev;
      energy_drift_percent = (float)((double)(energy_drift_absolute / prev) * 100.0);
    } else {
      energy_drift_absolute = 0.0f;
      energy_drift_percent = 0.0f;
    }
  }
};

float kinetic_energy(const State& s) {
  float ke = 0.0f;
  for (size_t i = 0; i < s.num_vertices(); ++i) {
    V3 v{s.velocities[3 * i], s.velocities[3 * i + 1], s.velocities[3 * i + 2]};
    ke = (float)((double)ke + 0.5 * (double)s.masses[i] * (double)dot3(v, v));
  }
  return ke;
}
V3 linear_momentum(const State& s) {
  V3 p{0, 0, 0};
  for (size_t i = 0; i < s.num_vertices(); ++i) {
    float m = s.masses[i];
    p = {p.x + m * s.velocities[3 * i], p.y + m * s.velocities[3 * i + 1], p.z + m * s.velocities[3 * i + 2]};
  }
  return p;
}
V3 angular_momentum(const State& s) {
  V3 L{0, 0, 0};
  for (size_t i = 0; i < s.num_vertices(); ++i) {
    float m = s.masses[i];
    V3 r{s.positions[3 * i], s.positions[3 * i + 1], s.positions[3 * i + 2]};
    V3 p{m * s.velocities[3 * i], m * s.velocities[3 * i + 1], m * s.velocities[3 * i + 2]};
    V3 c = cross(r, p);
    L = {L.x + c.x, L.y + c.y, L.z + c.z};
  }
  return L;
}
float max_velocity(const State& s) {
  float mv = 0.0f;
  for (size_t i = 0; i < s.num_vertices(); ++i) {
    V3 v{s.velocities[3 * i], s.velocities[3 * i + 1], s.velocities[3 * i + 2]};
    mv = std::max(mv, std::sqrt(dot3(v, v)));
  }
  return mv;
}

// collision_validator.h:12-54, collision_validator.cpp:7-131
struct CollisionMetrics {
  int num_point_triangle = 0, num_edge_edge = 0, num_wall = 0, num_total_contacts = 0;
  float min_gap = 0, max_gap = 0, avg_gap = 0;
  int num_penetrations = 0;
  float max_penetration = 0, avg_penetration = 0;
  bool ccd_enabled = false;
  int num_ccd_contacts = 0, num_broad_phase_contacts = 0;
  float ccd_effectiveness = 0;
  float max_relative_velocity = 0, avg_relative_velocity = 0;
  bool has_tunneling = false, has_major_penetration = false, is_stable = true;

  void compute_from_contacts(const std::vector<ContactPair>& cs, float gap_max, bool ccd) {
    ccd_enabled = ccd;
    num_total_contacts = (int)cs.size();
    if (cs.empty()) {
      min_gap = max_gap = avg_gap = 0.0f;
      num_penetrations = 0;
      max_penetration = avg_penetration = 0.0f;
      max_relative_velocity = avg_relative_velocity = 0.0f;
      has_tunneling = has_major_penetration = false;
      is_stable = true;
      return;
    }
    min_gap = max_gap = cs[0].gap;
    float sum_gap = 0.0f, sum_pen = 0.0f;
    num_point_triangle = num_edge_edge = num_wall = 0;
    for (const auto& c : cs) {
      if (c.type == ContactType::POINT_TRIANGLE) ++num_point_triangle;
      else if (c.type == ContactType::EDGE_EDGE) ++num_edge_edge;
      else if (c.type == ContactType::WALL) ++num_wall;
      float g = c.gap;
      min_gap = std::min(min_gap, g);
      max_gap = std::max(max_gap, g);
      sum_gap += g;
      if (g < 0.0f) {
        ++num_penetrations;
        float pen = -g;
        max_penetration = std::max(max_penetration, pen);
        sum_pen += pen;
        if ((double)pen > 0.001) has_major_penetration = true;
        if ((double)pen > 0.1 * (double)gap_max) has_tunneling = true;
      }
    }
    avg_gap = sum_gap / (float)num_total_contacts;
    if (num_penetrations > 0) {
      avg_penetration = sum_pen / (float)num_penetrations;
      is_stable = false;
    } else {
      avg_penetration = 0.0f;
      is_stable = true;
    }
    num_ccd_contacts = num_broad_phase_contacts = 0;
    for (const auto& c : cs) {
      if (ccd_enabled && (double)c.gap < (double)gap_max * 0.5) ++num_ccd_contacts;
      else ++num_broad_phase_contacts;
    }
    if (ccd_enabled && num_total_contacts > 0)
      ccd_effectiveness = (float)((double)((float)num_ccd_contacts / (float)num_total_contacts) * 100.0);
    else
      ccd_effectiveness = 0.0f;
  }
  int quality_level() const {
    if (has_tunneling) return 3;
    if (has_major_penetration) return 3;
    if ((double)max_penetration > 0.0001) return 2;
    if (num_penetrations > 0) return 1;
    return 0;
  }
  const char* quality_description() const {
    switch (quality_level()) {
      case 0: return "Excellent - No penetrations";
      case 1: return "Good - Tiny penetrations (<0.1mm)";
      case 2: return "Warning - Minor penetrations detected";
      case 3: return "Error - Major penetrations or tunneling";
      default: return "Unknown";
    }
  }
};

// The device holds this state (and this mesh, when one is given) with
// nothing newer on the host.
bool device_resident(const Mesh* mesh, const State& state) {
  if (!state.session || !state.session->ctx || state.host_dirty) return false;
  if (mesh && (state.session->mesh != mesh || state.session->mesh_version != mesh->version)) return false;
  return true;
}

// EnergyTracker::compute's sums on the device (ando_energy_diagnostics).
bool gpu_energy(const Mesh* mesh, const State& state, ando_energy& e) {
  if (!device_resident(mesh, state)) return false;
  abi_check(ando_energy_diagnostics(state.session->ctx, &e));
  return true;
}

// CollisionValidator::compute_metrics on the device when `cs` is exactly the
// session's last contact listing and the device still holds that table and
// the current velocities.
bool gpu_metrics(State& state, const std::vector<ContactPair>& cs, float gap_max, bool ccd, CollisionMetrics& M) {
  if (cs.empty() || !device_resident(nullptr, state)) return false;
  const Session& s = *state.session;
  if (s.listing == 0 || (int64_t)cs.size() != (int64_t)s.listed) return false;
  for (size_t i = 0; i < cs.size(); ++i)
    if (cs[i].listing != s.listing || cs[i].slot != (int32_t)i) return false;
  ando_collision_metrics r{};
  const int rc = ando_contact_metrics(s.ctx, nullptr, 0, gap_max, ccd ? 1 : 0, &r);
  if (rc == ANDO_ESTATE) return false;  // a later detection replaced the table
  abi_check(rc);
  M.ccd_enabled = ccd;
  M.num_point_triangle = r.num_point_triangle;
  M.num_edge_edge = r.num_edge_edge;
  M.num_wall = r.num_wall;
  M.num_total_contacts = r.num_total_contacts;
  M.min_gap = r.min_gap;
  M.max_gap = r.max_gap;
  M.avg_gap = r.avg_gap;
  M.num_penetrations = r.num_penetrations;
  M.max_penetration = r.max_penetration;
  M.avg_penetration = r.avg_penetration;
  M.has_major_penetration = r.has_major_penetration != 0;
  M.has_tunneling = r.has_tunneling != 0;
  M.is_stable = r.num_penetrations == 0;
  M.num_ccd_contacts = r.num_ccd_contacts;
  M.num_broad_phase_contacts = r.num_broad_phase_contacts;
  M.ccd_effectiveness = r.ccd_effectiveness;
  M.max_relative_velocity = r.max_relative_velocity;
  M.avg_relative_velocity = r.avg_relative_velocity;
  return true;
}

CollisionMetrics compute_metrics(const Mesh&, State& state, const std::vector<ContactPair>& cs, float gap_max,
                                 bool ccd) {
  CollisionMetrics M;
  if (gpu_metrics(state, cs, gap_max, ccd, M)) return M;
  state.pull();
  M.compute_from_contacts(cs, gap_max, ccd);
  std::vector<float> rel;
  auto vel = [&](int32_t i) {
    return V3{state.velocities[3 * i], state.velocities[3 * i + 1], state.velocities[3 * i + 2]};
  };
  for (const auto& c : cs) {
    if (c.type == ContactType::POINT_TRIANGLE) {
      V3 vp = vel(c.idx0), a = vel(c.idx1), b = vel(c.idx2), d = vel(c.idx3);
      V3 vt{(float)((double)((a.x + b.x) + d.x) / 3.0), (float)((double)((a.y + b.y) + d.y) / 3.0),
            (float)((double)((a.z + b.z) + d.z) / 3.0)};
      rel.push_back(std::fabs(dot3(sub(vp, vt), c.normal)));
    } else if (c.type == ContactType::EDGE_EDGE) {
      V3 a = vel(c.idx0), b = vel(c.idx1), d = vel(c.idx2), e = vel(c.idx3);
      V3 e0{(float)((double)(a.x + b.x) / 2.0), (float)((double)(a.y + b.y) / 2.0), (float)((double)(a.z + b.z) / 2.0)};
      V3 e1{(float)((double)(d.x + e.x) / 2.0), (float)((double)(d.y + e.y) / 2.0), (float)((double)(d.z + e.z) / 2.0)};
      rel.push_back(std::fabs(dot3(sub(e0, e1), c.normal)));
    } else if (c.type == ContactType::WALL) {
      rel.push_back(std::fabs(dot3(vel(c.idx0), c.normal)));
    }
  }
  if (!rel.empty()) {
    M.max_relative_velocity = *std::max_element(rel.begin(), rel.end());
    float s = 0.0f;
    for (float v : rel) s += v;
    M.avg_relative_velocity = s / (float)rel.size();
  }
  return M;
}

// adaptive_timestep.cpp:7-151
constexpr float kStaticVel = 1e-6f;
constexpr float kMinEdge = 1e-5f;
float at_max_velocity(const std::vector<float>& v) {
  float mv = 0.0f;
  for (size_t i = 0; i < v.size() / 3; ++i) {
    V3 q{v[3 * i], v[3 * i + 1], v[3 * i + 2]};
    mv = std::max(mv, dot3(q, q));
  }
  return std::sqrt(mv);
}
float at_min_edge(const Mesh& mesh) {
  const auto& E = mesh.rest.edges;
  if (E.empty()) return 0.0f;
  float mn = std::numeric_limits<float>::max();
  bool found = false;
  size_t nv = mesh.num_vertices();
  for (size_t e = 0; e + 1 < E.size(); e += 2) {
    int32_t a = E[e], b = E[e + 1];
    if (a < 0 || b < 0 || (size_t)a >= nv || (size_t)b >= nv) continue;
    V3 d = sub(mesh.vert(b), mesh.vert(a));
    float l2 = dot3(d, d);
    if (!std::isfinite(l2)) continue;
    mn = std::min(mn, l2);
    found = true;
  }
  if (!found) return 0.0f;
  float r = std::sqrt(std::max(mn, 0.0f));
  return std::isfinite(r) ? r : 0.0f;
}
float at_cfl(float vmax, float min_edge, float safety) {
  if (vmax < kStaticVel) return 1.0f;
  return safety * min_edge / vmax;
}
// compute_next_dt (adaptive_timestep.cpp:8-44) once max |v| is known; the
// minimum edge length comes from min_edge() only when it is needed.
template <class MinEdge>
float at_next_dt_from(float vmax, MinEdge min_edge, float cur, float dmin, float dmax, float safety) {
  if (vmax < kStaticVel) return dmax;
  float me = min_edge();
  if (me < kMinEdge) me = kMinEdge;
  float cfl = at_cfl(vmax, me, safety);
  float target = std::clamp(cfl, dmin, dmax);
  float next = target < cur ? target : std::min(target, cur * 1.5f);
  return std::clamp(next, dmin, dmax);
}
float at_next_dt(const std::vector<float>& v, const Mesh& mesh, float cur, float dmin, float dmax, float safety) {
  return at_next_dt_from(at_max_velocity(v), [&] { return at_min_edge(mesh); }, cur, dmin, dmax, safety);
}

std::vector<float> flat_floats(py::handle obj) {
  auto a = FArr::ensure(obj);
  if (!a) throw py::type_error("expected a float array");
  return std::vector<float>(a.data(), a.data() + a.size());
}

void parse_triangles(py::object triangles_obj, std::vector<int32_t>& out, const char* two_d_msg) {
  py::array triangles_array = py::array::ensure(triangles_obj);
  if (!triangles_array) throw py::value_error("triangles must be array-like");
  py::array reshaped;
  if (triangles_array.ndim() == 1) {
    if (triangles_array.shape(0) % 3 != 0) throw py::value_error("triangles must contain a multiple of 3 indices");
    reshaped = triangles_array.reshape({triangles_array.shape(0) / 3, (py::ssize_t)3});
  } else if (triangles_array.ndim() == 2) {
    if (triangles_array.shape(1) != 3) throw py::value_error("triangles must have shape (M, 3)");
    reshaped = triangles_array;
  } else {
    throw py::value_error(two_d_msg);
  }
  auto tri = IArr::ensure(reshaped);
  if (!tri) throw py::value_error("triangles must be convertible to int32");
  out.assign(tri.data(), tri.data() + tri.size());
}

py::list contacts_to_list(const std::vector<ContactPair>& cs) {
  py::list out;
  for (const auto& c : cs) out.append(py::cast(c));
  return out;
}

std::vector<ContactPair> gpu_contacts(const Mesh& mesh, State& state, py::object rigid_list) {
  sync_to_device(mesh, state, nullptr);
  auto handles = rigid_handles(rigid_list);
  std::vector<ando_rigid_body> bodies;
  for (RigidBody* b : handles) bodies.push_back(b->abi());
  int32_t count = 0;
  abi_check(ando_compute_contacts(state.session->ctx, bodies.data(), (int32_t)bodies.size(), nullptr, 0, &count));
  std::vector<ando_contact> raw(std::max(count, 1));
  abi_check(ando_compute_contacts(state.session->ctx, bodies.data(), (int32_t)bodies.size(), raw.data(), count,
                                  &count));
  static uint64_t listings = 0;
  Session& sess = *state.session;
  sess.listing = ++listings;
  sess.listed = count;
  std::vector<ContactPair> cs(count);
  for (int i = 0; i < count; ++i) {
    const ando_contact& r = raw[i];
    ContactPair& c = cs[i];
    c.listing = sess.listing;
    c.slot = i;
    c.type = (ContactType)r.type;
    c.idx0 = r.idx[0]; c.idx1 = r.idx[1]; c.idx2 = r.idx[2]; c.idx3 = r.idx[3];
    c.gap = r.gap;
    c.normal = {r.normal[0], r.normal[1], r.normal[2]};
    c.witness_p = {r.witness_p[0], r.witness_p[1], r.witness_p[2]};
    c.witness_q = {r.witness_q[0], r.witness_q[1], r.witness_q[2]};
    c.vertex_count = r.vertex_count;
    c.rigid_body_index = r.rigid_body_index;
  }
  return cs;
}

}  // namespace

PYBIND11_MODULE(ando_barrier_core, m) {
  m.doc() = "Ando 2024 Cubic Barrier with Elasticity-Inclusive Dynamic Stiffness (MI355X HIP backend)";
  m.def("version", []() { return std::string("ando_barrier_core v1.0.0"); }, "Get version string");
  m.def("backend", []() { return std::string("hip-gfx950"); }, "Name of the compute backend");
  m.def("_pinned_pool_bytes", &pinned_pool_bytes,
        "(live, idle) bytes of page-locked buffers behind get_positions / get_velocities arrays (test hook)");

  py::class_<Material>(m, "Material")
      .def(py::init<>())
      .def(py::init([](py::kwargs kwargs) {
        Material mat;
        for (auto item : kwargs) {
          const std::string key = py::cast<std::string>(item.first);
          if (key == "youngs_modulus") mat.youngs_modulus = py::cast<float>(item.second);
          else if (key == "poisson_ratio") mat.poisson_ratio = py::cast<float>(item.second);
          else if (key == "density") mat.density = py::cast<float>(item.second);
          else if (key == "thickness") mat.thickness = py::cast<float>(item.second);
          else if (key == "bending_stiffness") mat.bending_stiffness = py::cast<float>(item.second);
          else throw py::attribute_error(("Unknown material property: " + key).c_str());
        }
        return mat;
      }))
      .def_readwrite("youngs_modulus", &Material::youngs_modulus)
      .def_readwrite("poisson_ratio", &Material::poisson_ratio)
      .def_readwrite("density", &Material::density)
      .def_readwrite("thickness", &Material::thickness)
      .def_readwrite("bending_stiffness", &Material::bending_stiffness);

  py::class_<SimParams>(m, "SimParams")
      .def(py::init<>())
      .def_readwrite("dt", &SimParams::dt)
      .def_readwrite("beta_max", &SimParams::beta_max)
      .def_readwrite("min_newton_steps", &SimParams::min_newton_steps)
      .def_readwrite("max_newton_steps", &SimParams::max_newton_steps)
      .def_readwrite("pcg_tol", &SimParams::pcg_tol)
      .def_readwrite("pcg_max_iters", &SimParams::pcg_max_iters)
      .def_readwrite("contact_gap_max", &SimParams::contact_gap_max)
      .def_readwrite("wall_gap", &SimParams::wall_gap)
      .def_readwrite("enable_ccd", &SimParams::enable_ccd)
      .def_readwrite("enable_friction", &SimParams::enable_friction)
      .def_readwrite("friction_mu", &SimParams::friction_mu)
      .def_readwrite("friction_epsilon", &SimParams::friction_epsilon)
      .def_readwrite("velocity_damping", &SimParams::velocity_damping)
      .def_readwrite("contact_restitution", &SimParams::contact_restitution)
      .def_readwrite("enable_strain_limiting", &SimParams::enable_strain_limiting)
      .def_readwrite("strain_limit", &SimParams::strain_limit)
      .def_readwrite("strain_tau", &SimParams::strain_tau);

  py::class_<Triangle>(m, "Triangle")
      .def(py::init<>())
      .def(py::init([](int32_t a, int32_t b, int32_t c) {
        Triangle t;
        t.v[0] = a; t.v[1] = b; t.v[2] = c;
        return t;
      }))
      .def_property(
          "vertices", [](const Triangle& t) { return std::vector<int32_t>{t.v[0], t.v[1], t.v[2]}; },
          [](Triangle& t, const std::vector<int32_t>& v) {
            t.v[0] = v.at(0); t.v[1] = v.at(1); t.v[2] = v.at(2);
          });

  py::class_<Mesh>(m, "Mesh")
      .def(py::init<>())
      .def(
          "initialize",
          [](Mesh& mesh, py::object vertices_obj, py::object triangles_obj, const Material& mat) {
            auto vertices = FArr::ensure(vertices_obj);
            if (!vertices || vertices.ndim() != 2 || vertices.shape(1) != 3)
              throw py::value_error("vertices must be an array of shape (N, 3)");
            std::vector<int32_t> tris;
            parse_triangles(triangles_obj, tris, "triangles must be a 1D or 2D array");
            mesh.vertices.assign(vertices.data(), vertices.data() + vertices.size());
            mesh.triangles = tris;
            mesh.material = mat;
            mesh.compute_rest_state();
          },
          py::arg("vertices"), py::arg("triangles"), py::arg("material"),
          "Initialize mesh from array-like vertex and triangle data")
      .def("num_vertices", &Mesh::num_vertices)
      .def("num_triangles", &Mesh::num_triangles)
      .def("get_vertices", [](const Mesh& mesh) { return nx3(mesh.vertices.data(), mesh.num_vertices()); })
      .def("set_positions",
           [](Mesh& mesh, FArr positions) {
             if (positions.ndim() != 2 || positions.shape(1) < 3) throw py::value_error("positions must be (N, 3)");
             auto p = positions.unchecked<2>();
             mesh.vertices.resize(3 * (size_t)p.shape(0));
             for (py::ssize_t i = 0; i < p.shape(0); ++i)
               for (int k = 0; k < 3; ++k) mesh.vertices[3 * i + k] = p(i, k);
             ++mesh.version;
           })
      .def_property(
          "vertices", [](const Mesh& mesh) { return nx3(mesh.vertices.data(), mesh.num_vertices()); },
          [](Mesh& mesh, FArr positions) {
            if (positions.ndim() != 2 || positions.shape(1) < 3) throw py::value_error("positions must be (N, 3)");
            auto p = positions.unchecked<2>();
            for (py::ssize_t i = 0; i < p.shape(0) && (size_t)i < mesh.num_vertices(); ++i)
              for (int k = 0; k < 3; ++k) mesh.vertices[3 * i + k] = p(i, k);
            ++mesh.version;
          })
      .def_property(
          "triangles",
          [](const Mesh& mesh) {
            py::array_t<int32_t> r({py::ssize_t(mesh.num_triangles()), py::ssize_t(3)});
            if (!mesh.triangles.empty())
              std::memcpy(r.mutable_data(), mesh.triangles.data(), mesh.triangles.size() * sizeof(int32_t));
            return r;
          },
          [](Mesh& mesh, py::object triangles_obj) {
            std::vector<int32_t> tris;
            parse_triangles(triangles_obj, tris, "triangles must be 1D or 2D array-like");
            mesh.triangles = tris;
            mesh.compute_rest_state();
          });

  py::class_<State>(m, "State")
      .def(py::init<>())
      .def("initialize", &State::initialize)
      .def("num_vertices", &State::num_vertices)
      .def("get_positions", [](State& s) { return s.read(true); })
      .def("get_velocities", [](State& s) { return s.read(false); })
      .def("set_velocities",
           [](State& s, FArr velocities) {
             if (velocities.ndim() != 2 || velocities.shape(1) < 3)
               throw py::value_error("velocities must be (N, 3)");
             s.pull();
             auto v = velocities.unchecked<2>();
             for (py::ssize_t i = 0; i < v.shape(0) && (size_t)i < s.num_vertices(); ++i)
               for (int k = 0; k < 3; ++k) s.velocities[3 * i + k] = v(i, k);
             s.host_dirty = true;
           })
      .def("get_masses",
           [](const State& s) {
             py::array_t<float> r({py::ssize_t(s.num_vertices())});
             if (!s.masses.empty()) std::memcpy(r.mutable_data(), s.masses.data(), s.masses.size() * sizeof(float));
             return r;
           })
      .def(
          "apply_gravity",
          [](State& s, py::object gravity_obj, float dt) {
            if (s.num_vertices() == 0) throw std::runtime_error("State has not been initialised");
            auto g = py::array_t<float, py::array::forcecast>::ensure(gravity_obj);
            if (!g || g.ndim() != 1 || g.shape(0) != 3) throw py::value_error("Gravity must be a 3D vector");
            float gv[3] = {g.at(0), g.at(1), g.at(2)};
            if (s.session && s.session->ctx && s.session->mesh) {
              // Resident state: one kernel, no transfer (bindings.cpp:284-300 semantics).
              if (s.host_dirty) {
                abi_check(ando_set_state(s.session->ctx, s.positions.data(), s.velocities.data(), s.masses.data(),
                                         (int32_t)s.num_vertices()));
                s.host_dirty = false;
              }
              abi_check(ando_apply_gravity(s.session->ctx, gv, dt));
              s.device_dirty = true;
              return;
            }
            float gd[3] = {gv[0] * dt, gv[1] * dt, gv[2] * dt};
            for (size_t i = 0; i < s.num_vertices(); ++i)
              for (int k = 0; k < 3; ++k) {
                s.velocities[3 * i + k] += gd[k];
                s.positions[3 * i + k] += s.velocities[3 * i + k] * dt;
              }
            s.host_dirty = true;
          },
          "Apply gravity acceleration to all vertices");

  py::class_<Constraints>(m, "Constraints")
      .def(py::init<>())
      .def("add_pin",
           [](Constraints& c, int32_t vidx, py::object target_obj) {
             V3 t = to_v3(target_obj, "pin target must be a 3D vector");
             c.pins.push_back({vidx, t, true});
             ++c.version;
           })
      .def(
          "add_wall",
          [](Constraints& c, py::object normal_obj, float offset, float gap) {
            V3 n = to_v3(normal_obj, "wall normal must be a 3D vector");
            c.walls.push_back({normalized(n), offset, gap, true});
            ++c.version;
          },
          py::arg("normal"), py::arg("offset"), py::arg("gap"))
      .def("num_active_pins", &Constraints::num_active_pins)
      .def("num_active_contacts", [](const Constraints& c) { return c.num_active_walls(); })
      .def_property_readonly("_pins",
                             [](const Constraints& c) {
                               py::dict d;
                               for (const Pin& p : c.pins)
                                 if (p.active) d[py::cast(p.vertex_idx)] = v3_array(p.target);
                               return d;
                             })
      .def_property_readonly("_walls", [](const Constraints& c) {
        py::list l;
        for (const Wall& w : c.walls) {
          if (!w.active) continue;
          double off = std::round((double)w.offset * 1e6) / 1e6;
          double gap = std::round((double)w.gap * 1e6) / 1e6;
          l.append(py::make_tuple(v3_array(w.normal), off, gap));
        }
        return l;
      });

  py::enum_<ContactType>(m, "ContactType")
      .value("POINT_TRIANGLE", ContactType::POINT_TRIANGLE)
      .value("EDGE_EDGE", ContactType::EDGE_EDGE)
      .value("WALL", ContactType::WALL)
      .value("RIGID_POINT_TRIANGLE", ContactType::RIGID_POINT_TRIANGLE)
      .export_values();

  py::class_<ContactPair>(m, "Contact")
      .def(py::init<>())
      .def_property_readonly("type", [](const ContactPair& c) { return c.type; })
      .def_property_readonly("idx0", [](const ContactPair& c) { return c.idx0; })
      .def_property_readonly("idx1", [](const ContactPair& c) { return c.idx1; })
      .def_property_readonly("idx2", [](const ContactPair& c) { return c.idx2; })
      .def_property_readonly("idx3", [](const ContactPair& c) { return c.idx3; })
      .def_property_readonly("rigid_body_index", [](const ContactPair& c) { return c.rigid_body_index; })
      .def_property_readonly("gap", [](const ContactPair& c) { return c.gap; })
      .def_property_readonly("normal", [](const ContactPair& c) { return v3_array(c.normal); })
      .def_property_readonly("witness_p", [](const ContactPair& c) { return v3_array(c.witness_p); })
      .def_property_readonly("witness_q", [](const ContactPair& c) { return v3_array(c.witness_q); });

  py::class_<RigidBody>(m, "RigidBody")
      .def(py::init<>())
      .def("initialize",
           [](RigidBody& b, FArr vertices, IArr triangles, float density) {
             if (vertices.size() && (vertices.ndim() != 2 || vertices.shape(1) != 3))
               throw py::value_error("vertices must be (N, 3)");
             if (triangles.size() && (triangles.ndim() != 2 || triangles.shape(1) != 3))
               throw py::value_error("triangles must be (M, 3)");
             std::vector<float> v(vertices.data(), vertices.data() + vertices.size());
             std::vector<int32_t> t(triangles.data(), triangles.data() + triangles.size());
             b.initialize(v, t, density);
           })
      .def_property(
          "position", [](const RigidBody& b) { return std::vector<float>{b.position.x, b.position.y, b.position.z}; },
          [](RigidBody& b, const std::vector<float>& p) { b.position = {p.at(0), p.at(1), p.at(2)}; })
      .def_property(
          "linear_velocity", [](const RigidBody& b) { return std::vector<float>{b.lin.x, b.lin.y, b.lin.z}; },
          [](RigidBody& b, const std::vector<float>& v) { b.lin = {v.at(0), v.at(1), v.at(2)}; })
      .def_property_readonly("mass", [](const RigidBody& b) { return b.mass; })
      .def("world_vertices",
           [](const RigidBody& b) {
             auto w = b.world_vertices();
             return nx3(w.data(), w.size() / 3);
           })
      .def("apply_impulse",
           [](RigidBody& b, py::object p, py::object j) {
             b.apply_impulse(to_v3(p, "world_point must be a 3D vector"), to_v3(j, "impulse must be a 3D vector"));
           })
      .def("integrate", &RigidBody::integrate);

  py::class_<ando::RestState>(m, "_RestState");

  struct ElasticityTag {};
  py::class_<ElasticityTag>(m, "Elasticity")
      .def(py::init<>())
      .def_static(
          "compute_energy", [](const Mesh& mesh, const State&) { return elastic_energy(mesh); },
          "Compute total elastic energy")
      .def_static(
          "compute_gradient",
          [](const Mesh& mesh, const State&, py::array_t<float> gradient) {
            auto g = gradient.mutable_unchecked<1>();
            std::vector<float> tmp((size_t)g.shape(0), 0.0f);
            elastic_gradient(mesh, tmp.data(), tmp.size());
            for (py::ssize_t i = 0; i < g.shape(0); ++i) g(i) = tmp[i];
          },
          "Compute elastic gradient (forces)")
      .def_static(
          "compute_hessian",
          [](const Mesh& mesh, const State&, py::list triplets) {
            // Appends (row, col, value) for every entry of every clamped 3x3
            // face block (elasticity.cpp:63-91).
            float mu = ando::shear_modulus(mesh.material);
            for (size_t f = 0; f < mesh.num_triangles(); ++f) {
              const float* M = &mesh.rest.dm_inv[4 * f];
              float k = mesh.rest.area[f] * mesh.material.thickness * mu;
              float G00 = M[0] * M[0] + M[2] * M[2], G01 = M[0] * M[1] + M[2] * M[3], G11 = M[1] * M[1] + M[3] * M[3];
              float K00 = k * G00, K01 = k * G01, K11 = k * G11;
              float c[9] = {(float)((double)(K00 + K11) + 2.0 * (double)K01), -K00, -K11, -K00, K00, -K01, -K11,
                            -K01, K11};
              for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                  float cc = c[3 * a + b];
                  float s = std::fabs(cc) == 0.0f ? 1.0f : std::fabs(cc);
                  float d = (cc / s) * s;
                  d = d < 1e-8f ? 1e-8f : d;
                  float vals[9] = {d, 0, 0, 0, d, 0, 0, 0, 1e-8f};
                  int ia = mesh.triangles[3 * f + a], ib = mesh.triangles[3 * f + b];
                  for (int e = 0; e < 9; ++e) triplets.append(py::make_tuple(ia * 3 + e / 3, ib * 3 + e % 3, vals[e]));
                }
            }
          },
          "Compute elastic Hessian (explicit assembly)");

  m.def("barrier_energy", &barrier_energy, "Compute cubic barrier energy V_weak(g, g_max, k)");
  m.def("barrier_gradient", &barrier_gradient, "Compute cubic barrier gradient w.r.t. g");
  m.def("barrier_hessian", &barrier_hessian, "Compute cubic barrier Hessian w.r.t. g");

  m.def(
      "create_mesh_from_blender",
      [](FArr vertices, IArr triangles, const Material& mat) {
        Mesh mesh;
        mesh.vertices.assign(vertices.data(), vertices.data() + vertices.size());
        mesh.triangles.assign(triangles.data(), triangles.data() + triangles.size());
        mesh.material = mat;
        mesh.compute_rest_state();
        return mesh;
      },
      "Create mesh from numpy arrays (vertices Nx3, triangles Mx3)");

  py::class_<ando_stats>(m, "SolverStats")
      .def_readonly("newton_iterations", &ando_stats::newton_iterations)
      .def_readonly("pcg_iterations", &ando_stats::pcg_iterations)
      .def_readonly("pcg_solves", &ando_stats::pcg_solves)
      .def_readonly("pcg_failures", &ando_stats::pcg_failures)
      .def_readonly("line_search_failures", &ando_stats::line_search_failures)
      .def_readonly("contacts", &ando_stats::contacts)
      .def_readonly("rigid_contacts", &ando_stats::rigid_contacts)
      .def_readonly("strain_active", &ando_stats::strain_active)
      .def_readonly("beta", &ando_stats::beta)
      .def_readonly("last_alpha", &ando_stats::last_alpha)
      .def_readonly("step_ms", &ando_stats::step_ms)
      .def_readonly("pcg_kernel", &ando_stats::pcg_kernel)
      .def_readonly("spmv_lanes", &ando_stats::spmv_lanes)
      .def_readonly("spmv_wg_slices", &ando_stats::spmv_wg_slices)
      .def_readonly("pcg_vec_block", &ando_stats::pcg_vec_block)
      .def_readonly("pcg_autotuned", &ando_stats::pcg_autotuned)
      .def_readonly("long_hit_sorts", &ando_stats::long_hit_sorts)
      .def_readonly("max_candidates", &ando_stats::max_candidates)
      .def_readonly("held_iterations", &ando_stats::held_iterations)
      .def_readonly("capped_trials", &ando_stats::capped_trials)
      .def_readonly("timed_coupling_bytes", &ando_stats::timed_coupling_bytes)
      .def_readonly("learned_solves", &ando_stats::learned_solves)
      .def_readonly("learned_fallbacks", &ando_stats::learned_fallbacks)
      .def_readonly("fallback_iterations", &ando_stats::fallback_iterations)
      .def_property_readonly("pcg_per_solve", [](const ando_stats& s) {
        int k = std::min(s.pcg_solves, (int32_t)ANDO_MAX_RECORDED_SOLVES);
        return std::vector<int32_t>(s.pcg_per_solve, s.pcg_per_solve + k);
      });

  struct IntegratorTag {};
  py::class_<IntegratorTag>(m, "Integrator")
      .def(py::init<>())
      .def_static(
          "step",
          [](Mesh& mesh, State& state, Constraints& constraints, const SimParams& params, py::object rigid_list) {
            sync_to_device(mesh, state, &constraints);
            auto handles = rigid_handles(rigid_list);
            std::vector<ando_rigid_body> bodies;
            for (RigidBody* b : handles) bodies.push_back(b->abi());
            ando_params p = params.abi();
            int rc;
            {
              py::gil_scoped_release release;
              rc = ando_step(state.session->ctx, &p, bodies.data(), (int32_t)bodies.size());
            }
            state.device_dirty = true;
            abi_check(rc);
            for (size_t i = 0; i < handles.size(); ++i) handles[i]->take(bodies[i]);
            abi_check(ando_get_stats(state.session->ctx, &g_last_stats));
          },
          py::arg("mesh"), py::arg("state"), py::arg("constraints"), py::arg("params"),
          py::arg("rigid_bodies") = py::none(), "Take one simulation step using Newton integrator with β accumulation")
      .def_static(
          "compute_contacts",
          [](const Mesh& mesh, State& state, py::object rigid_list) {
            return contacts_to_list(gpu_contacts(mesh, state, rigid_list));
          },
          py::arg("mesh"), py::arg("state"), py::arg("rigid_bodies") = py::none(),
          "Detect all collision contacts for the current mesh/state")
      .def_static(
          "last_stats", []() { return g_last_stats; }, "Solver statistics of the most recent step");

  py::class_<EnergyDiagnostics>(m, "EnergyDiagnostics")
      .def(py::init<>())
      .def_readonly("kinetic_energy", &EnergyDiagnostics::kinetic_energy)
      .def_readonly("elastic_energy", &EnergyDiagnostics::elastic_energy)
      .def_readonly("barrier_energy", &EnergyDiagnostics::barrier_energy)
      .def_readonly("total_energy", &EnergyDiagnostics::total_energy)
      .def_readonly("energy_drift_percent", &EnergyDiagnostics::energy_drift_percent)
      .def_readonly("energy_drift_absolute", &EnergyDiagnostics::energy_drift_absolute)
      .def_property_readonly("linear_momentum",
                             [](const EnergyDiagnostics& d) {
                               return std::vector<float>{d.linear_momentum.x, d.linear_momentum.y, d.linear_momentum.z};
                             })
      .def_property_readonly("angular_momentum",
                             [](const EnergyDiagnostics& d) {
                               return std::vector<float>{d.angular_momentum.x, d.angular_momentum.y,
                                                         d.angular_momentum.z};
                             })
      .def_readonly("max_velocity", &EnergyDiagnostics::max_velocity)
      .def_readonly("num_contacts", &EnergyDiagnostics::num_contacts)
      .def_readonly("num_pins", &EnergyDiagnostics::num_pins)
      .def("update_drift", &EnergyDiagnostics::update_drift);

  struct EnergyTrackerTag {};
  py::class_<EnergyTrackerTag>(m, "EnergyTracker")
      .def(py::init<>())
      .def_static(
          "compute",
          [](const Mesh& mesh, State& state, const Constraints& constraints, const SimParams&) {
            EnergyDiagnostics d;
            ando_energy e;
            if (gpu_energy(&mesh, state, e)) {
              d.kinetic_energy = (float)e.kinetic_energy;
              d.elastic_energy = (float)e.elastic_energy;
              d.linear_momentum = {(float)e.linear_momentum[0], (float)e.linear_momentum[1],
                                   (float)e.linear_momentum[2]};
              d.angular_momentum = {(float)e.angular_momentum[0], (float)e.angular_momentum[1],
                                    (float)e.angular_momentum[2]};
              d.max_velocity = (float)e.max_velocity;
            } else {
              state.pull();
              d.kinetic_energy = kinetic_energy(state);
              d.elastic_energy = elastic_energy(mesh);
              d.linear_momentum = linear_momentum(state);
              d.angular_momentum = angular_momentum(state);
              d.max_velocity = max_velocity(state);
            }
            d.barrier_energy = 0.0f;
            d.total_energy = (d.kinetic_energy + d.elastic_energy) + d.barrier_energy;
            d.num_contacts = 0;
            d.num_pins = (int)constraints.num_active_pins();
            return d;
          },
          py::arg("mesh"), py::arg("state"), py::arg("constraints"), py::arg("params"),
          "Compute comprehensive energy diagnostics")
      .def_static(
          "compute_kinetic_energy",
          [](State& s) {
            ando_energy e;
            if (gpu_energy(nullptr, s, e)) return (float)e.kinetic_energy;
            s.pull();
            return kinetic_energy(s);
          },
          "Compute kinetic energy")
      .def_static(
          "compute_linear_momentum",
          [](State& s) {
            ando_energy e;
            if (gpu_energy(nullptr, s, e))
              return std::vector<float>{(float)e.linear_momentum[0], (float)e.linear_momentum[1],
                                        (float)e.linear_momentum[2]};
            s.pull();
            V3 p = linear_momentum(s);
            return std::vector<float>{p.x, p.y, p.z};
          },
          "Compute linear momentum")
      .def_static(
          "compute_angular_momentum",
          [](State& s) {
            ando_energy e;
            if (gpu_energy(nullptr, s, e))
              return std::vector<float>{(float)e.angular_momentum[0], (float)e.angular_momentum[1],
                                        (float)e.angular_momentum[2]};
            s.pull();
            V3 p = angular_momentum(s);
            return std::vector<float>{p.x, p.y, p.z};
          },
          "Compute angular momentum")
      .def_static(
          "compute_max_velocity",
          [](State& s) {
            ando_energy e;
            if (gpu_energy(nullptr, s, e)) return (float)e.max_velocity;
            s.pull();
            return max_velocity(s);
          },
          "Compute maximum velocity");

  py::class_<CollisionMetrics>(m, "CollisionMetrics")
      .def(py::init<>())
      .def_readonly("num_point_triangle", &CollisionMetrics::num_point_triangle)
      .def_readonly("num_edge_edge", &CollisionMetrics::num_edge_edge)
      .def_readonly("num_wall", &CollisionMetrics::num_wall)
      .def_readonly("num_total_contacts", &CollisionMetrics::num_total_contacts)
      .def_readonly("min_gap", &CollisionMetrics::min_gap)
      .def_readonly("max_gap", &CollisionMetrics::max_gap)
      .def_readonly("avg_gap", &CollisionMetrics::avg_gap)
      .def_readonly("num_penetrations", &CollisionMetrics::num_penetrations)
      .def_readonly("max_penetration", &CollisionMetrics::max_penetration)
      .def_readonly("avg_penetration", &CollisionMetrics::avg_penetration)
      .def_readonly("ccd_enabled", &CollisionMetrics::ccd_enabled)
      .def_readonly("num_ccd_contacts", &CollisionMetrics::num_ccd_contacts)
      .def_readonly("num_broad_phase_contacts", &CollisionMetrics::num_broad_phase_contacts)
      .def_readonly("ccd_effectiveness", &CollisionMetrics::ccd_effectiveness)
      .def_readonly("max_relative_velocity", &CollisionMetrics::max_relative_velocity)
      .def_readonly("avg_relative_velocity", &CollisionMetrics::avg_relative_velocity)
      .def_readonly("has_tunneling", &CollisionMetrics::has_tunneling)
      .def_readonly("has_major_penetration", &CollisionMetrics::has_major_penetration)
      .def_readonly("is_stable", &CollisionMetrics::is_stable)
      .def("quality_level", &CollisionMetrics::quality_level)
      .def("quality_description", &CollisionMetrics::quality_description);

  struct ValidatorTag {};
  py::class_<ValidatorTag>(m, "CollisionValidator")
      .def(py::init<>())
      .def_static("compute_metrics", &compute_metrics, py::arg("mesh"), py::arg("state"), py::arg("contacts"),
                  py::arg("gap_max"), py::arg("ccd_enabled"), "Compute comprehensive collision metrics")
      .def_static(
          "has_penetrations",
          [](const std::vector<ContactPair>& cs) {
            for (const auto& c : cs)
              if (c.gap < 0.0f) return true;
            return false;
          },
          "Check if any contacts have penetrations")
      .def_static(
          "max_penetration_depth",
          [](const std::vector<ContactPair>& cs) {
            float mp = 0.0f;
            for (const auto& c : cs)
              if (c.gap < 0.0f) mp = std::max(mp, -c.gap);
            return mp;
          },
          "Get maximum penetration depth");

  struct AdaptiveTag {};
  py::class_<AdaptiveTag>(m, "AdaptiveTimestep")
      .def(py::init<>())
      .def_static(
          "compute_next_dt",
          [](py::object velocities, const Mesh& mesh, double current_dt, double dt_min, double dt_max,
             double safety) {
            float dt;
            if (py::isinstance<State>(velocities)) {
              // Additive: the State itself may be passed; when it is resident on
              // the device, both reductions run there and nothing is read back.
              State& st = velocities.cast<State&>();
              float vmax = 0.0f, me = 0.0f;
              if (device_resident(&mesh, st)) {
                abi_check(ando_cfl_inputs(st.session->ctx, &vmax, &me));
                dt = at_next_dt_from(vmax, [&] { return me; }, (float)current_dt, (float)dt_min, (float)dt_max,
                                     (float)safety);
              } else {
                st.pull();
                dt = at_next_dt(st.velocities, mesh, (float)current_dt, (float)dt_min, (float)dt_max, (float)safety);
              }
            } else {
              std::vector<float> v = flat_floats(velocities);
              dt = at_next_dt(v, mesh, (float)current_dt, (float)dt_min, (float)dt_max, (float)safety);
            }
            return std::clamp((double)dt, dt_min, dt_max);
          },
          py::arg("velocities"), py::arg("mesh"), py::arg("current_dt"), py::arg("dt_min"), py::arg("dt_max"),
          py::arg("safety") = 0.5, "Compute next timestep using CFL condition")
      .def_static(
          "compute_cfl_timestep",
          [](float max_velocity, float min_edge_length, float safety) {
            return at_cfl(max_velocity, min_edge_length, safety);
          },
          py::arg("max_velocity"), py::arg("min_edge_length"), py::arg("safety") = 0.5,
          "Compute CFL timestep from velocity and mesh resolution")
      .def_static("compute_min_edge_length", &at_min_edge, py::arg("mesh"), "Compute minimum edge length in mesh")
      .def_static(
          "compute_max_velocity",
          [](py::object velocities) {
            if (py::isinstance<State>(velocities)) {
              State& st = velocities.cast<State&>();
              if (device_resident(nullptr, st)) {
                float vmax = 0.0f, me = 0.0f;
                abi_check(ando_cfl_inputs(st.session->ctx, &vmax, &me));
                return vmax;
              }
              st.pull();
              return at_max_velocity(st.velocities);
            }
            return at_max_velocity(flat_floats(velocities));
          },
          py::arg("velocities"), "Compute maximum velocity magnitude");
}
