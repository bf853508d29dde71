// Structures shared by the cloth-step kernels and their host driver.
#pragma once

#include "device.h"

namespace ando {

constexpr int kBlock = 256;   // threads per workgroup (4 x wave64)
constexpr int kGroup = 16;    // default rigid triangles per culling group
constexpr int kMaxBodies = 16;

// One detected rigid contact (ContactPair of type RIGID_POINT_TRIANGLE,
// collision.h:78-103) plus its step-constant barrier stiffness k-bar.
struct RigidContact {
  float gap;
  float nx, ny, nz;
  float qx, qy, qz;
  float k;
  int32_t body;
  int32_t tri;
  int32_t vertex;
};

// A self contact of the corrected mode (ContactPair, collision.h:78-103):
// type 0 PT (idx = v, t0, t1, t2; weights 1, -b0, -b1, -b2) or 1 EE (idx =
// e1 ends, e2 ends; weights 1/2, 1/2, -1/2, -1/2).
struct SelfContact {
  int type;
  int idx[4];
  float gap;
  float nx, ny, nz;
  float w[4];
  float qx, qy, qz;
};

// Step-constant barrier terms of one self contact: in_domain flag, the
// normalised normal, V' and V'' at the detected gap with the contact's
// stiffness (Barrier::compute_gradient/hessian, barrier.cpp:29-65;
// Stiffness::compute_contact_stiffness, stiffness.cpp:53-116).
struct SelfTerm {
  int active;
  float dV, d2;
  float nx, ny, nz;
  float kb;  // contact stiffness (also the friction normal force scale)
};

// Self contacts seen from the vertices: for vertex i, entries
// [ptr[i], ptr[i+1]) of code hold contact * 4 + slot in contact order.
// pt_of (faithful mode on a one-triangle mesh, else null): for vertex i the
// PT contact whose point it is, or -1; such contacts also carry friction.
struct SelfView {
  const int32_t* ptr;
  const int32_t* code;
  const SelfContact* sc;
  const SelfTerm* st;
  const int32_t* pt_of;
};

// Rigid body dynamic state on the device (rigid_body.h:72-85).
struct BodyState {
  float pos[3];
  float rot[9];
  float lin[3];
  float ang[3];
  float Ibinv[9];
  float mass;
};

struct DetectBody {
  int32_t nv, nt;
  int32_t grp;           // triangles per culling group
  const float* world;    // (nv, 3)
  const int32_t* tris;   // (nt, 3)
  const float* gbox;     // (ceil(nt/grp), 6): min xyz, max xyz, grown by 0.01 + 1e-5
  const float* bbox;     // (6): union of gbox
};
struct DetectBodies {
  int32_t count;
  DetectBody b[kMaxBodies];
};

struct StepParams {
  float dt;
  float gap_max;
  float normal_eps;
  float barrier_tol;
  float min_gap;
  int32_t num_walls;
  int32_t strain;
  int32_t fp64_near;
  int32_t friction;        // enable_friction && friction_mu > 0
  float mu;
  float friction_eps;
  float friction_tangent;
  // Corrected physics (ANDO_MODE_CORRECTED_PHYSICS): contact stiffness on the
  // wall's inertial scale m / max(gap_max, min_gap)^2, and attracting pins of
  // stiffness pin_k m / dt^2.
  int32_t cp;
  float pin_k;
};

// Friction of one contact on its first vertex (integrator.cpp:28-66,
// friction.cpp:6-90): tangential slip since the state's initial positions,
// stiffness mu |k gap| / max(eps, |slip|) capped at 1e8.
__device__ inline bool friction_data(f3 n, float k, float gap, f3 disp, const StepParams& P, float& kf, f3& tan) {
  float nc = dot(disp, n);
  tan = disp - nc * n;
  float tn = nrm(tan);
  if (!(tn > P.friction_tangent)) return false;
  float nf = k * fabsf(gap);
  float den = fmaxf(P.friction_eps, fabsf(tn));
  if (den <= 0.0f) den = P.friction_eps;
  float s = P.mu * fabsf(nf) / den;
  kf = fminf(s, 1e8f);
  return true;
}
__device__ inline bool friction_data(const RigidContact& c, f3 disp, const StepParams& P, float& kf, f3& tan) {
  return friction_data(mk(c.nx, c.ny, c.nz), c.k, c.gap, disp, P, kf, tan);
}

// Index of entry (a, b) of a symmetric 3x3 stored as 00 01 02 11 12 22.
__host__ __device__ inline int sym6(int a, int b) {
  const int lo = a < b ? a : b, hi = a < b ? b : a;
  return lo == 0 ? hi : (lo == 1 ? 2 + hi : 5);
}

struct StrainParams {
  float tau, eps, min_gap, svd_eps;
  float thick, rho;
  int32_t fp64_near;
  int32_t cp;  // corrected physics: stiffness scale eps instead of max(|gap|, min_gap)
};

}  // namespace ando
