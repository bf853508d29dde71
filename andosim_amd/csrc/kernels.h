// Structures shared by the cloth-step kernels and their host driver.
#pragma once

#include "device.h"

namespace ando {

constexpr int kBlock = 256;   // threads per workgroup (4 x wave64)
constexpr int kGroup = 32;    // rigid triangles per culling group
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
};

struct DetectBody {
  int32_t nv, nt;
  const float* world;    // (nv, 3)
  const int32_t* tris;   // (nt, 3)
  const float* gbox;     // (ceil(nt/kGroup), 6): min xyz, max xyz, grown by 0.01 + 1e-5
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
};

struct StrainParams {
  float tau, eps, min_gap, svd_eps;
  float thick, rho;
  int32_t fp64_near;
};

}  // namespace ando
