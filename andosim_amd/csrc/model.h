// Host-side data model shared by the HIP library and the ando_barrier_core
// module: rest state of a cloth mesh and lumped masses.
//
// Follows Mesh::compute_rest_state / compute_edges / build_topology
// (src/core/mesh.cpp:18-120) and State::compute_lumped_masses
// (src/core/state.cpp:19-33) of the reference, in float (the reference's Real).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <unordered_set>
#include <vector>

namespace ando {

struct Material {
  float youngs_modulus = 1e6f;
  float poisson_ratio = 0.3f;
  float density = 1000.0f;
  float thickness = 0.001f;
  float bending_stiffness = 0.0f;
};

namespace hm {
// Three-term reductions in the reference's unrolled order: e0 + (e1 + e2).
inline float dot3(const float* a, const float* b) { return a[0] * b[0] + (a[1] * b[1] + a[2] * b[2]); }
inline void cross3(const float* a, const float* b, float* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
}  // namespace hm

struct RestState {
  std::vector<int32_t> edges;      // 2 per edge, reference discovery order (mesh.cpp:77-98)
  std::vector<float> dm_inv;       // 4 per face: (0,0),(0,1),(1,0),(1,1)
  std::vector<float> area;         // per face
};

// mesh.cpp:18-98.  Degenerate faces get zero area and zero Dm_inv.
inline RestState compute_rest_state(const float* verts, int32_t n, const int32_t* tris, int32_t m) {
  RestState rs;
  std::unordered_set<uint64_t> seen;
  seen.reserve(size_t(m) * 2);
  for (int32_t f = 0; f < m; ++f)
    for (int i = 0; i < 3; ++i) {
      int32_t a = tris[3 * f + i], b = tris[3 * f + (i + 1) % 3];
      if (a < 0 || b < 0 || a >= n || b >= n) throw std::invalid_argument("triangle index out of range");
      int32_t lo = std::min(a, b), hi = std::max(a, b);
      uint64_t key = (uint64_t(uint32_t(lo)) << 32) | uint32_t(hi);
      if (seen.insert(key).second) {
        rs.edges.push_back(lo);
        rs.edges.push_back(hi);
      }
    }
  rs.dm_inv.assign(size_t(4) * m, 0.0f);
  rs.area.assign(m, 0.0f);
  for (int32_t f = 0; f < m; ++f) {
    const float* v0 = verts + 3 * tris[3 * f];
    const float* v1 = verts + 3 * tris[3 * f + 1];
    const float* v2 = verts + 3 * tris[3 * f + 2];
    float e1[3], e2[3];
    for (int k = 0; k < 3; ++k) { e1[k] = v1[k] - v0[k]; e2[k] = v2[k] - v0[k]; }
    const float eps = 1e-12f;
    float l1 = hm::dot3(e1, e1), l2 = hm::dot3(e2, e2);
    if (l1 < eps || l2 < eps) continue;
    float nn[3];
    hm::cross3(e1, e2, nn);
    float nl = std::sqrt(hm::dot3(nn, nn));
    if (nl < eps) continue;
    rs.area[f] = nl * 0.5f;
    float nu[3], t1[3], t2[3];
    float sl1 = std::sqrt(l1);
    for (int k = 0; k < 3; ++k) { nu[k] = nn[k] / nl; t1[k] = e1[k] / sl1; }
    hm::cross3(nu, t1, t2);
    float a00 = hm::dot3(e1, t1), a10 = hm::dot3(e1, t2), a01 = hm::dot3(e2, t1), a11 = hm::dot3(e2, t2);
    float invdet = 1.0f / (a00 * a11 - a10 * a01);
    rs.dm_inv[4 * f + 0] = a11 * invdet;
    rs.dm_inv[4 * f + 1] = -a01 * invdet;
    rs.dm_inv[4 * f + 2] = -a10 * invdet;
    rs.dm_inv[4 * f + 3] = a00 * invdet;
  }
  return rs;
}

// state.cpp:19-33: each face's mass split in thirds, accumulated in face order.
inline std::vector<float> lumped_masses(int32_t n, const int32_t* tris, int32_t m, const std::vector<float>& area,
                                        const Material& mat) {
  std::vector<float> masses(n, 0.0f);
  for (int32_t f = 0; f < m; ++f) {
    float face_mass = area[f] * mat.thickness * mat.density;
    float vm = (float)((double)face_mass / 3.0);
    for (int j = 0; j < 3; ++j) masses[tris[3 * f + j]] += vm;
  }
  return masses;
}

// mu = E / (2 (1 + nu)) evaluated in double, as the reference's literals 2.0
// and 1.0 promote it (elasticity.cpp:25).
inline float shear_modulus(const Material& mat) {
  return (float)((double)mat.youngs_modulus / (2.0 * (1.0 + (double)mat.poisson_ratio)));
}

}  // namespace ando
