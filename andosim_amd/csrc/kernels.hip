// gfx950 kernels of the per-step cloth solve.
//
// Layout in HBM (one context, n vertices, m faces, nb 3x3 blocks):
//   x, v, x_old, x_hat, g, d        3n f32, vertex-major xyz (the reference's flat VecX)
//   Hconst, H                       9 nb f32, row-major 3x3 blocks, block rows sorted by column
//   Helast3                         3 nb f32, diagonal of the elastic-only blocks
//   row_ptr (n+1), cols (nb)        block sparsity: diagonal + every mesh edge, both directions
//   blk_cptr / blk_contrib          per block, the (face, a, b) slots that feed it, ascending face
//   v2f_ptr / v2f                   per vertex, the (face, a) slots that feed it, ascending face
// Every "gather" kernel sums its contributions in that stored order, which is
// the order in which the reference appends triplets, so no atomics are needed
// and results are deterministic.
#include "kernels.h"

namespace ando {

// Elasticity::compute_gradient for one face (elasticity.cpp:18-61): F = Ds
// Dm_inv in the local frame (t1 along e1, t2 = n x t1) of the vertices given,
// P = 2k (F - I), forces of P Dm_inv^T mapped back through the frame; writes
// the three vertex contributions to out[0..8] and the unit normal to nn.
__device__ inline void membrane_grad(f3 v0, f3 v1, f3 v2, float M00, float M01, float M10, float M11, float k,
                                     float* __restrict__ out, f3& nn) {
  f3 e1 = v1 - v0, e2 = v2 - v0;
  nn = normalized(cross(e1, e2));
  f3 t1 = normalized(e1);
  f3 t2 = cross(nn, t1);
  float D00 = dot(e1, t1), D01 = dot(e2, t1), D10 = dot(e1, t2), D11 = dot(e2, t2);
  float F00 = D00 * M00 + D01 * M10, F01 = D00 * M01 + D01 * M11;
  float F10 = D10 * M00 + D11 * M10, F11 = D10 * M01 + D11 * M11;
  float twok = (float)(2.0 * (double)k);
  float P00 = twok * (F00 - 1.0f), P01 = twok * (F01 - 0.0f), P10 = twok * (F10 - 0.0f), P11 = twok * (F11 - 1.0f);
  float H00 = P00 * M00 + P01 * M01, H01 = P00 * M10 + P01 * M11;
  float H10 = P10 * M00 + P11 * M01, H11 = P10 * M10 + P11 * M11;
  f3 f1 = H00 * t1 + H10 * t2;
  f3 f2 = H01 * t1 + H11 * t2;
  f3 f0 = -(f1 + f2);
  st3(out, 0, f0);
  st3(out, 1, f1);
  st3(out, 2, f2);
}

// ---------------------------------------------------------------------------
// Setup: per-face elastic constants and the rest-state elastic gradient.
// Elasticity::face_hessian + enforce_spd (elasticity.cpp:63-91,135-181) and
// Elasticity::compute_gradient (elasticity.cpp:18-61) with F from the mesh
// buffer, which the reference never updates during a step.
// ---------------------------------------------------------------------------
__global__ void k_elastic_face(int m, const int32_t* __restrict__ tris, const float* __restrict__ mverts,
                               const float* __restrict__ dm, const float* __restrict__ area, float thick, float mu,
                               float* __restrict__ ec, float* __restrict__ fe) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= m) return;
  const float M00 = dm[4 * f], M01 = dm[4 * f + 1], M10 = dm[4 * f + 2], M11 = dm[4 * f + 3];
  const float k = area[f] * thick * mu;
  // K = k * (Dm_inv^T Dm_inv)
  float G00 = M00 * M00 + M10 * M10, G01 = M00 * M01 + M10 * M11, G11 = M01 * M01 + M11 * M11;
  float K00 = k * G00, K01 = k * G01, K11 = k * G11;
  float c[9];
  c[0] = (float)((double)(K00 + K11) + 2.0 * (double)K01);
  c[4] = K00;
  c[8] = K11;
  c[1] = c[3] = -K00;
  c[2] = c[6] = -K11;
  c[5] = c[7] = -K01;
  const float eps = 1e-8f;
  for (int i = 0; i < 9; ++i) {
    float s = fabsf(c[i]);
    if (s == 0.0f) s = 1.0f;
    float d = (c[i] / s) * s;
    ec[9 * f + i] = d < eps ? eps : d;
  }
  // Rest-state gradient: F = Ds Dm_inv in the face's local frame.
  int i0 = tris[3 * f], i1 = tris[3 * f + 1], i2 = tris[3 * f + 2];
  f3 nn;
  membrane_grad(ld3(mverts, i0), ld3(mverts, i1), ld3(mverts, i2), M00, M01, M10, M11, k, fe + 9 * (size_t)f, nn);
}

// egrad[i] = sum of the face forces on vertex i in face order.
__global__ void k_egrad_gather(int n, const int32_t* __restrict__ v2f_ptr, const int32_t* __restrict__ v2f,
                               const float* __restrict__ fe, float* __restrict__ egrad) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  f3 g = mk(0.0f, 0.0f, 0.0f);
  for (int e = v2f_ptr[i]; e < v2f_ptr[i + 1]; ++e) g = g + ld3(fe, v2f[e]);  // v2f holds 3*face + a
  st3(egrad, i, g);
}

// Mass/dt^2 followed by the elastic blocks, folded in triplet order
// (integrator.cpp:287-304; matrix_assembly.cpp:27-61).  Helast3 keeps the
// elastic-only diagonal (H_elastic).
__global__ void k_fold_const(int n, const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ cols,
                             const int32_t* __restrict__ blk_cptr, const int32_t* __restrict__ blk_contrib,
                             const float* __restrict__ ec, const float* __restrict__ masses, float dtinv2,
                             float* __restrict__ Hconst, float* __restrict__ Helast3) {
  int i = blockIdx.x;  // one block row per workgroup, one block per thread
  if (i >= n) return;
  for (int b = row_ptr[i] + threadIdx.x; b < row_ptr[i + 1]; b += blockDim.x) {
    int j = cols[b];
    float hx = 0.0f, hy = 0.0f, hz = 0.0f;
    if (i == j) {
      float mf = masses[i] * dtinv2;
      if (mf != 0.0f) { hx = hx + mf; hy = hy + mf; hz = hz + mf; }
    }
    float ex = 0.0f, ez = 0.0f;
    for (int e = blk_cptr[b]; e < blk_cptr[b + 1]; ++e) {
      float c = ec[blk_contrib[e]];
      hx = hx + c; hy = hy + c; hz = hz + 1e-8f;
      ex = ex + c; ez = ez + 1e-8f;
    }
    float* o = Hconst + 9 * (size_t)b;
    o[0] = hx; o[1] = 0.0f; o[2] = 0.0f;
    o[3] = 0.0f; o[4] = hy; o[5] = 0.0f;
    o[6] = 0.0f; o[7] = 0.0f; o[8] = hz;
    Helast3[3 * (size_t)b] = ex;
    Helast3[3 * (size_t)b + 1] = ex;
    Helast3[3 * (size_t)b + 2] = ez;
  }
}

// Stiffness::compute_wall_stiffness (stiffness.cpp:146-172) per vertex and wall;
// constant within a step (mass and the mass+elastic diagonal block only).
__global__ void k_wall_stiffness(int n, int nw, const int32_t* __restrict__ diag, const float* __restrict__ Hconst,
                                 const float* __restrict__ masses, const float* __restrict__ wall_n, float wall_gap,
                                 float min_gap, float* __restrict__ kw) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* B = Hconst + 9 * (size_t)diag[i];
  f3 Hd = spd_diag(mk(B[0], B[4], B[8]), 1e-8f);
  float gh = fmaxf(fmaxf(wall_gap, min_gap), 1e-12f);
  float kin = masses[i] / (gh * gh);
  for (int w = 0; w < nw; ++w) {
    f3 nn = ld3(wall_n, w);
    float nl = nrm(nn);
    if ((double)nl > 1e-9) nn = nn / nl; else nn = mk(0.0f, 1.0f, 0.0f);
    float kel = fmaxf(0.0f, quad_diag(Hd, nn));
    kw[(size_t)w * n + i] = kin + kel;
  }
}

// ---------------------------------------------------------------------------
// Step prologue / epilogue.
// ---------------------------------------------------------------------------
// State.apply_gravity (bindings.cpp:284-300): v += g dt, x += v dt.  With
// move_x = 0 (corrected physics) only v changes; the step's prediction
// x_hat = x + dt v carries the position update.
__global__ void k_gravity(int n, float gx, float gy, float gz, float dt, float* __restrict__ x,
                          float* __restrict__ v, int move_x) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  f3 gd = mk(gx * dt, gy * dt, gz * dt);
  f3 vi = ld3(v, i) + gd;
  st3(v, i, vi);
  if (move_x) st3(x, i, ld3(x, i) + dt * vi);
}

// x_old = x; x_hat = x + dt v (integrator.cpp:78-92).
__global__ void k_predict(int n3, float dt, const float* __restrict__ x, const float* __restrict__ v,
                          float* __restrict__ xold, float* __restrict__ xhat) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n3) return;
  float xi = x[i];
  xold[i] = xi;
  xhat[i] = xi + dt * v[i];
}

// v = (x - x_old) / (beta dt), then optional damping (integrator.cpp:124-140,557-563).
__global__ void k_velocity(int n3, float bdt, float damp_scale, int damp, const float* __restrict__ x,
                           const float* __restrict__ xold, float* __restrict__ v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n3) return;
  float vi = (x[i] - xold[i]) / bdt;
  if (damp) vi *= damp_scale;
  v[i] = vi;
}

// x <- x + s d with s = alpha * 1.25 (integrator.cpp:226-228).
__global__ void k_axpy(int n3, float s, const float* __restrict__ d, float* __restrict__ x) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n3) return;
  x[i] = x[i] + s * d[i];
}

// ---------------------------------------------------------------------------
// Rigid-body contact detection (collision.cpp:536-580): every cloth vertex
// against every rigid triangle, kept if the unsigned distance < 0.01.  Groups
// of consecutive triangles carry an AABB grown by 0.01 + 1e-5, so culling
// never drops a pair the distance test would keep; the surviving pairs are
// visited in (body, triangle) order, the reference's order.
// ---------------------------------------------------------------------------
__global__ void k_rigid_world(int nv, const float* __restrict__ local, const float* __restrict__ pose,
                              float* __restrict__ world) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  // pose: position[3], rotation[9] row-major.  world = p + R v (rigid_body.cpp:93-100).
  f3 v = ld3(local, i);
  f3 Rv = mk(pose[3] * v.x + (pose[4] * v.y + pose[5] * v.z), pose[6] * v.x + (pose[7] * v.y + pose[8] * v.z),
             pose[9] * v.x + (pose[10] * v.y + pose[11] * v.z));
  st3(world, i, mk(pose[0], pose[1], pose[2]) + Rv);
}

__global__ void k_rigid_groups(int nt, int grp, const int32_t* __restrict__ tris, const float* __restrict__ world,
                               float* __restrict__ gbox) {
  int gi = blockIdx.x * blockDim.x + threadIdx.x;
  int ng = (nt + grp - 1) / grp;
  if (gi >= ng) return;
  float mn[3] = {3.4e38f, 3.4e38f, 3.4e38f}, mx[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
  for (int t = gi * grp; t < min(nt, (gi + 1) * grp); ++t)
    for (int c = 0; c < 3; ++c) {
      f3 p = ld3(world, tris[3 * t + c]);
      mn[0] = fminf(mn[0], p.x); mn[1] = fminf(mn[1], p.y); mn[2] = fminf(mn[2], p.z);
      mx[0] = fmaxf(mx[0], p.x); mx[1] = fmaxf(mx[1], p.y); mx[2] = fmaxf(mx[2], p.z);
    }
  const float r = 0.01f + 1e-5f;
  for (int k = 0; k < 3; ++k) {
    gbox[6 * gi + k] = mn[k] - r;
    gbox[6 * gi + 3 + k] = mx[k] + r;
  }
}

// Contact stiffness of a one-vertex rigid contact (stiffness.cpp:53-116):
// m/s^2 + max(0, n^T E n), E the elastic-only diagonal block of the vertex,
// s = dt (the reference) or max(gap_max, min_gap) (corrected physics).
__device__ inline float rigid_contact_k(float mass, f3 E, f3 normal, float dt, float s) {
  if (dt <= 0.0f) return 0.0f;
  float nl = nrm(normal);
  if ((double)nl < 1e-12) return 0.0f;
  f3 nn = normal / nl;
  float inertial = 0.0f + mass * 1.0f * 1.0f;
  inertial /= (s * s);
  float h0 = E.x * nn.x, h1 = E.y * nn.y, h2 = E.z * nn.z;
  float el = nn.x * h0;
  el = el + nn.y * h1;
  el = el + nn.z * h2;
  if (el < 0.0f) el = 0.0f;
  return inertial + el;
}

// Whole-body box (union of the grown group boxes), one workgroup.
__global__ void k_rigid_bbox(int ng, const float* __restrict__ gbox, float* __restrict__ bbox) {
  __shared__ float lds[4];
  float mn[3] = {3.4e38f, 3.4e38f, 3.4e38f}, mx[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
  for (int g = threadIdx.x; g < ng; g += blockDim.x)
    for (int k = 0; k < 3; ++k) {
      mn[k] = fminf(mn[k], gbox[6 * g + k]);
      mx[k] = fmaxf(mx[k], gbox[6 * g + 3 + k]);
    }
  for (int k = 0; k < 3; ++k) {
    float a = block_reduce(mn[k], lds, [](float u, float v) { return fminf(u, v); });
    float b = block_reduce(mx[k], lds, [](float u, float v) { return fmaxf(u, v); });
    if (threadIdx.x == 0) {
      bbox[k] = a;
      bbox[3 + k] = b;
    }
  }
}

// Pass 0 counts, pass 1 writes records at ptr[i] in reference order.  Group
// boxes are staged through LDS 256 at a time; a workgroup skips a body when
// none of its 256 vertices lies in the body's grown box.  Pass 0 also keeps
// the (body, triangle) of a vertex's first kHitCache hits in `hits`; pass 1
// re-evaluates only those when the vertex had no more, instead of scanning
// again (the same arithmetic on the same inputs, so the same records).
constexpr int kGroupTile = 256;
constexpr int kHitCache = 16;

__device__ inline void write_rigid_contact(RigidContact& o, const PTResult& r, float mass, f3 E, float dt, int b,
                                           int t, int i) {
  o.gap = r.dist;
  o.nx = r.normal.x; o.ny = r.normal.y; o.nz = r.normal.z;
  o.qx = r.wq.x; o.qy = r.wq.y; o.qz = r.wq.z;
  o.k = rigid_contact_k(mass, E, r.normal, dt, dt);
  o.body = b;
  o.tri = t;
  o.vertex = i;
}

__global__ void __launch_bounds__(kBlock) k_rigid_detect(int n, int pass, const float* __restrict__ x,
                                                         DetectBodies bodies, const int32_t* __restrict__ ptr,
                                                         int32_t* __restrict__ count, RigidContact* __restrict__ out,
                                                         const float* __restrict__ masses,
                                                         const int32_t* __restrict__ diag,
                                                         const float* __restrict__ Helast3, float dt,
                                                         int32_t* __restrict__ hits) {
  __shared__ float sbox[6 * kGroupTile];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < n;
  f3 p = live ? ld3(x, i) : mk(0.0f, 0.0f, 0.0f);
  int c = 0;
  int base = (pass && live) ? ptr[i] : 0;
  f3 E = mk(0.0f, 0.0f, 0.0f);
  float mass = 0.0f;
  if (pass && live) {
    const float* e = Helast3 + 3 * (size_t)diag[i];
    E = mk(e[0], e[1], e[2]);
    mass = masses[i];
  }
  // Pass 1 of a vertex whose hits all fit the cache: only those triangles.
  const int cached = (pass && live) ? count[i] : kHitCache + 1;
  if (pass && cached <= kHitCache) {
    for (int h = 0; h < cached; ++h) {
      const int code = hits[(size_t)i * kHitCache + h];
      const int b = code >> 24, t = code & 0xffffff;
      const DetectBody& B = bodies.b[b];
      f3 a = ld3(B.world, B.tris[3 * t]), bq = ld3(B.world, B.tris[3 * t + 1]);
      f3 cq = ld3(B.world, B.tris[3 * t + 2]);
      PTResult r = point_triangle(p, a, bq, cq);
      write_rigid_contact(out[base + h], r, mass, E, dt, b, t, i);
    }
  }
  // Every thread of the workgroup takes part in the staging below; a vertex
  // served from the cache only skips its own tests.
  const bool scan = live && !(pass && cached <= kHitCache);
  for (int b = 0; b < bodies.count; ++b) {
    const DetectBody& B = bodies.b[b];
    const float* bb0 = B.bbox;
    bool in_body = scan && !(p.x < bb0[0] || p.y < bb0[1] || p.z < bb0[2] || p.x > bb0[3] || p.y > bb0[4] ||
                             p.z > bb0[5]);
    if (!__syncthreads_or(in_body)) continue;
    const int ng = (B.nt + B.grp - 1) / B.grp;
    for (int g0 = 0; g0 < ng; g0 += kGroupTile) {
      const int gn = min(kGroupTile, ng - g0);
      for (int k = threadIdx.x; k < 6 * gn; k += blockDim.x) sbox[k] = B.gbox[6 * g0 + k];
      __syncthreads();
      if (in_body)
        for (int gl = 0; gl < gn; ++gl) {
          const float* bx = sbox + 6 * gl;
          if (p.x < bx[0] || p.y < bx[1] || p.z < bx[2] || p.x > bx[3] || p.y > bx[4] || p.z > bx[5]) continue;
          const int gi = g0 + gl;
          const int t1 = min(B.nt, (gi + 1) * B.grp);
          for (int t = gi * B.grp; t < t1; ++t) {
            f3 a = ld3(B.world, B.tris[3 * t]), bq = ld3(B.world, B.tris[3 * t + 1]);
            f3 cq = ld3(B.world, B.tris[3 * t + 2]);
            // The triangle's own box, grown like the group boxes: a vertex
            // outside it is farther than 0.01 from the triangle.
            const float gr = 0.01f + 1e-5f;
            if (p.x < fminf(a.x, fminf(bq.x, cq.x)) - gr || p.x > fmaxf(a.x, fmaxf(bq.x, cq.x)) + gr ||
                p.y < fminf(a.y, fminf(bq.y, cq.y)) - gr || p.y > fmaxf(a.y, fmaxf(bq.y, cq.y)) + gr ||
                p.z < fminf(a.z, fminf(bq.z, cq.z)) - gr || p.z > fmaxf(a.z, fmaxf(bq.z, cq.z)) + gr)
              continue;
            PTResult r = point_triangle(p, a, bq, cq);
            if (r.dist < 0.01f) {
              if (pass) write_rigid_contact(out[base + c], r, mass, E, dt, b, t, i);
              else if (c < kHitCache) hits[(size_t)i * kHitCache + c] = (b << 24) | t;
              ++c;
            }
          }
        }
      __syncthreads();
    }
  }
  if (!pass && live) count[i] = c;
}

// ---------------------------------------------------------------------------
// Integrator::apply_rigid_coupling's force and torque sums (integrator.cpp:
// 676-754): each rigid contact at the final positions pushes its body with
// f = -V'(g) n (n normalised; skipped if |n| < 1e-8, |V'| < 1e-12 or the body
// has no mass) at the witness point.  Per-contact terms are the reference's
// float expressions; each workgroup sums them per body in double (fixed tree
// order), and k_rigid_forces_final adds the workgroup partials in order.
// ---------------------------------------------------------------------------
struct BodyFrames {
  float pos[kMaxBodies][3];
  float mass[kMaxBodies];
};

struct Wrench {
  double v[6];
  __device__ void add(const Wrench& o) {
    for (int k = 0; k < 6; ++k) v[k] += o.v[k];
  }
};

// sign: -1 for the reference's -V' n, +1 (corrected physics) for V' n, which
// pushes the body away from the cloth.
__global__ void __launch_bounds__(kBlock) k_rigid_forces(int count, const RigidContact* __restrict__ rc, int nbod,
                                                         BodyFrames bf, float gmax, float sign,
                                                         double* __restrict__ part) {
  __shared__ Wrench lds[kBlock];
  for (int b = 0; b < nbod; ++b) {
    Wrench a;
    for (int k = 0; k < 6; ++k) a.v[k] = 0.0;
    for (int e = blockIdx.x * kBlock + threadIdx.x; e < count; e += gridDim.x * kBlock) {
      const RigidContact& r = rc[e];
      if (r.body != b) continue;
      float nv[3] = {r.nx, r.ny, r.nz};
      const float nl = sqrtf(nv[0] * nv[0] + (nv[1] * nv[1] + nv[2] * nv[2]));
      if ((double)nl < 1e-8) continue;
      for (int k = 0; k < 3; ++k) nv[k] = nv[k] / nl;
      float dV = 0.0f;
      if (!(gmax <= 0.0f || r.gap >= gmax)) {
        const float dd = gmax - r.gap;
        const float inv = 1.0f / gmax;
        dV = -1.5f * r.k * inv * dd * dd;
      }
      if ((double)fabsf(dV) < 1e-12) continue;
      if (bf.mass[b] <= 0.0f) continue;
      const float f[3] = {sign * (dV * nv[0]), sign * (dV * nv[1]), sign * (dV * nv[2])};
      const float rr[3] = {r.qx - bf.pos[b][0], r.qy - bf.pos[b][1], r.qz - bf.pos[b][2]};
      for (int k = 0; k < 3; ++k) a.v[k] += (double)f[k];
      a.v[3] += (double)(rr[1] * f[2] - rr[2] * f[1]);
      a.v[4] += (double)(rr[2] * f[0] - rr[0] * f[2]);
      a.v[5] += (double)(rr[0] * f[1] - rr[1] * f[0]);
    }
    const int t = threadIdx.x;
    lds[t] = a;
    __syncthreads();
    for (int sft = kBlock / 2; sft > 0; sft >>= 1) {
      if (t < sft) lds[t].add(lds[t + sft]);
      __syncthreads();
    }
    if (t < 6) part[((size_t)blockIdx.x * nbod + b) * 6 + t] = lds[0].v[t];
    __syncthreads();
  }
}

// out[6 b + k] = sum over workgroups g (in order) of part[(g nbod + b) 6 + k].
__global__ void k_rigid_forces_final(int G, int nbod, const double* __restrict__ part, double* __restrict__ out) {
  const int t = threadIdx.x;
  if (t >= 6 * nbod) return;
  const int b = t / 6, k = t % 6;
  double s = 0.0;
  for (int g = 0; g < G; ++g) s += part[((size_t)g * nbod + b) * 6 + k];
  out[t] = s;
}

// RigidBody::integrate (rigid_body.cpp:151-176), one thread per body, with the
// force and torque of k_rigid_forces_final (w[6 b .. 6 b + 5], rounded to
// float as the reference's Vec3 accumulators): semi-implicit Euler on the
// velocities, then the position, then the rotation premultiplied by the
// axis-angle matrix of omega dt (AngleAxis::toRotationMatrix's formula).  The
// float expressions and their association are the host restatement's; sin and
// cos are evaluated in double and rounded once.  Bodies with mass <= 0 do not
// move.
__global__ void k_rigid_integrate(int nbod, const double* __restrict__ w, BodyState* __restrict__ bs, float dt) {
  const int b = threadIdx.x;
  if (b >= nbod) return;
  BodyState B = bs[b];
  if (B.mass <= 0.0f) return;
  auto mul33 = [](const float* A, const float* Bm, float* O) {
    for (int r = 0; r < 3; ++r)
      for (int cc = 0; cc < 3; ++cc)
        O[3 * r + cc] = A[3 * r] * Bm[cc] + (A[3 * r + 1] * Bm[3 + cc] + A[3 * r + 2] * Bm[6 + cc]);
  };
  float F[3], T[3];
  for (int k = 0; k < 3; ++k) {
    F[k] = (float)w[6 * b + k];
    T[k] = (float)w[6 * b + 3 + k];
  }
  for (int k = 0; k < 3; ++k) B.lin[k] = B.lin[k] + dt * (F[k] / B.mass);
  float RI[9], Rt[9], Iw[9];
  mul33(B.rot, B.Ibinv, RI);
  for (int r = 0; r < 3; ++r)
    for (int cc = 0; cc < 3; ++cc) Rt[3 * r + cc] = B.rot[3 * cc + r];
  mul33(RI, Rt, Iw);
  for (int k = 0; k < 3; ++k) {
    const float aa = Iw[3 * k] * T[0] + (Iw[3 * k + 1] * T[1] + Iw[3 * k + 2] * T[2]);
    B.ang[k] = B.ang[k] + dt * aa;
  }
  for (int k = 0; k < 3; ++k) B.pos[k] = B.pos[k] + dt * B.lin[k];
  const float* om3 = B.ang;
  const float om = sqrtf(om3[0] * om3[0] + (om3[1] * om3[1] + om3[2] * om3[2]));
  if ((double)om > 1e-8) {
    const float th = om * dt;
    const float ax[3] = {om3[0] / om, om3[1] / om, om3[2] / om};
    const float sn = (float)sin((double)th), cs = (float)cos((double)th);
    float sa[3], ca[3];
    for (int k = 0; k < 3; ++k) {
      sa[k] = sn * ax[k];
      ca[k] = (1.0f - cs) * ax[k];
    }
    float Rm[9];
    float tmp = ca[0] * ax[1];
    Rm[1] = tmp - sa[2];
    Rm[3] = tmp + sa[2];
    tmp = ca[0] * ax[2];
    Rm[2] = tmp + sa[1];
    Rm[6] = tmp - sa[1];
    tmp = ca[1] * ax[2];
    Rm[5] = tmp - sa[0];
    Rm[7] = tmp + sa[0];
    for (int k = 0; k < 3; ++k) Rm[4 * k] = ca[k] * ax[k] + cs;
    float Rn[9];
    mul33(Rm, B.rot, Rn);
    for (int e = 0; e < 9; ++e) B.rot[e] = Rn[e];
  }
  bs[b] = B;
}

// ---------------------------------------------------------------------------
// Exclusive scan of int32 counts (three phases, 1024 items per workgroup).
// ---------------------------------------------------------------------------
// Exclusive scan of one value per thread over the workgroup (blockDim a
// multiple of 64, at most 1024): wave-inclusive scans by shuffles, then the
// wave totals; returns the exclusive prefix and the workgroup total.
__device__ inline int32_t block_excl_scan(int32_t v, int32_t* lds, int32_t& total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  int32_t inc = v;
  for (int off = 1; off < 64; off <<= 1) {
    const int32_t u = __shfl_up(inc, off, 64);
    if (lane >= off) inc += u;
  }
  if (lane == 63) lds[wid] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t acc = 0;
    for (int w = 0; w < nw; ++w) {
      const int32_t t = lds[w];
      lds[w] = acc;
      acc += t;
    }
    lds[nw] = acc;
  }
  __syncthreads();
  const int32_t excl = inc - v + lds[wid];
  total = lds[nw];
  __syncthreads();
  return excl;
}

// Exclusive scan of 1024-element blocks (4 per thread at 256 threads);
// sums[block] = the block's total.
__global__ void k_scan_local(int n, const int32_t* __restrict__ in, int32_t* __restrict__ out,
                             int32_t* __restrict__ sums) {
  __shared__ int32_t lds[17];
  const int per = 1024 / blockDim.x;
  const int base = blockIdx.x * 1024 + threadIdx.x * per;
  int32_t v[4] = {0, 0, 0, 0};
  int32_t tsum = 0;
  for (int k = 0; k < per; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0;
    tsum += v[k];
  }
  int32_t total = 0;
  int32_t acc = block_excl_scan(tsum, lds, total);
  for (int k = 0; k < per; ++k) {
    if (base + k < n) out[base + k] = acc;
    acc += v[k];
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}
// Exclusive scan of the block sums in place (one workgroup, chunks of
// blockDim with a running carry); *total = the grand total.
__global__ void k_scan_sums(int nblk, int32_t* __restrict__ sums, int32_t* __restrict__ total) {
  __shared__ int32_t lds[17];
  int32_t carry = 0;
  for (int base = 0; base < nblk; base += blockDim.x) {
    const int k = base + threadIdx.x;
    const int32_t v = k < nblk ? sums[k] : 0;
    int32_t t = 0;
    const int32_t e = block_excl_scan(v, lds, t);
    if (k < nblk) sums[k] = carry + e;
    carry += t;
  }
  if (threadIdx.x == 0) *total = carry;
}
__global__ void k_scan_add(int n, int32_t* __restrict__ out, const int32_t* __restrict__ sums,
                           const int32_t* __restrict__ total) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] += sums[i / 1024];
  if (i == n) out[n] = *total;
}

// ---------------------------------------------------------------------------
// Strain limiting per face (strain_limiting.cpp:113-375): SVD of F = Ds Dm_inv,
// constraint set, barrier gradient increments and the J vector of the
// Gauss-Newton Hessian, once per face per Newton iteration.
// ---------------------------------------------------------------------------
__device__ inline bool svd32(const float F[3][2], float U[3][2], float sig[2], float V[2][2]) {
  float a = F[0][0] * F[0][0] + F[1][0] * F[1][0] + F[2][0] * F[2][0];
  float b = F[0][0] * F[0][1] + F[1][0] * F[1][1] + F[2][0] * F[2][1];
  float c = F[0][1] * F[0][1] + F[1][1] * F[1][1] + F[2][1] * F[2][1];
  float cs = 1.0f, sn = 0.0f;
  if (b != 0.0f) {
    float zeta = (c - a) / (2.0f * b);
    float t = (zeta >= 0.0f ? 1.0f : -1.0f) / (fabsf(zeta) + sqrtf(1.0f + zeta * zeta));
    cs = 1.0f / sqrtf(1.0f + t * t);
    sn = cs * t;
  }
  float v0[2] = {cs, -sn}, v1[2] = {sn, cs};
  float u0[3], u1[3];
  for (int r = 0; r < 3; ++r) {
    u0[r] = F[r][0] * v0[0] + F[r][1] * v0[1];
    u1[r] = F[r][0] * v1[0] + F[r][1] * v1[1];
  }
  float s0 = sqrtf(u0[0] * u0[0] + u0[1] * u0[1] + u0[2] * u0[2]);
  float s1 = sqrtf(u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
  if (s1 > s0) {
    float t = s0; s0 = s1; s1 = t;
    for (int r = 0; r < 3; ++r) { t = u0[r]; u0[r] = u1[r]; u1[r] = t; }
    t = v0[0]; v0[0] = v1[0]; v1[0] = t;
    t = v0[1]; v0[1] = v1[1]; v1[1] = t;
  }
  if (s0 > 0.0f) { for (int r = 0; r < 3; ++r) u0[r] /= s0; }
  else { u0[0] = 1.0f; u0[1] = 0.0f; u0[2] = 0.0f; }
  if (s1 > 0.0f) { for (int r = 0; r < 3; ++r) u1[r] /= s1; }
  else {
    float w[3] = {0.0f, 0.0f, 0.0f};
    int ax = fabsf(u0[0]) < 0.6f ? 0 : 1;
    w[ax] = 1.0f;
    float d = w[0] * u0[0] + w[1] * u0[1] + w[2] * u0[2];
    for (int r = 0; r < 3; ++r) w[r] -= d * u0[r];
    float l = sqrtf(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    for (int r = 0; r < 3; ++r) u1[r] = w[r] / l;
  }
  for (int r = 0; r < 3; ++r) { U[r][0] = u0[r]; U[r][1] = u1[r]; }
  V[0][0] = v0[0]; V[1][0] = v0[1]; V[0][1] = v1[0]; V[1][1] = v1[1];
  sig[0] = s0; sig[1] = s1;
  const float th = 1e-6f;
  if (sig[0] < th || sig[1] < th) { sig[0] = fmaxf(sig[0], th); sig[1] = fmaxf(sig[1], th); }
  return isfinite(sig[0]) && isfinite(sig[1]);
}

__global__ void k_strain_face(int m, StrainParams sp, const int32_t* __restrict__ tris, const float* __restrict__ x,
                              const float* __restrict__ dm, const float* __restrict__ area,
                              const int32_t* __restrict__ face_blk, const float* __restrict__ Helast3,
                              int32_t* __restrict__ flags, float* __restrict__ sgrad, float* __restrict__ shess,
                              uint32_t* __restrict__ active_count) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  int local_active = 0;
  if (f < m) {
    int fl = 0;
    if (area[f] > 1e-12f) {
      int i0 = tris[3 * f], i1 = tris[3 * f + 1], i2 = tris[3 * f + 2];
      f3 x0 = ld3(x, i0), x1 = ld3(x, i1), x2 = ld3(x, i2);
      f3 d1 = x1 - x0, d2 = x2 - x0;
      float Mi[2][2] = {{dm[4 * f], dm[4 * f + 1]}, {dm[4 * f + 2], dm[4 * f + 3]}};
      float F[3][2];
      for (int r = 0; r < 3; ++r) {
        float a = comp(d1, r), b = comp(d2, r);
        F[r][0] = a * Mi[0][0] + b * Mi[1][0];
        F[r][1] = a * Mi[0][1] + b * Mi[1][1];
      }
      float U[3][2], sig[2], V[2][2];
      if (svd32(F, U, sig, V)) {
        float fmass = area[f] * sp.thick * sp.rho;
        f3 cen = ((x0 + x1) + x2) / 3.0f;
        float w[9];
        f3 w0 = x0 - cen, w1 = x1 - cen, w2 = x2 - cen;
        w[0] = w0.x; w[1] = w0.y; w[2] = w0.z; w[3] = w1.x; w[4] = w1.y; w[5] = w1.z;
        w[6] = w2.x; w[7] = w2.y; w[8] = w2.z;
        // w^T H w with H the face's 9x9 elastic-only block (diagonal 3x3 blocks).
        float el = 0.0f;
        for (int a = 0; a < 3; ++a)
          for (int k = 0; k < 3; ++k) {
            float s = 0.0f;
            for (int b = 0; b < 3; ++b) s = s + Helast3[3 * (size_t)face_blk[9 * f + 3 * a + b] + k] * w[3 * b + k];
            el = el + w[3 * a + k] * s;
          }
        el = fmaxf(el, 0.0f);
        bool tie = fabsf(sig[0] - sig[1]) < sp.svd_eps;
        for (int s = 0; s < 2; ++s) {
          float gap = (1.0f + sp.tau + sp.eps) - sig[s];
          if (!in_domain(gap, sp.eps)) continue;
          float gc = sp.cp ? sp.eps : fmaxf(fabsf(gap), sp.min_gap);
          float kk = fmass / (gc * gc) + el;
          fl |= 1 << s;
          ++local_active;
          float dS[3][2];
          for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 2; ++c)
              dS[r][c] = tie ? (U[r][0] * V[c][0] + U[r][1] * V[c][1]) * 0.5f : U[r][s] * V[c][s];
          float dV = barrier_grad(gap, sp.eps, kk, sp.fp64_near);
          if (!(fabsf(dV) < 1e-12f)) {
            fl |= 4 << s;
            float* o = sgrad + 18 * (size_t)f + 9 * s;
            for (int r = 0; r < 3; ++r) {
              float a0 = -dV * dS[r][0], a1 = -dV * dS[r][1];
              float m0 = a0 * Mi[0][0] + a1 * Mi[0][1];
              float m1 = a0 * Mi[1][0] + a1 * Mi[1][1];
              o[3 + r] = m0;
              o[6 + r] = m1;
              o[r] = -(m0 + m1);
            }
          }
          float d2 = barrier_hess(gap, sp.eps, kk, sp.fp64_near);
          if (!(fabsf(d2) < 1e-12f)) {
            fl |= 16 << s;
            float* o = shess + 20 * (size_t)f + 10 * s;
            o[0] = d2;
            for (int r = 0; r < 3; ++r) {
              float x0_ = dS[r][0] * Mi[0][0] + dS[r][1] * Mi[0][1];
              float x1_ = dS[r][0] * Mi[1][0] + dS[r][1] * Mi[1][1];
              o[1 + r] = -(x0_ + x1_);
              o[4 + r] = x0_;
              o[7 + r] = x1_;
            }
          }
        }
      }
    }
    flags[f] = fl;
  }
  __shared__ int lds[4];
  int tot = block_reduce(local_active, lds, [](int a, int b) { return a + b; });
  if (threadIdx.x == 0 && tot) atomicAdd(active_count, (uint32_t)tot);
}

// ---------------------------------------------------------------------------
// Gradient (Integrator::compute_gradient, integrator.cpp:246-396), one vertex
// per thread, contributions in the reference's order: inertia, elastic,
// strain (face order), contacts (contact order), pins (pin order), walls.
// Also the max-norm for the Newton test (integrator.cpp:176-180).
// ---------------------------------------------------------------------------
__device__ inline f3 normalize_or_default(f3 n, float eps) {
  if (sqn(n) > eps * eps) return normalized(n);
  return mk(0.0f, 1.0f, 0.0f);
}

// Stiffness::compute_pin_stiffness (stiffness.cpp:118-144) on the mass+elastic diagonal block.
__device__ inline float pin_stiffness(float mass, f3 off, f3 Hdiag, float min_gap) {
  f3 Hd = spd_diag(Hdiag, 1e-8f);
  f3 dir = off;
  float len = nrm(dir);
  if ((double)len > 1e-9) dir = dir / len; else dir = mk(1.0f, 0.0f, 0.0f);
  float kel = fmaxf(0.0f, quad_diag(Hd, dir));
  float gh = fmaxf(fmaxf(len, min_gap), 1e-12f);
  return mass / (gh * gh) + kel;
}

__global__ void k_gradient(int n, StepParams P, const float* __restrict__ x, const float* __restrict__ xhat,
                           const float* __restrict__ masses, const float* __restrict__ egrad,
                           const int32_t* __restrict__ v2f_ptr, const int32_t* __restrict__ v2f,
                           const int32_t* __restrict__ sflags, const float* __restrict__ sgrad,
                           const int32_t* __restrict__ rc_ptr, const RigidContact* __restrict__ rc,
                           const int32_t* __restrict__ pin_ptr, const float* __restrict__ pin_t,
                           const int32_t* __restrict__ diag, const float* __restrict__ Hconst,
                           const float* __restrict__ wall_n, const float* __restrict__ wall_off,
                           const float* __restrict__ kw, const float* __restrict__ xprev, float* __restrict__ g,
                           uint32_t* __restrict__ gmax_bits, SelfView sv) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  float local = 0.0f;
  if (i < n) {
    f3 xi = ld3(x, i);
    float mf = masses[i] / (P.dt * P.dt);
    f3 xt = ld3(xhat, i);
    f3 gi = mk(0.0f + mf * (xi.x - xt.x), 0.0f + mf * (xi.y - xt.y), 0.0f + mf * (xi.z - xt.z));
    gi = gi + ld3(egrad, i);
    if (P.strain)
      for (int e = v2f_ptr[i]; e < v2f_ptr[i + 1]; ++e) {
        int slot = v2f[e];
        int f = slot / 3, a = slot - 3 * f;
        int fl = sflags[f];
        for (int s = 0; s < 2; ++s)
          if ((fl & (1 << s)) && (fl & (4 << s))) gi = gi + ld3(sgrad + 18 * (size_t)f + 9 * s, a);
      }
    // Self contacts (corrected mode) come first in the contact list.
    if (sv.ptr)
      for (int e = sv.ptr[i]; e < sv.ptr[i + 1]; ++e) {
        const int code = sv.code[e], c = code >> 2, slot = code & 3;
        const SelfTerm& t = sv.st[c];
        if (!t.active || t.dV == 0.0f) continue;
        const float s = t.dV * sv.sc[c].w[slot];
        gi = gi + s * mk(t.nx, t.ny, t.nz);
      }
    if (rc_ptr)
      for (int e = rc_ptr[i]; e < rc_ptr[i + 1]; ++e) {
        const RigidContact& c = rc[e];
        if (!in_domain(c.gap, P.gap_max)) continue;
        f3 nn = normalize_or_default(mk(c.nx, c.ny, c.nz), P.normal_eps);
        float dV = barrier_grad(c.gap, P.gap_max, c.k, P.fp64_near);
        if (dV == 0.0f) continue;
        float s = dV * 1.0f;
        gi = gi + s * nn;
      }
    if (pin_ptr && pin_ptr[i] < pin_ptr[i + 1] && P.cp) {
      // Corrected physics: attracting pins, (pin_k m / dt^2) (x - target).
      const float kp = P.pin_k * (masses[i] / (P.dt * P.dt));
      for (int e = pin_ptr[i]; e < pin_ptr[i + 1]; ++e) gi = gi + kp * (xi - ld3(pin_t, e));
    } else if (pin_ptr && pin_ptr[i] < pin_ptr[i + 1]) {
      const float* B = Hconst + 9 * (size_t)diag[i];
      f3 Hd = mk(B[0], B[4], B[8]);
      for (int e = pin_ptr[i]; e < pin_ptr[i + 1]; ++e) {
        f3 off = xi - ld3(pin_t, e);
        float kb = pin_stiffness(masses[i], off, Hd, P.min_gap);
        float gap = nrm(off);
        if (!in_domain(gap, P.gap_max) || gap <= P.normal_eps) continue;
        f3 nn = off / gap;
        float dV = barrier_grad(gap, P.gap_max, kb, P.fp64_near);
        gi = gi + dV * nn;
      }
    }
    for (int w = 0; w < P.num_walls; ++w) {
      float kb = kw[(size_t)w * n + i];
      f3 nn = normalize_or_default(ld3(wall_n, w), P.normal_eps);
      float gap = dot(nn, xi) - wall_off[w];
      if (!in_domain(gap, P.gap_max)) continue;
      float dV = barrier_grad(gap, P.gap_max, kb, P.fp64_near);
      gi = gi + dV * nn;
    }
    // Friction in contact order: the self PT contact of a one-triangle mesh
    // (faithful mode), then the rigid contacts.
    if (P.friction && (rc_ptr || sv.pt_of)) {
      f3 disp = xi - ld3(xprev, i);
      const int pc = sv.pt_of ? sv.pt_of[i] : -1;
      if (pc >= 0) {
        const SelfContact& C = sv.sc[pc];
        float kf;
        f3 tan;
        if (friction_data(mk(C.nx, C.ny, C.nz), sv.st[pc].kb, C.gap, disp, P, kf, tan)) gi = gi + kf * tan;
      }
      if (rc_ptr)
        for (int e = rc_ptr[i]; e < rc_ptr[i + 1]; ++e) {
          float kf;
          f3 tan;
          if (!friction_data(rc[e], disp, P, kf, tan)) continue;
          gi = gi + kf * tan;
        }
    }
    st3(g, i, gi);
    local = fmaxf(fabsf(gi.x), fmaxf(fabsf(gi.y), fabsf(gi.z)));
    if (local != local) local = __int_as_float(0x7f800000);  // NaN never passes the test
  }
  __shared__ float lds[4];
  float mx = block_reduce(local, lds, [](float a, float b) { return fmaxf(a, b); });
  if (threadIdx.x == 0) atomicMax(gmax_bits, __float_as_uint(mx));
}

// ---------------------------------------------------------------------------
// System matrix values (Integrator::assemble_system_matrix,
// integrator.cpp:398-544): constant mass+elastic fold, then strain blocks in
// constraint order, then (diagonal blocks) contacts, pins and walls.  The
// pattern is fixed, every contribution is symmetric, so the final
// (H + H^T)/2 is the identity and is not applied.
// ---------------------------------------------------------------------------
__global__ void k_assemble(int n, StepParams P, const int32_t* __restrict__ block_row, const int32_t* __restrict__ cols,
                           const int32_t* __restrict__ blk_cptr, const int32_t* __restrict__ blk_contrib,
                           const float* __restrict__ Hconst, const int32_t* __restrict__ sflags,
                           const float* __restrict__ shess, const float* __restrict__ x,
                           const float* __restrict__ masses, const int32_t* __restrict__ rc_ptr,
                           const RigidContact* __restrict__ rc, const int32_t* __restrict__ pin_ptr,
                           const float* __restrict__ pin_t, const float* __restrict__ wall_n,
                           const float* __restrict__ wall_off, const float* __restrict__ kw,
                           const float* __restrict__ xprev, float* __restrict__ H, int nb, SelfView sv,
                           const float* __restrict__ fgk, const float* __restrict__ fpi) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  float h[9];
  const float* hc = Hconst + 9 * (size_t)b;
  if (P.cp) {
    // Corrected physics: mass / dt^2, then each face's 2k G_ab (I - n n^T) in
    // face order.
#pragma unroll
    for (int e = 0; e < 9; ++e) h[e] = 0.0f;
    if (cols[b] == block_row[b]) {
      const float mf = masses[block_row[b]] * (1.0f / (P.dt * P.dt));
      if (mf != 0.0f) {
        h[0] = h[0] + mf;
        h[4] = h[4] + mf;
        h[8] = h[8] + mf;
      }
    }
    for (int e = blk_cptr[b]; e < blk_cptr[b + 1]; ++e) {
      const int slot = blk_contrib[e];
      const int f = slot / 9, ab = slot - 9 * f;
      const float g = fgk[6 * (size_t)f + sym6(ab / 3, ab % 3)];
      const float* pi = fpi + 6 * (size_t)f;
#pragma unroll
      for (int e2 = 0; e2 < 9; ++e2) h[e2] = h[e2] + g * pi[sym6(e2 / 3, e2 % 3)];
    }
  } else {
#pragma unroll
    for (int e = 0; e < 9; ++e) h[e] = hc[e];
  }
  if (P.strain)
    for (int e = blk_cptr[b]; e < blk_cptr[b + 1]; ++e) {
      int slot = blk_contrib[e];
      int f = slot / 9, ab = slot - 9 * f, a = ab / 3, bb = ab - 3 * a;
      int fl = sflags[f];
      for (int s = 0; s < 2; ++s) {
        if (!((fl & (1 << s)) && (fl & (16 << s)))) continue;
        const float* o = shess + 20 * (size_t)f + 10 * s;
        float d2 = o[0];
        float v[9];
        float mx = 0.0f;
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) {
            v[3 * r + c] = d2 * (o[1 + 3 * a + r] * o[1 + 3 * bb + c]);
            mx = fmaxf(mx, fabsf(v[3 * r + c]));
          }
        if (mx < 1e-12f) continue;
        for (int e2 = 0; e2 < 9; ++e2)
          if (!(fabsf(v[e2]) < 1e-12f)) h[e2] = h[e2] + v[e2];
      }
    }
  int i = block_row[b];
  if (cols[b] == i) {
    // Diagonal blocks of the self contacts (both slots on vertex i); the
    // blocks coupling two vertices are applied in the PCG product.
    if (sv.ptr)
      for (int e = sv.ptr[i]; e < sv.ptr[i + 1]; ++e) {
        const int code = sv.code[e], c = code >> 2, slot = code & 3;
        const SelfTerm& t = sv.st[c];
        if (!t.active || t.d2 == 0.0f) continue;
        const f3 gi = sv.sc[c].w[slot] * mk(t.nx, t.ny, t.nz);
        for (int e2 = 0; e2 < 9; ++e2) {
          const float val = t.d2 * (comp(gi, e2 / 3) * comp(gi, e2 % 3));
          if (fabsf(val) < P.barrier_tol) continue;
          h[e2] = h[e2] + val;
        }
      }
    if (rc_ptr)
      for (int e = rc_ptr[i]; e < rc_ptr[i + 1]; ++e) {
        const RigidContact& c = rc[e];
        if (!in_domain(c.gap, P.gap_max)) continue;
        f3 nn = normalize_or_default(mk(c.nx, c.ny, c.nz), P.normal_eps);
        float d2 = barrier_hess(c.gap, P.gap_max, c.k, P.fp64_near);
        if (d2 == 0.0f) continue;
        f3 gi = 1.0f * nn;
        for (int e2 = 0; e2 < 9; ++e2) {
          float val = d2 * (comp(gi, e2 / 3) * comp(gi, e2 % 3));
          if (fabsf(val) < P.barrier_tol) continue;
          h[e2] = h[e2] + val;
        }
      }
    if (pin_ptr && pin_ptr[i] < pin_ptr[i + 1] && P.cp) {
      const float kp = P.pin_k * (masses[i] / (P.dt * P.dt));
      for (int e = pin_ptr[i]; e < pin_ptr[i + 1]; ++e) {
        h[0] = h[0] + kp;
        h[4] = h[4] + kp;
        h[8] = h[8] + kp;
      }
    } else if (pin_ptr && pin_ptr[i] < pin_ptr[i + 1]) {
      f3 xi = ld3(x, i);
      f3 Hd = mk(hc[0], hc[4], hc[8]);
      for (int e = pin_ptr[i]; e < pin_ptr[i + 1]; ++e) {
        f3 off = xi - ld3(pin_t, e);
        float kb = pin_stiffness(masses[i], off, Hd, P.min_gap);
        float gap = nrm(off);
        if (!in_domain(gap, P.gap_max) || gap <= P.normal_eps) continue;
        f3 nn = off / gap;
        float dV = barrier_grad(gap, P.gap_max, kb, P.fp64_near);
        float d2 = barrier_hess(gap, P.gap_max, kb, P.fp64_near);
        float q = dV / gap;
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) {
            float nrc = comp(nn, r) * comp(nn, c);
            float val = d2 * nrc + q * ((r == c ? 1.0f : 0.0f) - nrc);
            if (fabsf(val) < P.barrier_tol) continue;
            h[3 * r + c] = h[3 * r + c] + val;
          }
      }
    }
    if (P.num_walls) {
      f3 xi = ld3(x, i);
      for (int w = 0; w < P.num_walls; ++w) {
        float kb = kw[(size_t)w * n + i];
        f3 nn = normalize_or_default(ld3(wall_n, w), P.normal_eps);
        float gap = dot(nn, xi) - wall_off[w];
        if (!in_domain(gap, P.gap_max)) continue;
        float d2 = barrier_hess(gap, P.gap_max, kb, P.fp64_near);
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) {
            float val = d2 * (comp(nn, r) * comp(nn, c));
            if (fabsf(val) < P.barrier_tol) continue;
            h[3 * r + c] = h[3 * r + c] + val;
          }
      }
    }
    if (P.friction && (rc_ptr || sv.pt_of)) {
      // k_f (I - n n^T) + 1e-8 I, entries above 1e-12 (integrator.cpp:515-536, friction.cpp:54-72),
      // in contact order: the one-triangle self PT contact, then the rigid ones.
      f3 disp = ld3(x, i) - ld3(xprev, i);
      auto add_friction = [&](f3 nn, float kf) {
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) {
            float val = r == c ? kf : 0.0f;
            val = val - kf * (comp(nn, r) * comp(nn, c));
            val = val + (r == c ? 1e-8f : 0.0f);
            if (!((double)fabsf(val) > 1e-12)) continue;
            h[3 * r + c] = h[3 * r + c] + val;
          }
      };
      const int pc = sv.pt_of ? sv.pt_of[i] : -1;
      if (pc >= 0) {
        const SelfContact& C = sv.sc[pc];
        const f3 nn = mk(C.nx, C.ny, C.nz);
        float kf;
        f3 tan;
        if (friction_data(nn, sv.st[pc].kb, C.gap, disp, P, kf, tan)) add_friction(nn, kf);
      }
      if (rc_ptr)
        for (int e = rc_ptr[i]; e < rc_ptr[i + 1]; ++e) {
          float kf;
          f3 tan;
          if (!friction_data(rc[e], disp, P, kf, tan)) continue;
          add_friction(mk(rc[e].nx, rc[e].ny, rc[e].nz), kf);
        }
    }
  }
  float* o = H + 9 * (size_t)b;
#pragma unroll
  for (int e = 0; e < 9; ++e) o[e] = h[e];
}

// ---------------------------------------------------------------------------
// Contact restitution (Integrator::apply_contact_restitution,
// integrator.cpp:565-674).  Rigid contacts first, one thread in contact order,
// because every impulse changes the body velocity the next contact sees; then
// every wall, per vertex.
// ---------------------------------------------------------------------------
__device__ inline void mul33(const float* A, const float* B, float* O) {
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) O[3 * r + c] = A[3 * r] * B[c] + (A[3 * r + 1] * B[3 + c] + A[3 * r + 2] * B[6 + c]);
}
__device__ inline f3 mulv33(const float* A, f3 v) {
  return mk(A[0] * v.x + (A[1] * v.y + A[2] * v.z), A[3] * v.x + (A[4] * v.y + A[5] * v.z),
            A[6] * v.x + (A[7] * v.y + A[8] * v.z));
}
__device__ inline void world_inv_inertia(const BodyState& b, float* O) {
  float RI[9], Rt[9];
  mul33(b.rot, b.Ibinv, RI);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) Rt[3 * r + c] = b.rot[3 * c + r];
  mul33(RI, Rt, O);
}

__global__ void k_restitution_rigid(int total, const RigidContact* __restrict__ rc, float* __restrict__ v,
                                    const float* __restrict__ masses, BodyState* __restrict__ bodies, int nbod,
                                    float e, float gap_limit) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (int k = 0; k < total; ++k) {
    const RigidContact& c = rc[k];
    if (c.gap > gap_limit) continue;
    if (c.body < 0 || c.body >= nbod) continue;
    f3 nn = mk(c.nx, c.ny, c.nz);
    float nl = nrm(nn);
    if ((double)nl < 1e-8) continue;
    nn = nn / nl;
    BodyState& B = bodies[c.body];
    f3 q = mk(c.qx, c.qy, c.qz);
    f3 pos = mk(B.pos[0], B.pos[1], B.pos[2]);
    f3 vv = ld3(v, c.vertex);
    f3 bv = mk(B.lin[0], B.lin[1], B.lin[2]) + cross(mk(B.ang[0], B.ang[1], B.ang[2]), q - pos);
    float vn = dot(vv - bv, nn);
    if (vn >= 0.0f) continue;
    float m = masses[c.vertex];
    float im = m > 0.0f ? 1.0f / m : 0.0f;
    f3 r = q - pos;
    f3 cr = cross(r, nn);
    float Iw[9];
    world_inv_inertia(B, Iw);
    float angular = dot(cr, mulv33(Iw, cr));
    float den = im + angular;
    if (den <= 0.0f) continue;
    float j = -(1.0f + e) * vn / den;
    f3 J = j * nn;
    vv = vv + mk(J.x * im, J.y * im, J.z * im);
    st3(v, c.vertex, vv);
    if (B.mass <= 0.0f) continue;
    f3 mJ = -J;
    B.lin[0] = B.lin[0] + mJ.x / B.mass;
    B.lin[1] = B.lin[1] + mJ.y / B.mass;
    B.lin[2] = B.lin[2] + mJ.z / B.mass;
    f3 dw = mulv33(Iw, cross(r, mJ));
    B.ang[0] = B.ang[0] + dw.x;
    B.ang[1] = B.ang[1] + dw.y;
    B.ang[2] = B.ang[2] + dw.z;
  }
}

__global__ void k_restitution_walls(int n, int nw, const float* __restrict__ wall_n, const float* __restrict__ wall_off,
                                    const float* __restrict__ wall_gap, float gap_max, float e,
                                    const float* __restrict__ x, float* __restrict__ v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  f3 xi = ld3(x, i);
  f3 vi = ld3(v, i);
  for (int w = 0; w < nw; ++w) {
    f3 nn = ld3(wall_n, w);
    float nl = nrm(nn);
    if ((double)nl < 1e-8) continue;
    nn = nn / nl;
    float sd = dot(nn, xi) - wall_off[w];
    if (!(sd <= wall_gap[w] + gap_max)) continue;
    f3 n2 = nn;
    float l2 = nrm(n2);
    if ((double)l2 < 1e-8) continue;
    n2 = n2 / l2;
    float vn = dot(vi, n2);
    if (vn < 0.0f) vi = vi - ((1.0f + e) * vn) * n2;
  }
  st3(v, i, vi);
}

// ---------------------------------------------------------------------------
// Line search (LineSearch::search / is_feasible, line_search.cpp:8-161) for
// the faithful contact set: trial t tests x + 2^-t (1.25 d) against the first
// active wall; each vertex reports the 20-bit mask of trials it passes and the
// masks are AND-reduced, so all 20 trials cost one pass.
// ---------------------------------------------------------------------------
__global__ void k_ls_mask(int n, f3 wn, float woff, int check_wall, const float* __restrict__ x,
                          const float* __restrict__ d, uint32_t* __restrict__ mask) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t mk_ = 0xFFFFFu;
  if (i < n && check_wall) {
    f3 xi = ld3(x, i);
    f3 di = ld3(d, i);
    f3 ed = 1.25f * di;
    float alpha = 1.0f;
    mk_ = 0;
    for (int t = 0; t < 20; ++t) {
      f3 xn = xi + alpha * ed;
      float sd = dot(wn, xn) - woff;
      if (!(sd < 0.0f)) mk_ |= 1u << t;
      alpha *= 0.5f;
    }
  }
  __shared__ uint32_t lds[4];
  uint32_t r = block_reduce(mk_, lds, [](uint32_t a, uint32_t b) { return a & b; });
  if (threadIdx.x == 0 && r != 0xFFFFFu) atomicAnd(mask, r);
}

}  // namespace ando
