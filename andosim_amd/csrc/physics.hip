// Corrected physics (ANDO_MODE_CORRECTED_PHYSICS): the kernels that differ
// from the faithful step.  The reference's step leaves the cloth without an
// elastic restoring force (F from mesh.vertices, elasticity.cpp:18-61), lets
// pins repel (barrier.cpp:161-215), moves vertices explicitly before the
// solve (the apply_gravity pre-step) and does not protect rigid contacts in
// its line search (line_search.cpp:70-132).  This mode:
//   * membrane gradient from the current positions, Gauss-Newton Hessian
//     2k (B B^T)_ab (I - n n^T) with B = S Dm_inv (S maps vertices to the two
//     edges) and n the face's current normal;
//   * contact candidates from per-vertex boxes (the points x at the start of a
//     step, the swept trial motion x -> x + 1.25 d after every solve), grown by
//     gap_max on the query side; their gaps, normals and closest-point weights
//     are refreshed at every Newton iteration;
//   * line search: it starts at the largest alpha = 2^-t0 whose trial motion
//     1.25 alpha d stays within L = max(4 x the mean rest edge length, 2 x the
//     longest remaining inertial displacement |x_hat - x|) -- the membrane
//     Hessian has no out-of-plane stiffness, so a contact force on a taut
//     region can give a direction tens of centimetres long whose swept boxes
//     cover most of the mesh; every candidate pair's additive-CCD time of
//     impact over the trial motion (FP64, keeping a separation of 32 FP32
//     ulps) bounds alpha, and the first wall may not get more violated.
// Layout: per face 6 floats for 2k B B^T (entries 00 01 02 11 12 22, `gk`)
// and for I - n n^T (`fpi`); per block 3 floats of isotropic elastic
// stiffness (`Hiso3`, the same sum on the three diagonal entries) and a 9-float
// block of mass / dt^2 plus that stiffness (`Hbase`, diagonal entries only).
#include "kernels.h"

namespace ando {

namespace cp {

// 2k B B^T per face (the Hessian's constant factor), B rows -(M0 + M1), M0, M1
// with M = Dm_inv (rows M0, M1).
__global__ void k_face_gk(int m, const float* __restrict__ dm, const float* __restrict__ area, float thick, float mu,
                          float* __restrict__ gk) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= m) return;
  const float Mi[2][2] = {{dm[4 * f], dm[4 * f + 1]}, {dm[4 * f + 2], dm[4 * f + 3]}};
  const float k = area[f] * thick * mu;
  const float twok = (float)(2.0 * (double)k);
  float B[3][2];
  for (int c = 0; c < 2; ++c) {
    B[0][c] = -(Mi[0][c] + Mi[1][c]);
    B[1][c] = Mi[0][c];
    B[2][c] = Mi[1][c];
  }
  for (int a = 0; a < 3; ++a)
    for (int b = a; b < 3; ++b) gk[6 * (size_t)f + sym6(a, b)] = twok * (B[a][0] * B[b][0] + B[a][1] * B[b][1]);
}

// Per block, in face order: Hiso3 = the sum of the contributing faces' 2k G_ab
// (all three entries), Hbase = mass / dt^2 on a diagonal block (when nonzero)
// followed by the same sum, on the block's diagonal.  One block row per
// workgroup.
__global__ void k_iso_fold(int n, const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ cols,
                           const int32_t* __restrict__ blk_cptr, const int32_t* __restrict__ blk_contrib,
                           const float* __restrict__ gk, const float* __restrict__ masses, float dtinv2,
                           float* __restrict__ Hbase, float* __restrict__ Hiso3) {
  const int i = blockIdx.x;
  if (i >= n) return;
  for (int b = row_ptr[i] + threadIdx.x; b < row_ptr[i + 1]; b += blockDim.x) {
    float h = 0.0f, e = 0.0f;
    if (cols[b] == i) {
      const float mf = masses[i] * dtinv2;
      if (mf != 0.0f) h = h + mf;
    }
    for (int q = blk_cptr[b]; q < blk_cptr[b + 1]; ++q) {
      const int slot = blk_contrib[q];
      const int f = slot / 9, ab = slot - 9 * f;
      const float g = gk[6 * (size_t)f + sym6(ab / 3, ab % 3)];
      h = h + g;
      e = e + g;
    }
    float* o = Hbase + 9 * (size_t)b;
    o[0] = h; o[1] = 0.0f; o[2] = 0.0f;
    o[3] = 0.0f; o[4] = h; o[5] = 0.0f;
    o[6] = 0.0f; o[7] = 0.0f; o[8] = h;
    Hiso3[3 * (size_t)b] = e;
    Hiso3[3 * (size_t)b + 1] = e;
    Hiso3[3 * (size_t)b + 2] = e;
  }
}

// Per face at the current positions: the membrane gradient contributions
// (9 floats into fe) and I - n n^T (6 floats into fpi).
__global__ void k_face_x(int m, const int32_t* __restrict__ tris, const float* __restrict__ x,
                         const float* __restrict__ dm, const float* __restrict__ area, float thick, float mu,
                         float* __restrict__ fe, float* __restrict__ fpi) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= m) return;
  const float k = area[f] * thick * mu;
  f3 nn;
  membrane_grad(ld3(x, tris[3 * f]), ld3(x, tris[3 * f + 1]), ld3(x, tris[3 * f + 2]), dm[4 * f], dm[4 * f + 1],
                dm[4 * f + 2], dm[4 * f + 3], k, fe + 9 * (size_t)f, nn);
  for (int a = 0; a < 3; ++a)
    for (int b = a; b < 3; ++b)
      fpi[6 * (size_t)f + sym6(a, b)] = (a == b ? 1.0f : 0.0f) - comp(nn, a) * comp(nn, b);
}

// ed = s d.
__global__ void k_scale(int n3, float s, const float* __restrict__ d, float* __restrict__ ed) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n3) ed[i] = s * d[i];
}

// Boxes of the motion x -> x + ed per vertex (ed null: the points x).
__global__ void k_vertex_boxes(int n, const float* __restrict__ x, const float* __restrict__ ed,
                               float* __restrict__ vbox) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int k = 0; k < 3; ++k) {
    const float a = x[3 * i + k];
    const float e = ed ? a + ed[3 * i + k] : a;
    vbox[6 * (size_t)i + k] = fminf(a, e);
    vbox[6 * (size_t)i + 3 + k] = fmaxf(a, e);
  }
}

// Gap, normal, witness and weights of each self candidate at x: PT weights
// 1, -b (barycentrics of the closest point), EE weights 1 - s, s, -(1 - t), -t
// (the closest points' parameters).
__global__ void k_self_refresh(int nc, const float* __restrict__ x, SelfContact* __restrict__ sc) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  SelfContact C = sc[c];
  if (C.type == 0) {
    const f3 p = ld3(x, C.idx[0]), a = ld3(x, C.idx[1]), b = ld3(x, C.idx[2]), cc = ld3(x, C.idx[3]);
    const PTResult r = point_triangle(p, a, b, cc);
    C.gap = r.dist;
    C.nx = r.normal.x; C.ny = r.normal.y; C.nz = r.normal.z;
    C.qx = r.wq.x; C.qy = r.wq.y; C.qz = r.wq.z;
    const f3 bc = lbvh::triangle_bary(r.wq, a, b, cc);
    C.w[0] = 1.0f; C.w[1] = -bc.x; C.w[2] = -bc.y; C.w[3] = -bc.z;
  } else {
    float dist, s_, t_;
    f3 nrm_, wp, wq;
    lbvh::edge_edge(ld3(x, C.idx[0]), ld3(x, C.idx[1]), ld3(x, C.idx[2]), ld3(x, C.idx[3]), dist, nrm_, wp, wq, &s_,
                    &t_);
    C.gap = dist;
    C.nx = nrm_.x; C.ny = nrm_.y; C.nz = nrm_.z;
    C.qx = wq.x; C.qy = wq.y; C.qz = wq.z;
    C.w[0] = 1.0f - s_; C.w[1] = s_; C.w[2] = -(1.0f - t_); C.w[3] = -t_;
  }
  sc[c] = C;
}

// Gap, normal, witness and stiffness of each rigid candidate at x (the
// bodies do not move during a step).  Stiffness m / max(gap_max, min_gap)^2 +
// max(0, n^T E n) with E the vertex's isotropic elastic stiffness.
__global__ void k_rigid_refresh(int total, const float* __restrict__ x, DetectBodies D, RigidContact* __restrict__ rc,
                                const float* __restrict__ masses, const int32_t* __restrict__ diag,
                                const float* __restrict__ Hiso3, float dt, float kscale) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  RigidContact& o = rc[e];
  const DetectBody& B = D.b[o.body];
  const int t = o.tri, i = o.vertex;
  const PTResult r = point_triangle(ld3(x, i), ld3(B.world, B.tris[3 * t]), ld3(B.world, B.tris[3 * t + 1]),
                                    ld3(B.world, B.tris[3 * t + 2]));
  const float* h = Hiso3 + 3 * (size_t)diag[i];
  o.gap = r.dist;
  o.nx = r.normal.x; o.ny = r.normal.y; o.nz = r.normal.z;
  o.qx = r.wq.x; o.qy = r.wq.y; o.qz = r.wq.z;
  o.k = rigid_contact_k(masses[i], mk(h[0], h[1], h[2]), r.normal, dt, kscale);
}

// Rigid candidates: vertex boxes grown by gm against the bodies' triangle
// boxes, in (body, triangle) order per vertex; pass 0 counts (and keeps the
// first kHitCache hits), pass 1 writes records at ptr[i].  The group boxes
// (grown by 0.01 + 1e-5) and the body box only cull.
__global__ void __launch_bounds__(kBlock) k_rigid_cand(int n, int pass, const float* __restrict__ vbox, float gm,
                                                       DetectBodies bodies, const int32_t* __restrict__ ptr,
                                                       int32_t* __restrict__ count, RigidContact* __restrict__ out,
                                                       int32_t* __restrict__ hits) {
  __shared__ float sbox[6 * kGroupTile];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < n;
  lbvh::Box q;
  for (int k = 0; k < 3; ++k) {
    q.lo[k] = live ? vbox[6 * (size_t)i + k] - gm : 0.0f;
    q.hi[k] = live ? vbox[6 * (size_t)i + 3 + k] + gm : 0.0f;
  }
  int c = 0;
  const int base = (pass && live) ? ptr[i] : 0;
  auto emit = [&](int b, int t) {
    if (pass) {
      RigidContact r;
      r.gap = 0.0f;
      r.nx = r.ny = r.nz = 0.0f;
      r.qx = r.qy = r.qz = 0.0f;
      r.k = 0.0f;
      r.body = b;
      r.tri = t;
      r.vertex = i;
      out[base + c] = r;
    } else if (c < kHitCache) {
      hits[(size_t)i * kHitCache + c] = (b << 24) | t;
    }
    ++c;
  };
  const int cached = (pass && live) ? count[i] : kHitCache + 1;
  if (pass && cached <= kHitCache)
    for (int h = 0; h < cached; ++h) {
      const int code = hits[(size_t)i * kHitCache + h];
      emit(code >> 24, code & 0xffffff);
    }
  const bool scan = live && !(pass && cached <= kHitCache);
  for (int b = 0; b < bodies.count; ++b) {
    const DetectBody& B = bodies.b[b];
    lbvh::Box bb;
    for (int k = 0; k < 3; ++k) {
      bb.lo[k] = B.bbox[k];
      bb.hi[k] = B.bbox[3 + k];
    }
    const bool in_body = scan && lbvh::overlaps(q, bb);
    if (!__syncthreads_or(in_body)) continue;
    const int ng = (B.nt + B.grp - 1) / B.grp;
    for (int g0 = 0; g0 < ng; g0 += kGroupTile) {
      const int gn = min(kGroupTile, ng - g0);
      for (int k = threadIdx.x; k < 6 * gn; k += blockDim.x) sbox[k] = B.gbox[6 * g0 + k];
      __syncthreads();
      if (in_body)
        for (int gl = 0; gl < gn; ++gl) {
          lbvh::Box gbx;
          for (int k = 0; k < 3; ++k) {
            gbx.lo[k] = sbox[6 * gl + k];
            gbx.hi[k] = sbox[6 * gl + 3 + k];
          }
          if (!lbvh::overlaps(q, gbx)) continue;
          const int gi = g0 + gl;
          const int t1 = min(B.nt, (gi + 1) * B.grp);
          for (int t = gi * B.grp; t < t1; ++t) {
            const f3 a = ld3(B.world, B.tris[3 * t]), bq = ld3(B.world, B.tris[3 * t + 1]),
                     cq = ld3(B.world, B.tris[3 * t + 2]);
            lbvh::Box tb;
            tb.lo[0] = fminf(fminf(a.x, bq.x), cq.x); tb.hi[0] = fmaxf(fmaxf(a.x, bq.x), cq.x);
            tb.lo[1] = fminf(fminf(a.y, bq.y), cq.y); tb.hi[1] = fmaxf(fmaxf(a.y, bq.y), cq.y);
            tb.lo[2] = fminf(fminf(a.z, bq.z), cq.z); tb.hi[2] = fmaxf(fmaxf(a.z, bq.z), cq.z);
            if (lbvh::overlaps(q, tb)) emit(b, t);
          }
        }
      __syncthreads();
    }
  }
  if (!pass && live) count[i] = c;
}

// ---------------------------------------------------------------------------
// Line-search time of impact in FP64 with a minimum separation.  Positions
// are FP32: an additive CCD with thickness 0 lets a pair pushed together over
// several Newton iterations close geometrically (s = 0.1 of the gap each
// time) until the gap is a few ulps of the coordinates, where rounding of the
// positions decides the side.  Here the pair keeps xi = 32 FP32 ulps of its
// largest coordinate (an upper bound 2^-23 m on ulp(m)): the accepted motion
// never brings it within xi, and the update x + alpha ed, rounded to FP32
// (alpha a power of two, so alpha ed is exact), moves each vertex by at most
// half an ulp, far less than xi.  A pair that rounding has put within xi may
// only move apart (band_ok); otherwise its vertices are held.
// ---------------------------------------------------------------------------
struct d3 {
  double x, y, z;
};
__device__ inline d3 dmk(double a, double b, double c) { return d3{a, b, c}; }
__device__ inline d3 dsub(d3 a, d3 b) { return dmk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline d3 dadd(d3 a, d3 b) { return dmk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ inline d3 dscale(double s, d3 a) { return dmk(s * a.x, s * a.y, s * a.z); }
__device__ inline double ddot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline double dnorm(d3 a) { return sqrt(ddot(a, a)); }
__device__ inline d3 dld(const float* v, int i) { return dmk(v[3 * i], v[3 * i + 1], v[3 * i + 2]); }

// Closest points of segments p1 q1 and p2 q2 (Ericson, Real-Time Collision
// Detection, 5.1.9); the distance between them.
__device__ inline double seg_seg_dist(d3 p1, d3 q1, d3 p2, d3 q2) {
  const d3 d1 = dsub(q1, p1), d2 = dsub(q2, p2), r = dsub(p1, p2);
  const double a = ddot(d1, d1), e = ddot(d2, d2), f = ddot(d2, r);
  double s = 0.0, t = 0.0;
  if (a <= 0.0 && e <= 0.0) {
  } else if (a <= 0.0) {
    t = fmin(fmax(f / e, 0.0), 1.0);
  } else {
    const double c = ddot(d1, r);
    if (e <= 0.0) {
      s = fmin(fmax(-c / a, 0.0), 1.0);
    } else {
      const double b = ddot(d1, d2), den = a * e - b * b;
      s = den != 0.0 ? fmin(fmax((b * f - c * e) / den, 0.0), 1.0) : 0.0;
      t = (b * s + f) / e;
      if (t < 0.0) {
        t = 0.0;
        s = fmin(fmax(-c / a, 0.0), 1.0);
      } else if (t > 1.0) {
        t = 1.0;
        s = fmin(fmax((b - c) / a, 0.0), 1.0);
      }
    }
  }
  return dnorm(dsub(dadd(p1, dscale(s, d1)), dadd(p2, dscale(t, d2))));
}

// Distance from p to triangle a b c (Ericson 5.1.5, Voronoi regions); a
// degenerate triangle falls back to its edges.
__device__ inline double pt_tri_dist(d3 p, d3 a, d3 b, d3 c) {
  const d3 ab = dsub(b, a), ac = dsub(c, a), ap = dsub(p, a);
  const double d1 = ddot(ab, ap), d2 = ddot(ac, ap);
  if (d1 <= 0.0 && d2 <= 0.0) return dnorm(ap);
  const d3 bp = dsub(p, b);
  const double d3_ = ddot(ab, bp), d4 = ddot(ac, bp);
  if (d3_ >= 0.0 && d4 <= d3_) return dnorm(bp);
  const double vc = d1 * d4 - d3_ * d2;
  if (vc <= 0.0 && d1 >= 0.0 && d3_ <= 0.0) return dnorm(dsub(p, dadd(a, dscale(d1 / (d1 - d3_), ab))));
  const d3 cp = dsub(p, c);
  const double d5 = ddot(ab, cp), d6 = ddot(ac, cp);
  if (d6 >= 0.0 && d5 <= d6) return dnorm(cp);
  const double vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) return dnorm(dsub(p, dadd(a, dscale(d2 / (d2 - d6), ac))));
  const double va = d3_ * d6 - d5 * d4;
  if (va <= 0.0 && (d4 - d3_) >= 0.0 && (d5 - d6) >= 0.0)
    return dnorm(dsub(p, dadd(b, dscale((d4 - d3_) / ((d4 - d3_) + (d5 - d6)), dsub(c, b)))));
  const double sum = va + vb + vc;
  if (!(sum > 0.0))
    return fmin(fmin(seg_seg_dist(p, p, a, b), seg_seg_dist(p, p, b, c)), seg_seg_dist(p, p, c, a));
  const double v = vb / sum, w = vc / sum;
  return dnorm(dsub(p, dadd(a, dadd(dscale(v, ab), dscale(w, ac)))));
}

__device__ inline double pair_dist_d(int type, const d3* q) {
  return type == 0 ? pt_tri_dist(q[0], q[1], q[2], q[3]) : seg_seg_dist(q[0], q[1], q[2], q[3]);
}

constexpr double kSepUlps = 32.0 * 1.1920928955078125e-07;  // 32 ulps per unit of coordinate magnitude

// The separation a pair keeps: 32 * 2^-23 * its largest coordinate over the
// motion x0 -> x0 + p.
__device__ inline double sep_xi(const d3* x0, const d3* p) {
  double m = 0.0;
  for (int k = 0; k < 4; ++k) {
    const d3 e = dadd(x0[k], p[k]);
    m = fmax(m, fmax(fmax(fabs(x0[k].x), fabs(x0[k].y)), fmax(fabs(x0[k].z), fmax(fabs(e.x), fmax(fabs(e.y), fabs(e.z))))));
  }
  return kSepUlps * m;
}

// Additive CCD (Li, Kaufman, Jiang et al., "Codimensional Incremental
// Potential Contact", 2021, Algorithm 1) in FP64 with thickness xi, s = 0.1,
// t_c = 1, at most 256 advances: 1 when the whole motion x0 -> x0 + p keeps
// the pair farther than xi, else a time up to which it provably does (0 when
// it starts within xi).
__device__ inline double accd_sep(int type, const d3* x0, const d3* p, double xi) {
  const d3 pbar = dscale(0.25, dadd(dadd(p[0], p[1]), dadd(p[2], p[3])));
  d3 q[4], x[4];
  for (int k = 0; k < 4; ++k) {
    q[k] = dsub(p[k], pbar);
    x[k] = x0[k];
  }
  const double lp = type == 0 ? dnorm(q[0]) + fmax(dnorm(q[1]), fmax(dnorm(q[2]), dnorm(q[3])))
                              : fmax(dnorm(q[0]), dnorm(q[1])) + fmax(dnorm(q[2]), dnorm(q[3]));
  if (lp == 0.0) return 1.0;
  double d = pair_dist_d(type, x);
  if (!(d > xi)) return 0.0;
  const double g = 0.1 * (d - xi);
  double t = 0.0, tl = 0.9 * (d - xi) / lp;
  for (int it = 0; it < 256; ++it) {
    for (int k = 0; k < 4; ++k) x[k] = dadd(x[k], dscale(tl, q[k]));
    d = pair_dist_d(type, x);
    if (t > 0.0 && d - xi < g) return t;
    t = t + tl;
    if (t > 1.0) return 1.0;
    tl = 0.9 * (d - xi) / lp;
  }
  return t;
}

// A pair already within its separation xi (rounding of the positions can put
// it there) may only move apart: the whole motion keeps it farther than half
// its gap and it ends no closer.  Otherwise its cloth vertices are held (their
// direction is zeroed for this iteration): a pair pushed together iteration
// after iteration would otherwise close geometrically until rounding decides
// its side, and with one alpha for the whole cloth a blocking test would stop
// every vertex instead of these.
__device__ inline bool band_ok(int type, const d3* x0, const d3* p, double d0) {
  d3 x1[4];
  for (int k = 0; k < 4; ++k) x1[k] = dadd(x0[k], p[k]);
  return pair_dist_d(type, x1) >= d0 && accd_sep(type, x0, p, 0.5 * d0) >= 1.0;
}

// One hold pass: band pairs that fail band_ok under the motion with the
// vertices held in fin set fout for their cloth vertices (fout starts as a
// copy of fin) and *changed.  pass_all: hold every band pair's vertices.
__global__ void k_band_self(int nc, const SelfContact* __restrict__ sc, const float* __restrict__ x,
                            const float* __restrict__ ed, const uint8_t* __restrict__ fin, uint8_t* __restrict__ fout,
                            uint32_t* __restrict__ changed, int pass_all) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const SelfContact& C = sc[c];
  d3 x0[4], p[4];
  for (int k = 0; k < 4; ++k) {
    x0[k] = dld(x, C.idx[k]);
    p[k] = fin[C.idx[k]] ? dmk(0.0, 0.0, 0.0) : dld(ed, C.idx[k]);
  }
  const double d0 = pair_dist_d(C.type, x0);
  if (!(d0 < sep_xi(x0, p))) return;
  if (!pass_all && band_ok(C.type, x0, p, d0)) return;
  for (int k = 0; k < 4; ++k)
    if (!fin[C.idx[k]]) {
      fout[C.idx[k]] = 1;
      atomicOr(changed, 1u);
    }
}

__global__ void k_band_rigid(int total, const RigidContact* __restrict__ rc, DetectBodies D,
                             const float* __restrict__ x, const float* __restrict__ ed,
                             const uint8_t* __restrict__ fin, uint8_t* __restrict__ fout,
                             uint32_t* __restrict__ changed, int pass_all) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  const RigidContact& r = rc[e];
  const DetectBody& B = D.b[r.body];
  d3 x0[4], p[4];
  x0[0] = dld(x, r.vertex);
  p[0] = fin[r.vertex] ? dmk(0.0, 0.0, 0.0) : dld(ed, r.vertex);
  for (int k = 1; k < 4; ++k) {
    x0[k] = dld(B.world, B.tris[3 * r.tri + k - 1]);
    p[k] = dmk(0.0, 0.0, 0.0);
  }
  const double d0 = pair_dist_d(0, x0);
  if (!(d0 < sep_xi(x0, p))) return;
  if (!pass_all && band_ok(0, x0, p, d0)) return;
  if (!fin[r.vertex]) {
    fout[r.vertex] = 1;
    atomicOr(changed, 1u);
  }
}

// Held vertices: d = ed = 0.
__global__ void k_apply_hold(int n, const uint8_t* __restrict__ hold, float* __restrict__ d, float* __restrict__ ed) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !hold[i]) return;
  for (int k = 0; k < 3; ++k) {
    d[3 * i + k] = 0.0f;
    ed[3 * i + k] = 0.0f;
  }
}

// Line search: bit t of a contact's mask is set when alpha_t = 2^-t is at most
// its time of impact (accd_sep) over the whole trial motion (x -> x + ed;
// rigid triangles stay put); masks are AND-reduced into *mask.
// The line search starts at trial t0: its motion x -> x + ed covers
// alpha = 2^-t0, and trial t >= t0 moves 2^-(t - t0) of it.
__device__ inline uint32_t toi_bits(double toi, int t0) {
  uint32_t m = 0u;
  double alpha = 1.0;
  for (int t = t0; t < 20; ++t) {
    if (alpha <= toi) m |= 1u << t;
    alpha *= 0.5;
  }
  return m;
}

// Squared lengths, in FP64, of the longest Newton displacement d_i and of the
// longest remaining inertial displacement x_hat_i - x_i; maxima kept as the
// bit patterns of the (non-negative) doubles in out[0], out[1] (start at 0).
__global__ void k_motion_max(int n, const float* __restrict__ d, const float* __restrict__ x,
                             const float* __restrict__ xhat, unsigned long long* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double dd = 0.0, hh = 0.0;
  if (i < n) {
    const double d0 = d[3 * i], d1 = d[3 * i + 1], d2 = d[3 * i + 2];
    dd = d0 * d0 + d1 * d1 + d2 * d2;
    const double h0 = (double)xhat[3 * i] - (double)x[3 * i], h1 = (double)xhat[3 * i + 1] - (double)x[3 * i + 1],
                 h2 = (double)xhat[3 * i + 2] - (double)x[3 * i + 2];
    hh = h0 * h0 + h1 * h1 + h2 * h2;
  }
  __shared__ double lds[4];
  const double md = block_reduce(dd, lds, [](double a, double b) { return fmax(a, b); });
  const double mh = block_reduce(hh, lds, [](double a, double b) { return fmax(a, b); });
  if (threadIdx.x == 0) {
    atomicMax(out, (unsigned long long)__double_as_longlong(md));
    atomicMax(out + 1, (unsigned long long)__double_as_longlong(mh));
  }
}

__global__ void k_ls_toi_self(int nc, int t0, const SelfContact* __restrict__ sc, const float* __restrict__ x,
                              const float* __restrict__ ed, uint32_t* __restrict__ mask) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t m = 0xFFFFFu;
  if (c < nc) {
    const SelfContact& C = sc[c];
    d3 x0[4], p[4];
    for (int k = 0; k < 4; ++k) {
      x0[k] = dld(x, C.idx[k]);
      p[k] = dld(ed, C.idx[k]);
    }
    const double xi = sep_xi(x0, p);
    // A pair within its separation passed band_ok or has its vertices held.
    m = pair_dist_d(C.type, x0) < xi ? toi_bits(1.0, t0) : toi_bits(accd_sep(C.type, x0, p, xi), t0);
  }
  __shared__ uint32_t lds[4];
  const uint32_t r = block_reduce(m, lds, [](uint32_t a, uint32_t b) { return a & b; });
  if (threadIdx.x == 0 && r != 0xFFFFFu) atomicAnd(mask, r);
}

__global__ void k_ls_toi_rigid(int total, int t0, const RigidContact* __restrict__ rc, DetectBodies D,
                               const float* __restrict__ x, const float* __restrict__ ed, uint32_t* __restrict__ mask) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t m = 0xFFFFFu;
  if (e < total) {
    const RigidContact& r = rc[e];
    const DetectBody& B = D.b[r.body];
    d3 x0[4], p[4];
    x0[0] = dld(x, r.vertex);
    p[0] = dld(ed, r.vertex);
    for (int k = 1; k < 4; ++k) {
      x0[k] = dld(B.world, B.tris[3 * r.tri + k - 1]);
      p[k] = dmk(0.0, 0.0, 0.0);
    }
    const double xi = sep_xi(x0, p);
    m = pair_dist_d(0, x0) < xi ? toi_bits(1.0, t0) : toi_bits(accd_sep(0, x0, p, xi), t0);
  }
  __shared__ uint32_t lds[4];
  const uint32_t rr = block_reduce(m, lds, [](uint32_t a, uint32_t b) { return a & b; });
  if (threadIdx.x == 0 && rr != 0xFFFFFu) atomicAnd(mask, rr);
}

// The first wall: trial t >= t0 is feasible for vertex i when its signed
// distance at x + 2^-(t - t0) ed is at least min(0, its distance at x).
__global__ void k_ls_wall(int n, int t0, f3 wn, float woff, const float* __restrict__ x, const float* __restrict__ ed,
                          uint32_t* __restrict__ mask) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t m = 0xFFFFFu;
  if (i < n) {
    const f3 xi = ld3(x, i), ei = ld3(ed, i);
    const float g0 = dot(wn, xi) - woff;
    const float lim = fminf(0.0f, g0);
    float alpha = 1.0f;
    m = 0u;
    for (int t = t0; t < 20; ++t) {
      const f3 xn = mk(xi.x + alpha * ei.x, xi.y + alpha * ei.y, xi.z + alpha * ei.z);
      if (dot(wn, xn) - woff >= lim) m |= 1u << t;
      alpha *= 0.5f;
    }
  }
  __shared__ uint32_t lds[4];
  const uint32_t r = block_reduce(m, lds, [](uint32_t a, uint32_t b) { return a & b; });
  if (threadIdx.x == 0 && r != 0xFFFFFu) atomicAnd(mask, r);
}

// v = (x - x_old) / dt, then optional damping.
__global__ void k_velocity_dt(int n3, float dt, float damp_scale, int damp, const float* __restrict__ x,
                              const float* __restrict__ xold, float* __restrict__ v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n3) return;
  float vi = (x[i] - xold[i]) / dt;
  if (damp) vi *= damp_scale;
  v[i] = vi;
}

}  // namespace cp

}  // namespace ando
