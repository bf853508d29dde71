// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of AndoSim's per-step cloth solve (the path behind
// `ando_barrier_core.Integrator.step`), written from the reference's C++
// sources without Eigen.  It is the checker for the HIP implementation and the
// CPU baseline timed by bench.py; nothing in the product links, loads or calls
// it.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// use it.
//
// Precision: templated on Real.  Real=float is the "faithful" restatement of
// the reference build (types.h:12-16 selects float).  Real=double is the FP64
// oracle ("the same code with Real=double").
//
// Expression order.  The reference evaluates through Eigen 3.4.  Where the
// order of floating-point operations is fixed by the reference's own source
// text, this file keeps it.  Where Eigen picks the order internally, this file
// assumes the following (none of it can be executed here, since Eigen is absent):
//   * fixed-size 3-vector reductions (dot, squaredNorm) are the unrolled tree
//     e0 + (e1 + e2);
//   * dynamic-size reductions shorter than one SIMD packet are sequential;
//   * dynamic-size VecX dot products follow the SSE2 packet order: two packet
//     accumulators, then (l0+l2)+(l1+l3) for float and l0+l1 for double;
//   * sparse row-major SpMV sums each row left to right in column order;
//   * setFromTriplets sums duplicates in triplet-list order.
// Parity unpinned: bit-level agreement with the real reference build wherever
// Eigen chooses the order.  The SVD (JacobiSVD), the general 3x3 eigen-solver
// and AngleAxis are restated with different algorithms (see svd32 and
// spd_general below) and agree with Eigen only to rounding; they are pinned by
// identities (U S V^T = F, ordering, derivatives by finite differences), not
// bit for bit.  Everything else is pinned by checks whose expected values come
// from the mathematics or the reference's own tests, not from this file's
// formulas: tests/test_oracle.py (the reference's known answers in
// tests/golden/, closed forms, brute force, finite differences, invariants);
// tests/test_oracle_pins.py (strain-limit gradient and Hessian against the
// singular-value derivatives, friction against the derivative of its
// potential, contact and wall stiffness closed forms, RigidBody::integrate
// against an exact spin and Euler's and Newton's laws, the reference's
// rigid-coupled known answer); tests/test_oracle_physics.py (the corrected
// physics mode: membrane gradient and Hessian closed forms, attracting pins,
// CCD line search, free fall).  The learned-preconditioner oracle
// (oracle/learned.py) is pinned by tests/test_learned_oracle.py, the
// two-level model (oracle/two_level.py) by dense solves in
// tests/test_two_level_model.py.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <unordered_set>
#include <vector>

#include <omp.h>

namespace orc {

template <class R> struct V3 {
  R x[3];
  V3() : x{R(0), R(0), R(0)} {}
  V3(R a, R b, R c) : x{a, b, c} {}
  R& operator[](int i) { return x[i]; }
  const R& operator[](int i) const { return x[i]; }
};
template <class R> V3<R> operator+(const V3<R>& a, const V3<R>& b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
template <class R> V3<R> operator-(const V3<R>& a, const V3<R>& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
template <class R> V3<R> operator-(const V3<R>& a) { return {-a[0], -a[1], -a[2]}; }
template <class R> V3<R> operator*(R s, const V3<R>& a) { return {s * a[0], s * a[1], s * a[2]}; }
template <class R> V3<R> operator/(const V3<R>& a, R s) { return {a[0] / s, a[1] / s, a[2] / s}; }
// Eigen fixed-size 3-term reduction: e0 + (e1 + e2).
template <class R> R dot3(const V3<R>& a, const V3<R>& b) { return a[0] * b[0] + (a[1] * b[1] + a[2] * b[2]); }
template <class R> R sqn3(const V3<R>& a) { return dot3(a, a); }
template <class R> R norm3(const V3<R>& a) { return std::sqrt(sqn3(a)); }
template <class R> V3<R> cross3(const V3<R>& a, const V3<R>& b) {
  return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
template <class R> V3<R> normalized3(const V3<R>& a) {
  R z = sqn3(a);
  if (z > R(0)) return a / std::sqrt(z);
  return a;
}

template <class R> struct M3 {
  R m[3][3];
  M3() { for (auto& r : m) for (auto& v : r) v = R(0); }
  static M3 identity() { M3 o; o.m[0][0] = o.m[1][1] = o.m[2][2] = R(1); return o; }
};
template <class R> V3<R> mul(const M3<R>& A, const V3<R>& v) {
  V3<R> o;
  for (int r = 0; r < 3; ++r) o[r] = A.m[r][0] * v[0] + (A.m[r][1] * v[1] + A.m[r][2] * v[2]);
  return o;
}
template <class R> M3<R> mul(const M3<R>& A, const M3<R>& B) {
  M3<R> o;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) o.m[r][c] = A.m[r][0] * B.m[0][c] + (A.m[r][1] * B.m[1][c] + A.m[r][2] * B.m[2][c]);
  return o;
}
template <class R> M3<R> transpose(const M3<R>& A) {
  M3<R> o;
  for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) o.m[r][c] = A.m[c][r];
  return o;
}
// Eigen determinant_impl<3>: bruteforce_det3_helper(0,1,2) - (1,0,2) + (2,0,1).
template <class R> R det_h(const M3<R>& M, int a, int b, int c) {
  return M.m[0][a] * (M.m[1][b] * M.m[2][c] - M.m[1][c] * M.m[2][b]);
}
template <class R> R det3(const M3<R>& M) { return (det_h(M, 0, 1, 2) - det_h(M, 1, 0, 2)) + det_h(M, 2, 0, 1); }
template <class R> R cof3(const M3<R>& M, int i, int j) {
  int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
  return M.m[i1][j1] * M.m[i2][j2] - M.m[i1][j2] * M.m[i2][j1];
}
// Eigen compute_inverse<3>: cofactor column 0 dotted with column 0, then adjugate * invdet.
template <class R> M3<R> inv3(const M3<R>& M) {
  R c0 = cof3(M, 0, 0), c1 = cof3(M, 1, 0), c2 = cof3(M, 2, 0);
  R det = c0 * M.m[0][0] + (c1 * M.m[1][0] + c2 * M.m[2][0]);
  R invdet = R(1) / det;
  M3<R> o;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) o.m[i][j] = cof3(M, j, i) * invdet;
  return o;
}

// ---------------------------------------------------------------------------
// enforce_spd (stiffness.cpp:236-254): symmetrize, eigen-decompose, clamp
// eigenvalues below eps, rebuild.  For an exactly diagonal input Eigen's
// SelfAdjointEigenSolver reduces to: scale by the largest |entry|, read the
// diagonal, scale back, so each entry becomes max-clamped (d/s)*s with exact
// permutation eigenvectors.  That case is restated exactly.  Every call on the
// reference's path (elastic blocks, pin/wall diagonal blocks) is diagonal; a
// non-diagonal input falls to a Jacobi solver and is counted.
// ---------------------------------------------------------------------------
static long g_spd_general_calls = 0;

template <class R> void spd_general(M3<R>& H, R eps) {
  ++g_spd_general_calls;
  double a[3][3], v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) a[i][j] = H.m[i][j];
  for (int sweep = 0; sweep < 50; ++sweep) {
    double off = std::fabs(a[0][1]) + std::fabs(a[0][2]) + std::fabs(a[1][2]);
    if (off < 1e-300) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        if (std::fabs(a[p][q]) < 1e-300) continue;
        double theta = (a[q][q] - a[p][p]) / (2 * a[p][q]);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
        double c = 1 / std::sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < 3; ++k) {
          double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - s * akq; a[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; ++k) {
          double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - s * aqk; a[q][k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 3; ++k) {
          double vkp = v[k][p], vkq = v[k][q];
          v[k][p] = c * vkp - s * vkq; v[k][q] = s * vkp + c * vkq;
        }
      }
  }
  double lam[3];
  for (int i = 0; i < 3; ++i) { lam[i] = (R)a[i][i]; if (lam[i] < (double)eps) lam[i] = eps; }
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += v[r][k] * lam[k] * v[c][k];
      H.m[r][c] = (R)s;
    }
}

template <class R> void enforce_spd(M3<R>& H, R eps = R(1e-8)) {
  M3<R> S;
  for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) S.m[r][c] = (H.m[r][c] + H.m[c][r]) * R(0.5);
  bool diagonal = S.m[0][1] == R(0) && S.m[0][2] == R(0) && S.m[1][2] == R(0);
  if (!diagonal) { H = S; spd_general(H, eps); return; }
  R s = std::max(std::fabs(S.m[0][0]), std::max(std::fabs(S.m[1][1]), std::fabs(S.m[2][2])));
  if (s == R(0)) s = R(1);
  M3<R> o;
  for (int i = 0; i < 3; ++i) {
    R d = (S.m[i][i] / s) * s;
    o.m[i][i] = d < eps ? eps : d;
  }
  H = o;
}

// Eigen VecX dot: SSE2 packet order (see header note).
template <class R> R vdot(const R* a, const R* b, size_t n) {
  constexpr size_t P = sizeof(R) == 4 ? 4 : 2;
  if (n == 0) return R(0);
  size_t aligned = (n / P) * P, aligned2 = (n / (2 * P)) * (2 * P);
  if (aligned == 0) {
    R res = a[0] * b[0];
    for (size_t i = 1; i < n; ++i) res = res + a[i] * b[i];
    return res;
  }
  R p0[P], p1[P];
  for (size_t l = 0; l < P; ++l) p0[l] = a[l] * b[l];
  if (aligned > P) {
    for (size_t l = 0; l < P; ++l) p1[l] = a[P + l] * b[P + l];
    for (size_t i = 2 * P; i < aligned2; i += 2 * P)
      for (size_t l = 0; l < P; ++l) {
        p0[l] = p0[l] + a[i + l] * b[i + l];
        p1[l] = p1[l] + a[i + P + l] * b[i + P + l];
      }
    for (size_t l = 0; l < P; ++l) p0[l] = p0[l] + p1[l];
    if (aligned > aligned2)
      for (size_t l = 0; l < P; ++l) p0[l] = p0[l] + a[aligned2 + l] * b[aligned2 + l];
  }
  R res = (P == 4) ? (p0[0] + p0[2]) + (p0[1] + p0[3]) : p0[0] + p0[1];
  for (size_t i = aligned; i < n; ++i) res = res + a[i] * b[i];
  return res;
}
template <class R> R vinf(const R* a, size_t n) {
  R m = R(0);
  for (size_t i = 0; i < n; ++i) m = std::max(m, std::fabs(a[i]));
  return m;
}

// ---------------------------------------------------------------------------
// Cubic barrier (barrier.cpp:29-65).
// ---------------------------------------------------------------------------
template <class R> R barrier_energy(R g, R gmax, R k) {
  if (gmax <= R(0) || g >= gmax) return R(0);
  R d = gmax - g;
  if (d <= R(0)) return R(0);
  R inv = R(1) / gmax;
  return R(0.5) * k * inv * d * d * d;
}
template <class R> R barrier_grad(R g, R gmax, R k) {
  if (gmax <= R(0) || g >= gmax) return R(0);
  R d = gmax - g;
  R inv = R(1) / gmax;
  return -R(1.5) * k * inv * d * d;
}
template <class R> R barrier_hess(R g, R gmax, R k) {
  if (gmax <= R(0) || g >= gmax) return R(0);
  R d = gmax - g;
  R inv = R(1) / gmax;
  return R(3.0) * k * inv * d;
}
template <class R> bool in_domain(R g, R gmax) { return gmax > R(0) && g <= gmax; }

// ---------------------------------------------------------------------------
// Corrected physics line search: time of impact in FP64 with a minimum
// separation of 32 FP32 ulps of the pair's largest coordinate (new; the
// reference samples 10 distances, line_search.cpp:163-247).  With thickness 0
// a pair pushed together over several Newton iterations closes geometrically
// until FP32 rounding of the positions decides its side.  Same operations in
// the same order as the device code, so both pick the same trial.
// ---------------------------------------------------------------------------
struct D3 {
  double x, y, z;
};
static inline D3 dsub(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline D3 dadd(D3 a, D3 b) { return D3{a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline D3 dscale(double s, D3 a) { return D3{s * a.x, s * a.y, s * a.z}; }
static inline double ddot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline double dnorm(D3 a) { return std::sqrt(ddot(a, a)); }

// Segment-segment distance (Ericson, Real-Time Collision Detection, 5.1.9).
static double seg_seg_dist(D3 p1, D3 q1, D3 p2, D3 q2) {
  const D3 d1 = dsub(q1, p1), d2 = dsub(q2, p2), r = dsub(p1, p2);
  const double a = ddot(d1, d1), e = ddot(d2, d2), f = ddot(d2, r);
  double s = 0.0, t = 0.0;
  if (a <= 0.0 && e <= 0.0) {
  } else if (a <= 0.0) {
    t = std::fmin(std::fmax(f / e, 0.0), 1.0);
  } else {
    const double c = ddot(d1, r);
    if (e <= 0.0) {
      s = std::fmin(std::fmax(-c / a, 0.0), 1.0);
    } else {
      const double b = ddot(d1, d2), den = a * e - b * b;
      s = den != 0.0 ? std::fmin(std::fmax((b * f - c * e) / den, 0.0), 1.0) : 0.0;
      t = (b * s + f) / e;
      if (t < 0.0) {
        t = 0.0;
        s = std::fmin(std::fmax(-c / a, 0.0), 1.0);
      } else if (t > 1.0) {
        t = 1.0;
        s = std::fmin(std::fmax((b - c) / a, 0.0), 1.0);
      }
    }
  }
  return dnorm(dsub(dadd(p1, dscale(s, d1)), dadd(p2, dscale(t, d2))));
}

// Point-triangle distance (Ericson 5.1.5); a degenerate triangle falls back
// to its edges.
static double pt_tri_dist(D3 p, D3 a, D3 b, D3 c) {
  const D3 ab = dsub(b, a), ac = dsub(c, a), ap = dsub(p, a);
  const double d1 = ddot(ab, ap), d2 = ddot(ac, ap);
  if (d1 <= 0.0 && d2 <= 0.0) return dnorm(ap);
  const D3 bp = dsub(p, b);
  const double d3 = ddot(ab, bp), d4 = ddot(ac, bp);
  if (d3 >= 0.0 && d4 <= d3) return dnorm(bp);
  const double vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) return dnorm(dsub(p, dadd(a, dscale(d1 / (d1 - d3), ab))));
  const D3 cp = dsub(p, c);
  const double d5 = ddot(ab, cp), d6 = ddot(ac, cp);
  if (d6 >= 0.0 && d5 <= d6) return dnorm(cp);
  const double vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) return dnorm(dsub(p, dadd(a, dscale(d2 / (d2 - d6), ac))));
  const double va = d3 * d6 - d5 * d4;
  if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0)
    return dnorm(dsub(p, dadd(b, dscale((d4 - d3) / ((d4 - d3) + (d5 - d6)), dsub(c, b)))));
  const double sum = va + vb + vc;
  if (!(sum > 0.0))
    return std::fmin(std::fmin(seg_seg_dist(p, p, a, b), seg_seg_dist(p, p, b, c)), seg_seg_dist(p, p, c, a));
  const double v = vb / sum, w = vc / sum;
  return dnorm(dsub(p, dadd(a, dadd(dscale(v, ab), dscale(w, ac)))));
}

static double pair_dist_d(int type, const D3* q) {
  return type == 0 ? pt_tri_dist(q[0], q[1], q[2], q[3]) : seg_seg_dist(q[0], q[1], q[2], q[3]);
}

// Corrected physics: the first line-search trial, the smallest t0 (at most 19)
// with 1.25 * 2^-t0 * max |d_i| <= L = max(4 * mean rest edge, 2 * max
// |x_hat_i - x_i|) (squared maxima given, FP64): the membrane Hessian has no
// out-of-plane stiffness, so a contact force on a taut region can give a
// direction tens of centimetres long whose swept boxes cover most of the mesh.
static int cp_first_trial(double mean_edge, double md2, double mh2) {
  const double L = std::max(4.0 * mean_edge, 2.0 * std::sqrt(mh2));
  const double len = 1.25 * std::sqrt(md2);
  int t0 = 0;
  while (t0 < 19 && std::ldexp(len, -t0) > L) ++t0;
  return t0;
}

// The separation a pair keeps: 32 * 2^-23 * its largest coordinate over the
// motion x0 -> x0 + p.
static double sep_xi(const D3* x0, const D3* p) {
  double m = 0.0;
  for (int k = 0; k < 4; ++k) {
    const D3 e = dadd(x0[k], p[k]);
    m = std::fmax(m, std::fmax(std::fmax(std::fabs(x0[k].x), std::fabs(x0[k].y)),
                               std::fmax(std::fabs(x0[k].z),
                                         std::fmax(std::fabs(e.x), std::fmax(std::fabs(e.y), std::fabs(e.z))))));
  }
  return 32.0 * 1.1920928955078125e-07 * m;
}

// Additive CCD (Li, Kaufman, Jiang et al. 2021, Algorithm 1) in FP64 with
// thickness xi, s = 0.1, t_c = 1, at most 256 advances: 1 when the whole
// motion x0 -> x0 + p keeps the pair farther than xi, else a time up to which
// it provably does (0 when it starts within xi).
static double accd_sep(int type, const D3* x0, const D3* p, double xi) {
  const D3 pbar = dscale(0.25, dadd(dadd(p[0], p[1]), dadd(p[2], p[3])));
  D3 q[4], x[4];
  for (int k = 0; k < 4; ++k) {
    q[k] = dsub(p[k], pbar);
    x[k] = x0[k];
  }
  const double lp = type == 0 ? dnorm(q[0]) + std::fmax(dnorm(q[1]), std::fmax(dnorm(q[2]), dnorm(q[3])))
                              : std::fmax(dnorm(q[0]), dnorm(q[1])) + std::fmax(dnorm(q[2]), dnorm(q[3]));
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

// A pair already within its separation (rounding of the positions can put it
// there) may only move apart: the whole motion keeps it farther than half its
// gap and it ends no closer; otherwise the line search holds its cloth
// vertices for the iteration.
static bool band_ok(int type, const D3* x0, const D3* p, double d0) {
  D3 x1[4];
  for (int k = 0; k < 4; ++k) x1[k] = dadd(x0[k], p[k]);
  return pair_dist_d(type, x1) >= d0 && accd_sep(type, x0, p, 0.5 * d0) >= 1.0;
}

// ---------------------------------------------------------------------------
// Point-triangle narrow phase (collision.cpp:279-387, Ericson regions).
// ---------------------------------------------------------------------------
template <class R>
void narrow_pt(const V3<R>& p, const V3<R>& a, const V3<R>& b, const V3<R>& c, R& dist, V3<R>& normal,
               V3<R>& wq) {
  auto fin = [&](const V3<R>& q) {
    wq = q;
    V3<R> diff = p - q;
    dist = norm3(diff);
    normal = (double)dist > 1e-10 ? diff / dist : V3<R>(0, 0, 1);
  };
  V3<R> ab = b - a, ac = c - a, ap = p - a;
  R d1 = dot3(ab, ap), d2 = dot3(ac, ap);
  if (d1 <= R(0) && d2 <= R(0)) { fin(a); return; }
  V3<R> bp = p - b;
  R d3 = dot3(ab, bp), d4 = dot3(ac, bp);
  if (d3 >= R(0) && d4 <= d3) { fin(b); return; }
  V3<R> cp = p - c;
  R d5 = dot3(ab, cp), d6 = dot3(ac, cp);
  if (d6 >= R(0) && d5 <= d6) { fin(c); return; }
  R vc = d1 * d4 - d3 * d2;
  if (vc <= R(0) && d1 >= R(0) && d3 <= R(0)) {
    R v = d1 / (d1 - d3);
    fin(a + v * ab);
    return;
  }
  R vb = d5 * d2 - d1 * d6;
  if (vb <= R(0) && d2 >= R(0) && d6 <= R(0)) {
    R w = d2 / (d2 - d6);
    fin(a + w * ac);
    return;
  }
  R va = d3 * d6 - d5 * d4;
  if (va <= R(0) && (d4 - d3) >= R(0) && (d5 - d6) >= R(0)) {
    R w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    fin(b + w * (c - b));
    return;
  }
  R denom = (R)(1.0 / (double)(va + vb + vc));
  R v = vb * denom, w = vc * denom;
  wq = (a + v * ab) + w * ac;
  V3<R> diff = p - wq;
  dist = norm3(diff);
  V3<R> tn = cross3(ab, ac);
  R ta = norm3(tn);
  if ((double)ta > 1e-10) {
    tn = tn / ta;
    if (dot3(tn, diff) < R(0)) tn = -tn;
    normal = tn;
  } else {
    normal = (double)dist > 1e-10 ? diff / dist : V3<R>(0, 0, 1);
  }
}

// Edge-edge narrow phase (collision.cpp:389-452).
// s_out / t_out (optional): the closest points' parameters on the two edges
// (0 for both when both edges are degenerate).
template <class R>
void narrow_ee(const V3<R>& p0, const V3<R>& p1, const V3<R>& q0, const V3<R>& q1, R& dist, V3<R>& normal,
               V3<R>& wp, V3<R>& wq, R* s_out = nullptr, R* t_out = nullptr) {
  V3<R> d1 = p1 - p0, d2 = q1 - q0, r = p0 - q0;
  R a = dot3(d1, d1), e = dot3(d2, d2), f = dot3(d2, r);
  R s = R(0), t = R(0);
  const R eps = R(1e-10);
  auto cl = [](R x) { return std::clamp(x, R(0), R(1)); };
  if (a <= eps && e <= eps) {
    wp = p0; wq = q0;
  } else {
    if (a <= eps) {
      s = R(0); t = cl(f / e);
    } else {
      R c = dot3(d1, r);
      if (e <= eps) {
        t = R(0); s = cl(-c / a);
      } else {
        R b = dot3(d1, d2);
        R denom = a * e - b * b;
        s = denom != R(0) ? cl((b * f - c * e) / denom) : R(0);
        t = (b * s + f) / e;
        if (t < R(0)) { t = R(0); s = cl(-c / a); }
        else if (t > R(1)) { t = R(1); s = cl((b - c) / a); }
      }
    }
    wp = p0 + s * d1;
    wq = q0 + t * d2;
  }
  if (s_out) *s_out = s;
  if (t_out) *t_out = t;
  V3<R> diff = wp - wq;
  dist = norm3(diff);
  normal = (double)dist > 1e-10 ? diff / dist : V3<R>(0, 0, 1);
}

// collision.cpp:11-43.
template <class R> V3<R> triangle_barycentric(const V3<R>& p, const V3<R>& a, const V3<R>& b, const V3<R>& c) {
  V3<R> v0 = b - a, v1 = c - a, v2 = p - a;
  R d00 = dot3(v0, v0), d01 = dot3(v0, v1), d11 = dot3(v1, v1), d20 = dot3(v2, v0), d21 = dot3(v2, v1);
  R denom = d00 * d11 - d01 * d01;
  R v = 0, w = 0;
  const R feps = std::numeric_limits<R>::epsilon();
  if (std::fabs(denom) > feps) {
    v = (d11 * d20 - d01 * d21) / denom;
    w = (d00 * d21 - d01 * d20) / denom;
  }
  v = std::clamp(v, R(0), R(1));
  w = std::clamp(w, R(0), R(1));
  if (v + w > R(1)) {
    R sum = v + w;
    if (sum > feps) { v /= sum; w /= sum; }
  }
  R u = R(1) - v - w;
  return {u, v, w};
}

enum CType { PT = 0, EE = 1, WALL = 2, RIGID_PT = 3 };

template <class R> struct Contact {
  int type = PT;
  int idx[4] = {-1, -1, -1, -1};
  R gap = 0;
  V3<R> normal, wp, wq, bary;
  R w[4] = {0, 0, 0, 0};
  int vcount = 0;
  int rigid = -1;
};

template <class R> struct Rigid {
  std::vector<V3<R>> local;
  std::vector<std::array<int, 3>> tris;
  R mass = 1;
  M3<R> Ib = M3<R>::identity(), Ibinv = M3<R>::identity();
  V3<R> pos, lin, ang, force, torque;
  M3<R> rot = M3<R>::identity();

  // rigid_body.cpp:20-41, 55-92.
  void initialize(const std::vector<V3<R>>& verts, const std::vector<std::array<int, 3>>& t, R density) {
    tris = t;
    if (verts.empty()) {
      local.clear(); mass = 0; Ib = M3<R>(); Ibinv = M3<R>(); return;
    }
    V3<R> cen;
    for (auto& v : verts) cen = cen + v;
    cen = cen / (R)verts.size();
    local.resize(verts.size());
    for (size_t i = 0; i < verts.size(); ++i) local[i] = verts[i] - cen;
    pos = cen; rot = M3<R>::identity(); lin = V3<R>(); ang = V3<R>();
    V3<R> mn(std::numeric_limits<R>::max(), std::numeric_limits<R>::max(), std::numeric_limits<R>::max());
    V3<R> mx = -mn;
    for (auto& v : local) {
      V3<R> w = v + pos;
      for (int k = 0; k < 3; ++k) { mn[k] = std::min(mn[k], w[k]); mx[k] = std::max(mx[k], w[k]); }
    }
    V3<R> ext = mx - mn;
    R vol = std::max(ext[0], R(1e-6)) * std::max(ext[1], R(1e-6)) * std::max(ext[2], R(1e-6));
    mass = std::max(density * vol, R(1e-3));
    R x2 = ext[0] * ext[0], y2 = ext[1] * ext[1], z2 = ext[2] * ext[2];
    Ib = M3<R>();
    Ib.m[0][0] = (mass / R(12)) * (y2 + z2);
    Ib.m[1][1] = (mass / R(12)) * (x2 + z2);
    Ib.m[2][2] = (mass / R(12)) * (x2 + y2);
    for (int i = 0; i < 3; ++i) if (Ib.m[i][i] < R(1e-6)) Ib.m[i][i] = R(1e-6);
    Ibinv = inv3(Ib);
    force = torque = V3<R>();
  }
  std::vector<V3<R>> world() const {
    std::vector<V3<R>> o;
    o.reserve(local.size());
    for (auto& v : local) o.push_back(pos + mul(rot, v));
    return o;
  }
  M3<R> inertia_world_inv() const { return mul(mul(rot, Ibinv), transpose(rot)); }
  // rigid_body.cpp:129-149 with dt = 0 (the only use on the path).
  void apply_force(const V3<R>& p, const V3<R>& f) {
    if (mass <= R(0)) return;
    force = force + f;
    V3<R> r = p - pos;
    torque = torque + cross3(r, f);
  }
  // rigid_body.cpp:151-176; Eigen AngleAxis::toRotationMatrix restated.
  void integrate(R dt) {
    if (mass <= R(0)) return;
    V3<R> acc = force / mass;
    lin = lin + dt * acc;
    M3<R> Iw = inertia_world_inv();
    V3<R> aa = mul(Iw, torque);
    ang = ang + dt * aa;
    pos = pos + dt * lin;
    R om = norm3(ang);
    if ((double)om > 1e-8) {
      R th = om * dt;
      V3<R> ax = ang / om;
      R s = std::sin(th), c = std::cos(th);
      V3<R> sa = s * ax, ca = (R(1) - c) * ax;
      M3<R> Rm;
      R tmp = ca[0] * ax[1];
      Rm.m[0][1] = tmp - sa[2]; Rm.m[1][0] = tmp + sa[2];
      tmp = ca[0] * ax[2];
      Rm.m[0][2] = tmp + sa[1]; Rm.m[2][0] = tmp - sa[1];
      tmp = ca[1] * ax[2];
      Rm.m[1][2] = tmp - sa[0]; Rm.m[2][1] = tmp + sa[0];
      for (int i = 0; i < 3; ++i) Rm.m[i][i] = ca[i] * ax[i] + c;
      rot = mul(Rm, rot);
    }
    force = torque = V3<R>();
  }
  V3<R> velocity_at_point(const V3<R>& p) const { return lin + cross3(ang, p - pos); }
  void apply_impulse(const V3<R>& p, const V3<R>& J) {
    if (mass <= R(0)) return;
    lin = lin + J / mass;
    V3<R> r = p - pos;
    ang = ang + mul(inertia_world_inv(), cross3(r, J));
  }
};

template <class R> struct Params {
  R dt = R(0.002), beta_max = R(0.25);
  int min_newton = 2, max_newton = 8, friction_min_newton = 32;
  R pcg_tol = R(1e-3);
  int pcg_max_iters = 1000;
  R gap_max = R(0.001), wall_gap = R(0.001);
  bool enable_ccd = true;
  R normal_eps = R(1e-8), barrier_tol = R(1e-12);
  bool enable_friction = false;
  R mu = R(0.1), friction_eps = R(1e-5), friction_tangent = R(1e-6);
  R damping = R(0), restitution = R(0);
  bool enable_strain = false;
  R strain_limit = R(0.05), strain_tau = R(0.05), strain_eps = R(0), strain_svd_eps = R(1e-6);
  R hessian_eps = R(1e-8), min_gap = R(1e-8);
};

struct Stats {
  int newton_iters = 0, pcg_iters = 0, pcg_solves = 0, pcg_failures = 0, line_search_failures = 0;
  int contacts = 0, rigid_contacts = 0, strain_active = 0;
  double beta = 0;
  std::vector<int> pcg_per_solve;
  std::vector<double> alphas;
};

// Strain constraint data computed once per face per Newton iteration.
template <class R> struct StrainFace {
  int active[2] = {0, 0};
  R k[2] = {0, 0};
  R gap[2] = {0, 0};
  R gradc[2][3][3];  // [s][vertex][component]: gradient increment (strain_limiting.cpp:263-269)
  int grad_on[2] = {0, 0};
  R d2V[2] = {0, 0};
  R J[2][9];
  int hess_on[2] = {0, 0};
};

template <class R> struct Sim {
  // Mesh (mesh.h:9-51).
  std::vector<V3<R>> mverts;
  std::vector<std::array<int, 3>> tris;
  std::vector<std::array<int, 2>> edges;
  std::vector<R> Dm[4];  // Dm_inv entries (0,0),(0,1),(1,0),(1,1)
  std::vector<R> area;
  std::vector<std::vector<int>> v2f;
  R E = R(1e6), nu = R(0.3), rho = R(1000), thick = R(0.001), bend = R(0);
  // State (state.h:10-38).
  std::vector<R> x, v, m, xprev;  // flat 3n
  int n = 0;
  // Constraints.
  struct Pin { int v; V3<R> t; bool active; };
  struct Wall { V3<R> n; R off, gap; bool active; };
  std::vector<Pin> pins;
  std::vector<Wall> walls;
  std::vector<Rigid<R>> rigids;
  Params<R> P;
  int mode = 0;  // 0 faithful
  Stats st;
  // Fixed BSR pattern: diagonal + mesh edges, columns sorted.
  std::vector<int> rp, ci, diag;
  // Elastic constants per face: 9 clamped diagonal blocks c_ab (x,y entries), z = eps.
  std::vector<std::array<R, 9>> ec;
  std::vector<R> Hconst;   // mass + elastic fold (H_total / H_base), 9 per block
  std::vector<R> Helast;   // elastic-only fold (H_elastic), 9 per block
  std::vector<R> egrad;    // elastic gradient (Elasticity::compute_gradient)
  std::vector<StrainFace<R>> sf;
  std::vector<R> H;        // system matrix values (last assembly)
  std::vector<Contact<R>> contacts;
  // Contact mode: 0 faithful (self broad phase empty, F1), 1 corrected (every
  // vertex/triangle and edge/edge pair tested by brute force with the
  // reference's intended predicate; PT and EE barriers both enter the energy).
  int contact_mode = 0;
  // Test hook: the system matrix and right-hand side of solve number
  // capture_solve (counted within a step) are kept for inspection.
  int capture_solve = -1;
  std::vector<double> cap_H, cap_b;
  int ccd_mode = 0;  // corrected mode line search: 0 the reference's 10 samples, 1 additive CCD
  R pt_box = R(1e-4), keep_gap = R(0.01);
  // Contact Hessian blocks that couple two different vertices.  The reference
  // folds them into a pattern built from the triplets; here the pattern is the
  // mesh's, so they are kept beside it and applied in the PCG product.
  struct Coupling { int i, j; R b[9]; };
  std::vector<Coupling> couplings;
  // Corrected physics (new; the intended behaviour that the reference's F3,
  // F7, F11 and F12 findings leave out).  With cp set:
  //   * apply_gravity changes only v; the step predicts x_hat = x + dt v from
  //     the positions of the previous step's end, which the previous line
  //     search left free of penetrations;
  //   * the membrane gradient reads the current positions (the reference's
  //     formula, elasticity.cpp:18-61, on x instead of mesh.vertices), its
  //     Hessian is the Gauss-Newton form 2k (B B^T)_ab (I - n n^T) of the
  //     face at the current x (B = S Dm_inv, S the edge-difference map);
  //   * pins attract: E = (pin_k m / dt^2) |x - p|^2 / 2;
  //   * contact candidates come from per-vertex boxes: PT pairs (a vertex box
  //     grown by gap_max against the union box of a triangle not incident to
  //     the vertex), EE pairs (the union box of one edge grown by gap_max
  //     against that of another sharing no vertex) and rigid PT pairs (a
  //     vertex box grown by gap_max against a rigid triangle's box).  At the
  //     start of a step the boxes are the points x, so the candidates are the
  //     pairs within gap_max (per axis); after every PCG solve they are the
  //     swept boxes of the trial motion [x, x + 1.25 d], which serve the line
  //     search and then the barrier at the next iterate (every iterate lies on
  //     that motion, so each pair within gap_max of it is a candidate);
  //   * every Newton iteration refreshes each candidate's gap, normal and
  //     weights at the current x (PT: 1, -b; EE: 1 - s, s, -(1 - t), -t from
  //     the closest points), so the barrier follows the iterate;
  //   * contact stiffness takes the wall's inertial scale, sum m w^2 /
  //     max(gap_max, min_gap)^2 (stiffness.cpp:146-172), in place of
  //     sum m w^2 / dt^2: the cubic barrier's force is bounded by
  //     1.5 k gap_max, which at m / dt^2 cannot stop a vertex faster than
  //     about 0.15 m/s within a step;
  //   * strain-limit stiffness takes the band width as its scale,
  //     m_face / eps^2, in place of m_face / max(|gap|, min_gap)^2, which
  //     grows without bound at the limit and stalls FP32 PCG;
  //   * the velocity is (x - x_old) / dt after the step's passes (the
  //     reference divides by beta dt, and beta = 0.5 after a capped pass
  //     doubles the velocity);
  //   * contact, wall and strain stiffness read the isotropic elastic
  //     stiffness sum_f 2k (B B^T)_ab I (an upper bound on the projected
  //     Hessian along any direction) in place of the clamped faithful blocks;
  //   * the line search accepts alpha only if the first wall is no more
  //     violated than at the iterate and every candidate pair's additive-CCD
  //     time of impact over the full trial motion (1.25 d) is at least alpha;
  //     pairs outside the candidate set cannot meet during that motion;
  //   * rigid coupling pushes the body away from the cloth (the reference
  //     applies -dV/dg n, toward it).
  bool cp = false;
  R pin_k = R(1e4);
  std::vector<std::array<R, 6>> gk, fpi;  // per face: 2k (B B^T) and I - n n^T, entries 00 01 02 11 12 22
  std::vector<R> Hiso, Hbase_iso;         // 9 per block: isotropic elastic stiffness; plus mass / dt^2
  std::vector<R> vbox;                    // 6 per vertex: candidate boxes (lo xyz, hi xyz)
  std::vector<std::vector<V3<R>>> rworld; // rigid world vertices during the step
  double mean_edge = 0.0;                 // mean rest edge length (FP64, edge order)
  static int sym6(int a, int b) {
    static const int t[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    return t[a][b];
  }
  const std::vector<R>& Hel() const { return cp ? Hiso : Helast; }
  bool assembled_symmetric = true;
  long asym_count = 0;

  int blk(int i, int j) const {
    auto b = ci.begin() + rp[i], e = ci.begin() + rp[i + 1];
    auto it = std::lower_bound(b, e, j);
    if (it == e || *it != j) return -1;
    return int(it - ci.begin());
  }
  V3<R> X(int i) const { return {x[3 * i], x[3 * i + 1], x[3 * i + 2]}; }
  V3<R> Xof(const std::vector<R>& a, int i) const { return {a[3 * i], a[3 * i + 1], a[3 * i + 2]}; }

  // Mesh::compute_rest_state (mesh.cpp:18-75), compute_edges (:77-98), build_topology (:100-120).
  void set_mesh(const std::vector<V3<R>>& verts, const std::vector<std::array<int, 3>>& t) {
    mverts = verts;
    tris = t;
    edges.clear();
    std::unordered_set<uint64_t> es;
    for (auto& tr : tris)
      for (int i = 0; i < 3; ++i) {
        int a = tr[i], b = tr[(i + 1) % 3];
        uint64_t key = a < b ? (uint64_t(uint32_t(a)) << 32) | uint32_t(b) : (uint64_t(uint32_t(b)) << 32) | uint32_t(a);
        if (es.insert(key).second) edges.push_back({std::min(a, b), std::max(a, b)});
      }
    mean_edge = 0.0;
    for (const auto& e : edges) {
      const double dx = (double)mverts[e[1]][0] - (double)mverts[e[0]][0];
      const double dy = (double)mverts[e[1]][1] - (double)mverts[e[0]][1];
      const double dz = (double)mverts[e[1]][2] - (double)mverts[e[0]][2];
      mean_edge += std::sqrt(dx * dx + dy * dy + dz * dz);
    }
    if (!edges.empty()) mean_edge /= (double)edges.size();
    int nv = (int)mverts.size();
    v2f.assign(nv, {});
    for (size_t f = 0; f < tris.size(); ++f)
      for (int j = 0; j < 3; ++j) v2f[tris[f][j]].push_back((int)f);
    size_t M = tris.size();
    for (auto& d : Dm) d.assign(M, R(0));
    area.assign(M, R(0));
    for (size_t f = 0; f < M; ++f) {
      V3<R> v0 = mverts[tris[f][0]], v1 = mverts[tris[f][1]], v2 = mverts[tris[f][2]];
      V3<R> e1 = v1 - v0, e2 = v2 - v0;
      const R eps = R(1e-12);
      R l1 = sqn3(e1), l2 = sqn3(e2);
      if (l1 < eps || l2 < eps) continue;
      V3<R> nn = cross3(e1, e2);
      R nl = norm3(nn);
      if (nl < eps) continue;
      area[f] = nl * R(0.5);
      V3<R> nu_ = nn / nl;
      V3<R> t1 = e1 / std::sqrt(l1);
      V3<R> t2 = cross3(nu_, t1);
      R a00 = dot3(e1, t1), a10 = dot3(e1, t2), a01 = dot3(e2, t1), a11 = dot3(e2, t2);
      R invdet = R(1) / (a00 * a11 - a10 * a01);
      Dm[0][f] = a11 * invdet;
      Dm[2][f] = -a10 * invdet;
      Dm[1][f] = -a01 * invdet;
      Dm[3][f] = a00 * invdet;
    }
    build_pattern();
  }

  void build_pattern() {
    std::vector<std::vector<int>> nb(mverts.size());
    for (size_t i = 0; i < mverts.size(); ++i) nb[i].push_back((int)i);
    for (auto& e : edges) { nb[e[0]].push_back(e[1]); nb[e[1]].push_back(e[0]); }
    rp.assign(mverts.size() + 1, 0);
    ci.clear();
    diag.assign(mverts.size(), -1);
    for (size_t i = 0; i < mverts.size(); ++i) {
      std::sort(nb[i].begin(), nb[i].end());
      nb[i].erase(std::unique(nb[i].begin(), nb[i].end()), nb[i].end());
      for (int j : nb[i]) { if (j == (int)i) diag[i] = (int)ci.size(); ci.push_back(j); }
      rp[i + 1] = (int)ci.size();
    }
  }

  // State::initialize + compute_lumped_masses (state.cpp:7-33).
  void init_state() {
    if (mverts.empty()) throw std::invalid_argument("Mesh must be initialised before creating a state");
    n = (int)mverts.size();
    x.assign(3 * n, R(0));
    for (int i = 0; i < n; ++i) for (int k = 0; k < 3; ++k) x[3 * i + k] = mverts[i][k];
    xprev = x;
    v.assign(3 * n, R(0));
    m.assign(n, R(0));
    for (size_t f = 0; f < tris.size(); ++f) {
      R fm = area[f] * thick * rho;
      R vm = (R)((double)fm / 3.0);
      for (int j = 0; j < 3; ++j) m[tris[f][j]] += vm;
    }
  }

  R mu_() const { return (R)((double)E / (2.0 * (1.0 + (double)nu))); }

  // Mesh::compute_F (mesh.cpp:126-150) on an arbitrary position set.
  void compute_F(const V3<R>& v0, const V3<R>& v1, const V3<R>& v2, size_t f, R F[2][2]) const {
    V3<R> e1 = v1 - v0, e2 = v2 - v0;
    V3<R> nn = normalized3(cross3(e1, e2));
    V3<R> t1 = normalized3(e1);
    V3<R> t2 = cross3(nn, t1);
    R D[2][2] = {{dot3(e1, t1), dot3(e2, t1)}, {dot3(e1, t2), dot3(e2, t2)}};
    R Mi[2][2] = {{Dm[0][f], Dm[1][f]}, {Dm[2][f], Dm[3][f]}};
    for (int r = 0; r < 2; ++r)
      for (int c = 0; c < 2; ++c) F[r][c] = D[r][0] * Mi[0][c] + D[r][1] * Mi[1][c];
  }

  // EnergyTracker::compute (energy_tracker.cpp:7-76) on caller-given x, v, m
  // with the elastic energy of Elasticity::compute_energy (elasticity.cpp:5-14,
  // 115-126; F from the mesh vertices).  Order of out: kinetic, elastic,
  // linear momentum (3), angular momentum (3), max |v|.  out[0..8] accumulate
  // as the reference does (in R; the kinetic sum as R(double(ke) + 0.5*m*|v|^2));
  // out[9..17] sum the same per-item R terms in double.
  void energy(const double* xd, const double* vd, const double* md, double* out) const {
    R ke = R(0), el = R(0), vmax = R(0);
    V3<R> lm{R(0), R(0), R(0)}, am{R(0), R(0), R(0)};
    double ke_d = 0.0, el_d = 0.0, lm_d[3] = {0, 0, 0}, am_d[3] = {0, 0, 0};
    for (int i = 0; i < n; ++i) {
      V3<R> vi{(R)vd[3 * i], (R)vd[3 * i + 1], (R)vd[3 * i + 2]};
      V3<R> ri{(R)xd[3 * i], (R)xd[3 * i + 1], (R)xd[3 * i + 2]};
      R mi = (R)md[i];
      R vsq = sqn3(vi);
      double term = 0.5 * (double)mi * (double)vsq;
      ke = (R)((double)ke + term);
      ke_d += term;
      V3<R> p{mi * vi[0], mi * vi[1], mi * vi[2]};
      V3<R> c = cross3(ri, p);
      for (int k = 0; k < 3; ++k) {
        lm[k] = lm[k] + p[k];
        am[k] = am[k] + c[k];
        lm_d[k] += (double)p[k];
        am_d[k] += (double)c[k];
      }
      vmax = std::max(vmax, (R)std::sqrt(vsq));
    }
    R mu = mu_();
    for (size_t f = 0; f < tris.size(); ++f) {
      R F[2][2];
      compute_F(mverts[tris[f][0]], mverts[tris[f][1]], mverts[tris[f][2]], f, F);
      R k = area[f] * thick * mu;
      R d00 = F[0][0] - R(1), d01 = F[0][1], d10 = F[1][0], d11 = F[1][1] - R(1);
      R e = k * ((d00 * d00 + d10 * d10) + (d01 * d01 + d11 * d11));
      el = el + e;
      el_d += (double)e;
    }
    const double a[9] = {(double)ke, (double)el, (double)lm[0], (double)lm[1], (double)lm[2],
                         (double)am[0], (double)am[1], (double)am[2], (double)vmax};
    const double b[9] = {ke_d, el_d, lm_d[0], lm_d[1], lm_d[2], am_d[0], am_d[1], am_d[2], (double)vmax};
    for (int k = 0; k < 9; ++k) {
      out[k] = a[k];
      out[9 + k] = b[k];
    }
  }

  // Elasticity::compute_gradient (elasticity.cpp:18-61).  Faithful: F from the
  // mesh buffer (mesh.vertices), which Integrator::step never updates.
  void elastic_gradient() {
    egrad.assign(3 * n, R(0));
    R mu = mu_();
    for (size_t f = 0; f < tris.size(); ++f) {
      const auto& tr = tris[f];
      V3<R> v0 = mverts[tr[0]], v1 = mverts[tr[1]], v2 = mverts[tr[2]];
      R F[2][2];
      compute_F(v0, v1, v2, f, F);
      R k = area[f] * thick * mu;
      R twok = (R)(2.0 * (double)k);
      R Pm[2][2] = {{twok * (F[0][0] - R(1)), twok * (F[0][1] - R(0))}, {twok * (F[1][0] - R(0)), twok * (F[1][1] - R(1))}};
      R Mi[2][2] = {{Dm[0][f], Dm[1][f]}, {Dm[2][f], Dm[3][f]}};
      R Hm[2][2];
      for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c) Hm[r][c] = Pm[r][0] * Mi[c][0] + Pm[r][1] * Mi[c][1];
      V3<R> e1 = v1 - v0, e2 = v2 - v0;
      V3<R> nn = normalized3(cross3(e1, e2));
      V3<R> t1 = normalized3(e1);
      V3<R> t2 = cross3(nn, t1);
      V3<R> f1 = Hm[0][0] * t1 + Hm[1][0] * t2;
      V3<R> f2 = Hm[0][1] * t1 + Hm[1][1] * t2;
      V3<R> f0 = -(f1 + f2);
      for (int k2 = 0; k2 < 3; ++k2) {
        egrad[tr[0] * 3 + k2] += f0[k2];
        egrad[tr[1] * 3 + k2] += f1[k2];
        egrad[tr[2] * 3 + k2] += f2[k2];
      }
    }
  }

  // Elasticity::face_hessian + enforce_spd per block (elasticity.cpp:63-91,135-181).
  void elastic_blocks() {
    ec.assign(tris.size(), {});
    R mu = mu_();
    const R eps = R(1e-8);
    for (size_t f = 0; f < tris.size(); ++f) {
      R k = area[f] * thick * mu;
      R Mi[2][2] = {{Dm[0][f], Dm[1][f]}, {Dm[2][f], Dm[3][f]}};
      R G[2][2];
      for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c) G[r][c] = Mi[0][r] * Mi[0][c] + Mi[1][r] * Mi[1][c];
      R K00 = k * G[0][0], K01 = k * G[0][1], K11 = k * G[1][1];
      R c[3][3];
      c[0][0] = (R)((double)(K00 + K11) + 2.0 * (double)K01);
      c[1][1] = K00; c[2][2] = K11;
      c[0][1] = c[1][0] = -K00;
      c[0][2] = c[2][0] = -K11;
      c[1][2] = c[2][1] = -K01;
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          M3<R> B;
          B.m[0][0] = B.m[1][1] = c[a][b];
          enforce_spd(B, eps);
          // Diagonal in, diagonal out: x,y entries equal, z entry = eps.
          ec[f][a * 3 + b] = B.m[0][0];
        }
    }
  }

  // Mass/dt^2 + elastic fold (integrator.cpp:287-304, matrix_assembly.cpp:27-61).
  void build_const() {
    R dt = P.dt;
    R dtinv2 = R(1) / (dt * dt);
    size_t nb = ci.size();
    Hconst.assign(9 * nb, R(0));
    Helast.assign(9 * nb, R(0));
    const R eps = R(1e-8);
    for (int i = 0; i < n; ++i) {
      R mf = m[i] * dtinv2;
      if (mf != R(0))
        for (int k = 0; k < 3; ++k) Hconst[9 * diag[i] + 4 * k] += mf;
    }
    for (size_t f = 0; f < tris.size(); ++f)
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          int bi = blk(tris[f][a], tris[f][b]);
          R cc = ec[f][a * 3 + b];
          R vals[9] = {cc, 0, 0, 0, cc, 0, 0, 0, eps};
          for (int e = 0; e < 9; ++e) {
            Hconst[9 * bi + e] += vals[e];
            Helast[9 * bi + e] += vals[e];
          }
        }
  }

  M3<R> block_of(const std::vector<R>& vals, int i) const {
    M3<R> B;
    int b = diag[i];
    for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) B.m[r][c] = vals[9 * b + 3 * r + c];
    return B;
  }

  // Stiffness::compute_contact_stiffness (stiffness.cpp:53-116).
  R contact_stiffness(const Contact<R>& c) const {
    R dt = P.dt;
    if (dt <= R(0)) return R(0);
    V3<R> nn = c.normal;
    R nl = norm3(nn);
    if ((double)nl < 1e-12) return R(0);
    nn = nn / nl;
    int vc = std::max(c.vcount, 1);
    std::vector<int> vv;
    std::vector<R> ww;
    for (int i = 0; i < vc; ++i) {
      int idx = c.idx[i];
      if (idx < 0 || idx >= n) continue;
      if (c.w[i] == R(0)) continue;
      vv.push_back(idx);
      ww.push_back(c.w[i]);
    }
    if (vv.empty()) return R(0);
    R inertial = R(0);
    for (size_t i = 0; i < vv.size(); ++i) inertial += m[vv[i]] * ww[i] * ww[i];
    if (cp) {
      // Corrected physics: the wall's inertial scale m / gap_max^2
      // (stiffness.cpp:146-172) instead of m / dt^2, so the barrier's force
      // near contact (1.5 k gap_max) stops what the wall stops.
      const R gh = std::max(std::max(P.gap_max, P.min_gap), R(1e-12));
      inertial /= (gh * gh);
    } else {
      inertial /= (dt * dt);
    }
    int bc = (int)vv.size();
    std::vector<R> Wn(3 * bc), Ht(3 * bc, R(0));
    for (int i = 0; i < bc; ++i) for (int k = 0; k < 3; ++k) Wn[3 * i + k] = ww[i] * nn[k];
    // H_sub * Wn (dynamic GEMV, column order) over the elastic-only matrix.
    for (int cb = 0; cb < bc; ++cb)
      for (int rb = 0; rb < bc; ++rb) {
        int bi = blk(vv[rb], vv[cb]);
        if (bi < 0) continue;
        for (int r = 0; r < 3; ++r)
          for (int k = 0; k < 3; ++k) Ht[3 * rb + r] += Hel()[9 * bi + 3 * r + k] * Wn[3 * cb + k];
      }
    R el = Wn[0] * Ht[0];
    for (int i = 1; i < 3 * bc; ++i) el = el + Wn[i] * Ht[i];
    if (el < R(0)) el = R(0);
    return inertial + el;
  }

  // Stiffness::compute_pin_stiffness (stiffness.cpp:118-144).
  R pin_stiffness(R mass, const V3<R>& off, M3<R> Hb) const {
    enforce_spd(Hb);
    V3<R> dir = off;
    R len = norm3(dir);
    if ((double)len > 1e-9) dir = dir / len; else dir = V3<R>(1, 0, 0);
    R kel = std::max(R(0), dot3(dir, mul(Hb, dir)));
    R gh = std::max(std::max(len, P.min_gap), R(1e-12));
    return mass / (gh * gh) + kel;
  }
  // Stiffness::compute_wall_stiffness (stiffness.cpp:146-172).
  R wall_stiffness(R mass, const V3<R>& normal, M3<R> Hb) const {
    R gh = std::max(std::max(P.wall_gap, P.min_gap), R(1e-12));
    R kin = mass / (gh * gh);
    enforce_spd(Hb);
    V3<R> nn = normal;
    R nl = norm3(nn);
    if ((double)nl > 1e-9) nn = nn / nl; else nn = V3<R>(0, 1, 0);
    R kel = std::max(R(0), dot3(nn, mul(Hb, nn)));
    return kin + kel;
  }
  V3<R> normalize_or_default(const V3<R>& nn) const {
    if (sqn3(nn) > P.normal_eps * P.normal_eps) return normalized3(nn);
    return V3<R>(0, 1, 0);
  }

  R to_fraction(R val) const { return val > R(1) ? val * R(0.01) : val; }

  // SVD of a 3x2 F (one-sided Jacobi on F^T F).  Restatement of the role of
  // Eigen::JacobiSVD (strain_limiting.cpp:28-50): singular values sorted
  // descending, non-negative, with the 1e-6 floor.
  static bool svd32(const R F[3][2], R U[3][2], R sig[2], R Vm[2][2]) {
    R a = F[0][0] * F[0][0] + F[1][0] * F[1][0] + F[2][0] * F[2][0];
    R b = F[0][0] * F[0][1] + F[1][0] * F[1][1] + F[2][0] * F[2][1];
    R c = F[0][1] * F[0][1] + F[1][1] * F[1][1] + F[2][1] * F[2][1];
    R cs = R(1), sn = R(0);
    if (b != R(0)) {
      R zeta = (c - a) / (R(2) * b);
      R t = (zeta >= R(0) ? R(1) : R(-1)) / (std::fabs(zeta) + std::sqrt(R(1) + zeta * zeta));
      cs = R(1) / std::sqrt(R(1) + t * t);
      sn = cs * t;
    }
    R v0[2] = {cs, -sn}, v1[2] = {sn, cs};
    R u0[3], u1[3];
    for (int r = 0; r < 3; ++r) {
      u0[r] = F[r][0] * v0[0] + F[r][1] * v0[1];
      u1[r] = F[r][0] * v1[0] + F[r][1] * v1[1];
    }
    R s0 = std::sqrt(u0[0] * u0[0] + u0[1] * u0[1] + u0[2] * u0[2]);
    R s1 = std::sqrt(u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
    if (s1 > s0) {
      std::swap(s0, s1);
      for (int r = 0; r < 3; ++r) std::swap(u0[r], u1[r]);
      std::swap(v0[0], v1[0]); std::swap(v0[1], v1[1]);
    }
    if (s0 > R(0)) for (int r = 0; r < 3; ++r) u0[r] /= s0;
    else { u0[0] = 1; u0[1] = 0; u0[2] = 0; }
    if (s1 > R(0)) for (int r = 0; r < 3; ++r) u1[r] /= s1;
    else {
      // Any unit vector orthogonal to u0.
      R w[3] = {R(0), R(0), R(0)};
      int ax = std::fabs(u0[0]) < R(0.6) ? 0 : 1;
      w[ax] = R(1);
      R d = w[0] * u0[0] + w[1] * u0[1] + w[2] * u0[2];
      for (int r = 0; r < 3; ++r) w[r] -= d * u0[r];
      R l = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
      for (int r = 0; r < 3; ++r) u1[r] = w[r] / l;
    }
    for (int r = 0; r < 3; ++r) { U[r][0] = u0[r]; U[r][1] = u1[r]; }
    Vm[0][0] = v0[0]; Vm[1][0] = v0[1]; Vm[0][1] = v1[0]; Vm[1][1] = v1[1];
    sig[0] = s0; sig[1] = s1;
    const R th = R(1e-6);
    if (sig[0] < th || sig[1] < th) { sig[0] = std::max(sig[0], th); sig[1] = std::max(sig[1], th); }
    return std::isfinite(sig[0]) && std::isfinite(sig[1]);
  }

  // StrainLimiting::rebuild_constraints / accumulate_gradient / accumulate_hessian
  // (strain_limiting.cpp:113-375), evaluated once per face at the current x.
  void strain_faces() {
    sf.assign(tris.size(), StrainFace<R>());
    st.strain_active = 0;
    if (!P.enable_strain) return;
    R tau = to_fraction(P.strain_tau);
    R eps = P.strain_eps > R(0) ? to_fraction(P.strain_eps) : tau;
    if (eps <= R(0)) return;
    const R min_gap = std::max(P.min_gap, R(1e-8));
    const R svd_eps = std::max(P.strain_svd_eps, R(1e-8));
    const R tiny = R(1e-12);
    int active_count = 0;
#pragma omp parallel for schedule(static) reduction(+ : active_count)
    for (size_t f = 0; f < tris.size(); ++f) {
      if (area[f] <= tiny) continue;
      const auto& tr = tris[f];
      V3<R> x0 = X(tr[0]), x1 = X(tr[1]), x2 = X(tr[2]);
      V3<R> d1 = x1 - x0, d2 = x2 - x0;
      R Mi[2][2] = {{Dm[0][f], Dm[1][f]}, {Dm[2][f], Dm[3][f]}};
      R F[3][2];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 2; ++c) F[r][c] = d1[r] * Mi[0][c] + d2[r] * Mi[1][c];
      R U[3][2], sig[2], Vm[2][2];
      if (!svd32(F, U, sig, Vm)) continue;
      R fmass = area[f] * thick * rho;
      V3<R> cen = ((x0 + x1) + x2) / R(3);
      R w[9];
      for (int k = 0; k < 3; ++k) { w[k] = x0[k] - cen[k]; w[3 + k] = x1[k] - cen[k]; w[6 + k] = x2[k] - cen[k]; }
      // w^T H_sym w with H the 9x9 elastic-only face block.  The elastic
      // matrix is exactly symmetric, so H_sym = H.  Sequential order.
      R Hw[9];
      for (int a = 0; a < 3; ++a)
        for (int k = 0; k < 3; ++k) {
          R s = R(0);
          for (int b = 0; b < 3; ++b) {
            int bi = blk(tr[a], tr[b]);
            for (int l = 0; l < 3; ++l) s += Hel()[9 * bi + 3 * k + l] * w[3 * b + l];
          }
          Hw[3 * a + k] = s;
        }
      R el = R(0);
      for (int i = 0; i < 9; ++i) el += w[i] * Hw[i];
      el = std::max(el, R(0));
      StrainFace<R>& S = sf[f];
      bool tie = std::fabs(sig[0] - sig[1]) < svd_eps;
      for (int s = 0; s < 2; ++s) {
        R gap = (R(1) + tau + eps) - sig[s];
        if (!in_domain(gap, eps)) continue;
        // Corrected physics: the band width as the stiffness scale (as the
        // wall uses its gap), so the stiffness stays bounded at the limit.
        R gc = cp ? eps : std::max(std::fabs(gap), min_gap);
        S.active[s] = 1;
        S.k[s] = fmass / (gc * gc) + el;
        S.gap[s] = gap;
        ++active_count;
        R dS[3][2];
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 2; ++c)
            dS[r][c] = tie ? (U[r][0] * Vm[c][0] + U[r][1] * Vm[c][1]) * R(0.5) : U[r][s] * Vm[c][s];
        // Gradient (strain_limiting.cpp:250-269).
        R dV = barrier_grad(gap, eps, S.k[s]);
        if (!(std::fabs(dV) < tiny)) {
          S.grad_on[s] = 1;
          R mapped[3][2];
          for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 2; ++c) {
              R a0 = -dV * dS[r][0], a1 = -dV * dS[r][1];
              mapped[r][c] = a0 * Mi[c][0] + a1 * Mi[c][1];
            }
          for (int r = 0; r < 3; ++r) {
            S.gradc[s][1][r] = mapped[r][0];
            S.gradc[s][2][r] = mapped[r][1];
            S.gradc[s][0][r] = -(mapped[r][0] + mapped[r][1]);
          }
        }
        // Hessian (strain_limiting.cpp:328-349).
        R dX[3][2];
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 2; ++c) dX[r][c] = dS[r][0] * Mi[c][0] + dS[r][1] * Mi[c][1];
        for (int r = 0; r < 3; ++r) {
          S.J[s][r] = -(dX[r][0] + dX[r][1]);
          S.J[s][3 + r] = dX[r][0];
          S.J[s][6 + r] = dX[r][1];
        }
        R d2 = barrier_hess(gap, eps, S.k[s]);
        if (!(std::fabs(d2) < tiny)) { S.hess_on[s] = 1; S.d2V[s] = d2; }
      }
    }
    st.strain_active = active_count;
  }

  // Collision::detect_all_collisions (collision.cpp:478-580).  Faithful: the
  // self BVH yields no primitive leaves unless the mesh has one triangle (F1).
  // Corrected self detection (collision.cpp:198-276, 478-534 as intended):
  // vertex v against every triangle not incident to v whose box overlaps
  // [p - pt_box, p + pt_box] (closed intervals); edges e1 < e2 sharing no
  // vertex whose boxes overlap; narrow phase, kept if gap < keep_gap.
  void detect_self_corrected(std::vector<Contact<R>>& out) const {
    const R eb = pt_box;
    for (int vi = 0; vi < n; ++vi) {
      V3<R> p = X(vi);
      for (size_t t = 0; t < tris.size(); ++t) {
        const auto& tr = tris[t];
        if (tr[0] == vi || tr[1] == vi || tr[2] == vi) continue;
        V3<R> a = X(tr[0]), b = X(tr[1]), c = X(tr[2]);
        bool ov = true;
        for (int k = 0; k < 3; ++k) {
          R mn = std::min(std::min(a[k], b[k]), c[k]), mx = std::max(std::max(a[k], b[k]), c[k]);
          ov = ov && (p[k] - eb <= mx && p[k] + eb >= mn);
        }
        if (!ov) continue;
        Contact<R> cp;
        cp.type = PT; cp.idx[0] = vi; cp.idx[1] = tr[0]; cp.idx[2] = tr[1]; cp.idx[3] = tr[2];
        narrow_pt(p, a, b, c, cp.gap, cp.normal, cp.wq);
        cp.wp = p;
        if (cp.gap < keep_gap) {
          cp.bary = triangle_barycentric(cp.wq, a, b, c);
          cp.w[0] = R(1); cp.w[1] = -cp.bary[0]; cp.w[2] = -cp.bary[1]; cp.w[3] = -cp.bary[2];
          cp.vcount = 4;
          out.push_back(cp);
        }
      }
    }
    for (size_t e1 = 0; e1 < edges.size(); ++e1)
      for (size_t e2 = e1 + 1; e2 < edges.size(); ++e2) {
        const auto& A = edges[e1];
        const auto& B = edges[e2];
        if (A[0] == B[0] || A[0] == B[1] || A[1] == B[0] || A[1] == B[1]) continue;
        V3<R> p0 = X(A[0]), p1 = X(A[1]), q0 = X(B[0]), q1 = X(B[1]);
        bool ov = true;
        for (int k = 0; k < 3; ++k) {
          R amn = std::min(p0[k], p1[k]), amx = std::max(p0[k], p1[k]);
          R bmn = std::min(q0[k], q1[k]), bmx = std::max(q0[k], q1[k]);
          ov = ov && (amn <= bmx && amx >= bmn);
        }
        if (!ov) continue;
        Contact<R> cp;
        cp.type = EE; cp.idx[0] = A[0]; cp.idx[1] = A[1]; cp.idx[2] = B[0]; cp.idx[3] = B[1];
        narrow_ee(p0, p1, q0, q1, cp.gap, cp.normal, cp.wp, cp.wq);
        if (cp.gap < keep_gap) {
          cp.vcount = 4;
          cp.w[0] = R(0.5); cp.w[1] = R(0.5); cp.w[2] = R(-0.5); cp.w[3] = R(-0.5);
          out.push_back(cp);
        }
      }
  }

  void detect(std::vector<Contact<R>>& out) {
    out.clear();
    if (contact_mode == 1) {
      detect_self_corrected(out);
    } else if (tris.size() == 1) {
      // Single-leaf BVH: vertex boxes (±1e-4) against the one triangle box,
      // incident vertices skipped (collision.cpp:198-244, 279-387).
      const auto& tr = tris[0];
      V3<R> a = X(tr[0]), b = X(tr[1]), c = X(tr[2]);
      V3<R> mn, mx;
      for (int k = 0; k < 3; ++k) {
        mn[k] = std::min(std::min(a[k], b[k]), c[k]);
        mx[k] = std::max(std::max(a[k], b[k]), c[k]);
      }
      for (int vi = 0; vi < n; ++vi) {
        if (vi == tr[0] || vi == tr[1] || vi == tr[2]) continue;
        V3<R> p = X(vi);
        bool ov = true;
        for (int k = 0; k < 3; ++k) {
          R lo = p[k] - R(1e-4), hi = p[k] + R(1e-4);
          ov = ov && (lo <= mx[k] && hi >= mn[k]);
        }
        if (!ov) continue;
        Contact<R> cp;
        cp.type = PT; cp.idx[0] = vi; cp.idx[1] = tr[0]; cp.idx[2] = tr[1]; cp.idx[3] = tr[2];
        narrow_pt(p, a, b, c, cp.gap, cp.normal, cp.wq);
        cp.wp = p;
        if (cp.gap < R(0.01)) {
          cp.bary = triangle_barycentric(cp.wq, a, b, c);
          cp.w[0] = R(1); cp.w[1] = -cp.bary[0]; cp.w[2] = -cp.bary[1]; cp.w[3] = -cp.bary[2];
          cp.vcount = 4;
          out.push_back(cp);
        }
      }
    }
    // Per-vertex lists filled in parallel, then appended in vertex order, so
    // the list order is the sequential (body, vertex, triangle) order.
    std::vector<std::vector<Contact<R>>> per(n);
    for (size_t rb = 0; rb < rigids.size(); ++rb) {
      auto W = rigids[rb].world();
#pragma omp parallel for schedule(dynamic, 64)
      for (int vi = 0; vi < n; ++vi) {
        per[vi].clear();
        V3<R> p = X(vi);
        for (const auto& t : rigids[rb].tris) {
          Contact<R> cp;
          cp.type = RIGID_PT; cp.idx[0] = vi; cp.idx[1] = t[0]; cp.idx[2] = t[1]; cp.idx[3] = t[2];
          cp.rigid = (int)rb;
          narrow_pt(p, W[t[0]], W[t[1]], W[t[2]], cp.gap, cp.normal, cp.wq);
          cp.wp = p;
          if (cp.gap < R(0.01)) { cp.vcount = 1; cp.w[0] = R(1); per[vi].push_back(cp); }
        }
      }
      for (int vi = 0; vi < n; ++vi) out.insert(out.end(), per[vi].begin(), per[vi].end());
    }
  }

  // Integrator::compute_gradient (integrator.cpp:246-396).
  void compute_gradient(const std::vector<R>& xt, std::vector<R>& g) {
    R dt = P.dt;
    g.assign(3 * n, R(0));
    for (int i = 0; i < n; ++i) {
      R mf = m[i] / (dt * dt);
      for (int j = 0; j < 3; ++j) g[3 * i + j] += mf * (x[3 * i + j] - xt[3 * i + j]);
    }
    if (cp) elastic_gradient_x();
    for (int i = 0; i < 3 * n; ++i) g[i] += egrad[i];
    strain_faces();
    for (size_t f = 0; f < tris.size(); ++f)
      for (int s = 0; s < 2; ++s)
        if (sf[f].active[s] && sf[f].grad_on[s])
          for (int a = 0; a < 3; ++a)
            for (int k = 0; k < 3; ++k) g[tris[f][a] * 3 + k] += sf[f].gradc[s][a][k];
    for (const auto& c : contacts) {
      if (c.type != PT && c.type != RIGID_PT && !(c.type == EE && contact_mode == 1)) continue;
      R kb = contact_stiffness(c);
      if (!in_domain(c.gap, P.gap_max)) continue;
      V3<R> nn = normalize_or_default(c.normal);
      R dV = barrier_grad(c.gap, P.gap_max, kb);
      if (dV == R(0)) continue;
      int cnt = std::max(c.vcount, 1);
      for (int i = 0; i < cnt; ++i) {
        int idx = c.idx[i];
        if (idx < 0 || c.w[i] == R(0)) continue;
        R s = dV * c.w[i];
        for (int k = 0; k < 3; ++k) g[idx * 3 + k] += s * nn[k];
      }
    }
    for (const auto& p : pins) {
      if (!p.active) continue;
      V3<R> off = X(p.v) - p.t;
      if (cp) {
        const R kp = pin_k * (m[p.v] / (dt * dt));
        for (int k = 0; k < 3; ++k) g[p.v * 3 + k] += kp * off[k];
        continue;
      }
      R kb = pin_stiffness(m[p.v], off, block_of(Hconst, p.v));
      R gap = norm3(off);
      if (!in_domain(gap, P.gap_max) || gap <= P.normal_eps) continue;
      V3<R> nn = off / gap;
      R dV = barrier_grad(gap, P.gap_max, kb);
      for (int k = 0; k < 3; ++k) g[p.v * 3 + k] += dV * nn[k];
    }
    for (const auto& w : walls) {
      if (!w.active) continue;
#pragma omp parallel for schedule(static)
      for (int i = 0; i < n; ++i) {
        R kb = wall_stiffness(m[i], w.n, block_of(cp ? Hbase_iso : Hconst, i));
        V3<R> nn = normalize_or_default(w.n);
        R gap = dot3(nn, X(i)) - w.off;
        if (!in_domain(gap, P.gap_max)) continue;
        R dV = barrier_grad(gap, P.gap_max, kb);
        for (int k = 0; k < 3; ++k) g[i * 3 + k] += dV * nn[k];
      }
    }
    if (P.enable_friction && P.mu > R(0))
      for (const auto& c : contacts) {
        if (contact_mode == 1 && c.type != RIGID_PT) continue;  // corrected mode: rigid contacts only
        R kf; V3<R> tan;
        if (!friction_data(c, kf, tan)) continue;
        for (int k = 0; k < 3; ++k) g[3 * c.idx[0] + k] += kf * tan[k];
      }
  }

  // compute_friction_data (integrator.cpp:28-66) with friction.cpp:6-90.
  bool friction_data(const Contact<R>& c, R& kf, V3<R>& tan) const {
    V3<R> d = X(c.idx[0]) - Xof(xprev, c.idx[0]);
    R nc = dot3(d, c.normal);
    tan = d - nc * c.normal;
    R tn = norm3(tan);
    if (!(tn > P.friction_tangent)) return false;
    R kc = contact_stiffness(c);
    R nf = kc * std::fabs(c.gap);
    R den = std::max(P.friction_eps, std::fabs(tn));
    if (den <= R(0)) den = P.friction_eps;
    R s = P.mu * std::fabs(nf) / den;
    kf = std::min(s, R(1e8));
    return true;
  }

  // Integrator::assemble_system_matrix (integrator.cpp:398-544): triplets
  // folded into the fixed pattern in list order, then (H + H^T) * 0.5.
  void assemble() {
    size_t nb = ci.size();
    H.assign(9 * nb, R(0));
    R dt = P.dt;
    R dti = R(1) / (dt * dt);
    auto add = [&](int r, int c, R val) {
      int bi = blk(r / 3, c / 3);
      H[9 * bi + 3 * (r % 3) + (c % 3)] += val;
    };
    for (int row = 0; row < 3 * n; ++row) {
      R val = m[row / 3] * dti;
      if (val == R(0)) continue;
      add(row, row, val);
    }
    const R eps = R(1e-8);
    if (cp) {
      cp_projectors();
      for (size_t f = 0; f < tris.size(); ++f)
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) {
            const R gab = gk[f][sym6(a, b)];
            for (int e = 0; e < 9; ++e)
              add(tris[f][a] * 3 + e / 3, tris[f][b] * 3 + e % 3, gab * fpi[f][sym6(e / 3, e % 3)]);
          }
    } else {
      for (size_t f = 0; f < tris.size(); ++f)
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) {
            R cc = ec[f][a * 3 + b];
            R vals[9] = {cc, 0, 0, 0, cc, 0, 0, 0, eps};
            for (int e = 0; e < 9; ++e) add(tris[f][a] * 3 + e / 3, tris[f][b] * 3 + e % 3, vals[e]);
          }
    }
    const R tiny = R(1e-12);
    if (P.enable_strain)
      for (size_t f = 0; f < tris.size(); ++f)
        for (int s = 0; s < 2; ++s) {
          const auto& S = sf[f];
          if (!S.active[s] || !S.hess_on[s]) continue;
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
              R blkv[9];
              R mx = R(0);
              for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                  blkv[3 * i + j] = S.d2V[s] * (S.J[s][3 * a + i] * S.J[s][3 * b + j]);
                  mx = std::max(mx, std::fabs(blkv[3 * i + j]));
                }
              if (mx < tiny) continue;
              for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                  R val = blkv[3 * i + j];
                  if (std::fabs(val) < tiny) continue;
                  add(tris[f][a] * 3 + i, tris[f][b] * 3 + j, val);
                }
            }
        }
    couplings.clear();
    for (const auto& c : contacts) {
      if (c.type != PT && c.type != RIGID_PT && !(c.type == EE && contact_mode == 1)) continue;
      R kb = contact_stiffness(c);
      if (!in_domain(c.gap, P.gap_max)) continue;
      V3<R> nn = normalize_or_default(c.normal);
      R d2 = barrier_hess(c.gap, P.gap_max, kb);
      if (d2 == R(0)) continue;
      int cnt = std::max(c.vcount, 1);
      for (int i = 0; i < cnt; ++i) {
        R wi = c.w[i];
        if (wi == R(0)) continue;
        V3<R> gi = wi * nn;
        for (int j = 0; j < cnt; ++j) {
          R wj = c.w[j];
          if (wj == R(0)) continue;
          V3<R> gj = wj * nn;
          if (c.idx[i] < 0 || c.idx[j] < 0) continue;
          if (c.idx[i] != c.idx[j]) {
            // Off-diagonal contact block (matrix_assembly.cpp:116-145 entry
            // rule): kept beside the pattern.
            Coupling cu{c.idx[i], c.idx[j], {}};
            bool any = false;
            for (int k = 0; k < 9; ++k) {
              R val = d2 * (gi[k / 3] * gj[k % 3]);
              cu.b[k] = std::fabs(val) < P.barrier_tol ? R(0) : val;
              any = any || cu.b[k] != R(0);
            }
            if (any) couplings.push_back(cu);
            continue;
          }
          for (int k = 0; k < 9; ++k) {
            R val = d2 * (gi[k / 3] * gj[k % 3]);
            if (std::fabs(val) < P.barrier_tol) continue;
            add(c.idx[i] * 3 + k / 3, c.idx[j] * 3 + k % 3, val);
          }
        }
      }
    }
    for (const auto& p : pins) {
      if (!p.active) continue;
      V3<R> off = X(p.v) - p.t;
      if (cp) {
        const R kp = pin_k * (m[p.v] / (dt * dt));
        for (int k = 0; k < 3; ++k) add(p.v * 3 + k, p.v * 3 + k, kp);
        continue;
      }
      R kb = pin_stiffness(m[p.v], off, block_of(Hconst, p.v));
      R gap = norm3(off);
      if (!in_domain(gap, P.gap_max) || gap <= P.normal_eps) continue;
      V3<R> nn = off / gap;
      R dV = barrier_grad(gap, P.gap_max, kb);
      R d2 = barrier_hess(gap, P.gap_max, kb);
      R q = dV / gap;
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          R nnrc = nn[r] * nn[c];
          R val = d2 * nnrc + q * ((r == c ? R(1) : R(0)) - nnrc);
          if (std::fabs(val) < P.barrier_tol) continue;
          add(p.v * 3 + r, p.v * 3 + c, val);
        }
    }
    for (const auto& w : walls) {
      if (!w.active) continue;
      for (int i = 0; i < n; ++i) {
        R kb = wall_stiffness(m[i], w.n, block_of(cp ? Hbase_iso : Hconst, i));
        V3<R> nn = normalize_or_default(w.n);
        R gap = dot3(nn, X(i)) - w.off;
        if (!in_domain(gap, P.gap_max)) continue;
        R d2 = barrier_hess(gap, P.gap_max, kb);
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) {
            R val = d2 * (nn[r] * nn[c]);
            if (std::fabs(val) < P.barrier_tol) continue;
            add(i * 3 + r, i * 3 + c, val);
          }
      }
    }
    if (P.enable_friction && P.mu > R(0))
      for (const auto& c : contacts) {
        if (contact_mode == 1 && c.type != RIGID_PT) continue;
        R kf; V3<R> tan;
        if (!friction_data(c, kf, tan)) continue;
        for (int r = 0; r < 3; ++r)
          for (int cc = 0; cc < 3; ++cc) {
            R val = (r == cc ? kf : R(0));
            val = val - kf * (c.normal[r] * c.normal[cc]);
            val = val + (r == cc ? R(1e-8) : R(0));
            if (!((double)std::fabs(val) > 1e-12)) continue;
            add(3 * c.idx[0] + r, 3 * c.idx[0] + cc, val);
          }
      }
    // (H + H^T) * 0.5
    std::vector<R> Hs(H.size());
    asym_count = 0;
    for (int i = 0; i < n; ++i)
      for (int b = rp[i]; b < rp[i + 1]; ++b) {
        int j = ci[b];
        int bt = blk(j, i);
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) {
            R a = H[9 * b + 3 * r + c], t = H[9 * bt + 3 * c + r];
            if (a != t) ++asym_count;
            Hs[9 * b + 3 * r + c] = (a + t) * R(0.5);
          }
      }
    H.swap(Hs);
  }

  // PCGSolver::solve (pcg_solver.cpp:8-146).  Returns converged flag; iters out.
  static bool pcg(int nv, const std::vector<int>& rp_, const std::vector<int>& ci_, const std::vector<int>& dg,
                  const std::vector<R>& A, const std::vector<R>& b, std::vector<R>& xs, R tol, int maxit,
                  int& iters, const std::vector<Coupling>* cpl = nullptr) {
    int N = 3 * nv;
    iters = 0;
    std::vector<M3<R>> Pc(nv);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nv; ++i) {
      M3<R> B;
      for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) B.m[r][c] = A[9 * dg[i] + 3 * r + c];
      R det = det3(B);
      if ((double)std::fabs(det) > 1e-16) Pc[i] = inv3(B);
      else Pc[i] = M3<R>::identity();
    }
    // Row-, vertex- and element-wise loops run on OpenMP threads (each
    // iteration writes its own outputs, so any thread count gives identical
    // results); the dot products stay sequential in the SSE2 order.
    auto spmv = [&](const std::vector<R>& p, std::vector<R>& out) {
#pragma omp parallel for schedule(static)
      for (int i = 0; i < nv; ++i)
        for (int r = 0; r < 3; ++r) {
          R tmp = R(0);
          for (int bb = rp_[i]; bb < rp_[i + 1]; ++bb) {
            int j = ci_[bb];
            for (int c = 0; c < 3; ++c) tmp += A[9 * bb + 3 * r + c] * p[3 * j + c];
          }
          out[3 * i + r] = tmp;
        }
      // Contact couplings outside the mesh pattern, in contact order.
      if (cpl)
        for (const auto& cu : *cpl)
          for (int r = 0; r < 3; ++r) {
            R t = R(0);
            for (int c = 0; c < 3; ++c) t += cu.b[3 * r + c] * p[3 * cu.j + c];
            out[3 * cu.i + r] += t;
          }
    };
    auto relres = [&](const std::vector<R>& r) {
      R ri = vinf(r.data(), N), bi = vinf(b.data(), N);
      if ((double)bi < 1e-16) return ri;
      return ri / bi;
    };
    auto precond = [&](const std::vector<R>& r, std::vector<R>& z) {
#pragma omp parallel for schedule(static)
      for (int i = 0; i < nv; ++i) {
        V3<R> rb(r[3 * i], r[3 * i + 1], r[3 * i + 2]);
        V3<R> zb = mul(Pc[i], rb);
        for (int k = 0; k < 3; ++k) z[3 * i + k] = zb[k];
      }
    };
    std::vector<R> Ax(N), r(N), z(N, R(0)), p(N), Ap(N);
    spmv(xs, Ax);
    for (int i = 0; i < N; ++i) r[i] = b[i] - Ax[i];
    R rel = relres(r);
    if (rel < tol) return true;
    precond(r, z);
    p = z;
    R rz = vdot(r.data(), z.data(), N);
    for (int it = 0; it < maxit; ++it) {
      spmv(p, Ap);
      R pAp = vdot(p.data(), Ap.data(), N);
      iters = it + 1;
      if ((double)std::fabs(pAp) < 1e-16) return false;
      R alpha = rz / pAp;
#pragma omp parallel for schedule(static)
      for (int i = 0; i < N; ++i) {
        xs[i] = xs[i] + alpha * p[i];
        r[i] = r[i] - alpha * Ap[i];
      }
      rel = relres(r);
      if (rel < tol) return true;
      precond(r, z);
      R rzn = vdot(r.data(), z.data(), N);
      R beta = rzn / rz;
      rz = rzn;
#pragma omp parallel for schedule(static)
      for (int i = 0; i < N; ++i) p[i] = z[i] + beta * p[i];
    }
    return false;
  }

  // LineSearch::search / is_feasible (line_search.cpp:8-161).
  bool feasible(const std::vector<R>& xn, const V3<R>& wn, R woff) const {
    const R gap_min = R(0);
    for (const auto& c : contacts) {
      if (c.type == PT) {
        V3<R> p0 = X(c.idx[0]), p1 = Xof(xn, c.idx[0]);
        V3<R> a0 = X(c.idx[1]), a1 = Xof(xn, c.idx[1]);
        V3<R> b0 = X(c.idx[2]), b1 = Xof(xn, c.idx[2]);
        V3<R> c0 = X(c.idx[3]), c1 = Xof(xn, c.idx[3]);
        if (ccd_mode == 1) {
          if (!still4(p0, p1, a0, a1, b0, b1, c0, c1) && accd_pair(0, p0, p1, a0, a1, b0, b1, c0, c1) < R(1))
            return false;
        } else if (ccd_pt(p0, p1, a0, a1, b0, b1, c0, c1) < R(1)) {
          return false;
        }
        R d; V3<R> nn, q;
        narrow_pt(p1, a1, b1, c1, d, nn, q);
        if (d < gap_min) return false;
      } else if (c.type == EE) {
        V3<R> a0 = X(c.idx[0]), a1 = Xof(xn, c.idx[0]);
        V3<R> b0 = X(c.idx[1]), b1 = Xof(xn, c.idx[1]);
        V3<R> c0 = X(c.idx[2]), c1 = Xof(xn, c.idx[2]);
        V3<R> d0 = X(c.idx[3]), d1 = Xof(xn, c.idx[3]);
        if (ccd_mode == 1) {
          if (!still4(a0, a1, b0, b1, c0, c1, d0, d1) && accd_pair(1, a0, a1, b0, b1, c0, c1, d0, d1) < R(1))
            return false;
        } else if (ccd_ee(a0, a1, b0, b1, c0, c1, d0, d1) < R(1)) {
          return false;
        }
        R d; V3<R> nn, wp, wq;
        narrow_ee(a1, b1, c1, d1, d, nn, wp, wq);
        if (d < gap_min) return false;
      }
    }
    for (const auto& p : pins) {
      if (!p.active) continue;
      if (norm3(Xof(xn, p.v) - p.t) < gap_min) return false;
    }
    if ((double)sqn3(wn) > 0.5) {
      bool ok = true;
#pragma omp parallel for schedule(static) reduction(&& : ok)
      for (int i = 0; i < n; ++i)
        if (dot3(wn, Xof(xn, i)) - woff < gap_min) ok = false;
      if (!ok) return false;
    }
    return true;
  }
  // No vertex of the pair moves (the reference's test at line_search.cpp:182-201).
  static bool still4(const V3<R>& p0, const V3<R>& p1, const V3<R>& a0, const V3<R>& a1, const V3<R>& b0,
                     const V3<R>& b1, const V3<R>& c0, const V3<R>& c1) {
    return (double)sqn3(p1 - p0) < 1e-12 && (double)sqn3(a1 - a0) < 1e-12 && (double)sqn3(b1 - b0) < 1e-12 &&
           (double)sqn3(c1 - c0) < 1e-12;
  }
  // Unsigned distance of a PT (type 0: point, triangle) or EE (type 1) pair.
  static R pair_distance(int type, const V3<R>* q) {
    R d;
    V3<R> nn, w1, w2;
    if (type == 0) narrow_pt(q[0], q[1], q[2], q[3], d, nn, w1);
    else narrow_ee(q[0], q[1], q[2], q[3], d, nn, w1, w2);
    return d;
  }
  // Additive CCD (Li, Kaufman, Jiang et al., "Codimensional Incremental
  // Potential Contact", 2021, Algorithm 1) with thickness 0, s = 0.1, t_c = 1,
  // on the motion x0 -> x1 of the pair's four vertices: 1 if the whole motion
  // keeps the pair apart, else a conservative time of impact.  Not part of the
  // reference (its line search samples 10 distances); corrected mode option.
  static R additive_ccd(int type, const V3<R>* x0, const V3<R>* p) {
    const R s = R(0.1f);
    V3<R> pbar;
    for (int k = 0; k < 3; ++k) pbar[k] = ((p[0][k] + p[1][k]) + (p[2][k] + p[3][k])) * R(0.25f);
    V3<R> q[4], x[4];
    for (int k = 0; k < 4; ++k) {
      q[k] = p[k] - pbar;
      x[k] = x0[k];
    }
    const R lp = type == 0 ? norm3(q[0]) + std::max(norm3(q[1]), std::max(norm3(q[2]), norm3(q[3])))
                           : std::max(norm3(q[0]), norm3(q[1])) + std::max(norm3(q[2]), norm3(q[3]));
    if (lp == R(0)) return R(1);
    R dist = pair_distance(type, x);
    if (!(dist > R(0))) return R(0);
    const R g = s * dist;
    R t = R(0), tl = (R(1) - s) * dist / lp;
    for (int it = 0; it < 256; ++it) {
      for (int k = 0; k < 4; ++k) x[k] = x[k] + tl * q[k];
      dist = pair_distance(type, x);
      if (t > R(0) && dist < g) return t;
      t = t + tl;
      if (t > R(1)) return R(1);
      tl = R(0.9f) * dist / lp;
    }
    return t;
  }
  static R accd_pair(int type, const V3<R>& a0, const V3<R>& a1, const V3<R>& b0, const V3<R>& b1, const V3<R>& c0,
                     const V3<R>& c1, const V3<R>& d0, const V3<R>& d1) {
    const V3<R> x0[4] = {a0, b0, c0, d0};
    const V3<R> p[4] = {a1 - a0, b1 - b0, c1 - c0, d1 - d0};
    return additive_ccd(type, x0, p);
  }
  R ccd_pt(const V3<R>& p0, const V3<R>& p1, const V3<R>& a0, const V3<R>& a1, const V3<R>& b0, const V3<R>& b1,
           const V3<R>& c0, const V3<R>& c1) const {
    if ((double)sqn3(p1 - p0) < 1e-12 && (double)sqn3(a1 - a0) < 1e-12 && (double)sqn3(b1 - b0) < 1e-12 &&
        (double)sqn3(c1 - c0) < 1e-12)
      return R(1);
    for (int i = 1; i <= 10; ++i) {
      R t = R(i) / R(10);
      V3<R> pt = p0 + t * (p1 - p0), at = a0 + t * (a1 - a0), bt = b0 + t * (b1 - b0), ct = c0 + t * (c1 - c0);
      R d; V3<R> nn, q;
      narrow_pt(pt, at, bt, ct, d, nn, q);
      if ((double)d < 1e-6) return t;
    }
    return R(1);
  }
  R ccd_ee(const V3<R>& p00, const V3<R>& p01, const V3<R>& p10, const V3<R>& p11, const V3<R>& q00,
           const V3<R>& q01, const V3<R>& q10, const V3<R>& q11) const {
    if ((double)sqn3(p01 - p00) < 1e-12 && (double)sqn3(p11 - p10) < 1e-12 && (double)sqn3(q01 - q00) < 1e-12 &&
        (double)sqn3(q11 - q10) < 1e-12)
      return R(1);
    for (int i = 1; i <= 10; ++i) {
      R t = R(i) / R(10);
      V3<R> a = p00 + t * (p01 - p00), b = p10 + t * (p11 - p10), c = q00 + t * (q01 - q00), d = q10 + t * (q11 - q10);
      R dd; V3<R> nn, wp, wq;
      narrow_ee(a, b, c, d, dd, nn, wp, wq);
      if ((double)dd < 1e-6) return t;
    }
    return R(1);
  }
  R line_search(const std::vector<R>& d, const V3<R>& wn, R woff, R ext) const {
    R alpha = R(1);
    std::vector<R> ed(3 * n), xn(3 * n);
    for (int i = 0; i < 3 * n; ++i) ed[i] = ext * d[i];
    for (int it = 0; it < 20; ++it) {
      for (int i = 0; i < 3 * n; ++i) xn[i] = x[i] + alpha * ed[i];
      if (feasible(xn, wn, woff)) return alpha;
      alpha *= R(0.5);
      if (alpha < R(1e-6)) return R(0);
    }
    return alpha;
  }

  // ---- Corrected physics ---------------------------------------------------
  // Per-face 2k (B B^T) and the isotropic stiffness folds (once per step, like
  // the faithful constants).
  void cp_constants() {
    const R mu = mu_();
    const size_t M = tris.size();
    gk.assign(M, {});
    for (size_t f = 0; f < M; ++f) {
      const R k = area[f] * thick * mu;
      const R twok = (R)(2.0 * (double)k);
      const R Mi[2][2] = {{Dm[0][f], Dm[1][f]}, {Dm[2][f], Dm[3][f]}};
      R B[3][2];
      for (int c = 0; c < 2; ++c) {
        B[0][c] = -(Mi[0][c] + Mi[1][c]);
        B[1][c] = Mi[0][c];
        B[2][c] = Mi[1][c];
      }
      for (int a = 0; a < 3; ++a)
        for (int b = a; b < 3; ++b) gk[f][sym6(a, b)] = twok * (B[a][0] * B[b][0] + B[a][1] * B[b][1]);
    }
    const size_t nb = ci.size();
    Hiso.assign(9 * nb, R(0));
    Hbase_iso.assign(9 * nb, R(0));
    const R dtinv2 = R(1) / (P.dt * P.dt);
    for (int i = 0; i < n; ++i) {
      const R mf = m[i] * dtinv2;
      if (mf != R(0))
        for (int k = 0; k < 3; ++k) Hbase_iso[9 * diag[i] + 4 * k] += mf;
    }
    for (size_t f = 0; f < M; ++f)
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          const int bi = blk(tris[f][a], tris[f][b]);
          const R g = gk[f][sym6(a, b)];
          for (int k = 0; k < 3; ++k) {
            Hiso[9 * bi + 4 * k] += g;
            Hbase_iso[9 * bi + 4 * k] += g;
          }
        }
  }

  // I - n n^T of every face at the current positions.
  void cp_projectors() {
    fpi.assign(tris.size(), {});
    for (size_t f = 0; f < tris.size(); ++f) {
      const V3<R> x0 = X(tris[f][0]);
      const V3<R> nn = normalized3(cross3(X(tris[f][1]) - x0, X(tris[f][2]) - x0));
      for (int a = 0; a < 3; ++a)
        for (int b = a; b < 3; ++b) fpi[f][sym6(a, b)] = (a == b ? R(1) : R(0)) - nn[a] * nn[b];
    }
  }

  // The membrane gradient of elastic_gradient with F from the current positions.
  void elastic_gradient_x() {
    egrad.assign(3 * n, R(0));
    R mu = mu_();
    for (size_t f = 0; f < tris.size(); ++f) {
      const auto& tr = tris[f];
      V3<R> v0 = X(tr[0]), v1 = X(tr[1]), v2 = X(tr[2]);
      R F[2][2];
      compute_F(v0, v1, v2, f, F);
      R k = area[f] * thick * mu;
      R twok = (R)(2.0 * (double)k);
      R Pm[2][2] = {{twok * (F[0][0] - R(1)), twok * (F[0][1] - R(0))}, {twok * (F[1][0] - R(0)), twok * (F[1][1] - R(1))}};
      R Mi[2][2] = {{Dm[0][f], Dm[1][f]}, {Dm[2][f], Dm[3][f]}};
      R Hm[2][2];
      for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c) Hm[r][c] = Pm[r][0] * Mi[c][0] + Pm[r][1] * Mi[c][1];
      V3<R> e1 = v1 - v0, e2 = v2 - v0;
      V3<R> nn = normalized3(cross3(e1, e2));
      V3<R> t1 = normalized3(e1);
      V3<R> t2 = cross3(nn, t1);
      V3<R> f1 = Hm[0][0] * t1 + Hm[1][0] * t2;
      V3<R> f2 = Hm[0][1] * t1 + Hm[1][1] * t2;
      V3<R> f0 = -(f1 + f2);
      for (int k2 = 0; k2 < 3; ++k2) {
        egrad[tr[0] * 3 + k2] += f0[k2];
        egrad[tr[1] * 3 + k2] += f1[k2];
        egrad[tr[2] * 3 + k2] += f2[k2];
      }
    }
  }

  // Per-vertex boxes of the motion x -> x + ed (ed null: the points x).
  void cp_boxes(const std::vector<R>* ed) {
    vbox.assign(6 * (size_t)n, R(0));
    for (int i = 0; i < n; ++i)
      for (int k = 0; k < 3; ++k) {
        const R a = x[3 * i + k];
        const R e = ed ? a + (*ed)[3 * i + k] : a;
        vbox[6 * (size_t)i + k] = std::min(a, e);
        vbox[6 * (size_t)i + 3 + k] = std::max(a, e);
      }
  }
  static bool box_overlap(const R* a, const R* b) {
    return (a[0] <= b[3] && a[3] >= b[0]) && (a[1] <= b[4] && a[4] >= b[1]) && (a[2] <= b[5] && a[5] >= b[2]);
  }
  // Union box of vertex boxes.
  void union_box(const int* vs, int cnt, R* o) const {
    for (int k = 0; k < 3; ++k) {
      o[k] = vbox[6 * (size_t)vs[0] + k];
      o[3 + k] = vbox[6 * (size_t)vs[0] + 3 + k];
    }
    for (int c = 1; c < cnt; ++c)
      for (int k = 0; k < 3; ++k) {
        o[k] = std::min(o[k], vbox[6 * (size_t)vs[c] + k]);
        o[3 + k] = std::max(o[3 + k], vbox[6 * (size_t)vs[c] + 3 + k]);
      }
  }
  // Every (q, t) with overlapping boxes (6 R per box), t ascending for each q
  // in ascending q.  A uniform grid (cell = the largest box extent) only
  // enumerates candidates; the predicate is the exact box overlap, so the
  // pairs are those of testing all of them.
  template <class F>
  static void grid_pairs(size_t nq, const std::vector<R>& qb, size_t nt, const std::vector<R>& tb, F visit) {
    if (nq == 0 || nt == 0) return;
    double h = 0.0, org[3] = {1e300, 1e300, 1e300};
    for (size_t t = 0; t < nt; ++t)
      for (int k = 0; k < 3; ++k) {
        h = std::max(h, (double)tb[6 * t + 3 + k] - (double)tb[6 * t + k]);
        org[k] = std::min(org[k], (double)tb[6 * t + k]);
      }
    for (size_t q = 0; q < nq; ++q)
      for (int k = 0; k < 3; ++k) h = std::max(h, (double)qb[6 * q + 3 + k] - (double)qb[6 * q + k]);
    if (!(h > 0.0)) h = 1.0;
    auto cell = [&](double v, int k) { return (int64_t)std::floor((v - org[k]) / h); };
    auto key = [](int64_t a, int64_t b, int64_t c) {
      return ((uint64_t)(a & 0x1fffff) << 42) | ((uint64_t)(b & 0x1fffff) << 21) | (uint64_t)(c & 0x1fffff);
    };
    std::unordered_map<uint64_t, std::vector<int>> grid;
    for (size_t t = 0; t < nt; ++t) {
      int64_t lo[3], hi[3];
      for (int k = 0; k < 3; ++k) {
        lo[k] = cell(tb[6 * t + k], k);
        hi[k] = cell(tb[6 * t + 3 + k], k);
      }
      for (int64_t a = lo[0]; a <= hi[0]; ++a)
        for (int64_t b = lo[1]; b <= hi[1]; ++b)
          for (int64_t c = lo[2]; c <= hi[2]; ++c) grid[key(a, b, c)].push_back((int)t);
    }
    std::vector<int> cand;
    for (size_t q = 0; q < nq; ++q) {
      cand.clear();
      int64_t lo[3], hi[3];
      for (int k = 0; k < 3; ++k) {
        lo[k] = cell(qb[6 * q + k], k);
        hi[k] = cell(qb[6 * q + 3 + k], k);
      }
      for (int64_t a = lo[0]; a <= hi[0]; ++a)
        for (int64_t b = lo[1]; b <= hi[1]; ++b)
          for (int64_t c = lo[2]; c <= hi[2]; ++c) {
            auto it = grid.find(key(a, b, c));
            if (it != grid.end()) cand.insert(cand.end(), it->second.begin(), it->second.end());
          }
      std::sort(cand.begin(), cand.end());
      cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
      for (int t : cand)
        if (box_overlap(&qb[6 * q], &tb[6 * (size_t)t])) visit((int)q, t);
    }
  }
  // The step's candidate pairs (PT by vertex then triangle, EE by e1 then e2,
  // rigid PT by body, vertex, triangle); gaps come from refresh_contacts.
  void detect_cp(const std::vector<R>* ed, std::vector<Contact<R>>& out) {
    out.clear();
    cp_boxes(ed);
    const R gm = P.gap_max;
    auto grow = [gm](std::vector<R>& b) {
      for (size_t i = 0; i < b.size() / 6; ++i)
        for (int k = 0; k < 3; ++k) {
          b[6 * i + k] = b[6 * i + k] - gm;
          b[6 * i + 3 + k] = b[6 * i + 3 + k] + gm;
        }
    };
    std::vector<R> qb(vbox), tb(6 * tris.size()), eb(6 * edges.size());
    grow(qb);
    for (size_t t = 0; t < tris.size(); ++t) union_box(tris[t].data(), 3, &tb[6 * t]);
    for (size_t e = 0; e < edges.size(); ++e) union_box(edges[e].data(), 2, &eb[6 * e]);
    std::vector<R> eq(eb);
    grow(eq);
    grid_pairs(n, qb, tris.size(), tb, [&](int v, int t) {
      const auto& tr = tris[t];
      if (tr[0] == v || tr[1] == v || tr[2] == v) return;
      Contact<R> c;
      c.type = PT; c.idx[0] = v; c.idx[1] = tr[0]; c.idx[2] = tr[1]; c.idx[3] = tr[2];
      c.vcount = 4;
      out.push_back(c);
    });
    grid_pairs(edges.size(), eq, edges.size(), eb, [&](int e1, int e2) {
      if (e2 <= e1) return;
      const auto& A = edges[e1];
      const auto& B = edges[e2];
      if (A[0] == B[0] || A[0] == B[1] || A[1] == B[0] || A[1] == B[1]) return;
      Contact<R> c;
      c.type = EE; c.idx[0] = A[0]; c.idx[1] = A[1]; c.idx[2] = B[0]; c.idx[3] = B[1];
      c.vcount = 4;
      out.push_back(c);
    });
    if (rworld.size() != rigids.size()) {
      rworld.clear();
      for (size_t rb = 0; rb < rigids.size(); ++rb) rworld.push_back(rigids[rb].world());
    }
    for (size_t rb = 0; rb < rigids.size(); ++rb) {
      const auto& W = rworld[rb];
      const auto& RT = rigids[rb].tris;
      std::vector<R> rtb(6 * RT.size());
      for (size_t t = 0; t < RT.size(); ++t)
        for (int k = 0; k < 3; ++k) {
          rtb[6 * t + k] = std::min(std::min(W[RT[t][0]][k], W[RT[t][1]][k]), W[RT[t][2]][k]);
          rtb[6 * t + 3 + k] = std::max(std::max(W[RT[t][0]][k], W[RT[t][1]][k]), W[RT[t][2]][k]);
        }
      grid_pairs(n, qb, RT.size(), rtb, [&](int v, int t) {
        Contact<R> c;
        c.type = RIGID_PT; c.idx[0] = v; c.idx[1] = RT[t][0]; c.idx[2] = RT[t][1]; c.idx[3] = RT[t][2];
        c.rigid = (int)rb;
        c.vcount = 1;
        out.push_back(c);
      });
    }
  }
  // Gap, normal, witness and weights of every candidate at the current x.
  void refresh_contacts(std::vector<Contact<R>>& cs) const {
#pragma omp parallel for schedule(static)
    for (size_t k = 0; k < cs.size(); ++k) {
      Contact<R>& c = cs[k];
      if (c.type == PT) {
        const V3<R> p = X(c.idx[0]), a = X(c.idx[1]), b = X(c.idx[2]), cc = X(c.idx[3]);
        narrow_pt(p, a, b, cc, c.gap, c.normal, c.wq);
        c.wp = p;
        c.bary = triangle_barycentric(c.wq, a, b, cc);
        c.w[0] = R(1); c.w[1] = -c.bary[0]; c.w[2] = -c.bary[1]; c.w[3] = -c.bary[2];
      } else if (c.type == EE) {
        R s_ = R(0), t_ = R(0);
        narrow_ee(X(c.idx[0]), X(c.idx[1]), X(c.idx[2]), X(c.idx[3]), c.gap, c.normal, c.wp, c.wq, &s_, &t_);
        c.w[0] = R(1) - s_; c.w[1] = s_; c.w[2] = -(R(1) - t_); c.w[3] = -t_;
      } else if (c.type == RIGID_PT) {
        const auto& W = rworld[c.rigid];
        const V3<R> p = X(c.idx[0]);
        narrow_pt(p, W[c.idx[1]], W[c.idx[2]], W[c.idx[3]], c.gap, c.normal, c.wq);
        c.wp = p;
        c.w[0] = R(1);
      }
    }
  }
  // LineSearch::search with the corrected feasibility rule; the candidates of
  // the trial motion x -> x + 1.25 d replace `contacts`.
  // The first trial t0 (cp_first_trial) from the longest |d_i| and the
  // longest remaining inertial displacement |xt_i - x_i|, both in FP64.
  // The endpoints of a candidate over the motion ed with the vertices in
  // `hold` not moving (rigid triangles stay put); returns the pair type.
  int pair_motion(const Contact<R>& c, const std::vector<R>& ed, const std::vector<uint8_t>& hold, D3* x0,
                  D3* p) const {
    auto d3of = [](const V3<R>& v) { return D3{(double)v[0], (double)v[1], (double)v[2]}; };
    auto mot = [&](int v) { return hold[v] ? D3{0.0, 0.0, 0.0} : d3of(Xof(ed, v)); };
    if (c.type == RIGID_PT) {
      const auto& W = rworld[c.rigid];
      x0[0] = d3of(X(c.idx[0]));
      p[0] = mot(c.idx[0]);
      for (int j = 1; j < 4; ++j) {
        x0[j] = d3of(W[c.idx[j]]);
        p[j] = D3{0.0, 0.0, 0.0};
      }
      return 0;
    }
    for (int j = 0; j < 4; ++j) {
      x0[j] = d3of(X(c.idx[j]));
      p[j] = mot(c.idx[j]);
    }
    return c.type == EE ? 1 : 0;
  }
  R line_search_cp(std::vector<R>& d, const V3<R>& wn, R woff, const std::vector<R>& xt) {
    double md2 = 0.0, mh2 = 0.0;
    for (int i = 0; i < n; ++i) {
      const double d0 = d[3 * i], d1 = d[3 * i + 1], d2 = d[3 * i + 2];
      md2 = std::fmax(md2, d0 * d0 + d1 * d1 + d2 * d2);
      const double h0 = (double)xt[3 * i] - (double)x[3 * i], h1 = (double)xt[3 * i + 1] - (double)x[3 * i + 1],
                   h2 = (double)xt[3 * i + 2] - (double)x[3 * i + 2];
      mh2 = std::fmax(mh2, h0 * h0 + h1 * h1 + h2 * h2);
    }
    const int t0 = cp_first_trial(mean_edge, md2, mh2);
    const R c0 = (R)std::ldexp(1.25, -t0);
    std::vector<R> ed(3 * n);
    for (int i = 0; i < 3 * n; ++i) ed[i] = c0 * d[i];
    detect_cp(&ed, contacts);
    // Pairs within their separation that would not move apart hold their
    // cloth vertices: up to four passes (each reads the holds of the last),
    // then every such pair's vertices.
    std::vector<uint8_t> hold(n, 0), next;
    bool any = false;
    for (int pass = 0; pass < 5; ++pass) {
      const bool all = pass == 4;
      next = hold;
      bool changed = false;
      for (const auto& c : contacts) {
        D3 x0[4], p[4];
        const int type = pair_motion(c, ed, hold, x0, p);
        const double d0 = pair_dist_d(type, x0);
        if (!(d0 < sep_xi(x0, p))) continue;
        if (!all && band_ok(type, x0, p, d0)) continue;
        const int nv = c.type == RIGID_PT ? 1 : 4;
        for (int j = 0; j < nv; ++j)
          if (!hold[c.idx[j]]) {
            next[c.idx[j]] = 1;
            changed = true;
          }
      }
      hold.swap(next);
      if (!changed) break;
      any = true;
    }
    if (any)
      for (int i = 0; i < n; ++i)
        if (hold[i])
          for (int k = 0; k < 3; ++k) {
            d[3 * i + k] = R(0);
            ed[3 * i + k] = R(0);
          }
    std::fill(hold.begin(), hold.end(), 0);
    double toi = 1.0;
#pragma omp parallel for schedule(dynamic, 256) reduction(min : toi)
    for (size_t k = 0; k < contacts.size(); ++k) {
      D3 x0[4], p[4];
      const int type = pair_motion(contacts[k], ed, hold, x0, p);
      const double xi = sep_xi(x0, p);
      // A pair within its separation passed band_ok or has its vertices held.
      const double t = pair_dist_d(type, x0) < xi ? 1.0 : accd_sep(type, x0, p, xi);
      toi = std::min(toi, t);
    }
    const bool wall = (double)sqn3(wn) > 0.5;
    R alpha = R(1);  // of the motion ed
    for (int it = t0; it < 20; ++it) {
      bool ok = (double)alpha <= toi;
      if (ok && wall) {
#pragma omp parallel for schedule(static) reduction(&& : ok)
        for (int i = 0; i < n; ++i) {
          V3<R> xi(x[3 * i] + alpha * ed[3 * i], x[3 * i + 1] + alpha * ed[3 * i + 1],
                   x[3 * i + 2] + alpha * ed[3 * i + 2]);
          const R g1 = dot3(wn, xi) - woff, g0 = dot3(wn, X(i)) - woff;
          if (!(g1 >= std::min(R(0), g0))) ok = false;
        }
      }
      if (ok) return (R)std::ldexp((double)alpha, -t0);
      alpha *= R(0.5);
    }
    return R(0);
  }

  // Integrator::inner_newton_step (integrator.cpp:152-236).
  R inner_newton(const std::vector<R>& xt) {
    int maxit = P.max_newton;
    if (P.enable_friction) maxit = std::max(maxit, P.friction_min_newton);
    std::vector<R> g, d;
    for (int it = 0; it < maxit; ++it) {
      ++st.newton_iters;
      if (cp) refresh_contacts(contacts);
      compute_gradient(xt, g);
      R gn = vinf(g.data(), g.size());
      if (gn < P.pcg_tol) return R(1);
      assemble();
      d.assign(3 * n, R(0));
      std::vector<R> b(3 * n);
      for (int i = 0; i < 3 * n; ++i) b[i] = -g[i];
      if (st.pcg_solves == capture_solve) {
        cap_H.assign(H.begin(), H.end());
        cap_b.assign(b.begin(), b.end());
      }
      int iters = 0;
      bool ok = pcg(n, rp, ci, diag, H, b, d, P.pcg_tol, P.pcg_max_iters, iters, &couplings);
      st.pcg_iters += iters;
      st.pcg_solves += 1;
      st.pcg_per_solve.push_back(iters);
      if (!ok) st.pcg_failures += 1;
      V3<R> wn(0, 0, 0);
      R woff = R(0);
      for (const auto& w : walls)
        if (w.active) { wn = w.n; woff = w.off; break; }
      R alpha = cp ? line_search_cp(d, wn, woff, xt) : line_search(d, wn, woff, R(1.25));
      st.alphas.push_back(alpha);
      if ((double)alpha < 1e-8) { st.line_search_failures += 1; return R(0); }
      R s = (R)((double)alpha * 1.25);
      for (int i = 0; i < 3 * n; ++i) x[i] = x[i] + s * d[i];
      if ((double)alpha > 0.99) return R(1);
    }
    return R(0.5);
  }

  // State.apply_gravity binding (bindings.cpp:284-300).
  void apply_gravity(const V3<R>& gv, R dt) {
    if (n == 0) throw std::runtime_error("State has not been initialised");
    V3<R> gd = dt * gv;
    for (int i = 0; i < n; ++i)
      for (int k = 0; k < 3; ++k) {
        v[3 * i + k] += gd[k];
        if (!cp) x[3 * i + k] += v[3 * i + k] * dt;  // corrected physics: x_hat carries it
      }
  }

  void refresh_constants() {
    elastic_blocks();
    build_const();
    elastic_gradient();
    if (cp) cp_constants();
  }

  // Integrator::step (integrator.cpp:70-150).
  void step() {
    st = Stats();
    R dt = P.dt;
    refresh_constants();
    std::vector<R> xold = x, xt(3 * n);
    for (int i = 0; i < 3 * n; ++i) xt[i] = x[i] + dt * v[i];
    if (cp) {
      rworld.clear();
      detect_cp(nullptr, contacts);
    } else {
      detect(contacts);
    }
    st.contacts = (int)contacts.size();
    for (auto& c : contacts) if (c.type == RIGID_PT) ++st.rigid_contacts;
    R beta = R(0);
    int bi = 0;
    while (beta < P.beta_max && bi < 20) {
      R a = inner_newton(xt);
      beta = (R)((double)beta + (1.0 - (double)beta) * (double)a);
      ++bi;
      if ((double)a < 1e-6) break;
    }
    if ((double)beta > 1e-6) inner_newton(xt);
    st.beta = beta;
    if (cp) {
      // Corrected physics: the implicit-Euler velocity of the positions
      // reached, whatever the Newton passes returned (the reference divides by
      // beta dt, which doubles the velocity after a pass that hit its cap).
      for (int i = 0; i < 3 * n; ++i) v[i] = (x[i] - xold[i]) / dt;
      if ((double)P.damping > 0.0) {
        R sc = R(1) - std::clamp(P.damping, R(0), R(1));
        for (auto& vv : v) vv *= sc;
      }
      if ((double)P.restitution > 0.0) restitution();
    } else if ((double)beta > 1e-6) {
      R bdt = beta * dt;
      for (int i = 0; i < 3 * n; ++i) v[i] = (x[i] - xold[i]) / bdt;
      if ((double)P.damping > 0.0) {
        R sc = R(1) - std::clamp(P.damping, R(0), R(1));
        for (auto& vv : v) vv *= sc;
      }
      if ((double)P.restitution > 0.0) restitution();
    }
    if (!rigids.empty()) rigid_coupling();
  }

  // Integrator::apply_contact_restitution (integrator.cpp:565-674).
  void restitution() {
    R e = std::clamp(P.restitution, R(0), R(1));
    if (e <= R(0)) return;
    auto impulse = [&](int idx, V3<R> nn) {
      if (idx < 0 || idx >= n) return;
      R nl = norm3(nn);
      if ((double)nl < 1e-8) return;
      nn = nn / nl;
      V3<R> vel = Xof(v, idx);
      R vn = dot3(vel, nn);
      if (vn < R(0)) {
        V3<R> nv = vel - ((R(1) + e) * vn) * nn;
        for (int k = 0; k < 3; ++k) v[3 * idx + k] = nv[k];
      }
    };
    std::vector<Contact<R>> cs;
    detect(cs);
    R glim = std::max(P.gap_max, R(1e-5));
    for (const auto& c : cs) {
      if (c.gap > glim) continue;
      if (c.type == PT) impulse(c.idx[0], c.normal);
      else if (c.type == EE) {
        impulse(c.idx[0], c.normal); impulse(c.idx[1], c.normal);
        impulse(c.idx[2], -c.normal); impulse(c.idx[3], -c.normal);
      } else if (c.type == WALL) impulse(c.idx[0], c.normal);
      else if (c.type == RIGID_PT && c.rigid >= 0 && c.rigid < (int)rigids.size()) {
        V3<R> nn = c.normal;
        R nl = norm3(nn);
        if ((double)nl < 1e-8) continue;
        nn = nn / nl;
        int vi = c.idx[0];
        if (vi < 0 || vi >= n) continue;
        Rigid<R>& body = rigids[c.rigid];
        V3<R> vv = Xof(v, vi);
        V3<R> rel = vv - body.velocity_at_point(c.wq);
        R vn = dot3(rel, nn);
        if (vn >= R(0)) continue;
        R im = m[vi] > R(0) ? R(1) / m[vi] : R(0);
        V3<R> r = c.wq - body.pos;
        V3<R> cr = cross3(r, nn);
        R ang = dot3(cr, mul(body.inertia_world_inv(), cr));
        R den = im + ang;
        if (den <= R(0)) continue;
        R j = -(R(1) + e) * vn / den;
        V3<R> J = j * nn;
        vv = vv + im * J;
        for (int k = 0; k < 3; ++k) v[3 * vi + k] = vv[k];
        body.apply_impulse(c.wq, -J);
      }
    }
    for (const auto& w : walls) {
      if (!w.active) continue;
      V3<R> nn = w.n;
      R nl = norm3(nn);
      if ((double)nl < 1e-8) continue;
      nn = nn / nl;
      for (int i = 0; i < n; ++i) {
        R sd = dot3(nn, X(i)) - w.off;
        if (sd <= w.gap + P.gap_max) impulse(i, nn);
      }
    }
  }

  // Integrator::apply_rigid_coupling (integrator.cpp:676-754).
  void rigid_coupling() {
    std::vector<Contact<R>> cs;
    if (cp) {
      // The step's candidates at the final positions; the body is pushed
      // away from the cloth.
      refresh_contacts(contacts);
      for (auto& b : rigids) b.force = b.torque = V3<R>();
      for (const auto& c : contacts) {
        if (c.type != RIGID_PT || c.rigid < 0 || c.rigid >= (int)rigids.size()) continue;
        R kb = contact_stiffness(c);
        V3<R> nn = c.normal;
        R nl = norm3(nn);
        if ((double)nl < 1e-8) continue;
        nn = nn / nl;
        R dV = barrier_grad(c.gap, P.gap_max, kb);
        if ((double)std::fabs(dV) < 1e-12) continue;
        rigids[c.rigid].apply_force(c.wq, dV * nn);
      }
      for (auto& b : rigids) b.integrate(P.dt);
      return;
    }
    detect(cs);
    for (auto& b : rigids) b.force = b.torque = V3<R>();
    for (const auto& c : cs) {
      if (c.type != RIGID_PT) continue;
      if (c.rigid < 0 || c.rigid >= (int)rigids.size()) continue;
      R kb = contact_stiffness(c);
      V3<R> nn = c.normal;
      R nl = norm3(nn);
      if ((double)nl < 1e-8) continue;
      nn = nn / nl;
      R dV = barrier_grad(c.gap, P.gap_max, kb);
      if ((double)std::fabs(dV) < 1e-12) continue;
      V3<R> f = dV * nn;
      rigids[c.rigid].apply_force(c.wq, -f);
    }
    for (auto& b : rigids) b.integrate(P.dt);
  }
};

}  // namespace orc

// ---------------------------------------------------------------------------
// C interface for ctypes (tests and the CPU baseline).
// ---------------------------------------------------------------------------
using namespace orc;

struct OSim {
  int prec;
  Sim<float>* f = nullptr;
  Sim<double>* d = nullptr;
};

static thread_local std::string g_err;

template <class F> static int guard(F&& fn) {
  try {
    fn();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

#define DISPATCH(s, ...)                                  \
  do {                                                     \
    if ((s)->prec == 32) { auto& S = *(s)->f; __VA_ARGS__; } \
    else { auto& S = *(s)->d; __VA_ARGS__; }               \
  } while (0)

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

void* orc_create(int prec) {
  OSim* s = new OSim;
  s->prec = prec == 64 ? 64 : 32;
  if (s->prec == 32) s->f = new Sim<float>; else s->d = new Sim<double>;
  return s;
}
void orc_destroy(void* h) {
  OSim* s = (OSim*)h;
  delete s->f; delete s->d; delete s;
}

// material: E, nu, density, thickness, bending
// Intersection validator, brute force: (edge, triangle) pairs sharing no
// vertex where the edge properly crosses the triangle (orientation tests in
// double on the float coordinates), and cloth vertices with generalized
// winding number above 1/2 w.r.t. a rigid body's surface.
// Sign of orient3d with Shewchuk's static error bound ("Adaptive Precision
// Floating-Point Arithmetic and Fast Robust Geometric Predicates", 1997,
// o3derrboundA = (7 + 56 eps) eps, eps = 2^-53): 0 when rounding could have
// decided the sign, so near-coplanar configurations count as degenerate
// rather than as crossings.
static int orient3d_sign(const double* a, const double* b, const double* c, const double* d) {
  const double ax = a[0] - d[0], ay = a[1] - d[1], az = a[2] - d[2];
  const double bx = b[0] - d[0], by = b[1] - d[1], bz = b[2] - d[2];
  const double cx = c[0] - d[0], cy = c[1] - d[1], cz = c[2] - d[2];
  const double v = ax * (by * cz - bz * cy) - ay * (bx * cz - bz * cx) + az * (bx * cy - by * cx);
  const double perm = std::fabs(ax) * (std::fabs(by * cz) + std::fabs(bz * cy)) +
                      std::fabs(ay) * (std::fabs(bx * cz) + std::fabs(bz * cx)) +
                      std::fabs(az) * (std::fabs(bx * cy) + std::fabs(by * cx));
  const double bound = 7.771561172376103e-16 * perm;
  return v > bound ? 1 : (v < -bound ? -1 : 0);
}

int orc_count_intersections(void* h, int* edge_tri, int* inside) {
  OSim* s = (OSim*)h;
  int et = 0, in = 0;
  DISPATCH(s, {
    auto P3 = [&](int i, double* o) { for (int k = 0; k < 3; ++k) o[k] = (double)(float)S.x[3 * i + k]; };
    for (const auto& e : S.edges)
      for (const auto& t : S.tris) {
        if (t[0] == e[0] || t[0] == e[1] || t[1] == e[0] || t[1] == e[1] || t[2] == e[0] || t[2] == e[1]) continue;
        double p[3], q[3], a[3], b[3], c[3];
        P3(e[0], p); P3(e[1], q); P3(t[0], a); P3(t[1], b); P3(t[2], c);
        const int sp = orient3d_sign(a, b, c, p), sq = orient3d_sign(a, b, c, q);
        if (sp == 0 || sq == 0 || sp == sq) continue;
        const int s1 = orient3d_sign(p, q, a, b), s2 = orient3d_sign(p, q, b, c), s3 = orient3d_sign(p, q, c, a);
        if (s1 != 0 && s1 == s2 && s2 == s3) ++et;
      }
    for (int i = 0; i < S.n; ++i) {
      double p[3];
      P3(i, p);
      for (const auto& rb : S.rigids) {
        auto W = rb.world();
        double w = 0.0;
        for (const auto& t : rb.tris) {
          double a[3], b[3], c[3];
          for (int k = 0; k < 3; ++k) {
            a[k] = (double)(float)W[t[0]][k] - p[k];
            b[k] = (double)(float)W[t[1]][k] - p[k];
            c[k] = (double)(float)W[t[2]][k] - p[k];
          }
          const double la = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
          const double lb = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
          const double lc = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
          const double det = a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) +
                             a[2] * (b[0] * c[1] - b[1] * c[0]);
          const double den = la * lb * lc + (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) * lc +
                             (b[0] * c[0] + b[1] * c[1] + b[2] * c[2]) * la +
                             (c[0] * a[0] + c[1] * a[1] + c[2] * a[2]) * lb;
          w += 2.0 * std::atan2(det, den);
        }
        if (std::fabs(w) / (4.0 * 3.141592653589793) > 0.5) ++in;
      }
    }
  });
  *edge_tri = et;
  *inside = in;
  return 0;
}

// Contact mode: 0 faithful, 1 corrected (brute-force self PT/EE with the
// reference's intended predicate, PT and EE barriers in the energy).
int orc_energy(void* h, const double* x, const double* v, const double* m, double* out) {
  OSim* s = (OSim*)h;
  return guard([&] { DISPATCH(s, S.energy(x, v, m, out)); });
}

// CollisionValidator::compute_metrics (collision_validator.cpp:7-157) with
// compute_relative_velocities (:178-230) on a contact list and velocities v
// (3 per vertex).  type: 0 PT, 1 EE, 2 wall, 3 rigid; idx 4 per contact;
// normal 3 per contact.  out: num_point_triangle, num_edge_edge, num_wall,
// num_total, num_penetrations, num_ccd, num_broad_phase, num_rel_velocities,
// has_major_penetration, has_tunneling, min_gap, max_gap, avg_gap,
// max_penetration, avg_penetration, ccd_effectiveness, max_rel_velocity,
// avg_rel_velocity (all as the reference computes them in R), then the gap,
// penetration and relative-velocity sums of the same R terms in double.
}  // extern "C"

template <class R>
static void contact_metrics(int count, const int* type, const int* idx, const double* gap, const double* nrm,
                            const double* vel, double gap_max_d, int ccd, double* out) {
  const R gap_max = (R)gap_max_d;
  auto V = [&](int i) { return V3<R>{(R)vel[3 * i], (R)vel[3 * i + 1], (R)vel[3 * i + 2]}; };
  int npt = 0, nee = 0, nwall = 0, npen = 0, nccd = 0, nbroad = 0, nrel = 0;
  bool major = false, tunnel = false;
  R min_gap = R(0), max_gap = R(0), avg_gap = R(0), max_pen = R(0), avg_pen = R(0), eff = R(0);
  R max_rel = R(0), avg_rel = R(0);
  double sgap_d = 0.0, spen_d = 0.0, srel_d = 0.0;
  if (count > 0) {
    min_gap = max_gap = (R)gap[0];
    R sum_gap = R(0), sum_pen = R(0);
    for (int e = 0; e < count; ++e) {
      if (type[e] == 0) ++npt;
      else if (type[e] == 1) ++nee;
      else if (type[e] == 2) ++nwall;
      R g = (R)gap[e];
      min_gap = std::min(min_gap, g);
      max_gap = std::max(max_gap, g);
      sum_gap = sum_gap + g;
      sgap_d += (double)g;
      if (g < R(0)) {
        ++npen;
        R pen = -g;
        max_pen = std::max(max_pen, pen);
        sum_pen = sum_pen + pen;
        spen_d += (double)pen;
        if ((double)pen > 0.001) major = true;
        if ((double)pen > 0.1 * (double)gap_max) tunnel = true;
      }
    }
    avg_gap = sum_gap / (R)count;
    avg_pen = npen > 0 ? sum_pen / (R)npen : R(0);
    for (int e = 0; e < count; ++e) {
      if (ccd && (double)(R)gap[e] < (double)gap_max * 0.5) ++nccd;
      else ++nbroad;
    }
    eff = ccd ? (R)((double)((R)nccd / (R)count) * 100.0) : R(0);
    R sum_rel = R(0);
    for (int e = 0; e < count; ++e) {
      const int* q = idx + 4 * e;
      V3<R> n{(R)nrm[3 * e], (R)nrm[3 * e + 1], (R)nrm[3 * e + 2]};
      R rel;
      if (type[e] == 0) {
        V3<R> vp = V(q[0]), a = V(q[1]), b = V(q[2]), d = V(q[3]);
        V3<R> vt;
        for (int k = 0; k < 3; ++k) vt[k] = (R)((double)((a[k] + b[k]) + d[k]) / 3.0);
        rel = std::fabs(dot3(vp - vt, n));
      } else if (type[e] == 1) {
        V3<R> a = V(q[0]), b = V(q[1]), d = V(q[2]), f = V(q[3]);
        V3<R> e0, e1;
        for (int k = 0; k < 3; ++k) {
          e0[k] = (R)((double)(a[k] + b[k]) / 2.0);
          e1[k] = (R)((double)(d[k] + f[k]) / 2.0);
        }
        rel = std::fabs(dot3(e0 - e1, n));
      } else if (type[e] == 2) {
        rel = std::fabs(dot3(V(q[0]), n));
      } else {
        continue;
      }
      ++nrel;
      max_rel = nrel == 1 ? rel : std::max(max_rel, rel);
      sum_rel = sum_rel + rel;
      srel_d += (double)rel;
    }
    avg_rel = nrel ? sum_rel / (R)nrel : R(0);
  }
  const double r[21] = {(double)npt, (double)nee, (double)nwall, (double)count, (double)npen, (double)nccd,
                        (double)(count > 0 ? nbroad : 0), (double)nrel, major ? 1.0 : 0.0, tunnel ? 1.0 : 0.0,
                        (double)min_gap, (double)max_gap, (double)avg_gap, (double)max_pen, (double)avg_pen,
                        (double)eff, (double)max_rel, (double)avg_rel, sgap_d, spen_d, srel_d};
  for (int k = 0; k < 21; ++k) out[k] = r[k];
}

extern "C" {

int orc_contact_metrics(int prec, int count, const int* type, const int* idx, const double* gap, const double* nrm,
                        const double* vel, double gap_max, int ccd, double* out) {
  return guard([&] {
    if (prec == 64) contact_metrics<double>(count, type, idx, gap, nrm, vel, gap_max, ccd, out);
    else contact_metrics<float>(count, type, idx, gap, nrm, vel, gap_max, ccd, out);
  });
}

int orc_set_capture(void* h, int solve) {
  OSim* s = (OSim*)h;
  return guard([&] { DISPATCH(s, S.capture_solve = solve; S.cap_H.clear(); S.cap_b.clear()); });
}
// Copies the captured system (9 nb values, 3 n rhs); returns 1 if one was captured.
int orc_get_capture(void* h, double* H, double* b) {
  OSim* s = (OSim*)h;
  int got = 0;
  DISPATCH(s, {
    if (!S.cap_H.empty()) {
      std::copy(S.cap_H.begin(), S.cap_H.end(), H);
      std::copy(S.cap_b.begin(), S.cap_b.end(), b);
      got = 1;
    }
  });
  return got;
}

// Corrected physics on (1) or off (0); on implies the corrected contact mode.
int orc_set_physics(void* h, int on) {
  OSim* s = (OSim*)h;
  return guard([&] {
    DISPATCH(s, {
      S.cp = on != 0;
      if (S.cp) S.contact_mode = 1;
    });
  });
}

int orc_set_ccd_mode(void* h, int mode) {
  OSim* s = (OSim*)h;
  return guard([&] { DISPATCH(s, S.ccd_mode = mode == 1 ? 1 : 0); });
}

// Additive CCD on one pair: type 0 PT / 1 EE, x0 and p 12 doubles each.
// The corrected physics line search's FP64 time of impact with its minimum
// separation (accd_sep).
double orc_accd_sep(int type, const double* x0, const double* p) {
  D3 xs[4], ps[4];
  for (int k = 0; k < 4; ++k) {
    xs[k] = D3{x0[3 * k], x0[3 * k + 1], x0[3 * k + 2]};
    ps[k] = D3{p[3 * k], p[3 * k + 1], p[3 * k + 2]};
  }
  return accd_sep(type, xs, ps, sep_xi(xs, ps));
}
double orc_additive_ccd(int prec, int type, const double* x0, const double* p) {
  double r = 0.0;
  auto run = [&](auto tag) {
    using R = decltype(tag);
    V3<R> xs[4], ps[4];
    for (int k = 0; k < 4; ++k)
      for (int c = 0; c < 3; ++c) {
        xs[k][c] = (R)x0[3 * k + c];
        ps[k][c] = (R)p[3 * k + c];
      }
    r = (double)Sim<R>::additive_ccd(type, xs, ps);
  };
  if (prec == 64) run(0.0); else run(0.0f);
  return r;
}

int orc_set_contact_mode(void* h, int mode, double pt_box, double keep_gap) {
  OSim* s = (OSim*)h;
  DISPATCH(s, S.contact_mode = mode; S.pt_box = (decltype(S.pt_box))pt_box; S.keep_gap = (decltype(S.keep_gap))keep_gap);
  return 0;
}

// Threads for the oracle's parallel loops (results do not depend on it).
void orc_set_threads(int k) { omp_set_num_threads(k > 0 ? k : 1); }
int orc_get_threads() { return omp_get_max_threads(); }

int orc_set_mesh(void* h, int nv, const double* verts, int nt, const int32_t* tris, const double* mat) {
  OSim* s = (OSim*)h;
  return guard([&] {
    DISPATCH(s, {
      using R = std::remove_reference_t<decltype(S.E)>;
      std::vector<V3<R>> vv(nv);
      for (int i = 0; i < nv; ++i) vv[i] = V3<R>((R)verts[3 * i], (R)verts[3 * i + 1], (R)verts[3 * i + 2]);
      std::vector<std::array<int, 3>> tt(nt);
      for (int i = 0; i < nt; ++i) tt[i] = {tris[3 * i], tris[3 * i + 1], tris[3 * i + 2]};
      S.E = (R)mat[0]; S.nu = (R)mat[1]; S.rho = (R)mat[2]; S.thick = (R)mat[3]; S.bend = (R)mat[4];
      S.set_mesh(vv, tt);
    });
  });
}
int orc_init_state(void* h) {
  OSim* s = (OSim*)h;
  return guard([&] { DISPATCH(s, S.init_state()); });
}

// Params order: dt, beta_max, min_newton, max_newton, pcg_tol, pcg_max_iters,
// gap_max, wall_gap, enable_ccd, enable_friction, mu, friction_eps, damping,
// restitution, enable_strain, strain_limit, strain_tau
int orc_set_params(void* h, const double* p) {
  OSim* s = (OSim*)h;
  return guard([&] {
    DISPATCH(s, {
      using R = std::remove_reference_t<decltype(S.E)>;
      auto& P = S.P;
      P.dt = (R)p[0]; P.beta_max = (R)p[1]; P.min_newton = (int)p[2]; P.max_newton = (int)p[3];
      P.pcg_tol = (R)p[4]; P.pcg_max_iters = (int)p[5]; P.gap_max = (R)p[6]; P.wall_gap = (R)p[7];
      P.enable_ccd = p[8] != 0; P.enable_friction = p[9] != 0; P.mu = (R)p[10]; P.friction_eps = (R)p[11];
      P.damping = (R)p[12]; P.restitution = (R)p[13]; P.enable_strain = p[14] != 0;
      P.strain_limit = (R)p[15]; P.strain_tau = (R)p[16];
    });
  });
}
int orc_add_pin(void* h, int vi, const double* t) {
  OSim* s = (OSim*)h;
  return guard([&] {
    DISPATCH(s, {
      using R = std::remove_reference_t<decltype(S.E)>;
      S.pins.push_back({vi, V3<R>((R)t[0], (R)t[1], (R)t[2]), true});
    });
  });
}
// constraints.cpp:13-20: the stored normal is normalized().
int orc_add_wall(void* h, const double* nrm, double off, double gap) {
  OSim* s = (OSim*)h;
  return guard([&] {
    DISPATCH(s, {
      using R = std::remove_reference_t<decltype(S.E)>;
      V3<R> nn((R)nrm[0], (R)nrm[1], (R)nrm[2]);
      S.walls.push_back({normalized3(nn), (R)off, (R)gap, true});
    });
  });
}
int orc_add_rigid(void* h, int nv, const double* verts, int nt, const int32_t* tris, double density) {
  OSim* s = (OSim*)h;
  return guard([&] {
    DISPATCH(s, {
      using R = std::remove_reference_t<decltype(S.E)>;
      std::vector<V3<R>> vv(nv);
      for (int i = 0; i < nv; ++i) vv[i] = V3<R>((R)verts[3 * i], (R)verts[3 * i + 1], (R)verts[3 * i + 2]);
      std::vector<std::array<int, 3>> tt(nt);
      for (int i = 0; i < nt; ++i) tt[i] = {tris[3 * i], tris[3 * i + 1], tris[3 * i + 2]};
      Rigid<R> b;
      b.initialize(vv, tt, (R)density);
      S.rigids.push_back(b);
    });
  });
}
int orc_apply_gravity(void* h, const double* g, double dt) {
  OSim* s = (OSim*)h;
  return guard([&] {
    DISPATCH(s, {
      using R = std::remove_reference_t<decltype(S.E)>;
      S.apply_gravity(V3<R>((R)g[0], (R)g[1], (R)g[2]), (R)dt);
    });
  });
}
int orc_step(void* h) {
  OSim* s = (OSim*)h;
  return guard([&] { DISPATCH(s, S.step()); });
}
int orc_num_vertices(void* h) { OSim* s = (OSim*)h; int r = 0; DISPATCH(s, r = S.n); return r; }
int orc_num_blocks(void* h) { OSim* s = (OSim*)h; int r = 0; DISPATCH(s, r = (int)S.ci.size()); return r; }

// which: 0 positions, 1 velocities, 2 masses (n), 3 mesh vertices, 4 elastic gradient
int orc_get(void* h, int which, double* out) {
  OSim* s = (OSim*)h;
  return guard([&] {
    DISPATCH(s, {
      if (which == 0) for (int i = 0; i < 3 * S.n; ++i) out[i] = S.x[i];
      else if (which == 1) for (int i = 0; i < 3 * S.n; ++i) out[i] = S.v[i];
      else if (which == 2) for (int i = 0; i < S.n; ++i) out[i] = S.m[i];
      else if (which == 3) for (int i = 0; i < S.n; ++i) for (int k = 0; k < 3; ++k) out[3 * i + k] = S.mverts[i][k];
      else if (which == 4) { S.refresh_constants(); for (int i = 0; i < 3 * S.n; ++i) out[i] = S.egrad[i]; }
      else throw std::invalid_argument("bad selector");
    });
  });
}
// which: 0 positions, 1 velocities, 2 masses, 3 mesh vertices, 4 positions_prev
int orc_set(void* h, int which, const double* in) {
  OSim* s = (OSim*)h;
  return guard([&] {
    DISPATCH(s, {
      using R = std::remove_reference_t<decltype(S.E)>;
      if (which == 0) for (int i = 0; i < 3 * S.n; ++i) S.x[i] = (R)in[i];
      else if (which == 1) for (int i = 0; i < 3 * S.n; ++i) S.v[i] = (R)in[i];
      else if (which == 2) for (int i = 0; i < S.n; ++i) S.m[i] = (R)in[i];
      else if (which == 3)  // Mesh.vertices setter: the rest state is kept (mesh.cpp:122-124)
        for (int i = 0; i < S.n; ++i) for (int k = 0; k < 3; ++k) S.mverts[i][k] = (R)in[3 * i + k];
      else if (which == 4)  // State::positions_prev (state.h:18), read by friction (integrator.cpp:39)
        for (int i = 0; i < 3 * S.n; ++i) S.xprev[i] = (R)in[i];
      else throw std::invalid_argument("bad selector");
    });
  });
}
// stats: newton_iters, pcg_iters, pcg_solves, pcg_failures, ls_failures, contacts,
// rigid_contacts, strain_active, beta, asym_count, spd_general_calls
int orc_stats(void* h, double* out) {
  OSim* s = (OSim*)h;
  DISPATCH(s, {
    out[0] = S.st.newton_iters; out[1] = S.st.pcg_iters; out[2] = S.st.pcg_solves; out[3] = S.st.pcg_failures;
    out[4] = S.st.line_search_failures; out[5] = S.st.contacts; out[6] = S.st.rigid_contacts;
    out[7] = S.st.strain_active; out[8] = S.st.beta; out[9] = (double)S.asym_count;
    out[10] = (double)g_spd_general_calls;
  });
  return 0;
}
int orc_pcg_per_solve(void* h, int* out, int cap) {
  OSim* s = (OSim*)h;
  int cnt = 0;
  DISPATCH(s, {
    cnt = (int)S.st.pcg_per_solve.size();
    for (int i = 0; i < cnt && i < cap; ++i) out[i] = S.st.pcg_per_solve[i];
  });
  return cnt;
}
int orc_rigid_state(void* h, int rb, double* out /*pos3, lin3, ang3, rot9, mass*/) {
  OSim* s = (OSim*)h;
  return guard([&] {
    DISPATCH(s, {
      auto& b = S.rigids.at(rb);
      for (int k = 0; k < 3; ++k) { out[k] = b.pos[k]; out[3 + k] = b.lin[k]; out[6 + k] = b.ang[k]; }
      for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) out[9 + 3 * r + c] = b.rot.m[r][c];
      out[18] = b.mass;
    });
  });
}

int orc_set_rigid_state(void* h, int rb, const double* in /*pos3, lin3, ang3, rot9*/) {
  OSim* s = (OSim*)h;
  return guard([&] {
    DISPATCH(s, {
      using R = std::remove_reference_t<decltype(S.E)>;
      auto& b = S.rigids.at(rb);
      for (int k = 0; k < 3; ++k) { b.pos[k] = (R)in[k]; b.lin[k] = (R)in[3 + k]; b.ang[k] = (R)in[6 + k]; }
      for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) b.rot.m[r][c] = (R)in[9 + 3 * r + c];
    });
  });
}

// RigidBody::apply_force accumulation then integrate (rigid_body.cpp:129-176)
// on body rb with the given total force and torque (test hook).
int orc_rigid_integrate(void* h, int rb, const double* force, const double* torque, double dt) {
  OSim* s = (OSim*)h;
  return guard([&] {
    DISPATCH(s, {
      using R = std::remove_reference_t<decltype(S.E)>;
      auto& b = S.rigids.at(rb);
      b.force = V3<R>((R)force[0], (R)force[1], (R)force[2]);
      b.torque = V3<R>((R)torque[0], (R)torque[1], (R)torque[2]);
      b.integrate((R)dt);
    });
  });
}

// Pattern (row pointers n+1, block columns nb).
int orc_pattern(void* h, int32_t* rp, int32_t* ci) {
  OSim* s = (OSim*)h;
  DISPATCH(s, {
    for (size_t i = 0; i < S.rp.size(); ++i) rp[i] = S.rp[i];
    for (size_t i = 0; i < S.ci.size(); ++i) ci[i] = S.ci[i];
  });
  return 0;
}

// Evaluate one Newton linearisation at the current positions with target
// x_hat: gradient (3n), system matrix values (9 nb), contacts are detected at
// the current positions first.  Used for per-stage parity.
int orc_linearize(void* h, const double* xhat, double* g_out, double* H_out) {
  OSim* s = (OSim*)h;
  return guard([&] {
    DISPATCH(s, {
      using R = std::remove_reference_t<decltype(S.E)>;
      S.refresh_constants();
      std::vector<R> xt(3 * S.n), g;
      for (int i = 0; i < 3 * S.n; ++i) xt[i] = (R)xhat[i];
      if (S.cp) {
        S.rworld.clear();
        S.detect_cp(nullptr, S.contacts);
        S.refresh_contacts(S.contacts);
      } else {
        S.detect(S.contacts);
      }
      S.compute_gradient(xt, g);
      S.assemble();
      for (int i = 0; i < 3 * S.n; ++i) g_out[i] = g[i];
      for (size_t i = 0; i < S.H.size(); ++i) H_out[i] = S.H[i];
    });
  });
}

// Integrator::apply_contact_restitution on the current state (the pass step
// runs after its velocity update when contact_restitution > 0).
int orc_restitution(void* h) {
  OSim* s = (OSim*)h;
  return guard([&] { DISPATCH(s, S.restitution()); });
}

// Block-Jacobi PCG on a BSR matrix (9 values per block).
int orc_pcg(int prec, int nv, const int32_t* rp, const int32_t* ci, const double* vals, const double* b,
            double* x, double tol, int maxit, int* iters) {
  return guard([&] {
    auto run = [&](auto tag) {
      using R = decltype(tag);
      std::vector<int> rp_(rp, rp + nv + 1), ci_(ci, ci + rp[nv]), dg(nv, -1);
      for (int i = 0; i < nv; ++i)
        for (int k = rp_[i]; k < rp_[i + 1]; ++k) if (ci_[k] == i) dg[i] = k;
      std::vector<R> A(9 * (size_t)rp[nv]), bb(3 * nv), xx(3 * nv);
      for (size_t i = 0; i < A.size(); ++i) A[i] = (R)vals[i];
      for (int i = 0; i < 3 * nv; ++i) { bb[i] = (R)b[i]; xx[i] = (R)x[i]; }
      int it = 0;
      bool ok = Sim<R>::pcg(nv, rp_, ci_, dg, A, bb, xx, (R)tol, maxit, it);
      for (int i = 0; i < 3 * nv; ++i) x[i] = xx[i];
      *iters = ok ? it : -it - 1;
    };
    if (prec == 64) run(double()); else run(float());
  });
}

// Contacts at the current positions: per contact type, idx[4], rigid, gap,
// normal[3], witness_q[3], weights[4] -> 17 doubles per contact.
int orc_contacts(void* h, double* out, int cap) {
  OSim* s = (OSim*)h;
  int cnt = 0;
  DISPATCH(s, {
    using R = std::remove_reference_t<decltype(S.E)>;
    std::vector<Contact<R>> cs;
    S.detect(cs);
    cnt = (int)cs.size();
    for (int i = 0; i < cnt && i < cap; ++i) {
      double* o = out + 17 * i;
      const auto& c = cs[i];
      o[0] = c.type;
      for (int k = 0; k < 4; ++k) o[1 + k] = c.idx[k];
      o[5] = c.rigid; o[6] = c.gap;
      for (int k = 0; k < 3; ++k) { o[7 + k] = c.normal[k]; o[10 + k] = c.wq[k]; }
      for (int k = 0; k < 4; ++k) o[13 + k] = c.w[k];
    }
  });
  return cnt;
}

// Scalar kernels for known-answer tests.
double orc_barrier(int prec, int which, double g, double gmax, double k) {
  if (prec == 64) {
    if (which == 0) return barrier_energy<double>(g, gmax, k);
    if (which == 1) return barrier_grad<double>(g, gmax, k);
    return barrier_hess<double>(g, gmax, k);
  }
  if (which == 0) return barrier_energy<float>((float)g, (float)gmax, (float)k);
  if (which == 1) return barrier_grad<float>((float)g, (float)gmax, (float)k);
  return barrier_hess<float>((float)g, (float)gmax, (float)k);
}
// out: dist, normal[3], witness_q[3]
void orc_narrow_pt(int prec, const double* p, const double* a, const double* b, const double* c, double* out) {
  auto run = [&](auto tag) {
    using R = decltype(tag);
    auto mk = [](const double* q) { return V3<R>((R)q[0], (R)q[1], (R)q[2]); };
    R d; V3<R> nn, wq;
    narrow_pt(mk(p), mk(a), mk(b), mk(c), d, nn, wq);
    out[0] = d;
    for (int k = 0; k < 3; ++k) { out[1 + k] = nn[k]; out[4 + k] = wq[k]; }
  };
  if (prec == 64) run(double()); else run(float());
}
// out: dist, normal[3], wp[3], wq[3]
void orc_narrow_ee(int prec, const double* p0, const double* p1, const double* q0, const double* q1, double* out) {
  auto run = [&](auto tag) {
    using R = decltype(tag);
    auto mk = [](const double* q) { return V3<R>((R)q[0], (R)q[1], (R)q[2]); };
    R d; V3<R> nn, wp, wq;
    narrow_ee(mk(p0), mk(p1), mk(q0), mk(q1), d, nn, wp, wq);
    out[0] = d;
    for (int k = 0; k < 3; ++k) { out[1 + k] = nn[k]; out[4 + k] = wp[k]; out[7 + k] = wq[k]; }
  };
  if (prec == 64) run(double()); else run(float());
}
// enforce_spd on a 3x3 (row-major in/out).
void orc_enforce_spd(int prec, double* h9) {
  auto run = [&](auto tag) {
    using R = decltype(tag);
    M3<R> M;
    for (int i = 0; i < 9; ++i) M.m[i / 3][i % 3] = (R)h9[i];
    enforce_spd(M, R(1e-8));
    for (int i = 0; i < 9; ++i) h9[i] = M.m[i / 3][i % 3];
  };
  if (prec == 64) run(double()); else run(float());
}
// Pin stiffness (stiffness.cpp:118-144) with min_gap.
double orc_pin_stiffness(int prec, double mass, const double* off, const double* h9, double min_gap) {
  double r = 0;
  auto run = [&](auto tag) {
    using R = decltype(tag);
    Sim<R> S;
    S.P.min_gap = (R)min_gap;
    M3<R> M;
    for (int i = 0; i < 9; ++i) M.m[i / 3][i % 3] = (R)h9[i];
    r = S.pin_stiffness((R)mass, V3<R>((R)off[0], (R)off[1], (R)off[2]), M);
  };
  if (prec == 64) run(double()); else run(float());
  return r;
}
// 3x2 SVD: F row-major (6) -> U (6, row-major 3x2), sig (2), V (4, row-major)
int orc_svd32(int prec, const double* F6, double* U6, double* s2, double* V4) {
  int ok = 0;
  auto run = [&](auto tag) {
    using R = decltype(tag);
    R F[3][2], U[3][2], sg[2], Vm[2][2];
    for (int r = 0; r < 3; ++r) for (int c = 0; c < 2; ++c) F[r][c] = (R)F6[2 * r + c];
    ok = Sim<R>::svd32(F, U, sg, Vm);
    for (int r = 0; r < 3; ++r) for (int c = 0; c < 2; ++c) U6[2 * r + c] = U[r][c];
    s2[0] = sg[0]; s2[1] = sg[1];
    for (int r = 0; r < 2; ++r) for (int c = 0; c < 2; ++c) V4[2 * r + c] = Vm[r][c];
  };
  if (prec == 64) run(double()); else run(float());
  return ok;
}
// Line search on the current state with a given direction; returns alpha and
// (d_out, nullable) the direction after the corrected physics' holds.
double orc_line_search(void* h, const double* d, double ext, double* d_out) {
  OSim* s = (OSim*)h;
  double r = 0;
  DISPATCH(s, {
    using R = std::remove_reference_t<decltype(S.E)>;
    std::vector<R> dd(3 * S.n);
    for (int i = 0; i < 3 * S.n; ++i) dd[i] = (R)d[i];
    V3<R> wn(0, 0, 0);
    R woff = 0;
    for (const auto& w : S.walls) if (w.active) { wn = w.n; woff = w.off; break; }
    // Corrected physics: its own rule over the trial motion x -> x + 1.25 d
    // (ext is not used), with the candidates of that motion.
    r = S.cp ? S.line_search_cp(dd, wn, woff, S.x) : S.line_search(dd, wn, woff, (R)ext);
    if (d_out)
      for (int i = 0; i < 3 * S.n; ++i) d_out[i] = (double)dd[i];
  });
  return r;
}
// Add an explicit contact to the oracle's contact list (for line-search tests).
int orc_set_contacts_pt(void* h, int count, const int32_t* idx4) {
  OSim* s = (OSim*)h;
  DISPATCH(s, {
    S.contacts.clear();
    for (int i = 0; i < count; ++i) {
      typename std::remove_reference_t<decltype(S.contacts)>::value_type c;
      c.type = PT;
      for (int k = 0; k < 4; ++k) c.idx[k] = idx4[4 * i + k];
      S.contacts.push_back(c);
    }
  });
  return 0;
}
}
