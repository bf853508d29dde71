// Device helpers for the gfx950 cloth-step kernels.
//
// All HIP sources are compiled with -ffp-contract=off and correctly rounded
// fp32 division and square root, so every expression below rounds exactly as
// the reference's x86-64 float build does for the same operation order.
// Three-term sums use the reference's fixed-size reduction order
// e0 + (e1 + e2) (Eigen's unrolled Vector3 dot/squaredNorm).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ando {

struct f3 {
  float x, y, z;
};

__host__ __device__ inline f3 mk(float a, float b, float c) { return f3{a, b, c}; }
__device__ inline f3 ld3(const float* p, int i) { return f3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
__device__ inline void st3(float* p, int i, f3 v) {
  p[3 * i] = v.x;
  p[3 * i + 1] = v.y;
  p[3 * i + 2] = v.z;
}
__device__ inline f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
__device__ inline f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
__device__ inline f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
__device__ inline float dot(f3 a, f3 b) { return a.x * b.x + (a.y * b.y + a.z * b.z); }
__device__ inline float sqn(f3 a) { return dot(a, a); }
__device__ inline float nrm(f3 a) { return sqrtf(sqn(a)); }
__device__ inline f3 cross(f3 a, f3 b) {
  return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ inline f3 normalized(f3 a) {
  float z = sqn(a);
  return z > 0.0f ? a / sqrtf(z) : a;
}
__device__ inline float comp(f3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

// ---------------------------------------------------------------------------
// Cubic barrier V(g) = k/(2 gmax) (gmax - g)^3 and its g-derivatives
// (barrier.cpp:29-65).  Near contact (g < 1e-2 gmax) the derivatives are
// evaluated in FP64 and rounded once.
// ---------------------------------------------------------------------------
__device__ inline float barrier_grad(float g, float gmax, float k, bool fp64_near = true) {
  if (gmax <= 0.0f || g >= gmax) return 0.0f;
  if (fp64_near && g < 1e-2f * gmax) {
    double d = (double)gmax - (double)g;
    return (float)(-1.5 * (double)k * (1.0 / (double)gmax) * d * d);
  }
  float d = gmax - g;
  float inv = 1.0f / gmax;
  return -1.5f * k * inv * d * d;
}
__device__ inline float barrier_hess(float g, float gmax, float k, bool fp64_near = true) {
  if (gmax <= 0.0f || g >= gmax) return 0.0f;
  if (fp64_near && g < 1e-2f * gmax) {
    double d = (double)gmax - (double)g;
    return (float)(3.0 * (double)k * (1.0 / (double)gmax) * d);
  }
  float d = gmax - g;
  float inv = 1.0f / gmax;
  return 3.0f * k * inv * d;
}
__device__ inline bool in_domain(float g, float gmax) { return gmax > 0.0f && g <= gmax; }

// enforce_spd on an exactly diagonal 3x3 (stiffness.cpp:236-254): the
// eigen-solver scales by the largest |entry| and scales back, then clamps.
__device__ inline f3 spd_diag(f3 d, float eps) {
  float s = fmaxf(fabsf(d.x), fmaxf(fabsf(d.y), fabsf(d.z)));
  if (s == 0.0f) s = 1.0f;
  float a = (d.x / s) * s, b = (d.y / s) * s, c = (d.z / s) * s;
  return f3{a < eps ? eps : a, b < eps ? eps : b, c < eps ? eps : c};
}
// Diagonal matrix times vector, then dot with the same vector (n^T H n).
__device__ inline float quad_diag(f3 Hd, f3 n) {
  f3 Hn = f3{Hd.x * n.x, Hd.y * n.y, Hd.z * n.z};
  return dot(n, Hn);
}

// ---------------------------------------------------------------------------
// Closest point on a triangle (collision.cpp:279-387).
// ---------------------------------------------------------------------------
struct PTResult {
  float dist;
  f3 normal;
  f3 wq;
};

__device__ inline void pt_finish(f3 p, f3 q, PTResult& r) {
  r.wq = q;
  f3 diff = p - q;
  r.dist = nrm(diff);
  r.normal = (double)r.dist > 1e-10 ? diff / r.dist : f3{0.0f, 0.0f, 1.0f};
}

__device__ inline PTResult point_triangle(f3 p, f3 a, f3 b, f3 c) {
  PTResult r;
  f3 ab = b - a, ac = c - a, ap = p - a;
  float d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0f && d2 <= 0.0f) { pt_finish(p, a, r); return r; }
  f3 bp = p - b;
  float d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0f && d4 <= d3) { pt_finish(p, b, r); return r; }
  f3 cp = p - c;
  float d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0f && d5 <= d6) { pt_finish(p, c, r); return r; }
  float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
    float v = d1 / (d1 - d3);
    pt_finish(p, a + v * ab, r);
    return r;
  }
  float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
    float w = d2 / (d2 - d6);
    pt_finish(p, a + w * ac, r);
    return r;
  }
  float va = d3 * d6 - d5 * d4;
  if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
    float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    pt_finish(p, b + w * (c - b), r);
    return r;
  }
  // 1.0 / float promotes to a double division, then rounds to float.
  float denom = (float)(1.0 / (double)(va + vb + vc));
  float v = vb * denom, w = vc * denom;
  r.wq = (a + v * ab) + w * ac;
  f3 diff = p - r.wq;
  r.dist = nrm(diff);
  f3 tn = cross(ab, ac);
  float ta = nrm(tn);
  if ((double)ta > 1e-10) {
    tn = tn / ta;
    if (dot(tn, diff) < 0.0f) tn = -tn;
    r.normal = tn;
  } else {
    r.normal = (double)r.dist > 1e-10 ? diff / r.dist : f3{0.0f, 0.0f, 1.0f};
  }
  return r;
}

// Block reductions for 256-thread workgroups (wave64 x 4).
template <class T, class Op>
__device__ inline T block_reduce(T v, T* lds, Op op) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int off = 32; off > 0; off >>= 1) v = op(v, __shfl_down(v, off, 64));
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  T r = lds[0];
  for (int w = 1; w < (int)(blockDim.x >> 6); ++w) r = op(r, lds[w]);
  __syncthreads();
  return r;
}

// A sum (double) and a maximum (float) reduced together: one shuffle tree
// and one LDS round instead of two.  Each result equals what block_reduce
// returns for that value alone (same association).
__device__ inline void block_reduce_sum_max(double& sum, float& mx, double* ldsd, float* ldsf) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int off = 32; off > 0; off >>= 1) {
    sum += __shfl_down(sum, off, 64);
    mx = fmaxf(mx, __shfl_down(mx, off, 64));
  }
  if (lane == 0) {
    ldsd[wid] = sum;
    ldsf[wid] = mx;
  }
  __syncthreads();
  double s = ldsd[0];
  float m = ldsf[0];
  for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
    s = s + ldsd[w];
    m = fmaxf(m, ldsf[w]);
  }
  __syncthreads();
  sum = s;
  mx = m;
}

}  // namespace ando
