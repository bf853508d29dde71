// Corrected-mode self contact: a linear BVH over triangles and over edges,
// built at a step's first detection (refitted for the swept boxes of its line
// searches), and the point-triangle / edge-edge queries.
//
// The reference's self broad phase (collision.cpp:63-276) never reports a
// pair for meshes with two or more primitives (its BVH only stores a
// primitive in size-1 ranges).  This restates the predicate it intends:
//   PT: vertex v, triangle t not incident to v, box [p - e, p + e] overlaps
//       t's box (closed intervals, e = 1e-4 in the reference);
//   EE: edges e1 < e2 sharing no vertex, boxes overlapping;
// then the narrow phases (collision.cpp:279-452), kept if gap < 0.01.
//
// Build: primitive boxes -> centroid bounds -> 30-bit Morton codes -> stable
// LSD radix sort (8-bit digits, per-workgroup LDS histograms and wave-ballot
// ranking) -> Karras hierarchy (one thread per internal node) -> child boxes
// from a sparse table of box unions over the sorted primitives (every node
// covers a contiguous sorted range, so its box is the union of two
// power-of-two windows; no atomics, deterministic).  A refit keeps the sorted
// order and the hierarchy and recomputes the boxes.  Query: one thread per
// vertex (or edge), stack traversal, a count pass and a fill pass into an
// exclusive-scanned list, each query's hits sorted by primitive index so the
// list order is the brute-force order (vertex, triangle) / (e1, e2).
#include <stdint.h>

#include "device.h"
#include "kernels.h"

namespace ando {

namespace lbvh {

constexpr int kStack = 64;

struct Box {
  float lo[3], hi[3];
};

__device__ inline Box load_box(const float* b, int i) {
  Box r;
  for (int k = 0; k < 3; ++k) {
    r.lo[k] = b[6 * (size_t)i + k];
    r.hi[k] = b[6 * (size_t)i + 3 + k];
  }
  return r;
}
__device__ inline void store_box(float* b, int i, const Box& r) {
  for (int k = 0; k < 3; ++k) {
    b[6 * (size_t)i + k] = r.lo[k];
    b[6 * (size_t)i + 3 + k] = r.hi[k];
  }
}
__device__ inline Box unite(const Box& a, const Box& b) {
  Box r;
  for (int k = 0; k < 3; ++k) {
    r.lo[k] = fminf(a.lo[k], b.lo[k]);
    r.hi[k] = fmaxf(a.hi[k], b.hi[k]);
  }
  return r;
}
// AABB::overlaps (collision.h:37-41): closed intervals.
__device__ inline bool overlaps(const Box& a, const Box& b) {
  return (a.lo[0] <= b.hi[0] && a.hi[0] >= b.lo[0]) && (a.lo[1] <= b.hi[1] && a.hi[1] >= b.lo[1]) &&
         (a.lo[2] <= b.hi[2] && a.hi[2] >= b.lo[2]);
}

// Primitive boxes: min / max over the primitive's `kind` vertices (3 for
// triangles, 2 for edges), as AABB::expand (compute_triangle_aabb /
// compute_edge_aabb).
__global__ void k_prim_boxes(int P, int kind, const int32_t* __restrict__ prims, const float* __restrict__ x,
                             float* __restrict__ boxes) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  Box b;
  for (int k = 0; k < 3; ++k) {
    b.lo[k] = x[3 * prims[kind * i] + k];
    b.hi[k] = b.lo[k];
  }
  for (int c = 1; c < kind; ++c)
    for (int k = 0; k < 3; ++k) {
      const float v = x[3 * prims[kind * i + c] + k];
      b.lo[k] = fminf(b.lo[k], v);
      b.hi[k] = fmaxf(b.hi[k], v);
    }
  store_box(boxes, i, b);
}

// Primitive boxes as unions of per-vertex boxes (6 floats per vertex).
__global__ void k_prim_boxes_union(int P, int kind, const int32_t* __restrict__ prims, const float* __restrict__ vbox,
                                   float* __restrict__ boxes) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  Box b = load_box(vbox, prims[kind * i]);
  for (int c = 1; c < kind; ++c) {
    const Box v = load_box(vbox, prims[kind * i + c]);
    for (int k = 0; k < 3; ++k) {
      b.lo[k] = fminf(b.lo[k], v.lo[k]);
      b.hi[k] = fmaxf(b.hi[k], v.hi[k]);
    }
  }
  store_box(boxes, i, b);
}

// Centroid bounds: per-workgroup partial min / max, then one workgroup.
__global__ void k_centroid_bounds(int P, const float* __restrict__ boxes, float* __restrict__ part) {
  __shared__ float s[6][256];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  float c[3] = {INFINITY, INFINITY, INFINITY}, d[3] = {-INFINITY, -INFINITY, -INFINITY};
  if (i < P) {
    const Box b = load_box(boxes, i);
    for (int k = 0; k < 3; ++k) c[k] = d[k] = 0.5f * (b.lo[k] + b.hi[k]);
  }
  for (int k = 0; k < 3; ++k) {
    s[k][threadIdx.x] = c[k];
    s[3 + k][threadIdx.x] = d[k];
  }
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st)
      for (int k = 0; k < 3; ++k) {
        s[k][threadIdx.x] = fminf(s[k][threadIdx.x], s[k][threadIdx.x + st]);
        s[3 + k][threadIdx.x] = fmaxf(s[3 + k][threadIdx.x], s[3 + k][threadIdx.x + st]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}
// One 256-thread workgroup: each thread folds a strided share of the partials,
// then a tree in LDS (minima and maxima are exact in any order).
__global__ void __launch_bounds__(256) k_centroid_bounds_final(int G, const float* __restrict__ part,
                                                              float* __restrict__ out) {
  __shared__ float s[6][256];
  float c[3] = {INFINITY, INFINITY, INFINITY}, d[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int g = threadIdx.x; g < G; g += blockDim.x)
    for (int k = 0; k < 3; ++k) {
      c[k] = fminf(c[k], part[6 * g + k]);
      d[k] = fmaxf(d[k], part[6 * g + 3 + k]);
    }
  for (int k = 0; k < 3; ++k) {
    s[k][threadIdx.x] = c[k];
    s[3 + k][threadIdx.x] = d[k];
  }
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st)
      for (int k = 0; k < 3; ++k) {
        s[k][threadIdx.x] = fminf(s[k][threadIdx.x], s[k][threadIdx.x + st]);
        s[3 + k][threadIdx.x] = fmaxf(s[3 + k][threadIdx.x], s[3 + k][threadIdx.x + st]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) out[threadIdx.x] = s[threadIdx.x][0];
}

__device__ inline uint32_t spread10(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// 30-bit Morton code of the box centroid in the centroid bounds; the value is
// the primitive index (the sort is stable, so equal codes keep index order).
__global__ void k_morton(int P, const float* __restrict__ boxes, const float* __restrict__ bounds,
                         uint32_t* __restrict__ keys, int32_t* __restrict__ vals) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  const Box b = load_box(boxes, i);
  uint32_t q[3];
  for (int k = 0; k < 3; ++k) {
    const float c = 0.5f * (b.lo[k] + b.hi[k]);
    const float ext = bounds[3 + k] - bounds[k];
    float u = ext > 0.0f ? (c - bounds[k]) / ext : 0.0f;
    u = fminf(fmaxf(u, 0.0f), 1.0f);
    q[k] = (uint32_t)fminf(u * 1024.0f, 1023.0f);
  }
  keys[i] = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
  vals[i] = i;
}

// Radix sort, pass `shift`: digit histogram per workgroup (256 keys), stored
// digit-major (counts[d * G + g]) so that an exclusive scan gives each
// (digit, workgroup) its output offset.
__global__ void __launch_bounds__(256) k_radix_count(int P, const uint32_t* __restrict__ keys, int shift,
                                                     int32_t* __restrict__ counts) {
  __shared__ int32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < P) atomicAdd(&h[(keys[i] >> shift) & 0xffu], 1);
  __syncthreads();
  counts[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter: rank of each key among the workgroup's keys with the same
// digit that precede it = (same-digit keys in earlier waves) + (same-digit
// lanes below it in its wave, found with 9 ballots: 8 digit bits + validity).
__global__ void __launch_bounds__(256) k_radix_scatter(int P, const uint32_t* __restrict__ keys,
                                                       const int32_t* __restrict__ vals, int shift,
                                                       const int32_t* __restrict__ offsets,
                                                       uint32_t* __restrict__ keys_out,
                                                       int32_t* __restrict__ vals_out) {
  __shared__ int32_t wc[4][256];
  for (int k = threadIdx.x; k < 4 * 256; k += 256) wc[k / 256][k % 256] = 0;
  __syncthreads();
  const int i = blockIdx.x * 256 + threadIdx.x;
  const bool ok = i < P;
  const uint32_t key = ok ? keys[i] : 0u;
  const uint32_t d = ok ? ((key >> shift) & 0xffu) : 0x100u;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t m = ~0ull;
  for (int bit = 0; bit < 9; ++bit) {
    const uint64_t bb = __ballot((d >> bit) & 1u);
    m &= ((d >> bit) & 1u) ? bb : ~bb;
  }
  const int rank = __popcll(m & ((1ull << lane) - 1ull));
  if (ok && rank == 0) wc[wave][d] = __popcll(m);
  __syncthreads();
  if (!ok) return;
  int before = 0;
  for (int w = 0; w < wave; ++w) before += wc[w][d];
  const int dst = offsets[(size_t)d * gridDim.x + blockIdx.x] + before + rank;
  keys_out[dst] = key;
  vals_out[dst] = vals[i];
}

// Karras (2012) hierarchy over sorted codes.  Internal node i in [0, P-1);
// child c >= 0 is internal node c, c < 0 is leaf ~c (sorted position).
// delta(i, j): common prefix of (code, position) pairs, -1 outside [0, P).
__device__ inline int delta(const uint32_t* keys, int P, int i, int j) {
  if (j < 0 || j >= P) return -1;
  const uint32_t a = keys[i], b = keys[j];
  if (a != b) return __clz(a ^ b);
  return 32 + __clz((uint32_t)i ^ (uint32_t)j);
}

__global__ void k_karras(int P, const uint32_t* __restrict__ keys, int32_t* __restrict__ left,
                         int32_t* __restrict__ right, int32_t* __restrict__ first, int32_t* __restrict__ last) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P - 1) return;
  const int dl = delta(keys, P, i, i - 1), dr = delta(keys, P, i, i + 1);
  const int d = dr > dl ? 1 : -1;
  const int dmin = d > 0 ? dl : dr;
  int lmax = 2;
  while (delta(keys, P, i, i + lmax * d) > dmin) lmax *= 2;
  int l = 0;
  for (int t = lmax / 2; t >= 1; t /= 2)
    if (delta(keys, P, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dnode = delta(keys, P, i, j);
  int s = 0;
  for (int div = 2;; div *= 2) {
    const int t = (l + div - 1) / div;
    if (delta(keys, P, i, i + (s + t) * d) > dnode) s += t;
    if (t <= 1) break;
  }
  const int g = i + s * d + (d < 0 ? -1 : 0);
  const int lo = min(i, j), hi = max(i, j);
  left[i] = (lo == g) ? ~g : g;
  right[i] = (hi == g + 1) ? ~(g + 1) : g + 1;
  first[i] = lo;
  last[i] = hi;
}

// Sparse table of box unions: level 0 = boxes in sorted order; level k holds
// the union of [s, s + 2^k).
__global__ void k_table_level0(int P, const float* __restrict__ boxes, const int32_t* __restrict__ order,
                               float* __restrict__ T0) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= P) return;
  store_box(T0, s, load_box(boxes, order[s]));
}
// Levels k0 + 1 .. k0 + L (L <= 3) in one pass, from level k0 (h = 2^k0
// boxes wide): level k0 + j at s is formed from the 2^j level-k0 boxes
// s + t h with the pairing of the level-by-level rule (level k at s = level
// k - 1 at s united with level k - 1 at s + 2^(k-1)), so every entry equals
// the one that rule gives.  Level k at s exists for s + 2^k <= P.
__global__ void k_table_levels(int P, int k0, int L, float* __restrict__ table) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int h = 1 << k0;
  if (s + 2 * h > P) return;
  const float* base = table + (size_t)k0 * 6 * P;
  Box cur = unite(load_box(base, s), load_box(base, s + h));
  store_box(table + (size_t)(k0 + 1) * 6 * P, s, cur);
  if (L < 2 || s + 4 * h > P) return;
  cur = unite(cur, unite(load_box(base, s + 2 * h), load_box(base, s + 3 * h)));
  store_box(table + (size_t)(k0 + 2) * 6 * P, s, cur);
  if (L < 3 || s + 8 * h > P) return;
  const Box o1 = unite(load_box(base, s + 4 * h), load_box(base, s + 5 * h));
  const Box o2 = unite(load_box(base, s + 6 * h), load_box(base, s + 7 * h));
  cur = unite(cur, unite(o1, o2));
  store_box(table + (size_t)(k0 + 3) * 6 * P, s, cur);
}
// Box of the sorted range [a, b] = the union of two overlapping power-of-two
// windows of the table.
__device__ inline Box range_box(const float* __restrict__ table, int P, int a, int b) {
  const int len = b - a + 1;
  const int k = 31 - __clz((uint32_t)len);
  const float* Tk = table + (size_t)k * 6 * P;
  return unite(load_box(Tk, a), load_box(Tk, b - (1 << k) + 1));
}

// Boxes of both children of every internal node side by side (12 floats per
// node: left, right), so a traversal tests both children of a node with one
// contiguous read: a leaf child takes its primitive's box (table level 0), an
// internal child the box of its sorted range.
__global__ void k_child_boxes(int P, const int32_t* __restrict__ left, const int32_t* __restrict__ right,
                              const int32_t* __restrict__ first, const int32_t* __restrict__ last,
                              const float* __restrict__ table, float* __restrict__ cbox) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P - 1) return;
  const int l = left[i], r = right[i];
  store_box(cbox, 2 * i, l < 0 ? load_box(table, ~l) : range_box(table, P, first[l], last[l]));
  store_box(cbox, 2 * i + 1, r < 0 ? load_box(table, ~r) : range_box(table, P, first[r], last[r]));
}

// Refit: the primitive boxes from the per-vertex boxes (as k_prim_boxes_union)
// stored both by primitive and, through the kept sorted order, as table
// level 0.
__global__ void k_refit_leaves(int P, int kind, const int32_t* __restrict__ prims, const int32_t* __restrict__ order,
                               const float* __restrict__ vbox, float* __restrict__ boxes, float* __restrict__ T0) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= P) return;
  const int i = order[s];
  Box b = load_box(vbox, prims[kind * i]);
  for (int c = 1; c < kind; ++c) {
    const Box v = load_box(vbox, prims[kind * i + c]);
    for (int k = 0; k < 3; ++k) {
      b.lo[k] = fminf(b.lo[k], v.lo[k]);
      b.hi[k] = fmaxf(b.hi[k], v.hi[k]);
    }
  }
  store_box(boxes, i, b);
  store_box(T0, s, b);
}

// Triangle barycentrics of p (collision.cpp:11-43): clamped, renormalised.
__device__ inline f3 triangle_bary(f3 p, f3 a, f3 b, f3 c) {
  const f3 v0 = b - a, v1 = c - a, v2 = p - a;
  const float d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1), d20 = dot(v2, v0), d21 = dot(v2, v1);
  const float denom = d00 * d11 - d01 * d01;
  float v = 0.0f, w = 0.0f;
  const float feps = 1.1920928955078125e-07f;
  if (fabsf(denom) > feps) {
    v = (d11 * d20 - d01 * d21) / denom;
    w = (d00 * d21 - d01 * d20) / denom;
  }
  v = fminf(fmaxf(v, 0.0f), 1.0f);
  w = fminf(fmaxf(w, 0.0f), 1.0f);
  if (v + w > 1.0f) {
    const float sum = v + w;
    if (sum > feps) {
      v /= sum;
      w /= sum;
    }
  }
  return f3{1.0f - v - w, v, w};
}

// Segment-segment closest points (collision.cpp:389-452); s_out / t_out
// (optional) receive the closest points' parameters (0 for both when both
// edges are degenerate).
__device__ inline void edge_edge(f3 p0, f3 p1, f3 q0, f3 q1, float& dist, f3& normal, f3& wp, f3& wq,
                                 float* s_out = nullptr, float* t_out = nullptr) {
  const f3 d1 = p1 - p0, d2 = q1 - q0, r = p0 - q0;
  const float a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
  const float eps = 1e-10f;
  auto cl = [](float v) { return fminf(fmaxf(v, 0.0f), 1.0f); };
  float s = 0.0f, t = 0.0f;
  if (a <= eps && e <= eps) {
    wp = p0;
    wq = q0;
  } else {
    if (a <= eps) {
      s = 0.0f;
      t = cl(f / e);
    } else {
      const float c = dot(d1, r);
      if (e <= eps) {
        t = 0.0f;
        s = cl(-c / a);
      } else {
        const float b = dot(d1, d2);
        const float denom = a * e - b * b;
        s = denom != 0.0f ? cl((b * f - c * e) / denom) : 0.0f;
        t = (b * s + f) / e;
        if (t < 0.0f) {
          t = 0.0f;
          s = cl(-c / a);
        } else if (t > 1.0f) {
          t = 1.0f;
          s = cl((b - c) / a);
        }
      }
    }
    wp = p0 + s * d1;
    wq = q0 + t * d2;
  }
  if (s_out) *s_out = s;
  if (t_out) *t_out = t;
  const f3 diff = wp - wq;
  dist = nrm(diff);
  normal = (double)dist > 1e-10 ? diff / dist : f3{0.0f, 0.0f, 1.0f};
}

// Stack traversal of one BVH for query box q; visit(sorted position) for every
// overlapping leaf.  Returns false if the stack would overflow (the caller
// reports an error; nothing is written out of bounds).
template <class F>
__device__ inline bool traverse(int P, const Box& q, const int32_t* left, const int32_t* right, const float* cbox,
                                const float* leaves, F&& visit) {
  if (P == 1) {
    if (overlaps(q, load_box(leaves, 0))) visit(0);
    return true;
  }
  // A popped node's two children are tested together (cbox: k_child_boxes):
  // overlapping leaves are visited at once, overlapping internal children are
  // pushed (right first, so the left subtree is entered first).  Callers sort
  // each query's hits, so the visit order does not reach any result.
  int stack[kStack];
  int sp = 0;
  stack[sp++] = 0;
  while (sp > 0) {
    const int node = stack[--sp];
    const int l = left[node], r = right[node];
    const bool ol = overlaps(q, load_box(cbox, 2 * node));
    const bool orr = overlaps(q, load_box(cbox, 2 * node + 1));
    if (ol && l < 0) visit(~l);
    if (orr && r < 0) visit(~r);
    if (sp + 2 > kStack) return false;
    if (orr && r >= 0) stack[sp++] = r;
    if (ol && l >= 0) stack[sp++] = l;
  }
  return true;
}

// Insertion sort of one query's hits out[base .. base+k) by key (the
// traversal visits them in Morton order; the lists are short).
// A query sorts up to kSortLocal hits itself (insertion sort).  A longer list
// (a swept box over a large part of the mesh can hit 10^5 primitives, and an
// insertion sort of that many records by one thread runs for minutes) stays
// in traversal order and sets bit 1 of err; the caller then sorts all the
// hits of the pass by (query, key) with sort_long_hits.
constexpr int kSortLocal = 64;

__device__ inline void sort_hits(SelfContact* out, int32_t* key, int base, int k, uint32_t* err) {
  if (k > kSortLocal) {
    atomicOr(err, 2u);
    return;
  }
  for (int a = 1; a < k; ++a) {
    const SelfContact cur = out[base + a];
    const int ck = key[base + a];
    int b = a - 1;
    while (b >= 0 && key[base + b] > ck) {
      out[base + b + 1] = out[base + b];
      key[base + b + 1] = key[base + b];
      --b;
    }
    out[base + b + 1] = cur;
    key[base + b + 1] = ck;
  }
}

// PT query, pass 0 counts hits per vertex, pass 1 writes them at ptr[v] and
// sorts the vertex's hits by triangle index.
// Hits a query keeps from its count pass for its fill pass.
constexpr int kQueryCache = 16;

__global__ void __launch_bounds__(64) k_query_pt(int pass, int n, const float* __restrict__ x, float pt_box, float keep,
                           const int32_t* __restrict__ tris, int P, const int32_t* __restrict__ left,
                           const int32_t* __restrict__ right, const float* __restrict__ nodes,
                           const float* __restrict__ leaves, const int32_t* __restrict__ order,
                           int32_t* __restrict__ count, const int32_t* __restrict__ ptr,
                           SelfContact* __restrict__ out, int32_t* __restrict__ key, uint32_t* __restrict__ err,
                           int32_t* __restrict__ cache) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n || P == 0) {
    if (pass == 0 && v < n) count[v] = 0;
    return;
  }
  const f3 p = ld3(x, v);
  Box q;
  q.lo[0] = p.x - pt_box; q.lo[1] = p.y - pt_box; q.lo[2] = p.z - pt_box;
  q.hi[0] = p.x + pt_box; q.hi[1] = p.y + pt_box; q.hi[2] = p.z + pt_box;
  int k = 0;
  const int base = pass ? ptr[v] : 0;
  auto visit = [&](int t) {
    const int a = tris[3 * t], b = tris[3 * t + 1], c = tris[3 * t + 2];
    if (a == v || b == v || c == v) return;
    const f3 xa = ld3(x, a), xb = ld3(x, b), xc = ld3(x, c);
    const PTResult r = point_triangle(p, xa, xb, xc);
    if (!(r.dist < keep)) return;
    if (!pass && k < kQueryCache) cache[(size_t)v * kQueryCache + k] = t;
    if (pass) {
      SelfContact sc;
      sc.type = 0;
      sc.idx[0] = v; sc.idx[1] = a; sc.idx[2] = b; sc.idx[3] = c;
      sc.gap = r.dist;
      sc.nx = r.normal.x; sc.ny = r.normal.y; sc.nz = r.normal.z;
      const f3 bc = triangle_bary(r.wq, xa, xb, xc);
      sc.w[0] = 1.0f; sc.w[1] = -bc.x; sc.w[2] = -bc.y; sc.w[3] = -bc.z;
      sc.qx = r.wq.x; sc.qy = r.wq.y; sc.qz = r.wq.z;
      out[base + k] = sc;
      key[base + k] = t;
    }
    ++k;
  };
  // The fill pass of a query whose hits all fit the cache re-evaluates only
  // those, in the order the count pass found them (the same records).
  const int cached = pass ? count[v] : kQueryCache + 1;
  bool ok = true;
  if (cached <= kQueryCache) {
    for (int h = 0; h < cached; ++h) visit(cache[(size_t)v * kQueryCache + h]);
  } else {
    ok = traverse(P, q, left, right, nodes, leaves, [&](int s) { visit(order[s]); });
  }
  if (!ok) atomicOr(err, 1u);
  if (pass == 0) {
    count[v] = k;
    return;
  }
  sort_hits(out, key, base, k, err);
}

// EE query for edge e1 against edges e2 > e1, same two passes; hits sorted by e2.
__global__ void __launch_bounds__(64) k_query_ee(int pass, int E, const float* __restrict__ x, float keep, const int32_t* __restrict__ edges,
                           int P, const int32_t* __restrict__ left, const int32_t* __restrict__ right,
                           const float* __restrict__ nodes, const float* __restrict__ leaves,
                           const int32_t* __restrict__ order, int32_t* __restrict__ count,
                           const int32_t* __restrict__ ptr, SelfContact* __restrict__ out,
                           int32_t* __restrict__ key, uint32_t* __restrict__ err, int32_t* __restrict__ cache) {
  const int e1 = blockIdx.x * blockDim.x + threadIdx.x;
  if (e1 >= E) return;
  const int a0 = edges[2 * e1], a1 = edges[2 * e1 + 1];
  const f3 p0 = ld3(x, a0), p1 = ld3(x, a1);
  Box q;
  q.lo[0] = fminf(p0.x, p1.x); q.lo[1] = fminf(p0.y, p1.y); q.lo[2] = fminf(p0.z, p1.z);
  q.hi[0] = fmaxf(p0.x, p1.x); q.hi[1] = fmaxf(p0.y, p1.y); q.hi[2] = fmaxf(p0.z, p1.z);
  int k = 0;
  const int base = pass ? ptr[e1] : 0;
  auto visit = [&](int e2) {
    if (e2 <= e1) return;
    const int b0 = edges[2 * e2], b1 = edges[2 * e2 + 1];
    if (a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1) return;
    float dist;
    f3 nrm_, wp, wq;
    edge_edge(p0, p1, ld3(x, b0), ld3(x, b1), dist, nrm_, wp, wq);
    if (!(dist < keep)) return;
    if (!pass && k < kQueryCache) cache[(size_t)e1 * kQueryCache + k] = e2;
    if (pass) {
      SelfContact sc;
      sc.type = 1;
      sc.idx[0] = a0; sc.idx[1] = a1; sc.idx[2] = b0; sc.idx[3] = b1;
      sc.gap = dist;
      sc.nx = nrm_.x; sc.ny = nrm_.y; sc.nz = nrm_.z;
      sc.w[0] = 0.5f; sc.w[1] = 0.5f; sc.w[2] = -0.5f; sc.w[3] = -0.5f;
      sc.qx = wq.x; sc.qy = wq.y; sc.qz = wq.z;
      out[base + k] = sc;
      key[base + k] = e2;
    }
    ++k;
  };
  const int cached = pass ? count[e1] : kQueryCache + 1;
  bool ok = true;
  if (cached <= kQueryCache) {
    for (int h = 0; h < cached; ++h) visit(cache[(size_t)e1 * kQueryCache + h]);
  } else {
    ok = traverse(P, q, left, right, nodes, leaves, [&](int s) { visit(order[s]); });
  }
  if (!ok) atomicOr(err, 1u);
  if (pass == 0) {
    count[e1] = k;
    return;
  }
  sort_hits(out, key, base, k, err);
}

// Corrected-physics candidates: box overlap alone (no distance filter; the
// gaps are refreshed at every Newton iteration).  PT: vertex box grown by gm
// against the union boxes of triangles not incident to the vertex; EE: an
// edge's union box grown by gm against those of edges e2 > e1 sharing no
// vertex.  Same two passes and hit cache as the queries above; records carry
// only type and indices.
__global__ void __launch_bounds__(64) k_query_pt_box(int pass, int n, const float* __restrict__ vbox, float gm,
                                                     const int32_t* __restrict__ tris, int P,
                                                     const int32_t* __restrict__ left, const int32_t* __restrict__ right,
                                                     const float* __restrict__ nodes, const float* __restrict__ leaves,
                                                     const int32_t* __restrict__ order, int32_t* __restrict__ count,
                                                     const int32_t* __restrict__ ptr, SelfContact* __restrict__ out,
                                                     int32_t* __restrict__ key, uint32_t* __restrict__ err,
                                                     int32_t* __restrict__ cache) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n || P == 0) {
    if (pass == 0 && v < n) count[v] = 0;
    return;
  }
  Box q = load_box(vbox, v);
  for (int k = 0; k < 3; ++k) {
    q.lo[k] = q.lo[k] - gm;
    q.hi[k] = q.hi[k] + gm;
  }
  int k = 0;
  const int base = pass ? ptr[v] : 0;
  auto visit = [&](int t) {
    const int a = tris[3 * t], b = tris[3 * t + 1], c = tris[3 * t + 2];
    if (a == v || b == v || c == v) return;
    if (!pass && k < kQueryCache) cache[(size_t)v * kQueryCache + k] = t;
    if (pass) {
      SelfContact sc = {};
      sc.type = 0;
      sc.idx[0] = v; sc.idx[1] = a; sc.idx[2] = b; sc.idx[3] = c;
      out[base + k] = sc;
      key[base + k] = t;
    }
    ++k;
  };
  const int cached = pass ? count[v] : kQueryCache + 1;
  bool ok = true;
  if (cached <= kQueryCache) {
    for (int h = 0; h < cached; ++h) visit(cache[(size_t)v * kQueryCache + h]);
  } else {
    ok = traverse(P, q, left, right, nodes, leaves, [&](int s) { visit(order[s]); });
  }
  if (!ok) atomicOr(err, 1u);
  if (pass == 0) {
    count[v] = k;
    return;
  }
  sort_hits(out, key, base, k, err);
}

__global__ void __launch_bounds__(64) k_query_ee_box(int pass, int E, const float* __restrict__ eboxes, float gm,
                                                     const int32_t* __restrict__ edges, int P,
                                                     const int32_t* __restrict__ left, const int32_t* __restrict__ right,
                                                     const float* __restrict__ nodes, const float* __restrict__ leaves,
                                                     const int32_t* __restrict__ order, int32_t* __restrict__ count,
                                                     const int32_t* __restrict__ ptr, SelfContact* __restrict__ out,
                                                     int32_t* __restrict__ key, uint32_t* __restrict__ err,
                                                     int32_t* __restrict__ cache) {
  const int e1 = blockIdx.x * blockDim.x + threadIdx.x;
  if (e1 >= E) return;
  const int a0 = edges[2 * e1], a1 = edges[2 * e1 + 1];
  Box q = load_box(eboxes, e1);
  for (int k = 0; k < 3; ++k) {
    q.lo[k] = q.lo[k] - gm;
    q.hi[k] = q.hi[k] + gm;
  }
  int k = 0;
  const int base = pass ? ptr[e1] : 0;
  auto visit = [&](int e2) {
    if (e2 <= e1) return;
    const int b0 = edges[2 * e2], b1 = edges[2 * e2 + 1];
    if (a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1) return;
    if (!pass && k < kQueryCache) cache[(size_t)e1 * kQueryCache + k] = e2;
    if (pass) {
      SelfContact sc = {};
      sc.type = 1;
      sc.idx[0] = a0; sc.idx[1] = a1; sc.idx[2] = b0; sc.idx[3] = b1;
      out[base + k] = sc;
      key[base + k] = e2;
    }
    ++k;
  };
  const int cached = pass ? count[e1] : kQueryCache + 1;
  bool ok = true;
  if (cached <= kQueryCache) {
    for (int h = 0; h < cached; ++h) visit(cache[(size_t)e1 * kQueryCache + h]);
  } else {
    ok = traverse(P, q, left, right, nodes, leaves, [&](int s) { visit(order[s]); });
  }
  if (!ok) atomicOr(err, 1u);
  if (pass == 0) {
    count[e1] = k;
    return;
  }
  sort_hits(out, key, base, k, err);
}

// ---------------------------------------------------------------------------
// Self-contact physics (corrected mode): step-constant barrier terms, the
// vertex view of the contact list, the PCG product of the blocks that couple
// two vertices, and the line-search CCD.
// ---------------------------------------------------------------------------

__device__ inline int find_block(const int32_t* rp, const int32_t* cols, int i, int j) {
  int lo = rp[i], hi = rp[i + 1];
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (cols[mid] < j) lo = mid + 1;
    else hi = mid;
  }
  return (lo < rp[i + 1] && cols[lo] == j) ? lo : -1;
}

// Per contact: k-bar = sum m_i w_i^2 / dt^2 (corrected physics:
// / max(gap_max, min_gap)^2) + max(0, Wn^T H_el Wn) over the
// elastic-only blocks linking its vertices (Stiffness::compute_contact_stiffness,
// stiffness.cpp:53-116; H_el blocks are diagonal, stored as 3 floats), then
// V' and V'' at the detected gap (integrator.cpp:315-338, 450-480).
__global__ void k_self_terms(int nc, int n, StepParams P, const SelfContact* __restrict__ sc,
                             const float* __restrict__ masses, const int32_t* __restrict__ rp,
                             const int32_t* __restrict__ cols, const float* __restrict__ Hel3,
                             SelfTerm* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const SelfContact& C = sc[c];
  float kb = 0.0f;
  const f3 n0 = mk(C.nx, C.ny, C.nz);
  const float nl = nrm(n0);
  if (P.dt > 0.0f && !((double)nl < 1e-12)) {
    const f3 nn = n0 / nl;
    int vv[4];
    float ww[4];
    int bc = 0;
    for (int i = 0; i < 4; ++i) {
      const int idx = C.idx[i];
      if (idx < 0 || idx >= n || C.w[i] == 0.0f) continue;
      vv[bc] = idx;
      ww[bc] = C.w[i];
      ++bc;
    }
    if (bc > 0) {
      float inertial = 0.0f;
      for (int i = 0; i < bc; ++i) inertial += masses[vv[i]] * ww[i] * ww[i];
      // Corrected physics: the wall's inertial scale m / max(gap_max, min_gap)^2.
      const float sc = P.cp ? fmaxf(fmaxf(P.gap_max, P.min_gap), 1e-12f) : P.dt;
      inertial /= (sc * sc);
      float Wn[12], Ht[12];
      for (int i = 0; i < bc; ++i) {
        Wn[3 * i] = ww[i] * nn.x;
        Wn[3 * i + 1] = ww[i] * nn.y;
        Wn[3 * i + 2] = ww[i] * nn.z;
      }
      for (int i = 0; i < 3 * bc; ++i) Ht[i] = 0.0f;
      for (int cb = 0; cb < bc; ++cb)
        for (int rb = 0; rb < bc; ++rb) {
          const int bi = find_block(rp, cols, vv[rb], vv[cb]);
          if (bi < 0) continue;
          for (int r = 0; r < 3; ++r) Ht[3 * rb + r] += Hel3[3 * (size_t)bi + r] * Wn[3 * cb + r];
        }
      float el = Wn[0] * Ht[0];
      for (int i = 1; i < 3 * bc; ++i) el = el + Wn[i] * Ht[i];
      if (el < 0.0f) el = 0.0f;
      kb = inertial + el;
    }
  }
  SelfTerm t;
  t.kb = kb;
  t.active = in_domain(C.gap, P.gap_max) ? 1 : 0;
  const f3 nn = normalize_or_default(n0, P.normal_eps);
  t.nx = nn.x; t.ny = nn.y; t.nz = nn.z;
  t.dV = t.active ? barrier_grad(C.gap, P.gap_max, kb, P.fp64_near) : 0.0f;
  t.d2 = t.active ? barrier_hess(C.gap, P.gap_max, kb, P.fp64_near) : 0.0f;
  out[c] = t;
}

// Vertex view: key = vertex of slot (c, s) or 0xffffffff when the slot carries
// no vertex (idx < 0 or weight 0) or the contact adds nothing (term inactive,
// or V' and V'' both zero: the gradient, the diagonal blocks and the coupling
// blocks all skip such a contact); value = c * 4 + s.  A stable sort by key
// then groups the slots by vertex in contact order.
__global__ void k_self_slots(int nc, const SelfContact* __restrict__ sc, const SelfTerm* __restrict__ st,
                             uint32_t* __restrict__ keys, int32_t* __restrict__ vals) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 4 * nc) return;
  const int c = t >> 2, s = t & 3;
  const int idx = sc[c].idx[s];
  const SelfTerm& T = st[c];
  const bool adds = T.active && (T.dV != 0.0f || T.d2 != 0.0f);
  keys[t] = (adds && idx >= 0 && sc[c].w[s] != 0.0f) ? (uint32_t)idx : 0xffffffffu;
  vals[t] = t;
}
__global__ void k_count_keys(int count, int n, const uint32_t* __restrict__ keys, int32_t* __restrict__ hist) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const uint32_t k = keys[t];
  if (k < (uint32_t)n) atomicAdd(&hist[k], 1);
}

// Rows with self-contact coupling blocks (vertex view entries), listed in
// ascending order: flags, then the rows at their exclusive-scan positions.
__global__ void k_couple_row_flags(int n, const int32_t* __restrict__ vptr, int32_t* __restrict__ flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flag[i] = vptr[i + 1] > vptr[i] ? 1 : 0;
}
__global__ void k_couple_row_fill(int n, const int32_t* __restrict__ vptr, const int32_t* __restrict__ pos,
                                  int32_t* __restrict__ crow) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && vptr[i + 1] > vptr[i]) crow[pos[i]] = i;
}

// Faithful mode on a one-triangle mesh: each vertex is the point of at most
// one PT contact (there is one triangle); pt_of[v] = that contact, else -1
// (cleared by the caller).
__global__ void k_pt_of(int npt, int n, const SelfContact* __restrict__ sc, int32_t* __restrict__ pt_of) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= npt || sc[c].type != 0) return;
  const int v = sc[c].idx[0];
  if (v >= 0 && v < n) pt_of[v] = c;
}

// Restitution of the self contacts (Integrator::apply_contact_restitution,
// integrator.cpp:565-674): every contact within gap_limit, in list order,
// gives its impulse vertices -- a PT contact's point, an EE contact's four
// vertices, the second edge's with the normal negated -- the velocity
// v - (1 + e) (v.n) n when v.n < 0.  Impulses on different vertices are
// independent and those on one vertex run in contact order, so slots are
// grouped by vertex (key = vertex of an impulse slot, else 0xffffffff; a
// stable sort keeps contact order) and vertices run in parallel.
__global__ void k_rest_slots(int nc, int n, const SelfContact* __restrict__ sc, float gap_limit,
                             uint32_t* __restrict__ keys, int32_t* __restrict__ vals) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 4 * nc) return;
  const int c = t >> 2, s = t & 3;
  const SelfContact& C = sc[c];
  const bool use = !(C.gap > gap_limit) && (C.type == 1 || s == 0);
  const int idx = C.idx[s];
  keys[t] = (use && idx >= 0 && idx < n) ? (uint32_t)idx : 0xffffffffu;
  vals[t] = t;
}
__global__ void k_rest_self(int n, const int32_t* __restrict__ ptr, const int32_t* __restrict__ code,
                            const SelfContact* __restrict__ sc, float e, float* __restrict__ v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int beg = ptr[i], end = ptr[i + 1];
  if (beg == end) return;
  f3 vel = ld3(v, i);
  for (int k = beg; k < end; ++k) {
    const int c = code[k] >> 2, s = code[k] & 3;
    const SelfContact& C = sc[c];
    f3 nn = mk(C.nx, C.ny, C.nz);
    if (C.type == 1 && s >= 2) nn = -nn;
    const float nl = nrm(nn);
    if ((double)nl < 1e-8) continue;
    nn = nn / nl;
    const float vn = dot(vel, nn);
    if (vn < 0.0f) vel = vel - ((1.0f + e) * vn) * nn;
  }
  st3(v, i, vel);
}

// The blocks coupling two vertices, materialised once per step: for vertex
// view entry e (contact c, slot si on row vertex i) and the contact's other
// slots sj in slot order, block 3 e + jj = d2 (w_si n)(w_sj n)^T with entries
// below barrier_tol dropped (zero when the term is inactive or the slot has
// no vertex); column = that slot's vertex (the row vertex when it has none).
__global__ void k_self_blocks(int entries, SelfView sv, float barrier_tol, int32_t* __restrict__ ccol,
                              float* __restrict__ cval) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * entries) return;
  const int e = t / 3, jj = t - 3 * e;
  const int code = sv.code[e], c = code >> 2, si = code & 3;
  const int sj = jj < si ? jj : jj + 1;
  const SelfContact& C = sv.sc[c];
  const SelfTerm& st = sv.st[c];
  float b[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  int col = C.idx[si];
  if (st.active && st.d2 != 0.0f && C.w[sj] != 0.0f && C.idx[sj] >= 0) {
    const f3 nn = mk(st.nx, st.ny, st.nz);
    const f3 gi = C.w[si] * nn, gj = C.w[sj] * nn;
    for (int k = 0; k < 9; ++k) {
      const float val = st.d2 * (comp(gi, k / 3) * comp(gj, k % 3));
      b[k] = fabsf(val) < barrier_tol ? 0.0f : val;
    }
    col = C.idx[sj];
  }
  ccol[t] = col;
  for (int k = 0; k < 9; ++k) cval[9 * (size_t)t + k] = b[k];
}


// Segmented sort of a pass's hits by (query, key): two stable radix sorts of
// positions, by key and then by query.  k_hit_keys: key of position j;
// k_hit_queries: the query of the position each sorted slot holds (the q with
// ptr[q] <= pos < ptr[q + 1], by bisection); k_gather_hits: the records in
// sorted order.
__global__ void k_hit_keys(int count, const int32_t* __restrict__ key, uint32_t* __restrict__ keys,
                           int32_t* __restrict__ pos) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  keys[j] = (uint32_t)key[j];
  pos[j] = j;
}
__global__ void k_hit_queries(int count, int nq, const int32_t* __restrict__ ptr, const int32_t* __restrict__ pos,
                              uint32_t* __restrict__ keys) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  const int e = pos[j];
  int lo = 0, hi = nq - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (ptr[mid] <= e) lo = mid;
    else hi = mid - 1;
  }
  keys[j] = (uint32_t)lo;
}
__global__ void k_gather_hits(int count, const int32_t* __restrict__ pos, const SelfContact* __restrict__ in,
                              const int32_t* __restrict__ kin, SelfContact* __restrict__ out,
                              int32_t* __restrict__ kout) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  out[j] = in[pos[j]];
  kout[j] = kin[pos[j]];
}

// Sum of n hit counts in 64 bits (the int32 scan of a runaway detection would
// wrap); *out must start at zero.
__global__ void __launch_bounds__(256) k_sum_counts64(int n, const int32_t* __restrict__ count,
                                                      unsigned long long* __restrict__ out) {
  unsigned long long v = 0ull;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    v += (unsigned long long)(uint32_t)count[i];
  __shared__ unsigned long long lds[4];
  const unsigned long long r = block_reduce(v, lds, [](unsigned long long a, unsigned long long b) { return a + b; });
  if (threadIdx.x == 0 && r) atomicAdd(out, r);
}

// Line-search feasibility of the self contacts (LineSearch::is_feasible,
// line_search.cpp:70-132, with ccd_point_triangle / ccd_edge_edge :163-247):
// for trial t (alpha = 2^-t) the contact fails if its 10-sample CCD finds a
// distance below 1e-6 at a sample time t' < 1.  Bit t of the AND-reduced mask
// is set when every contact passes trial t.
// Additive CCD (Li, Kaufman, Jiang et al., "Codimensional Incremental
// Potential Contact", 2021, Algorithm 1), thickness 0, t_c = 1: the pair moves
// by p (per vertex) over t in [0, 1]; with the mean displacement removed, l_p
// bounds how fast the distance can shrink, so advancing by (1 - s) d / l_p,
// then 0.9 d / l_p, never crosses.  Returns 1 when the whole motion keeps the
// pair apart, otherwise the last time at which the distance was still at
// least s times its initial value (a conservative time of impact).
constexpr float kAccdS = 0.1f;
constexpr int kAccdMaxIter = 256;

__device__ inline float pair_distance(int type, const f3* q) {
  if (type == 0) return point_triangle(q[0], q[1], q[2], q[3]).dist;
  float dist;
  f3 nn, wp, wq;
  edge_edge(q[0], q[1], q[2], q[3], dist, nn, wp, wq);
  return dist;
}

__device__ inline float additive_ccd(int type, const f3* x0, const f3* p) {
  const f3 pbar = mk(((p[0].x + p[1].x) + (p[2].x + p[3].x)) * 0.25f, ((p[0].y + p[1].y) + (p[2].y + p[3].y)) * 0.25f,
                     ((p[0].z + p[1].z) + (p[2].z + p[3].z)) * 0.25f);
  f3 q[4], x[4];
  for (int k = 0; k < 4; ++k) {
    q[k] = p[k] - pbar;
    x[k] = x0[k];
  }
  const float lp = type == 0 ? nrm(q[0]) + fmaxf(nrm(q[1]), fmaxf(nrm(q[2]), nrm(q[3])))
                             : fmaxf(nrm(q[0]), nrm(q[1])) + fmaxf(nrm(q[2]), nrm(q[3]));
  if (lp == 0.0f) return 1.0f;
  float dist = pair_distance(type, x);
  if (!(dist > 0.0f)) return 0.0f;
  const float g = kAccdS * dist;
  float t = 0.0f, tl = (1.0f - kAccdS) * dist / lp;
  for (int it = 0; it < kAccdMaxIter; ++it) {
    for (int k = 0; k < 4; ++k) x[k] = x[k] + tl * q[k];
    dist = pair_distance(type, x);
    if (t > 0.0f && dist < g) return t;
    t = t + tl;
    if (t > 1.0f) return 1.0f;
    tl = 0.9f * dist / lp;
  }
  return t;
}

// accd = 0: the reference's 10-sample test; 1: additive CCD over each trial's
// motion (the trial is feasible iff the whole motion is free).
__global__ void k_ls_self(int nc, const SelfContact* __restrict__ sc, const float* __restrict__ x,
                          const float* __restrict__ d, float ext, int accd, uint32_t* __restrict__ mask) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t m = 0xFFFFFu;
  if (c < nc) {
    const SelfContact& C = sc[c];
    f3 x0[4], ed[4];
    for (int k = 0; k < 4; ++k) {
      x0[k] = ld3(x, C.idx[k]);
      const f3 dk = ld3(d, C.idx[k]);
      ed[k] = mk(ext * dk.x, ext * dk.y, ext * dk.z);
    }
    float alpha = 1.0f;
    m = 0u;
    for (int tr = 0; tr < 20; ++tr) {
      f3 x1[4];
      bool still = true;
      for (int k = 0; k < 4; ++k) {
        x1[k] = x0[k] + alpha * ed[k];
        still = still && (double)sqn(x1[k] - x0[k]) < 1e-12;
      }
      bool ok = true;
      if (accd) {
        f3 pm[4];
        for (int k = 0; k < 4; ++k) pm[k] = x1[k] - x0[k];
        ok = still || additive_ccd(C.type, x0, pm) >= 1.0f;
      } else if (!still)
        for (int smp = 1; smp <= 10; ++smp) {
          const float tt = (float)smp / 10.0f;
          f3 q[4];
          for (int k = 0; k < 4; ++k) q[k] = x0[k] + tt * (x1[k] - x0[k]);
          float dist;
          if (C.type == 0) {
            dist = point_triangle(q[0], q[1], q[2], q[3]).dist;
          } else {
            f3 nn, wp, wq;
            edge_edge(q[0], q[1], q[2], q[3], dist, nn, wp, wq);
          }
          if ((double)dist < 1e-6) {
            ok = tt >= 1.0f;
            break;
          }
        }
      if (ok) m |= 1u << tr;
      alpha *= 0.5f;
    }
  }
  __shared__ uint32_t lds[4];
  const uint32_t r = block_reduce(m, lds, [](uint32_t a, uint32_t b) { return a & b; });
  if (threadIdx.x == 0 && r != 0xFFFFFu) atomicAnd(mask, r);
}

// ---------------------------------------------------------------------------
// Intersection validator (new; the reference's CollisionValidator counts
// gap < 0 on unsigned distances and never sees a penetration, F12).
// ---------------------------------------------------------------------------

// Sign of the volume of (a, b, c, d), in double on the float coordinates.
// 0 when rounding could have decided the sign: Shewchuk's static error bound
// for orient3d (o3derrboundA = (7 + 56 eps) eps, eps = 2^-53), so a
// near-coplanar configuration counts as degenerate, not as a crossing.
__device__ inline int orient3(f3 a, f3 b, f3 c, f3 d) {
  const double ax = (double)a.x - d.x, ay = (double)a.y - d.y, az = (double)a.z - d.z;
  const double bx = (double)b.x - d.x, by = (double)b.y - d.y, bz = (double)b.z - d.z;
  const double cx = (double)c.x - d.x, cy = (double)c.y - d.y, cz = (double)c.z - d.z;
  const double v = ax * (by * cz - bz * cy) - ay * (bx * cz - bz * cx) + az * (bx * cy - by * cx);
  const double perm = fabs(ax) * (fabs(by * cz) + fabs(bz * cy)) + fabs(ay) * (fabs(bx * cz) + fabs(bz * cx)) +
                      fabs(az) * (fabs(bx * cy) + fabs(by * cx));
  const double bound = 7.771561172376103e-16 * perm;
  return v > bound ? 1 : (v < -bound ? -1 : 0);
}

// Segment p q properly crosses triangle a b c: p and q strictly on opposite
// sides of the plane and the line through them strictly inside the three
// edges (degenerate touching contacts are not counted).
__device__ inline bool segment_crosses_triangle(f3 p, f3 q, f3 a, f3 b, f3 c) {
  const int sp = orient3(a, b, c, p), sq = orient3(a, b, c, q);
  if (sp == 0 || sq == 0 || sp == sq) return false;
  const int s1 = orient3(p, q, a, b), s2 = orient3(p, q, b, c), s3 = orient3(p, q, c, a);
  return s1 != 0 && s1 == s2 && s2 == s3;
}

// Count of (mesh edge, triangle) pairs sharing no vertex where the edge
// crosses the triangle; one thread per edge over the triangle BVH.
__global__ void __launch_bounds__(64) k_count_edge_tri(int E, const float* __restrict__ x, const int32_t* __restrict__ edges,
                                 const int32_t* __restrict__ tris, int P, const int32_t* __restrict__ left,
                                 const int32_t* __restrict__ right, const float* __restrict__ nodes,
                                 const float* __restrict__ leaves, const int32_t* __restrict__ order,
                                 int32_t* __restrict__ count, uint32_t* __restrict__ err) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  const int a0 = edges[2 * e], a1 = edges[2 * e + 1];
  const f3 p = ld3(x, a0), q = ld3(x, a1);
  Box bq;
  bq.lo[0] = fminf(p.x, q.x); bq.lo[1] = fminf(p.y, q.y); bq.lo[2] = fminf(p.z, q.z);
  bq.hi[0] = fmaxf(p.x, q.x); bq.hi[1] = fmaxf(p.y, q.y); bq.hi[2] = fmaxf(p.z, q.z);
  int k = 0;
  const bool ok = traverse(P, bq, left, right, nodes, leaves, [&](int s) {
    const int t = order[s];
    const int a = tris[3 * t], b = tris[3 * t + 1], c = tris[3 * t + 2];
    if (a == a0 || a == a1 || b == a0 || b == a1 || c == a0 || c == a1) return;
    if (segment_crosses_triangle(p, q, ld3(x, a), ld3(x, b), ld3(x, c))) ++k;
  });
  if (!ok) atomicOr(err, 1u);
  count[e] = k;
}

// Cloth vertices inside a closed rigid surface: generalized winding number
// (sum of signed solid angles / 4 pi, Van Oosterom-Strackee, in double) above
// 1/2.  One thread per vertex, all bodies' world triangles.
__global__ void k_count_inside(int n, const float* __restrict__ x, int nbod, DetectBodies D,
                               int32_t* __restrict__ count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const f3 p = ld3(x, i);
  int inside = 0;
  for (int b = 0; b < nbod; ++b) {
    const DetectBody& B = D.b[b];
    double w = 0.0;
    for (int t = 0; t < B.nt; ++t) {
      const int* T = B.tris + 3 * t;
      double a[3], bb[3], c[3];
      for (int k = 0; k < 3; ++k) {
        a[k] = (double)B.world[3 * T[0] + k] - comp(p, k);
        bb[k] = (double)B.world[3 * T[1] + k] - comp(p, k);
        c[k] = (double)B.world[3 * T[2] + k] - comp(p, k);
      }
      const double la = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
      const double lb = sqrt(bb[0] * bb[0] + bb[1] * bb[1] + bb[2] * bb[2]);
      const double lc = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
      const double det = a[0] * (bb[1] * c[2] - bb[2] * c[1]) - a[1] * (bb[0] * c[2] - bb[2] * c[0]) +
                         a[2] * (bb[0] * c[1] - bb[1] * c[0]);
      const double ab = a[0] * bb[0] + a[1] * bb[1] + a[2] * bb[2];
      const double bc = bb[0] * c[0] + bb[1] * c[1] + bb[2] * c[2];
      const double ca = c[0] * a[0] + c[1] * a[1] + c[2] * a[2];
      const double den = la * lb * lc + ab * lc + bc * la + ca * lb;
      w += 2.0 * atan2(det, den);
    }
    if (fabs(w) / (4.0 * 3.141592653589793) > 0.5) ++inside;
  }
  count[i] = inside;
}

}  // namespace lbvh
}  // namespace ando
