// Block-Jacobi PCG on the BSR system matrix (PCGSolver::solve,
// pcg_solver.cpp:8-146), FP32 vectors and matrix, FP64-accumulated dot
// products.
//
// Two kernels per iteration, no host round trip inside a chunk of iterations:
//   k_pcg_spmv(it):   reduce the previous iteration's r.z and max|r| partials,
//                     test convergence, form p = z + beta p on the fly for every
//                     column it touches (double-buffered p), Ap = A p with 8
//                     lanes per block row (one block each, summed by a shuffle
//                     tree), and partial p.Ap.
//   k_pcg_update(it): reduce p.Ap, alpha = r.z / p.Ap, x += alpha p,
//                     r -= alpha Ap, z = P^-1 r, partial r.z and max|r|.
// Every workgroup reduces the partial arrays itself in a fixed order, so all
// workgroups reach the same scalar and the same decision without atomics.
// A device-side done flag turns the remaining launches of a chunk into no-ops.

#include "kernels.h"

namespace ando {

constexpr int kLanesPerRow = 8;
constexpr int kRowsPerPass = kBlock / kLanesPerRow;  // 32 block rows per workgroup pass

struct PcgStatus {
  int32_t done;
  int32_t iters;
  int32_t failed;
  int32_t rel_it;  // the iteration whose convergence test last wrote rel (0: none yet)
  float bmax;
  float rel;       // max-norm relative residual: the verdict once done, else the latest test's
};

// Reductions of per-workgroup partials.  Each thread issues its kUnroll loads
// of a chunk back to back before summing them (a plain strided loop would wait
// for every load in turn), then the workgroup combines the thread sums; the
// order is fixed, so every workgroup obtains the same value.
constexpr int kUnroll = 4;

// The per-thread parts (loads and the thread's own sum) are separate so a
// kernel can issue them before it waits on anything else.
// Loads are unconditional (indices past the end read the last partial and
// are then dropped), so they issue back to back without branches; count >= 1.
__device__ inline double thread_sum_partials(const double* p, int count) {
  double s = 0.0;
  for (int base = 0; base < count; base += kUnroll * (int)blockDim.x) {
    double v[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int k = base + u * (int)blockDim.x + (int)threadIdx.x;
      v[u] = p[k < count ? k : count - 1];
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int k = base + u * (int)blockDim.x + (int)threadIdx.x;
      if (k >= count) v[u] = 0.0;
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) s += v[u];
  }
  return s;
}
__device__ inline float thread_max_partials(const float* p, int count) {
  float s = 0.0f;
  for (int base = 0; base < count; base += kUnroll * (int)blockDim.x) {
    float v[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int k = base + u * (int)blockDim.x + (int)threadIdx.x;
      v[u] = p[k < count ? k : count - 1];
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int k = base + u * (int)blockDim.x + (int)threadIdx.x;
      if (k >= count) v[u] = 0.0f;
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) s = fmaxf(s, v[u]);
  }
  return s;
}
// Split form of thread_sum_partials / thread_max_partials: the first
// kUnroll * blockDim partials are loaded into registers by prefetch_partials
// (issued before a kernel's other loads, since they head the longest chain),
// and summed in the same order by finish_*.
// (tid, nt): this thread's index among the nt threads that share the
// partials; the plain forms use the whole workgroup.
// The prefetch holds N partials per thread (N = kUnroll, or more where a
// kernel reads many); the sums add them in index order either way, so the
// result does not depend on N.
template <class T, int N>
__device__ inline void prefetch_partials(const T* p, int count, T (&v)[N], int tid, int nt) {
#pragma unroll
  for (int u = 0; u < N; ++u) {
    const int k = u * nt + tid;
    v[u] = p[k < count ? k : count - 1];
  }
}
template <class T, int N>
__device__ inline void prefetch_partials(const T* p, int count, T (&v)[N]) {
  prefetch_partials(p, count, v, (int)threadIdx.x, (int)blockDim.x);
}
template <int N>
__device__ inline double finish_sum_partials(const double* p, int count, const double (&v)[N], int tid, int nt) {
  double s = 0.0;
#pragma unroll
  for (int u = 0; u < N; ++u) s += (u * nt + tid) < count ? v[u] : 0.0;
  for (int base = N * nt; base < count; base += kUnroll * nt) {
    double w[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int k = base + u * nt + tid;
      w[u] = p[k < count ? k : count - 1];
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) s += (base + u * nt + tid) < count ? w[u] : 0.0;
  }
  return s;
}
template <int N>
__device__ inline double finish_sum_partials(const double* p, int count, const double (&v)[N]) {
  return finish_sum_partials(p, count, v, (int)threadIdx.x, (int)blockDim.x);
}
template <int N>
__device__ inline float finish_max_partials(const float* p, int count, const float (&v)[N], int tid, int nt) {
  float s = 0.0f;
#pragma unroll
  for (int u = 0; u < N; ++u) s = fmaxf(s, (u * nt + tid) < count ? v[u] : 0.0f);
  for (int base = N * nt; base < count; base += kUnroll * nt) {
    float w[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int k = base + u * nt + tid;
      w[u] = p[k < count ? k : count - 1];
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) s = fmaxf(s, (base + u * nt + tid) < count ? w[u] : 0.0f);
  }
  return s;
}
template <int N>
__device__ inline float finish_max_partials(const float* p, int count, const float (&v)[N]) {
  return finish_max_partials(p, count, v, (int)threadIdx.x, (int)blockDim.x);
}

__device__ inline double reduce_sum_partials(const double* p, int count, double* lds) {
  return block_reduce(thread_sum_partials(p, count), lds, [](double a, double b) { return a + b; });
}
__device__ inline float reduce_max_partials(const float* p, int count, float* lds) {
  return block_reduce(thread_max_partials(p, count), lds, [](float a, float b) { return fmaxf(a, b); });
}

// 3x3 determinant (Eigen's bruteforce formula) and cofactor inverse.
__device__ inline float det3(const float* B) {
  auto h = [&](int a, int b, int c) { return B[a] * (B[3 + b] * B[6 + c] - B[3 + c] * B[6 + b]); };
  return (h(0, 1, 2) - h(1, 0, 2)) + h(2, 0, 1);
}
__device__ inline float cof3(const float* B, int i, int j) {
  int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
  return B[3 * i1 + j1] * B[3 * i2 + j2] - B[3 * i1 + j2] * B[3 * i2 + j1];
}

// Preconditioner, r = b = -g, x = 0, z = P^-1 r, p = z (pcg_solver.cpp:13-34,77-135).
// Entry e of vertex i's preconditioner block sits at Pinv[e * pe + i * pi]
// (pe = 1, pi = 9: block per vertex; pe = n, pi = 1: one plane per entry, which
// the update kernel reads with coalesced loads).
__device__ __forceinline__ void pcg_init_body(int wg, int n, const float* __restrict__ A, const int32_t* __restrict__ diag,
                           const float* __restrict__ g, float* __restrict__ x, float* __restrict__ r,
                           float* __restrict__ z, float* __restrict__ p, float* __restrict__ Pinv, int pe, int pi,
                           double* __restrict__ part_rz, float* __restrict__ part_max, int negate) {
  int i = wg * blockDim.x + threadIdx.x;
  double rz = 0.0;
  float bm = 0.0f;
  if (i < n) {
    const float* B = A + 9 * (size_t)diag[i];
    float P[9];
    float det = det3(B);
    if ((double)fabsf(det) > 1e-16) {
      float c0 = cof3(B, 0, 0), c1 = cof3(B, 1, 0), c2 = cof3(B, 2, 0);
      float d = c0 * B[0] + (c1 * B[3] + c2 * B[6]);
      float inv = 1.0f / d;
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) P[3 * a + b] = cof3(B, b, a) * inv;
    } else {
      for (int e = 0; e < 9; ++e) P[e] = (e % 4 == 0) ? 1.0f : 0.0f;
    }
    for (int e = 0; e < 9; ++e) Pinv[(size_t)e * pe + (size_t)i * pi] = P[e];
    float rv[3];
    for (int k = 0; k < 3; ++k) {
      float bk = negate ? -g[3 * i + k] : g[3 * i + k];
      rv[k] = bk - 0.0f;
      r[3 * i + k] = rv[k];
      x[3 * i + k] = 0.0f;
      bm = fmaxf(bm, fabsf(bk));
    }
    for (int k = 0; k < 3; ++k) {
      float zk = P[3 * k] * rv[0] + (P[3 * k + 1] * rv[1] + P[3 * k + 2] * rv[2]);
      z[3 * i + k] = zk;
      p[3 * i + k] = zk;
      rz += (double)rv[k] * (double)zk;
    }
  }
  __shared__ double ldsd[4];
  __shared__ float ldsf[4];
  double s = block_reduce(rz, ldsd, [](double a, double b) { return a + b; });
  float mx = block_reduce(bm, ldsf, [](float a, float b) { return fmaxf(a, b); });
  if (threadIdx.x == 0) {
    part_rz[wg] = s;
    part_max[wg] = mx;
  }
}

__global__ void k_pcg_init(int n, const float* __restrict__ A, const int32_t* __restrict__ diag,
                           const float* __restrict__ g, float* __restrict__ x, float* __restrict__ r,
                           float* __restrict__ z, float* __restrict__ p, float* __restrict__ Pinv, int pe, int pi,
                           double* __restrict__ part_rz, float* __restrict__ part_max, int negate) {
  pcg_init_body(blockIdx.x, n, A, diag, g, x, r, z, p, Pinv, pe, pi, part_rz, part_max, negate);
}

// With a starting guess, part_bmax holds the max|b| partials and part_max the
// max|r0| ones (pcg_solver.cpp:16-21 tests the initial residual); without
// one (part_bmax == nullptr) r0 = b.
// Set-up with a given preconditioner P (9 floats per vertex) and starting
// guess x0: r = b - A x0, x = x0, z = P r, p = z, partial r.z, max|r| and
// max|b| (the learned preconditioner's warm start).
__global__ void k_pcg_init_given(int n, const float* __restrict__ A, const int32_t* __restrict__ rp,
                                 const int32_t* __restrict__ cols, const float* __restrict__ P,
                                 const float* __restrict__ g, int negate, const float* __restrict__ x0,
                                 float* __restrict__ x, float* __restrict__ r, float* __restrict__ z,
                                 float* __restrict__ p, double* __restrict__ part_rz, float* __restrict__ part_max,
                                 float* __restrict__ part_bmax, const float* __restrict__ Cx0) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double rz = 0.0;
  float rm = 0.0f, bm = 0.0f;
  if (i < n) {
    float ax[3] = {0.0f, 0.0f, 0.0f};
    if (x0)
      for (int b = rp[i]; b < rp[i + 1]; ++b) {
        const int j = cols[b];
        const float* B = A + 9 * (size_t)b;
        for (int k = 0; k < 3; ++k)
          ax[k] = ax[k] + (B[3 * k] * x0[3 * j] + (B[3 * k + 1] * x0[3 * j + 1] + B[3 * k + 2] * x0[3 * j + 2]));
      }
    if (x0 && Cx0)
      for (int k = 0; k < 3; ++k) ax[k] = ax[k] + Cx0[3 * i + k];
    float rv[3];
    for (int k = 0; k < 3; ++k) {
      const float bk = negate ? -g[3 * i + k] : g[3 * i + k];
      rv[k] = bk - ax[k];
      r[3 * i + k] = rv[k];
      x[3 * i + k] = x0 ? x0[3 * i + k] : 0.0f;
      rm = fmaxf(rm, fabsf(rv[k]));
      bm = fmaxf(bm, fabsf(bk));
    }
    const float* Pi = P + 9 * (size_t)i;
    for (int k = 0; k < 3; ++k) {
      float zk = Pi[3 * k] * rv[0] + (Pi[3 * k + 1] * rv[1] + Pi[3 * k + 2] * rv[2]);
      z[3 * i + k] = zk;
      p[3 * i + k] = zk;
      rz += (double)rv[k] * (double)zk;
    }
    if (rm != rm) rm = __int_as_float(0x7f800000);
  }
  __shared__ double ldsd[4];
  __shared__ float ldsf[4];
  double s = block_reduce(rz, ldsd, [](double a, double b) { return a + b; });
  float mx = block_reduce(rm, ldsf, [](float a, float b) { return fmaxf(a, b); });
  float mb = block_reduce(bm, ldsf, [](float a, float b) { return fmaxf(a, b); });
  if (threadIdx.x == 0) {
    part_rz[blockIdx.x] = s;
    part_max[blockIdx.x] = mx;
    part_bmax[blockIdx.x] = mb;
  }
}

__device__ __forceinline__ void pcg_start_body(int nwg, const double* __restrict__ part_rz,
                                               const float* __restrict__ part_max, const float* __restrict__ part_bmax,
                                               double* __restrict__ rz_hist, PcgStatus* __restrict__ st, float tol) {
  __shared__ double ldsd[4];
  __shared__ float ldsf[4];
  double rz = reduce_sum_partials(part_rz, nwg, ldsd);
  float rmax = reduce_max_partials(part_max, nwg, ldsf);
  float bmax = part_bmax ? reduce_max_partials(part_bmax, nwg, ldsf) : rmax;
  if (threadIdx.x == 0) {
    rz_hist[0] = rz;
    float rel = ((double)bmax < 1e-16) ? rmax : rmax / bmax;
    st->bmax = bmax;
    st->rel = rel;
    st->rel_it = 0;
    st->failed = 0;
    st->iters = 0;
    st->done = rel < tol ? 1 : 0;
  }
}
__global__ void k_pcg_start(int nwg, const double* __restrict__ part_rz, const float* __restrict__ part_max,
                            const float* __restrict__ part_bmax, double* __restrict__ rz_hist,
                            PcgStatus* __restrict__ st, float tol) {
  pcg_start_body(nwg, part_rz, part_max, part_bmax, rz_hist, st, tol);
}

// Workgroup numbering that gives each XCD a contiguous run of logical
// workgroups: hardware places workgroup w on XCD w % 8, and XCD x receives
// logical groups [x*q + min(x, rem), ...) in order, so neighbouring row or
// block ranges (which read the same z and p entries) share one L2.
__device__ inline int xcd_logical_group(int w, int G) {
  const int x = w % 8, q = G / 8, rem = G % 8;
  return x * q + min(x, rem) + w / 8;
}

// Self-contact coupling blocks (corrected modes).  They change with every
// Newton iteration and cluster on a few rows, so they are not part of the
// sliced layout: workgroups added to every SpMV grid (dispatched before the
// mesh workgroups) take the rows that have such blocks (crow, ncrow of them), 8
// lanes per row, and write the row's coupling product to Ap2 (zero on the
// other rows); k_pcg_update adds Ap + Ap2, as the row's mesh sum plus its
// coupling sum.  Row i's blocks are [3 cptr[i], 3 cptr[i + 1]) with columns
// ccol and 3x3 values cval; lane l sums blocks l, l + 8, ... in order and the
// lanes meet by an xor-shuffle tree.  Their p.Ap2 partials follow the mesh
// workgroups' partials.
struct Coupling {
  int ncrow;              // rows with coupling blocks (0: none)
  int nwg_mesh;           // mesh workgroups of the grid (their partials come first)
  int nwg_couple;         // coupling workgroups, dispatched first: blockIdx < nwg_couple
  const int32_t* crow;    // those rows
  const int32_t* cptr;
  const int32_t* ccol;
  const float* cval;
  float* Ap2;
};

__device__ __forceinline__ void couple_rows_body(int w, const Coupling& C, int n, int it, const float* __restrict__ z,
                                              const float* __restrict__ p_old, const double* __restrict__ part_rz,
                                              const float* __restrict__ part_max, int nwg_v,
                                              const double* __restrict__ rz_hist, double* __restrict__ part_pAp,
                                              PcgStatus* __restrict__ st, float tol) {
  __shared__ double ldsd[16];
  __shared__ float ldsf[16];
  const int lane = threadIdx.x & 7;
  const int k = w * (int)(blockDim.x >> 3) + (int)(threadIdx.x >> 3);
  const int i = k < C.ncrow ? C.crow[k] : -1;
  const int cb = i >= 0 ? 3 * C.cptr[i] : 0, ce = i >= 0 ? 3 * C.cptr[i + 1] : 0;
  if (st->done) return;
  float beta = 0.0f;
  if (it > 0) {
    const float bmax = st->bmax;
    const double rz_prev = rz_hist[it - 1];
    double rz = reduce_sum_partials(part_rz, nwg_v, ldsd);
    float rmax = reduce_max_partials(part_max, nwg_v, ldsf);
    float rel = ((double)bmax < 1e-16) ? rmax : rmax / bmax;
    if (rel < tol) return;  // the mesh workgroups record the verdict
    beta = (float)(rz / rz_prev);
  }
  float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
  // Four blocks per lane are requested before any is used (the sums keep
  // the block order).
  constexpr int kU = 4;
  for (int b0 = cb + lane; b0 < ce; b0 += 8 * kU) {
    int j[kU];
    float B[kU][9], pv[kU][3];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int b = b0 + 8 * u;
      if (b < ce) {
        j[u] = C.ccol[b];
#pragma unroll
        for (int e = 0; e < 9; ++e) B[u][e] = C.cval[9 * (size_t)b + e];
      }
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      if (b0 + 8 * u < ce) {
        if (it > 0) {
          pv[u][0] = z[3 * j[u]] + beta * p_old[3 * j[u]];
          pv[u][1] = z[3 * j[u] + 1] + beta * p_old[3 * j[u] + 1];
          pv[u][2] = z[3 * j[u] + 2] + beta * p_old[3 * j[u] + 2];
        } else {
          pv[u][0] = p_old[3 * j[u]]; pv[u][1] = p_old[3 * j[u] + 1]; pv[u][2] = p_old[3 * j[u] + 2];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      if (b0 + 8 * u < ce) {
        s0 = s0 + (B[u][0] * pv[u][0] + (B[u][1] * pv[u][1] + B[u][2] * pv[u][2]));
        s1 = s1 + (B[u][3] * pv[u][0] + (B[u][4] * pv[u][1] + B[u][5] * pv[u][2]));
        s2 = s2 + (B[u][6] * pv[u][0] + (B[u][7] * pv[u][1] + B[u][8] * pv[u][2]));
      }
    }
  }
  for (int off = 4; off > 0; off >>= 1) {
    s0 += __shfl_xor(s0, off, 8);
    s1 += __shfl_xor(s1, off, 8);
    s2 += __shfl_xor(s2, off, 8);
  }
  double pap = 0.0;
  if (i >= 0 && lane == 0) {
    float q0, q1, q2;
    if (it > 0) {
      q0 = z[3 * i] + beta * p_old[3 * i];
      q1 = z[3 * i + 1] + beta * p_old[3 * i + 1];
      q2 = z[3 * i + 2] + beta * p_old[3 * i + 2];
    } else {
      q0 = p_old[3 * i]; q1 = p_old[3 * i + 1]; q2 = p_old[3 * i + 2];
    }
    C.Ap2[3 * i] = s0;
    C.Ap2[3 * i + 1] = s1;
    C.Ap2[3 * i + 2] = s2;
    pap = (double)q0 * s0 + (double)q1 * s1 + (double)q2 * s2;
  }
  const double sp = block_reduce(pap, ldsd, [](double a, double b) { return a + b; });
  if (threadIdx.x == 0) part_pAp[C.nwg_mesh + w] = sp;
}

// out_i = (coupling blocks of row i) x for the coupling rows (out must be zero
// elsewhere): the starting residual of a warm-started solve.  8 lanes per row,
// 32 rows per 256-thread workgroup.
__global__ void __launch_bounds__(256) k_couple_mul(Coupling C, const float* __restrict__ x, float* __restrict__ out) {
  const int lane = threadIdx.x & 7;
  const int k = blockIdx.x * 32 + (int)(threadIdx.x >> 3);
  const int i = k < C.ncrow ? C.crow[k] : -1;
  const int cb = i >= 0 ? 3 * C.cptr[i] : 0, ce = i >= 0 ? 3 * C.cptr[i + 1] : 0;
  float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
  for (int b = cb + lane; b < ce; b += 8) {
    const int j = C.ccol[b];
    const float* B = C.cval + 9 * (size_t)b;
    s0 = s0 + (B[0] * x[3 * j] + (B[1] * x[3 * j + 1] + B[2] * x[3 * j + 2]));
    s1 = s1 + (B[3] * x[3 * j] + (B[4] * x[3 * j + 1] + B[5] * x[3 * j + 2]));
    s2 = s2 + (B[6] * x[3 * j] + (B[7] * x[3 * j + 1] + B[8] * x[3 * j + 2]));
  }
  for (int off = 4; off > 0; off >>= 1) {
    s0 += __shfl_xor(s0, off, 8);
    s1 += __shfl_xor(s1, off, 8);
    s2 += __shfl_xor(s2, off, 8);
  }
  if (i >= 0 && lane == 0) {
    out[3 * i] = s0;
    out[3 * i + 1] = s1;
    out[3 * i + 2] = s2;
  }
}

// One block row per group of kLanesPerRow lanes, one 3x3 block per lane per
// pass, so a row's blocks (contiguous in BSR order) are read as one run; the
// three row sums are combined by an xor-shuffle tree inside the lane group.
// Each lane group handles R rows (kRowsPerPass apart), so a workgroup covers
// R * kRowsPerPass rows; R is chosen so that the whole grid is resident at
// once (no second round of waves).  Row ranges follow xcd_logical_group.  Every
// load that does not depend on beta (the rows' first pass of blocks, and the
// z, p_old entries they touch) is issued before the reduction of the previous
// iteration's partials, so the memory latency chains overlap.  Each row holds
// its diagonal block, so row_ptr[i] is a valid block index for masked lanes.
template <int R>
__global__ void __launch_bounds__(kBlock) k_pcg_spmv(int n, int it, const float* __restrict__ A,
                           const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ cols,
                           const float* __restrict__ z, const float* __restrict__ p_old, float* __restrict__ p_new,
                           float* __restrict__ Ap, const double* __restrict__ part_rz,
                           const float* __restrict__ part_max, int nwg_v, double* __restrict__ rz_hist,
                           double* __restrict__ part_pAp, PcgStatus* __restrict__ st, float tol,
                           Coupling cpl, int spec) {
  // A launch the host queued past the predicted convergence (spec) reads the
  // done flag before anything else and returns at once once the solve is over.
  if (spec && st->done) return;
  constexpr int kRowsPerGroup = kRowsPerPass * R;
  if ((int)blockIdx.x < cpl.nwg_couple) {
    couple_rows_body((int)blockIdx.x, cpl, n, it, z, p_old, part_rz, part_max, nwg_v, rz_hist, part_pAp, st, tol);
    return;
  }
  __shared__ double ldsd[4];
  __shared__ float ldsf[4];
  const int lane = threadIdx.x % kLanesPerRow;
  const int g = xcd_logical_group((int)blockIdx.x - cpl.nwg_couple, cpl.nwg_mesh);
  int row[R], ri[R], beg[R], end[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    row[r] = g * kRowsPerGroup + r * kRowsPerPass + threadIdx.x / kLanesPerRow;
    ri[r] = row[r] < n ? row[r] : n - 1;
    beg[r] = row_ptr[ri[r]];  // in flight while the done flag is read
    end[r] = row_ptr[ri[r] + 1];
  }
  if (st->done) return;
  bool act[R];
  int bl[R];
  float Ab[R][9], po[R][3], zo[R][3], pi[R][3], zi[R][3];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    act[r] = row[r] < n && beg[r] + lane < end[r];
    bl[r] = act[r] ? beg[r] + lane : beg[r];
    const int j = cols[bl[r]];
#pragma unroll
    for (int e = 0; e < 9; ++e) Ab[r][e] = A[9 * (size_t)bl[r] + e];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      po[r][c] = p_old[3 * j + c];
      pi[r][c] = p_old[3 * ri[r] + c];
      zo[r][c] = it > 0 ? z[3 * j + c] : 0.0f;
      zi[r][c] = it > 0 ? z[3 * ri[r] + c] : 0.0f;
    }
  }
  float beta = 0.0f;
  if (it > 0) {
    const float bmax = st->bmax;
    const double rz_prev = rz_hist[it - 1];
    double rz = reduce_sum_partials(part_rz, nwg_v, ldsd);
    float rmax = reduce_max_partials(part_max, nwg_v, ldsf);
    float rel = ((double)bmax < 1e-16) ? rmax : rmax / bmax;
    if (rel < tol) {
      if (threadIdx.x == 0) {
        st->rel = rel;
        st->iters = it;
        st->done = 1;
      }
      return;
    }
    beta = (float)(rz / rz_prev);
    if (g == 0 && threadIdx.x == 0) {
      rz_hist[it] = rz;
      st->rel = rel;  // progress for the host's chunk sizing
      st->rel_it = it;
    }
  }
  double pap = 0.0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
    if (act[r]) {
      float pv0, pv1, pv2;
      if (it > 0) {
        pv0 = zo[r][0] + beta * po[r][0];
        pv1 = zo[r][1] + beta * po[r][1];
        pv2 = zo[r][2] + beta * po[r][2];
      } else {
        pv0 = po[r][0]; pv1 = po[r][1]; pv2 = po[r][2];
      }
      s0 = s0 + (Ab[r][0] * pv0 + (Ab[r][1] * pv1 + Ab[r][2] * pv2));
      s1 = s1 + (Ab[r][3] * pv0 + (Ab[r][4] * pv1 + Ab[r][5] * pv2));
      s2 = s2 + (Ab[r][6] * pv0 + (Ab[r][7] * pv1 + Ab[r][8] * pv2));
    }
    // Rows with more than kLanesPerRow blocks.
    for (int bb = beg[r] + lane + kLanesPerRow; row[r] < n && bb < end[r]; bb += kLanesPerRow) {
      const int j = cols[bb];
      float pv0, pv1, pv2;
      if (it > 0) {
        pv0 = z[3 * j] + beta * p_old[3 * j];
        pv1 = z[3 * j + 1] + beta * p_old[3 * j + 1];
        pv2 = z[3 * j + 2] + beta * p_old[3 * j + 2];
      } else {
        pv0 = p_old[3 * j];
        pv1 = p_old[3 * j + 1];
        pv2 = p_old[3 * j + 2];
      }
      const float* B = A + 9 * (size_t)bb;
      s0 = s0 + (B[0] * pv0 + (B[1] * pv1 + B[2] * pv2));
      s1 = s1 + (B[3] * pv0 + (B[4] * pv1 + B[5] * pv2));
      s2 = s2 + (B[6] * pv0 + (B[7] * pv1 + B[8] * pv2));
    }
    for (int off = kLanesPerRow / 2; off > 0; off >>= 1) {
      s0 += __shfl_xor(s0, off, kLanesPerRow);
      s1 += __shfl_xor(s1, off, kLanesPerRow);
      s2 += __shfl_xor(s2, off, kLanesPerRow);
    }
    const int i = row[r];
    if (i < n && lane == 0) {
      float q0, q1, q2;
      if (it > 0) {
        q0 = zi[r][0] + beta * pi[r][0];
        q1 = zi[r][1] + beta * pi[r][1];
        q2 = zi[r][2] + beta * pi[r][2];
        p_new[3 * i] = q0;
        p_new[3 * i + 1] = q1;
        p_new[3 * i + 2] = q2;
      } else {
        q0 = pi[r][0]; q1 = pi[r][1]; q2 = pi[r][2];
      }
      Ap[3 * i] = s0;
      Ap[3 * i + 1] = s1;
      Ap[3 * i + 2] = s2;
      pap += (double)q0 * s0 + (double)q1 * s1 + (double)q2 * s2;
    }
  }
  double s = block_reduce(pap, ldsd, [](double a, double b) { return a + b; });
  if (threadIdx.x == 0) part_pAp[g] = s;
}

// Sliced-ELL SpMV.  Rows are cut into slices of 64 (one wave, one row per
// lane); slice s holds W_s slots (its longest row), slot k of a row being its
// k-th block in BSR order (padding: a zero block on the row's own column).
// Values are stored slot-major and entry-major inside a slice,
//   vals[9 * off_s + (9 k + e) * 64 + lane],   cols[off_s + 64 k + lane],
// so every load of the kernel's first pass is 64 consecutive words.  A row's
// sum runs over its blocks in BSR order.  The slots of a chunk (kSellChunk)
// are all requested before any is used; rows longer than that loop.
constexpr int kSellChunk = 8;  // = the 8 lanes per row of k_pcg_spmv (summation order)

// Copies BSR values into the sliced layout (perm: BSR block of each slot
// position, -1 for padding); one wave per slice.
__device__ __forceinline__ void sell_fill_body(int wg, int slices, const int32_t* __restrict__ slice_off,
                                                     const int32_t* __restrict__ slice_w,
                                                     const int32_t* __restrict__ perm, const float* __restrict__ A,
                                                     float* __restrict__ vals) {
  const int sl = wg * (kBlock / 64) + (int)(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (sl >= slices) return;
  const int off = slice_off[sl], W = slice_w[sl];
  for (int k = 0; k < W; ++k) {
    const int b = perm[off + 64 * k + lane];
    float* dst = vals + 9 * (size_t)off + (size_t)(9 * k) * 64 + lane;
#pragma unroll
    for (int e = 0; e < 9; ++e) dst[64 * e] = b >= 0 ? A[9 * (size_t)b + e] : 0.0f;
  }
}

__global__ void __launch_bounds__(kBlock) k_sell_fill(int slices, const int32_t* __restrict__ slice_off,
                                                     const int32_t* __restrict__ slice_w,
                                                     const int32_t* __restrict__ perm, const float* __restrict__ A,
                                                     float* __restrict__ vals) {
  sell_fill_body(blockIdx.x, slices, slice_off, slice_w, perm, A, vals);
}

// Same contract as k_pcg_spmv (convergence test, beta, p = z + beta p on the
// fly, Ap, partial p.Ap per workgroup), over the sliced layout; a workgroup
// covers kBlock / 64 consecutive slices, numbered XCD-contiguously.
__global__ void __launch_bounds__(kBlock) k_pcg_spmv_sell(int n, int slices, int it, const float* __restrict__ vals,
                           const int32_t* __restrict__ scol, const int32_t* __restrict__ slice_off,
                           const int32_t* __restrict__ slice_w, int uniform_w, const float* __restrict__ z,
                           const float* __restrict__ p_old, float* __restrict__ p_new, float* __restrict__ Ap,
                           const double* __restrict__ part_rz, const float* __restrict__ part_max, int nwg_v,
                           double* __restrict__ rz_hist, double* __restrict__ part_pAp, PcgStatus* __restrict__ st,
                           float tol, Coupling cpl, int spec) {
  // A launch the host queued past the predicted convergence (spec) reads the
  // done flag before anything else and returns at once once the solve is over.
  if (spec && st->done) return;
  if ((int)blockIdx.x < cpl.nwg_couple) {
    couple_rows_body((int)blockIdx.x, cpl, n, it, z, p_old, part_rz, part_max, nwg_v, rz_hist, part_pAp, st, tol);
    return;
  }
  __shared__ double ldsd[4];
  __shared__ float ldsf[4];
  const int g = xcd_logical_group((int)blockIdx.x - cpl.nwg_couple, cpl.nwg_mesh);
  const int lane = threadIdx.x & 63;
  // The slice index is wave-uniform; saying so makes its metadata scalar loads
  // and the slot tests scalar branches.
  const int sl0 = __builtin_amdgcn_readfirstlane(g * (kBlock / 64) + (int)(threadIdx.x >> 6));
  const bool live = sl0 < slices;
  const int sl = live ? sl0 : slices - 1;
  const int row = sl0 * 64 + lane;
  const int ri = row < n ? row : n - 1;
  // A pattern whose slices all have the same width (uniform_w > 0) needs no
  // slice metadata: offsets follow from the slice index.
  const int off = uniform_w > 0 ? sl * 64 * uniform_w : slice_off[sl];
  const int W = !live ? 0 : (uniform_w > 0 ? uniform_w : slice_w[sl]);
  // The previous update's partials head the longest chain (reduction,
  // convergence test, beta): requested first.
  double pre_rz[kUnroll];
  float pre_max[kUnroll];
  if (it > 0) {
    prefetch_partials(part_rz, nwg_v, pre_rz);
    prefetch_partials(part_max, nwg_v, pre_max);
  }
  // First chunk of slots, the row's own z and p entries: none depend on beta.
  int jc[kSellChunk];
  float Ab[kSellChunk][9];
  const float* vbase = vals + 9 * (size_t)off + lane;
  const int32_t* cbase = scol + off + lane;
#pragma unroll
  for (int k = 0; k < kSellChunk; ++k) {
    if (k < W) {
      jc[k] = cbase[64 * k];
#pragma unroll
      for (int e = 0; e < 9; ++e) Ab[k][e] = vbase[(size_t)(9 * k + e) * 64];
    }
  }
  float zi[3], pi[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    zi[c] = z[3 * ri + c];
    pi[c] = p_old[3 * ri + c];
  }
  float zo[kSellChunk][3], po[kSellChunk][3];
#pragma unroll
  for (int k = 0; k < kSellChunk; ++k) {
    if (k < W) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        zo[k][c] = z[3 * jc[k] + c];
        po[k][c] = p_old[3 * jc[k] + c];
      }
    }
  }
  if (st->done) return;
  float beta = 0.0f;
  if (it > 0) {
    const float bmax = st->bmax;
    const double rz_prev = rz_hist[it - 1];
    double rz = finish_sum_partials(part_rz, nwg_v, pre_rz);
    float rmax = finish_max_partials(part_max, nwg_v, pre_max);
    block_reduce_sum_max(rz, rmax, ldsd, ldsf);
    float rel = ((double)bmax < 1e-16) ? rmax : rmax / bmax;
    if (rel < tol) {
      if (threadIdx.x == 0) {
        st->rel = rel;
        st->iters = it;
        st->done = 1;
      }
      return;
    }
    beta = (float)(rz / rz_prev);
    if (g == 0 && threadIdx.x == 0) {
      rz_hist[it] = rz;
      st->rel = rel;  // progress for the host's chunk sizing
      st->rel_it = it;
    }
  }
  // Slot k's product goes to partial t[k % 8]; the row sum combines the eight
  // partials as ((t0 + t4) + (t2 + t6)) + ((t1 + t5) + (t3 + t7)), the order in
  // which k_pcg_spmv's 8-lane shuffle tree adds its lanes (lane l holds blocks
  // l, l + 8, ... of the row), so both kernels produce the same Ap bit for bit.
  float t[kSellChunk][3];
#pragma unroll
  for (int k = 0; k < kSellChunk; ++k) {
    t[k][0] = t[k][1] = t[k][2] = 0.0f;
    if (k < W) {
      float pv0, pv1, pv2;
      if (it > 0) {
        pv0 = zo[k][0] + beta * po[k][0];
        pv1 = zo[k][1] + beta * po[k][1];
        pv2 = zo[k][2] + beta * po[k][2];
      } else {
        pv0 = po[k][0]; pv1 = po[k][1]; pv2 = po[k][2];
      }
      t[k][0] = Ab[k][0] * pv0 + (Ab[k][1] * pv1 + Ab[k][2] * pv2);
      t[k][1] = Ab[k][3] * pv0 + (Ab[k][4] * pv1 + Ab[k][5] * pv2);
      t[k][2] = Ab[k][6] * pv0 + (Ab[k][7] * pv1 + Ab[k][8] * pv2);
    }
  }
  // Rows longer than one chunk (slot k + u adds to partial u).
  for (int k0 = kSellChunk; k0 < W; k0 += kSellChunk) {
#pragma unroll
    for (int u = 0; u < kSellChunk; ++u) {
      const int k = k0 + u;
      if (k < W) {
        const int j = cbase[64 * k];
        float B[9];
#pragma unroll
        for (int e = 0; e < 9; ++e) B[e] = vbase[(size_t)(9 * k + e) * 64];
        float pv0, pv1, pv2;
        if (it > 0) {
          pv0 = z[3 * j] + beta * p_old[3 * j];
          pv1 = z[3 * j + 1] + beta * p_old[3 * j + 1];
          pv2 = z[3 * j + 2] + beta * p_old[3 * j + 2];
        } else {
          pv0 = p_old[3 * j]; pv1 = p_old[3 * j + 1]; pv2 = p_old[3 * j + 2];
        }
        t[u][0] = t[u][0] + (B[0] * pv0 + (B[1] * pv1 + B[2] * pv2));
        t[u][1] = t[u][1] + (B[3] * pv0 + (B[4] * pv1 + B[5] * pv2));
        t[u][2] = t[u][2] + (B[6] * pv0 + (B[7] * pv1 + B[8] * pv2));
      }
    }
  }
  float sr[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) sr[c] = ((t[0][c] + t[4][c]) + (t[2][c] + t[6][c])) + ((t[1][c] + t[5][c]) + (t[3][c] + t[7][c]));
  const float s0 = sr[0], s1 = sr[1], s2 = sr[2];
  double pap = 0.0;
  if (live && row < n) {
    float q0, q1, q2;
    if (it > 0) {
      q0 = zi[0] + beta * pi[0];
      q1 = zi[1] + beta * pi[1];
      q2 = zi[2] + beta * pi[2];
      p_new[3 * row] = q0;
      p_new[3 * row + 1] = q1;
      p_new[3 * row + 2] = q2;
    } else {
      q0 = pi[0]; q1 = pi[1]; q2 = pi[2];
    }
    Ap[3 * row] = s0;
    Ap[3 * row + 1] = s1;
    Ap[3 * row + 2] = s2;
    pap = (double)q0 * s0 + (double)q1 * s1 + (double)q2 * s2;
  }
  double sp = block_reduce(pap, ldsd, [](double a, double b) { return a + b; });
  if (threadIdx.x == 0) part_pAp[g] = sp;
}

// k_pcg_spmv_sell with L lanes per row (L = 2 or 4), for more waves in flight
// per slot; same contract and the same results bit for bit.  A workgroup of
// 256 L threads still covers four slices (L waves per slice, 64 / L rows per
// wave), so the partial counts are unchanged.  Lane group q of a row takes the
// slots k with k % L == q; their products go to the partials t[k % 8] as in
// the one-lane kernel, and the groups' sums meet in that kernel's tree order
// ((t0 + t4) + (t2 + t6)) + ((t1 + t5) + (t3 + t7)) through shuffles.  The
// r.z / max|r| reduction runs on the first 256 threads (the rest add exact
// zeros) and the p.Ap partial is reduced from the one-lane layout restored in
// LDS, so both associate as in the one-lane kernel.
// S (slices per workgroup, default 4) below 4 gives a finer grid, S = 0 one
// wave per workgroup (64 / L rows of a slice); the r.z / max|r| and p.Ap
// reductions then run over the workgroup's threads, another association than
// the one-lane kernel's.
template <int L, int S>
__device__ __forceinline__ void sell_split_body(int g, int n, int slices, int it,
                           const float* __restrict__ vals, const int32_t* __restrict__ scol,
                           const int32_t* __restrict__ slice_off, const int32_t* __restrict__ slice_w, int uniform_w,
                           const float* __restrict__ z, const float* __restrict__ p_old, float* __restrict__ p_new,
                           float* __restrict__ Ap, const double* __restrict__ part_rz,
                           const float* __restrict__ part_max, int nwg_v, double* __restrict__ rz_hist,
                           double* __restrict__ part_pAp, PcgStatus* __restrict__ st, float tol) {
  static_assert(L == 2 || L == 4, "two or four lanes per row");
  static_assert(S == 0 || S == 1 || S == 2 || S == 4, "a wave, or one, two or four slices per workgroup");
  constexpr int R = 64 / L;             // rows per wave
  constexpr int KC = kSellChunk / L;    // slots per lane per chunk
  constexpr int WPW = S == 0 ? 1 : S * L;         // waves per workgroup
  constexpr int RT = 64 * WPW < 256 ? 64 * WPW : 256;  // threads that reduce the partials
  __shared__ double ldsd[WPW];
  __shared__ float ldsf[WPW];
  __shared__ double ldsp[S == 0 ? 1 : S][64];
  const int tid = threadIdx.x;
  const int wv = tid >> 6, lane = tid & 63;
  const int q = lane / R, rl = lane % R;
  const int wslice = S == 0 ? 0 : wv / L, sub = S == 0 ? g % L : wv % L;
  const int sl0 = __builtin_amdgcn_readfirstlane(S == 0 ? g / L : g * S + wslice);
  const bool live = sl0 < slices;
  const int sl = live ? sl0 : slices - 1;
  const int rs = sub * R + rl;  // row within the slice
  const int row = sl0 * 64 + rs;
  const int ri = row < n ? row : n - 1;
  const int off = uniform_w > 0 ? sl * 64 * uniform_w : slice_off[sl];
  const int W = !live ? 0 : (uniform_w > 0 ? uniform_w : slice_w[sl]);
  const bool red = tid < RT;
  double pre_rz[kUnroll];
  float pre_max[kUnroll];
  if (it > 0 && red) {
    prefetch_partials(part_rz, nwg_v, pre_rz, tid, RT);
    prefetch_partials(part_max, nwg_v, pre_max, tid, RT);
  }
  int jc[KC];
  float Ab[KC][9];
  const float* vbase = vals + 9 * (size_t)off + rs;
  const int32_t* cbase = scol + off + rs;
#pragma unroll
  for (int m = 0; m < KC; ++m) {
    const int k = L * m + q;
    if (k < W) {
      jc[m] = cbase[64 * k];
#pragma unroll
      for (int e = 0; e < 9; ++e) Ab[m][e] = vbase[(size_t)(9 * k + e) * 64];
    }
  }
  float zi[3], pi[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    zi[c] = z[3 * ri + c];
    pi[c] = p_old[3 * ri + c];
  }
  float zo[KC][3], po[KC][3];
#pragma unroll
  for (int m = 0; m < KC; ++m) {
    if (L * m + q < W) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        zo[m][c] = z[3 * jc[m] + c];
        po[m][c] = p_old[3 * jc[m] + c];
      }
    }
  }
  if (st->done) return;
  float beta = 0.0f;
  if (it > 0) {
    const float bmax = st->bmax;
    const double rz_prev = rz_hist[it - 1];
    double rz = red ? finish_sum_partials(part_rz, nwg_v, pre_rz, tid, RT) : 0.0;
    float rmax = red ? finish_max_partials(part_max, nwg_v, pre_max, tid, RT) : 0.0f;
    block_reduce_sum_max(rz, rmax, ldsd, ldsf);
    float rel = ((double)bmax < 1e-16) ? rmax : rmax / bmax;
    if (rel < tol) {
      if (tid == 0) {
        st->rel = rel;
        st->iters = it;
        st->done = 1;
      }
      return;
    }
    beta = (float)(rz / rz_prev);
    if (g == 0 && tid == 0) {
      rz_hist[it] = rz;
      st->rel = rel;  // progress for the host's chunk sizing
      st->rel_it = it;
    }
  }
  // T[m] is partial t[L m + q] of the one-lane kernel.
  float T[KC][3];
#pragma unroll
  for (int m = 0; m < KC; ++m) {
    T[m][0] = T[m][1] = T[m][2] = 0.0f;
    if (L * m + q < W) {
      float pv0, pv1, pv2;
      if (it > 0) {
        pv0 = zo[m][0] + beta * po[m][0];
        pv1 = zo[m][1] + beta * po[m][1];
        pv2 = zo[m][2] + beta * po[m][2];
      } else {
        pv0 = po[m][0]; pv1 = po[m][1]; pv2 = po[m][2];
      }
      T[m][0] = Ab[m][0] * pv0 + (Ab[m][1] * pv1 + Ab[m][2] * pv2);
      T[m][1] = Ab[m][3] * pv0 + (Ab[m][4] * pv1 + Ab[m][5] * pv2);
      T[m][2] = Ab[m][6] * pv0 + (Ab[m][7] * pv1 + Ab[m][8] * pv2);
    }
  }
  for (int k0 = kSellChunk; k0 < W; k0 += kSellChunk) {
#pragma unroll
    for (int m = 0; m < KC; ++m) {
      const int k = k0 + L * m + q;
      if (k < W) {
        const int j = cbase[64 * k];
        float B[9];
#pragma unroll
        for (int e = 0; e < 9; ++e) B[e] = vbase[(size_t)(9 * k + e) * 64];
        float pv0, pv1, pv2;
        if (it > 0) {
          pv0 = z[3 * j] + beta * p_old[3 * j];
          pv1 = z[3 * j + 1] + beta * p_old[3 * j + 1];
          pv2 = z[3 * j + 2] + beta * p_old[3 * j + 2];
        } else {
          pv0 = p_old[3 * j]; pv1 = p_old[3 * j + 1]; pv2 = p_old[3 * j + 2];
        }
        T[m][0] = T[m][0] + (B[0] * pv0 + (B[1] * pv1 + B[2] * pv2));
        T[m][1] = T[m][1] + (B[3] * pv0 + (B[4] * pv1 + B[5] * pv2));
        T[m][2] = T[m][2] + (B[6] * pv0 + (B[7] * pv1 + B[8] * pv2));
      }
    }
  }
  float sr[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float h;
    if constexpr (L == 2) {
      // q = 0 holds t0, t2, t4, t6 and q = 1 holds t1, t3, t5, t7.
      h = (T[0][c] + T[2][c]) + (T[1][c] + T[3][c]);
      sr[c] = h + __shfl_down(h, 32, 64);
    } else {
      // q holds t[q] and t[q + 4]; pairs (q0 + q2) and (q1 + q3), then their sum.
      h = T[0][c] + T[1][c];
      h = h + __shfl_down(h, 32, 64);
      sr[c] = h + __shfl_down(h, 16, 64);
    }
  }

  double pap = 0.0;
  if (q == 0 && live && row < n) {
    float q0, q1, q2;
    if (it > 0) {
      q0 = zi[0] + beta * pi[0];
      q1 = zi[1] + beta * pi[1];
      q2 = zi[2] + beta * pi[2];
      p_new[3 * row] = q0;
      p_new[3 * row + 1] = q1;
      p_new[3 * row + 2] = q2;
    } else {
      q0 = pi[0]; q1 = pi[1]; q2 = pi[2];
    }
    Ap[3 * row] = sr[0];
    Ap[3 * row + 1] = sr[1];
    Ap[3 * row + 2] = sr[2];
    pap = (double)q0 * sr[0] + (double)q1 * sr[1] + (double)q2 * sr[2];
  }
  double pv = pap;
  if constexpr (S != 0) {
    if (q == 0) ldsp[wslice][rs] = pap;
    __syncthreads();
    pv = tid < 64 * S ? ldsp[wv][lane] : 0.0;
  }
  double sp = block_reduce(pv, ldsd, [](double a, double b) { return a + b; });
  if (tid == 0) part_pAp[g] = sp;
}

template <int L, int S = 4>
__global__ void __launch_bounds__(64 * (S == 0 ? 1 : S * L)) k_pcg_spmv_sell_split(int n, int slices, int it,
                           const float* __restrict__ vals, const int32_t* __restrict__ scol,
                           const int32_t* __restrict__ slice_off, const int32_t* __restrict__ slice_w, int uniform_w,
                           const float* __restrict__ z, const float* __restrict__ p_old, float* __restrict__ p_new,
                           float* __restrict__ Ap, const double* __restrict__ part_rz,
                           const float* __restrict__ part_max, int nwg_v, double* __restrict__ rz_hist,
                           double* __restrict__ part_pAp, PcgStatus* __restrict__ st, float tol,
                           Coupling cpl, int spec) {
  // A launch the host queued past the predicted convergence (spec) reads the
  // done flag before anything else and returns at once once the solve is over.
  if (spec && st->done) return;
  if ((int)blockIdx.x < cpl.nwg_couple) {
    couple_rows_body((int)blockIdx.x, cpl, n, it, z, p_old, part_rz, part_max, nwg_v, rz_hist, part_pAp, st, tol);
    return;
  }
  sell_split_body<L, S>(xcd_logical_group((int)blockIdx.x - cpl.nwg_couple, cpl.nwg_mesh), n, slices, it, vals, scol, slice_off, slice_w,
                        uniform_w, z, p_old, p_new, Ap, part_rz, part_max, nwg_v, rz_hist, part_pAp, st, tol);
}

// alpha = r.z / p.Ap, x += alpha p, r -= alpha Ap, z = P^-1 r; one thread per
// vertex.  The vertex's operands are loaded before the p.Ap reduction.
__device__ __forceinline__ void update_body(int wg, int n, int it, const float* __restrict__ p,
                             const float* __restrict__ Ap, const float* __restrict__ Ap2, const float* __restrict__ Pinv, int pe,
                             int pi, float* __restrict__ x, float* __restrict__ r, float* __restrict__ z,
                             const double* __restrict__ part_pAp, int nwg_r, const double* __restrict__ rz_hist,
                             double* __restrict__ part_rz, float* __restrict__ part_max, PcgStatus* __restrict__ st) {
  __shared__ double ldsd[4];
  __shared__ float ldsf[4];
  const int i = wg * blockDim.x + threadIdx.x;
  const int ii = i < n ? i : n - 1;
  float pv[3], av[3], xv[3], rv0[3], P[9];
  // The p.Ap partials head the longest dependency chain (reduction, alpha,
  // update): they are requested first, then every other load that does not
  // depend on alpha.
  // Eight per thread: with one SpMV workgroup per slice there are about
  // n / 64 partials, which then take one pass.
  double pre[2 * kUnroll];
  prefetch_partials(part_pAp, nwg_r, pre);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    pv[k] = p[3 * ii + k];
    av[k] = Ap2 ? Ap[3 * ii + k] + Ap2[3 * ii + k] : Ap[3 * ii + k];
    xv[k] = x[3 * ii + k];
    rv0[k] = r[3 * ii + k];
  }
#pragma unroll
  for (int e = 0; e < 9; ++e) P[e] = Pinv[(size_t)e * pe + (size_t)ii * pi];
  if (st->done) return;
  const double pAp_t = finish_sum_partials(part_pAp, nwg_r, pre);
  const double rz_it = rz_hist[it];
  const double pAp = block_reduce(pAp_t, ldsd, [](double a, double b) { return a + b; });
  if (fabs(pAp) < 1e-16) {
    if (threadIdx.x == 0) {
      st->failed = 1;
      st->iters = it + 1;
      st->done = 1;
    }
    return;
  }
  float alpha = (float)(rz_it / pAp);
  double rz = 0.0;
  float rm = 0.0f;
  if (i < n) {
    float rv[3];
    for (int k = 0; k < 3; ++k) {
      int e = 3 * i + k;
      x[e] = xv[k] + alpha * pv[k];
      rv[k] = rv0[k] - alpha * av[k];
      r[e] = rv[k];
      rm = fmaxf(rm, fabsf(rv[k]));
    }
    for (int k = 0; k < 3; ++k) {
      float zk = P[3 * k] * rv[0] + (P[3 * k + 1] * rv[1] + P[3 * k + 2] * rv[2]);
      z[3 * i + k] = zk;
      rz += (double)rv[k] * (double)zk;
    }
    if (rm != rm) rm = __int_as_float(0x7f800000);
  }
  double s = rz;
  float mx = rm;
  block_reduce_sum_max(s, mx, ldsd, ldsf);
  if (threadIdx.x == 0) {
    part_rz[wg] = s;
    part_max[wg] = mx;
  }
}

__global__ void __launch_bounds__(kBlock) k_pcg_update(int n, int it, const float* __restrict__ p,
                             const float* __restrict__ Ap, const float* __restrict__ Ap2, const float* __restrict__ Pinv, int pe,
                             int pi, float* __restrict__ x, float* __restrict__ r, float* __restrict__ z,
                             const double* __restrict__ part_pAp, int nwg_r, const double* __restrict__ rz_hist,
                             double* __restrict__ part_rz, float* __restrict__ part_max, PcgStatus* __restrict__ st,
                             int spec) {
  if (spec && st->done) return;  // queued past the predicted convergence: flag first
  update_body(blockIdx.x, n, it, p, Ap, Ap2, Pinv, pe, pi, x, r, z, part_pAp, nwg_r, rz_hist, part_rz, part_max, st);
}

// ---------------------------------------------------------------------------
// Batched PCG (ando_step_batch): the iterations of many independent systems in
// one SpMV launch and one update launch.  Workgroup w of a batched grid runs
// workgroup w - wg0 of system wg_sys[w] with that system's buffers, the same
// row range and partial slot as in the system's own launch, so every system's
// results equal those of solving it alone, bit for bit.  All systems of a
// batch start together, so the iteration index is common; a converged system's
// workgroups return at its done flag.
// ---------------------------------------------------------------------------
struct PcgSys {
  int n, slices, uniform_w, nwg_v, nwg_r;
  int wg0_r, wg0_v, wg0_f;  // first workgroup of the system in the SpMV, vector and fill grids
  const float* A;           // the system's BSR values
  const int32_t* diag;      // BSR index of each diagonal block
  const float* rhs;         // b = -rhs
  const int32_t* perm;      // BSR block of each sliced slot
  float* vals;
  const int32_t* scol;
  const int32_t* slice_off;
  const int32_t* slice_w;
  float* Pinv;
  int pe, pi;
  float* z;
  float* p0;
  float* p1;
  float* Ap;
  float* x;
  float* r;
  double* part_rz;
  float* part_max;
  double* rz_hist;
  double* part_pAp;
  PcgStatus* st;
  float tol;
};

__global__ void __launch_bounds__(128) k_pcg_spmv_sell_batch(const PcgSys* __restrict__ sys,
                                                            const int32_t* __restrict__ wg_sys, int it) {
  const PcgSys& S = sys[wg_sys[blockIdx.x]];
  float* pold = it == 0 ? S.p0 : (((it - 1) & 1) ? S.p1 : S.p0);
  float* pcur = (it & 1) ? S.p1 : S.p0;
  sell_split_body<2, 1>((int)blockIdx.x - S.wg0_r, S.n, S.slices, it, S.vals, S.scol, S.slice_off, S.slice_w,
                        S.uniform_w, S.z, pold, pcur, S.Ap, S.part_rz, S.part_max, S.nwg_v, S.rz_hist, S.part_pAp,
                        S.st, S.tol);
}

__global__ void __launch_bounds__(kBlock) k_pcg_update_batch(const PcgSys* __restrict__ sys,
                                                            const int32_t* __restrict__ wg_sys, int it) {
  const PcgSys& S = sys[wg_sys[blockIdx.x]];
  float* pcur = (it & 1) ? S.p1 : S.p0;
  update_body((int)blockIdx.x - S.wg0_v, S.n, it, pcur, S.Ap, nullptr, S.Pinv, S.pe, S.pi, S.x, S.r, S.z,
              S.part_pAp, S.nwg_r, S.rz_hist, S.part_rz, S.part_max, S.st);
}

// Set-up of every system of a batch: sliced-ELL values, then x0 = 0, r = b,
// z = p = P^-1 r with the block-Jacobi inverse and the partials, then the
// initial residual test (one workgroup per system).
__global__ void __launch_bounds__(kBlock) k_sell_fill_batch(const PcgSys* __restrict__ sys,
                                                           const int32_t* __restrict__ wg_sys) {
  const PcgSys& S = sys[wg_sys[blockIdx.x]];
  sell_fill_body((int)blockIdx.x - S.wg0_f, S.slices, S.slice_off, S.slice_w, S.perm, S.A, S.vals);
}
__global__ void __launch_bounds__(kBlock) k_pcg_init_batch(const PcgSys* __restrict__ sys,
                                                          const int32_t* __restrict__ wg_sys) {
  const PcgSys& S = sys[wg_sys[blockIdx.x]];
  pcg_init_body((int)blockIdx.x - S.wg0_v, S.n, S.A, S.diag, S.rhs, S.x, S.r, S.z, S.p0, S.Pinv, S.pe, S.pi,
                S.part_rz, S.part_max, 1);
}
__global__ void __launch_bounds__(kBlock) k_pcg_start_batch(const PcgSys* __restrict__ sys) {
  const PcgSys& S = sys[blockIdx.x];
  pcg_start_body(S.nwg_v, S.part_rz, S.part_max, nullptr, S.rz_hist, S.st, S.tol);
}

// The systems' status words side by side, for one read-back per chunk.
__global__ void k_pcg_status_gather(int nsys, const PcgSys* __restrict__ sys, PcgStatus* __restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nsys) out[s] = *sys[s].st;
}

// Convergence test after the last permitted iteration (pcg_solver.cpp:55-74).
__global__ void k_pcg_final(int it, const float* __restrict__ part_max, int nwg_v, PcgStatus* __restrict__ st,
                            float tol) {
  if (st->done) return;
  __shared__ float ldsf[4];
  float rmax = reduce_max_partials(part_max, nwg_v, ldsf);
  if (threadIdx.x == 0) {
    float bmax = st->bmax;
    float rel = ((double)bmax < 1e-16) ? rmax : rmax / bmax;
    st->rel = rel;
    st->iters = it;
    st->failed = rel < tol ? 0 : 1;
    st->done = 1;
  }
}

}  // namespace ando
