// Block-Jacobi PCG on the BSR system matrix (PCGSolver::solve,
// pcg_solver.cpp:8-146), FP32 vectors and matrix, FP64-accumulated dot
// products.
//
// Two kernels per iteration, no host round trip inside a chunk of iterations:
//   k_pcg_spmv(it):   reduce the previous iteration's r.z and max|r| partials,
//                     test convergence, form p = z + beta p on the fly for every
//                     column it touches (double-buffered p), Ap = A p with each
//                     row summed left to right in column order (the order of
//                     Eigen's row-major SpMV), and partial p.Ap.
//   k_pcg_update(it): reduce p.Ap, alpha = r.z / p.Ap, x += alpha p,
//                     r -= alpha Ap, z = P^-1 r, partial r.z and max|r|.
// Every workgroup reduces the partial arrays itself in a fixed order, so all
// workgroups reach the same scalar and the same decision without atomics.
// A device-side done flag turns the remaining launches of a chunk into no-ops.
#include <hip/hip_cooperative_groups.h>

#include "kernels.h"

namespace ando {

struct PcgStatus {
  int32_t done;
  int32_t iters;
  int32_t failed;
  int32_t pad;
  float bmax;
  float rel;
};

__device__ inline double reduce_sum_partials(const double* p, int count, double* lds) {
  double s = 0.0;
  for (int k = threadIdx.x; k < count; k += blockDim.x) s += p[k];
  return block_reduce(s, lds, [](double a, double b) { return a + b; });
}
__device__ inline float reduce_max_partials(const float* p, int count, float* lds) {
  float s = 0.0f;
  for (int k = threadIdx.x; k < count; k += blockDim.x) s = fmaxf(s, p[k]);
  return block_reduce(s, lds, [](float a, float b) { return fmaxf(a, b); });
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
__global__ void k_pcg_init(int n, const float* __restrict__ A, const int32_t* __restrict__ diag,
                           const float* __restrict__ g, float* __restrict__ x, float* __restrict__ r,
                           float* __restrict__ z, float* __restrict__ p, float* __restrict__ Pinv,
                           double* __restrict__ part_rz, float* __restrict__ part_max, int negate) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
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
    for (int e = 0; e < 9; ++e) Pinv[9 * (size_t)i + e] = P[e];
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
    part_rz[blockIdx.x] = s;
    part_max[blockIdx.x] = mx;
  }
}

__global__ void k_pcg_start(int nwg, const double* __restrict__ part_rz, const float* __restrict__ part_max,
                            double* __restrict__ rz_hist, PcgStatus* __restrict__ st, float tol) {
  __shared__ double ldsd[4];
  __shared__ float ldsf[4];
  double rz = reduce_sum_partials(part_rz, nwg, ldsd);
  float bmax = reduce_max_partials(part_max, nwg, ldsf);
  if (threadIdx.x == 0) {
    rz_hist[0] = rz;
    float rel = ((double)bmax < 1e-16) ? bmax : bmax / bmax;
    st->bmax = bmax;
    st->rel = rel;
    st->failed = 0;
    st->iters = 0;
    st->done = rel < tol ? 1 : 0;
  }
}

__global__ void k_pcg_spmv(int n, int it, const float* __restrict__ A, const int32_t* __restrict__ row_ptr,
                           const int32_t* __restrict__ cols, const float* __restrict__ z,
                           const float* __restrict__ p_old, float* __restrict__ p_new, float* __restrict__ Ap,
                           const double* __restrict__ part_rz, const float* __restrict__ part_max, int nwg_v,
                           double* __restrict__ rz_hist, double* __restrict__ part_pAp, PcgStatus* __restrict__ st,
                           float tol) {
  if (st->done) return;
  __shared__ double ldsd[4];
  __shared__ float ldsf[4];
  float beta = 0.0f;
  if (it > 0) {
    double rz = reduce_sum_partials(part_rz, nwg_v, ldsd);
    float rmax = reduce_max_partials(part_max, nwg_v, ldsf);
    float bmax = st->bmax;
    float rel = ((double)bmax < 1e-16) ? rmax : rmax / bmax;
    if (rel < tol) {
      if (threadIdx.x == 0) {
        st->rel = rel;
        st->iters = it;
        st->done = 1;
      }
      return;
    }
    beta = (float)(rz / rz_hist[it - 1]);
    if (blockIdx.x == 0 && threadIdx.x == 0) rz_hist[it] = rz;
  }
  int row = blockIdx.x * blockDim.x + threadIdx.x;
  double pap = 0.0;
  if (row < 3 * n) {
    int i = row / 3, rr = row - 3 * i;
    float tmp = 0.0f;
    for (int bb = row_ptr[i]; bb < row_ptr[i + 1]; ++bb) {
      int j = cols[bb];
      const float* Ab = A + 9 * (size_t)bb + 3 * rr;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        int col = 3 * j + c;
        float pv = it > 0 ? z[col] + beta * p_old[col] : p_old[col];
        tmp = tmp + Ab[c] * pv;
      }
    }
    float pr = it > 0 ? z[row] + beta * p_old[row] : p_old[row];
    if (it > 0) p_new[row] = pr;
    Ap[row] = tmp;
    pap = (double)pr * (double)tmp;
  }
  double s = block_reduce(pap, ldsd, [](double a, double b) { return a + b; });
  if (threadIdx.x == 0) part_pAp[blockIdx.x] = s;
}

__global__ void k_pcg_update(int n, int it, const float* __restrict__ p, const float* __restrict__ Ap,
                             const float* __restrict__ Pinv, float* __restrict__ x, float* __restrict__ r,
                             float* __restrict__ z, const double* __restrict__ part_pAp, int nwg_r,
                             const double* __restrict__ rz_hist, double* __restrict__ part_rz,
                             float* __restrict__ part_max, PcgStatus* __restrict__ st) {
  if (st->done) return;
  __shared__ double ldsd[4];
  __shared__ float ldsf[4];
  double pAp = reduce_sum_partials(part_pAp, nwg_r, ldsd);
  if (fabs(pAp) < 1e-16) {
    if (threadIdx.x == 0) {
      st->failed = 1;
      st->iters = it + 1;
      st->done = 1;
    }
    return;
  }
  float alpha = (float)(rz_hist[it] / pAp);
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double rz = 0.0;
  float rm = 0.0f;
  if (i < n) {
    float rv[3];
    for (int k = 0; k < 3; ++k) {
      int e = 3 * i + k;
      x[e] = x[e] + alpha * p[e];
      rv[k] = r[e] - alpha * Ap[e];
      r[e] = rv[k];
      rm = fmaxf(rm, fabsf(rv[k]));
    }
    const float* P = Pinv + 9 * (size_t)i;
    for (int k = 0; k < 3; ++k) {
      float zk = P[3 * k] * rv[0] + (P[3 * k + 1] * rv[1] + P[3 * k + 2] * rv[2]);
      z[3 * i + k] = zk;
      rz += (double)rv[k] * (double)zk;
    }
    if (rm != rm) rm = __int_as_float(0x7f800000);
  }
  double s = block_reduce(rz, ldsd, [](double a, double b) { return a + b; });
  float mx = block_reduce(rm, ldsf, [](float a, float b) { return fmaxf(a, b); });
  if (threadIdx.x == 0) {
    part_rz[blockIdx.x] = s;
    part_max[blockIdx.x] = mx;
  }
}

// ---------------------------------------------------------------------------
// Persistent PCG: the whole solve in one cooperative launch.
//
// Every workgroup owns a contiguous range of block rows (and the same range of
// vertices in the vector update), numbered so that consecutive ranges sit on
// the same XCD: hardware dispatches workgroup w to XCD w % 8, and the logical
// range index is (w % 8) * (G / 8) + w / 8.  Neighbouring rows read the same p
// and z entries, so an XCD's L2 serves most of those reads.
//
// A block row is handled by 8 lanes, one 3x3 block per lane per pass, so the
// 36-byte blocks of a row are read as one contiguous run; the three row sums
// are combined by an xor-shuffle tree inside the 8-lane group.
// Per iteration: SpMV + partial p.Ap | grid barrier | alpha, x/r/z update,
// partial r.z and max|r| | grid barrier | convergence test and beta.
// ---------------------------------------------------------------------------
struct PcgPersistentArgs {
  int n;
  const float* A;
  const int32_t* row_ptr;
  const int32_t* cols;
  const int32_t* diag;
  const float* rhs;
  int negate;
  float* x;
  float* r;
  float* z;
  float* p0;
  float* p1;
  float* Ap;
  float* Pinv;
  double* part_a;   // r.z partials (one per workgroup)
  double* part_b;   // p.Ap partials
  float* part_max;  // max |r| partials
  PcgStatus* st;
  float tol;
  int maxit;
};

constexpr int kLanesPerRow = 8;
constexpr int kRowsPerPass = kBlock / kLanesPerRow;  // 32 block rows per workgroup pass

__device__ inline void wg_range(int n, int& lo, int& hi) {
  const int G = gridDim.x, w = blockIdx.x;
  const int xcds = (G % 8 == 0) ? 8 : 1;
  const int logical = (w % xcds) * (G / xcds) + w / xcds;
  const long long per = (n + (long long)G - 1) / G;
  lo = (int)min((long long)n, logical * per);
  hi = (int)min((long long)n, lo + per);
}

__global__ void __launch_bounds__(kBlock) k_pcg_persistent(PcgPersistentArgs a) {
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  __shared__ double ldsd[4];
  __shared__ float ldsf[4];
  const int G = gridDim.x;
  int lo, hi;
  wg_range(a.n, lo, hi);

  // Initialisation: preconditioner, r = b, x = 0, z = P^-1 r, p = z.
  double rz_part = 0.0;
  float bmax_part = 0.0f;
  for (int i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    const float* B = a.A + 9 * (size_t)a.diag[i];
    float P[9];
    float det = det3(B);
    if ((double)fabsf(det) > 1e-16) {
      float c0 = cof3(B, 0, 0), c1 = cof3(B, 1, 0), c2 = cof3(B, 2, 0);
      float d = c0 * B[0] + (c1 * B[3] + c2 * B[6]);
      float inv = 1.0f / d;
      for (int q = 0; q < 3; ++q)
        for (int b = 0; b < 3; ++b) P[3 * q + b] = cof3(B, b, q) * inv;
    } else {
      for (int e = 0; e < 9; ++e) P[e] = (e % 4 == 0) ? 1.0f : 0.0f;
    }
    for (int e = 0; e < 9; ++e) a.Pinv[9 * (size_t)i + e] = P[e];
    float rv[3];
    for (int k = 0; k < 3; ++k) {
      float bk = a.negate ? -a.rhs[3 * i + k] : a.rhs[3 * i + k];
      rv[k] = bk - 0.0f;
      a.r[3 * i + k] = rv[k];
      a.x[3 * i + k] = 0.0f;
      bmax_part = fmaxf(bmax_part, fabsf(bk));
    }
    for (int k = 0; k < 3; ++k) {
      float zk = P[3 * k] * rv[0] + (P[3 * k + 1] * rv[1] + P[3 * k + 2] * rv[2]);
      a.z[3 * i + k] = zk;
      a.p0[3 * i + k] = zk;
      rz_part += (double)rv[k] * (double)zk;
    }
  }
  {
    double s = block_reduce(rz_part, ldsd, [](double u, double v) { return u + v; });
    float m = block_reduce(bmax_part, ldsf, [](float u, float v) { return fmaxf(u, v); });
    if (threadIdx.x == 0) {
      a.part_a[blockIdx.x] = s;
      a.part_max[blockIdx.x] = m;
    }
  }
  grid.sync();
  double rz = reduce_sum_partials(a.part_a, G, ldsd);
  const float bmax = reduce_max_partials(a.part_max, G, ldsf);
  {
    float rel = ((double)bmax < 1e-16) ? bmax : bmax / bmax;
    if (rel < a.tol) {
      if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.st->bmax = bmax; a.st->rel = rel; a.st->iters = 0; a.st->failed = 0; a.st->done = 1;
      }
      return;
    }
  }

  const int lane = threadIdx.x % kLanesPerRow;
  const int slot = threadIdx.x / kLanesPerRow;
  float beta = 0.0f;
  float* pold = a.p0;
  float* pnew = a.p1;
  for (int it = 0; it < a.maxit; ++it) {
    const float* pcur = it > 0 ? pnew : a.p0;
    // SpMV with p = z + beta p_old formed on the fly (it > 0).
    double pap_part = 0.0;
    for (int base = lo; base < hi; base += kRowsPerPass) {
      int i = base + slot;
      float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
      if (i < hi) {
        for (int bb = a.row_ptr[i] + lane; bb < a.row_ptr[i + 1]; bb += kLanesPerRow) {
          int j = a.cols[bb];
          float pv0, pv1, pv2;
          if (it > 0) {
            pv0 = a.z[3 * j] + beta * pold[3 * j];
            pv1 = a.z[3 * j + 1] + beta * pold[3 * j + 1];
            pv2 = a.z[3 * j + 2] + beta * pold[3 * j + 2];
            if (j == i) {
              pnew[3 * i] = pv0;
              pnew[3 * i + 1] = pv1;
              pnew[3 * i + 2] = pv2;
            }
          } else {
            pv0 = pold[3 * j];
            pv1 = pold[3 * j + 1];
            pv2 = pold[3 * j + 2];
          }
          const float* Ab = a.A + 9 * (size_t)bb;
          s0 = s0 + (Ab[0] * pv0 + (Ab[1] * pv1 + Ab[2] * pv2));
          s1 = s1 + (Ab[3] * pv0 + (Ab[4] * pv1 + Ab[5] * pv2));
          s2 = s2 + (Ab[6] * pv0 + (Ab[7] * pv1 + Ab[8] * pv2));
        }
      }
      for (int off = kLanesPerRow / 2; off > 0; off >>= 1) {
        s0 += __shfl_xor(s0, off, kLanesPerRow);
        s1 += __shfl_xor(s1, off, kLanesPerRow);
        s2 += __shfl_xor(s2, off, kLanesPerRow);
      }
      if (i < hi && lane == 0) {
        a.Ap[3 * i] = s0;
        a.Ap[3 * i + 1] = s1;
        a.Ap[3 * i + 2] = s2;
        float q0, q1, q2;
        if (it > 0) {
          q0 = a.z[3 * i] + beta * pold[3 * i];
          q1 = a.z[3 * i + 1] + beta * pold[3 * i + 1];
          q2 = a.z[3 * i + 2] + beta * pold[3 * i + 2];
        } else {
          q0 = pold[3 * i]; q1 = pold[3 * i + 1]; q2 = pold[3 * i + 2];
        }
        pap_part += (double)q0 * s0 + (double)q1 * s1 + (double)q2 * s2;
      }
    }
    {
      double s = block_reduce(pap_part, ldsd, [](double u, double v) { return u + v; });
      if (threadIdx.x == 0) a.part_b[blockIdx.x] = s;
    }
    grid.sync();
    double pAp = reduce_sum_partials(a.part_b, G, ldsd);
    if (fabs(pAp) < 1e-16) {
      if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.st->bmax = bmax; a.st->iters = it + 1; a.st->failed = 1; a.st->done = 1;
      }
      return;
    }
    const float alpha = (float)(rz / pAp);
    const float* p = it > 0 ? pnew : a.p0;
    double rz_p = 0.0;
    float rm = 0.0f;
    for (int i = lo + threadIdx.x; i < hi; i += blockDim.x) {
      float rv[3];
      for (int k = 0; k < 3; ++k) {
        int e = 3 * i + k;
        a.x[e] = a.x[e] + alpha * p[e];
        rv[k] = a.r[e] - alpha * a.Ap[e];
        a.r[e] = rv[k];
        rm = fmaxf(rm, fabsf(rv[k]));
      }
      const float* P = a.Pinv + 9 * (size_t)i;
      for (int k = 0; k < 3; ++k) {
        float zk = P[3 * k] * rv[0] + (P[3 * k + 1] * rv[1] + P[3 * k + 2] * rv[2]);
        a.z[3 * i + k] = zk;
        rz_p += (double)rv[k] * (double)zk;
      }
      if (rm != rm) rm = __int_as_float(0x7f800000);
    }
    {
      double s = block_reduce(rz_p, ldsd, [](double u, double v) { return u + v; });
      float m = block_reduce(rm, ldsf, [](float u, float v) { return fmaxf(u, v); });
      if (threadIdx.x == 0) {
        a.part_a[blockIdx.x] = s;
        a.part_max[blockIdx.x] = m;
      }
    }
    grid.sync();
    double rz_new = reduce_sum_partials(a.part_a, G, ldsd);
    float rmax = reduce_max_partials(a.part_max, G, ldsf);
    float rel = ((double)bmax < 1e-16) ? rmax : rmax / bmax;
    if (rel < a.tol) {
      if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.st->bmax = bmax; a.st->rel = rel; a.st->iters = it + 1; a.st->failed = 0; a.st->done = 1;
      }
      return;
    }
    beta = (float)(rz_new / rz);
    rz = rz_new;
    // p_new of this iteration becomes p_old of the next.
    if (it > 0) {
      float* t = pold;
      pold = pnew;
      pnew = t;
    } else {
      pold = a.p0;
      pnew = a.p1;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.st->bmax = bmax; a.st->iters = a.maxit; a.st->failed = 1; a.st->done = 1;
  }
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
