// Two-level Schwarz preconditioned CG (new; the reference's PCG is
// block-Jacobi, pcg_solver.cpp:8-146).
//
// The block rows are renumbered in Morton order of the rest positions and cut
// into patches of `pr` consecutive rows (chosen so that there are at most 256
// patches up to 131,072 rows: one workgroup per CU).  The preconditioner is
//
//   M = sum_patches R_p^T C_p R_p + P Ac^-1 P^T,
//
// where C_p is a fixed odd-degree Chebyshev polynomial in D_p^-1 A_pp (the
// patch's own blocks, block-Jacobi scaling D), which is symmetric positive
// definite for every symmetric positive definite A_pp whatever the interval
// estimate, and P sums each patch's rows into three coarse unknowns (a patch
// moving rigidly), Ac = P^T A P, inverted densely.
//
// Two launches per iteration, as plain PCG, one workgroup per patch:
//   k_tl_spmv:   convergence test and beta, p = z + beta p on the fly, Ap (two
//                lanes per row), the partial p.Ap and the patch sum of Ap;
//   k_tl_update: alpha, x += alpha p, r -= alpha Ap, then z = M r directly
//                (one thread per row): the patch's rows, read into registers,
//                run the Chebyshev steps through LDS, and the coarse part is
//                Ac^-1 P^T r with P^T r = (the patch sums of the previous r) -
//                alpha P^T A p (both summed in FP64 from the vectors
//                themselves, so nothing accumulates over iterations); partial
//                r.z, max|r| and the patch sum of the new r.
// z is formed from r every iteration: a recurrence z -= alpha M A p carries
// the FP32 rounding of every application along and stalls CG on stiff
// systems once r has fallen by a few orders of magnitude.
//
// Vectors inside a solve are in the permuted order; the right-hand side is
// gathered at the start and the solution scattered at the end.

namespace ando {
namespace tl {

constexpr int kMaxPatchRows = 512;
constexpr int kLanes = 2;                        // lanes per row in the SpMV
constexpr int kSpmvSlots = kSellChunk / kLanes;  // slots per lane held in registers by the SpMV
constexpr int kPowerIters = 10;                  // power iterations for the patch's largest eigenvalue
constexpr float kLmaxSafety = 1.2f;              // applied to the power estimate
constexpr int kMaxCheb = 15;
constexpr int kMaxWaves = 2 * kMaxPatchRows / 64;  // waves of the widest patch workgroup (the SpMV's)
// Rounds of the SpMV's coupling pass: blockDim / 8 rows per round, and
// blockDim = spmv_threads(pr) >= 2 pr, so four rounds cover the patch.
constexpr int kCplRounds = 4;

// Rows per patch for n block rows: at least 256, at most 512, and no more than
// 256 patches while that allows.
inline int patch_rows(int n) { return std::min(kMaxPatchRows, std::max(256, (n + 255) / 256)); }
// Workgroup size of the SpMV: 32 rows (two lanes each) per wave.
inline int spmv_threads(int pr) { return 64 * ((pr + 31) / 32); }

struct TlArgs {
  int n;                     // block rows
  int pr;                    // rows per patch
  int npatch;                // patches (workgroups of every patch kernel)
  int slices;                // sliced-ELL slices over the permuted rows
  int uniform_w;             // common slice width, or 0
  int nc;                    // coarse unknowns (3 per patch), 0: no coarse level
  int cheb;                  // Chebyshev degree (odd)
  float lmin_ratio;          // interval [lmax / lmin_ratio, lmax]
  const int32_t* perm;       // permuted row -> original row
  const int32_t* iperm;      // original row -> permuted row
  const float* vals;         // sliced values (9 per slot, slot-major, entry-major in a slice)
  const int32_t* scol;       // permuted column per slot
  const int32_t* slice_off;
  const int32_t* slice_w;
  const float* Dinv;         // block-Jacobi blocks, permuted rows, plane e at Dinv[e * n + i]
  const float* lmax;         // per patch: largest eigenvalue bound of D^-1 A_pp
  const float* Acinv;        // coarse inverse, nc x nc row-major (FP32)
  // Self-contact coupling blocks (original numbering): row i's blocks are
  // [3 cptr[i], 3 cptr[i + 1]) with columns ccol and values cval; cptr null: none.
  const int32_t* cptr;
  const int32_t* ccol;
  const float* cval;
  // The same blocks for the SpMV's coupling pass in the permuted numbering:
  // cpr[2 i], cpr[2 i + 1] = first and end block of permuted row i, ccolp =
  // permuted column per block (built by k_tl_couple_perm for each solve).
  const int32_t* cpr;
  const int32_t* ccolp;
};

// Coupling blocks in the permuted numbering (see TlArgs::cpr, ccolp).
__global__ void k_tl_couple_perm(int n, int nblk3, const int32_t* __restrict__ perm,
                                 const int32_t* __restrict__ iperm, const int32_t* __restrict__ cptr,
                                 const int32_t* __restrict__ ccol, int32_t* __restrict__ cpr,
                                 int32_t* __restrict__ ccolp) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) {
    const int o = perm[t];
    cpr[2 * t] = 3 * cptr[o];
    cpr[2 * t + 1] = 3 * cptr[o + 1];
  }
  if (t < nblk3) ccolp[t] = iperm[ccol[t]];
}

// A row's place in the sliced layout: the slice's first slot position, the
// row within the slice and the slice width (0 past the last slice).
struct SliceRow {
  int soff, rs, W;
};
__device__ __forceinline__ SliceRow slice_row(const TlArgs& a, int row) {
  SliceRow R;
  const int sl0 = row >> 6;
  const bool slive = sl0 < a.slices;
  const int sl = slive ? sl0 : a.slices - 1;
  R.soff = a.uniform_w > 0 ? sl * 64 * a.uniform_w : a.slice_off[sl];
  R.rs = row & 63;
  R.W = !slive ? 0 : (a.uniform_w > 0 ? a.uniform_w : a.slice_w[sl]);
  return R;
}
__device__ __forceinline__ int slot_col(const TlArgs& a, const SliceRow& R, int k) {
  return a.scol[R.soff + 64 * k + R.rs];
}
__device__ __forceinline__ float slot_val(const TlArgs& a, const SliceRow& R, int k, int e) {
  return a.vals[9 * (size_t)R.soff + (size_t)(9 * k + e) * 64 + R.rs];
}

// The preconditioner's own arithmetic (block-Jacobi blocks, patch products)
// uses fused multiply-adds: it has no counterpart in the reference to match,
// and the Chebyshev steps are bound by these instructions.
__device__ __forceinline__ void mul3(const float* D, const float* v, float* o) {
  o[0] = fmaf(D[0], v[0], fmaf(D[1], v[1], D[2] * v[2]));
  o[1] = fmaf(D[3], v[0], fmaf(D[4], v[1], D[5] * v[2]));
  o[2] = fmaf(D[6], v[0], fmaf(D[7], v[1], D[8] * v[2]));
}

// The patch solves: kPatchLanes lanes per row of the patch (one: thread t
// takes row t; two: as in the SpMV, wave w takes rows 32 w .. 32 w + 31 and
// lane group q the slots of parity q); each lane keeps its first slots in
// registers (local column, or -1 when the column is outside the patch or the
// slot is past the row's width).
constexpr int kPatchLanes = 2;
constexpr int kPatchSlots = kSellChunk / kPatchLanes;
inline int patch_threads(int pr) { return kPatchLanes == 2 ? spmv_threads(pr) : 64 * ((pr + 63) / 64); }

struct PatchRow {
  int q, lr, row;
  bool live;
  SliceRow R;
  int lc[kPatchSlots];
  float Ab[kPatchSlots][9];
};

__device__ __forceinline__ void load_patch_row(const TlArgs& a, int patch, PatchRow& P) {
  const int p0 = patch * a.pr;
  const int tid = threadIdx.x;
  if (kPatchLanes == 2) {
    P.q = (tid >> 5) & 1;
    P.lr = (tid >> 6) * 32 + (tid & 31);
  } else {
    P.q = 0;
    P.lr = tid;
  }
  P.row = p0 + P.lr;
  P.live = P.lr < a.pr && P.row < a.n;
  P.R = slice_row(a, P.live ? P.row : a.n - 1);
  if (!P.live) P.R.W = 0;
#pragma unroll
  for (int m = 0; m < kPatchSlots; ++m) {
    const int k = kPatchLanes * m + P.q;
    P.lc[m] = -1;
    if (k < P.R.W) {
      const int lc = slot_col(a, P.R, k) - p0;
      P.lc[m] = (lc >= 0 && lc < a.pr) ? lc : -1;
#pragma unroll
      for (int e = 0; e < 9; ++e) P.Ab[m][e] = slot_val(a, P.R, k, e);
    }
  }
}

// o = sum over the row's slots inside the patch of A_ij v_j (v in LDS, local
// numbering); every lane of the row gets the sum.
__device__ __forceinline__ void local_product(const TlArgs& a, int patch, const PatchRow& P,
                                              const float* __restrict__ lv, float (&o)[3]) {
  o[0] = o[1] = o[2] = 0.0f;
#pragma unroll
  for (int m = 0; m < kPatchSlots; ++m) {
    if (P.lc[m] >= 0) {
      const int lc = P.lc[m];
      const float v0 = lv[3 * lc], v1 = lv[3 * lc + 1], v2 = lv[3 * lc + 2];
      o[0] = fmaf(P.Ab[m][0], v0, fmaf(P.Ab[m][1], v1, fmaf(P.Ab[m][2], v2, o[0])));
      o[1] = fmaf(P.Ab[m][3], v0, fmaf(P.Ab[m][4], v1, fmaf(P.Ab[m][5], v2, o[1])));
      o[2] = fmaf(P.Ab[m][6], v0, fmaf(P.Ab[m][7], v1, fmaf(P.Ab[m][8], v2, o[2])));
    }
  }
  // Rows longer than the register chunk read their further slots again.
  const int p0 = patch * a.pr;
  for (int k = kSellChunk + P.q; k < P.R.W; k += kPatchLanes) {
    const int lc = slot_col(a, P.R, k) - p0;
    if (lc >= 0 && lc < a.pr) {
      float B[9];
      for (int e = 0; e < 9; ++e) B[e] = slot_val(a, P.R, k, e);
      const float v0 = lv[3 * lc], v1 = lv[3 * lc + 1], v2 = lv[3 * lc + 2];
      o[0] = fmaf(B[0], v0, fmaf(B[1], v1, fmaf(B[2], v2, o[0])));
      o[1] = fmaf(B[3], v0, fmaf(B[4], v1, fmaf(B[5], v2, o[1])));
      o[2] = fmaf(B[6], v0, fmaf(B[7], v1, fmaf(B[8], v2, o[2])));
    }
  }
  if (kPatchLanes == 2)
    for (int c = 0; c < 3; ++c) o[c] = o[c] + __shfl_xor(o[c], 32, 64);
}

// y = C_p f: `a.cheb` steps of Chebyshev iteration for A_pp y = f from y = 0
// with preconditioner D on [lmax / ratio, lmax].  lv: two LDS buffers of
// 3 kMaxPatchRows floats.  Every thread of the workgroup takes part (barriers).
__device__ __forceinline__ void patch_chebyshev(const TlArgs& a, int patch, const PatchRow& P, float lmax,
                                                const float (&D)[9], const float (&f)[3], float* lv,
                                                float (&y)[3]) {
  const float lmin = lmax / a.lmin_ratio;
  const float th = 0.5f * (lmax + lmin), de = 0.5f * (lmax - lmin);
  const float sigma = th / de;
  float rho = 1.0f / sigma;
  float d[3];
  mul3(D, f, d);
  for (int c = 0; c < 3; ++c) {
    d[c] = d[c] / th;
    y[c] = d[c];
  }
  for (int j = 1; j < a.cheb; ++j) {
    float* buf = lv + (j & 1) * 3 * kMaxPatchRows;
    if (P.q == 0 && P.lr < a.pr)
      for (int c = 0; c < 3; ++c) buf[3 * P.lr + c] = P.live ? y[c] : 0.0f;
    __syncthreads();
    float ay[3];
    local_product(a, patch, P, buf, ay);
    const float rho_n = 1.0f / (2.0f * sigma - rho);
    const float c1 = rho_n * rho, c2 = 2.0f * rho_n / de;
    float res[3] = {f[0] - ay[0], f[1] - ay[1], f[2] - ay[2]};
    float dr[3];
    mul3(D, res, dr);
    for (int c = 0; c < 3; ++c) {
      d[c] = fmaf(c1, d[c], c2 * dr[c]);
      y[c] = y[c] + d[c];
    }
    rho = rho_n;
  }
  if (!P.live) y[0] = y[1] = y[2] = 0.0f;
}

// Workgroup sums of N doubles, in a fixed order; every thread gets the
// results.  lds: N rows of kMaxWaves.
template <int N>
__device__ __forceinline__ void wg_sum(double (&v)[N], double (*lds)[kMaxWaves]) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
  for (int c = 0; c < N; ++c)
    for (int o = 32; o > 0; o >>= 1) v[c] += __shfl_down(v[c], o, 64);
  if (lane == 0)
    for (int c = 0; c < N; ++c) lds[c][wid] = v[c];
  __syncthreads();
  for (int c = 0; c < N; ++c) {
    double s = lds[c][0];
    for (int w = 1; w < nw; ++w) s += lds[c][w];
    v[c] = s;
  }
  __syncthreads();
}

// wg_sum of N doubles together with a float maximum (one LDS round).
template <int N>
__device__ __forceinline__ void wg_sum_max(double (&v)[N], float& mx, double (*lds)[kMaxWaves],
                                           float* ldsf) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
  for (int o = 32; o > 0; o >>= 1) {
    for (int c = 0; c < N; ++c) v[c] += __shfl_down(v[c], o, 64);
    mx = fmaxf(mx, __shfl_down(mx, o, 64));
  }
  if (lane == 0) {
    for (int c = 0; c < N; ++c) lds[c][wid] = v[c];
    ldsf[wid] = mx;
  }
  __syncthreads();
  for (int c = 0; c < N; ++c) {
    double s = lds[c][0];
    for (int w = 1; w < nw; ++w) s += lds[c][w];
    v[c] = s;
  }
  float m = ldsf[0];
  for (int w = 1; w < nw; ++w) m = fmaxf(m, ldsf[w]);
  mx = m;
  __syncthreads();
}

// Workgroup sums of N doubles (and a float maximum) that only the first
// threads need: thread c < N returns sum c, thread N gets the maximum in mx.
// Each wave sums down to its lane 0, one LDS round, then thread c adds the
// waves' values in wave order (the order of wg_sum).  kHalf: only lanes 0-31
// of every wave hold values, the others zeros.
template <int N, bool kHalf>
__device__ __forceinline__ double wg_sum_max_first(double (&v)[N], float& mx, double (*lds)[kMaxWaves],
                                                   float* ldsf) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (int)(blockDim.x >> 6);
  for (int o = kHalf ? 16 : 32; o > 0; o >>= 1) {
    for (int c = 0; c < N; ++c) v[c] += __shfl_down(v[c], o, 64);
    mx = fmaxf(mx, __shfl_down(mx, o, 64));
  }
  if (lane == 0) {
    for (int c = 0; c < N; ++c) lds[c][wid] = v[c];
    ldsf[wid] = mx;
  }
  __syncthreads();
  const int t = threadIdx.x;
  double s = 0.0;
  if (t < N) {
    s = lds[t][0];
    for (int w = 1; w < nw; ++w) s += lds[t][w];
  } else if (t == N) {
    float m = ldsf[0];
    for (int w = 1; w < nw; ++w) m = fmaxf(m, ldsf[w]);
    mx = m;
  }
  return s;
}

// Rows 3 patch .. 3 patch + 2 of Ac^-1 times s (nc values): this thread's
// share of the three dot products (reduced by the caller).
__device__ __forceinline__ void coarse_rows(const TlArgs& a, int patch, const double* __restrict__ s,
                                            double (&e)[3]) {
  e[0] = e[1] = e[2] = 0.0;
  if (a.nc == 0) return;
  for (int col = threadIdx.x; col < a.nc; col += blockDim.x) {
    const double sv = s[col];
#pragma unroll
    for (int c = 0; c < 3; ++c) e[c] += (double)a.Acinv[(size_t)(3 * patch + c) * a.nc + col] * sv;
  }
}

// Warm start: x0 (permuted) = s * x0_orig[perm].
__global__ void k_tl_gather_x0(int n, const int32_t* __restrict__ perm, const float* __restrict__ x0, float s,
                               float* __restrict__ xp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int o = perm[i];
  for (int c = 0; c < 3; ++c) xp[3 * i + c] = s * x0[3 * o + c];
}

// out = A x (permuted rows and columns; the mesh blocks from the sliced
// layout in slot order, then the row's self-contact coupling blocks), one
// thread per row: the starting residual of a warm-started solve.
__global__ void __launch_bounds__(kBlock) k_tl_ax0(TlArgs a, const float* __restrict__ xp, float* __restrict__ out) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= a.n) return;
  const SliceRow R = slice_row(a, row);
  float T[3] = {0.0f, 0.0f, 0.0f};
  for (int k = 0; k < R.W; ++k) {
    const int j = slot_col(a, R, k);
    float B[9];
    for (int e = 0; e < 9; ++e) B[e] = slot_val(a, R, k, e);
    T[0] = T[0] + (B[0] * xp[3 * j] + (B[1] * xp[3 * j + 1] + B[2] * xp[3 * j + 2]));
    T[1] = T[1] + (B[3] * xp[3 * j] + (B[4] * xp[3 * j + 1] + B[5] * xp[3 * j + 2]));
    T[2] = T[2] + (B[6] * xp[3 * j] + (B[7] * xp[3 * j + 1] + B[8] * xp[3 * j + 2]));
  }
  if (a.cptr) {
    const int o = a.perm[row];
    for (int b = 3 * a.cptr[o]; b < 3 * a.cptr[o + 1]; ++b) {
      const int j = a.iperm[a.ccol[b]];
      const float* B = a.cval + 9 * (size_t)b;
      T[0] = T[0] + (B[0] * xp[3 * j] + (B[1] * xp[3 * j + 1] + B[2] * xp[3 * j + 2]));
      T[1] = T[1] + (B[3] * xp[3 * j] + (B[4] * xp[3 * j + 1] + B[5] * xp[3 * j + 2]));
      T[2] = T[2] + (B[6] * xp[3 * j] + (B[7] * xp[3 * j + 1] + B[8] * xp[3 * j + 2]));
    }
  }
  for (int c = 0; c < 3; ++c) out[3 * row + c] = T[c];
}

// With a warm start the relative residual is measured against max|b|, not
// max|r0|: the status written by the start gets that denominator.
__global__ void k_tl_fix_bmax(int nwg, const float* __restrict__ part_bmax, const float* __restrict__ part_max,
                              int nmax, PcgStatus* __restrict__ st, float tol) {
  __shared__ float ldsf[4];
  const float bmax = reduce_max_partials(part_bmax, nwg, ldsf);
  const float rmax = reduce_max_partials(part_max, nmax, ldsf);
  if (threadIdx.x == 0) {
    const float rel = ((double)bmax < 1e-16) ? rmax : rmax / bmax;
    st->bmax = bmax;
    st->rel = rel;
    st->done = rel < tol ? 1 : 0;
  }
}

// Solve set-up, per permuted row: r = +-g[perm] (minus A x0 when ax0 is
// given, keeping x = x0; else x = 0), the block-Jacobi inverse of the row's
// diagonal block (pcg_solver.cpp:77-135), and with part_bmax the
// workgroup's max|b|.
__global__ void __launch_bounds__(kBlock) k_tl_init(TlArgs a, const float* __restrict__ A,
                                                    const int32_t* __restrict__ diag, const float* __restrict__ g,
                                                    int negate, float* __restrict__ x, float* __restrict__ r,
                                                    float* __restrict__ Dinv, const float* __restrict__ ax0,
                                                    float* __restrict__ part_bmax) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (part_bmax) {
    __shared__ float ldsf[4];
    float bm = 0.0f;
    if (i < a.n)
      for (int c = 0; c < 3; ++c) bm = fmaxf(bm, fabsf(g[3 * a.perm[i] + c]));
    bm = block_reduce(bm, ldsf, [](float u, float w) { return fmaxf(u, w); });
    if (threadIdx.x == 0) part_bmax[blockIdx.x] = bm;
  }
  if (i >= a.n) return;
  const int o = a.perm[i];
  const float* B = A + 9 * (size_t)diag[o];
  float P[9];
  const float det = det3(B);
  if ((double)fabsf(det) > 1e-16) {
    const float c0 = cof3(B, 0, 0), c1 = cof3(B, 1, 0), c2 = cof3(B, 2, 0);
    const float dd = c0 * B[0] + (c1 * B[3] + c2 * B[6]);
    const float inv = 1.0f / dd;
    for (int u = 0; u < 3; ++u)
      for (int w = 0; w < 3; ++w) P[3 * u + w] = cof3(B, w, u) * inv;
  } else {
    for (int e = 0; e < 9; ++e) P[e] = (e % 4 == 0) ? 1.0f : 0.0f;
  }
  for (int e = 0; e < 9; ++e) Dinv[(size_t)e * a.n + i] = P[e];
  for (int c = 0; c < 3; ++c) {
    const float gv = g[3 * o + c];
    const float bv = negate ? -gv : gv;
    if (ax0) {
      r[3 * i + c] = bv - ax0[3 * i + c];
    } else {
      r[3 * i + c] = bv;
      x[3 * i + c] = 0.0f;
    }
  }
}

// Per patch: the largest eigenvalue of D^-1 A_pp by power iteration (times
// kLmaxSafety), then qf = C_p r and sr = P^T r (the patch's sum of r).
__global__ void __launch_bounds__(kPatchLanes * kMaxPatchRows) k_tl_prep(TlArgs a, float* __restrict__ lmax_out,
                                                           const float* __restrict__ r, float* __restrict__ qf,
                                                           double* __restrict__ sr) {
  __shared__ float lv[2 * 3 * kMaxPatchRows];
  __shared__ double ldd[3][kMaxWaves];
  const int patch = xcd_logical_group((int)blockIdx.x, a.npatch);
  PatchRow P;
  load_patch_row(a, patch, P);
  const int ri = P.live ? P.row : a.n - 1;
  float D[9];
  for (int e = 0; e < 9; ++e) D[e] = a.Dinv[(size_t)e * a.n + ri];
  // Power iteration from a fixed pseudo-random start.
  float v[3];
  for (int c = 0; c < 3; ++c) {
    const uint32_t h = (uint32_t)(3 * P.row + c) * 2654435761u;
    v[c] = P.live ? 0.5f + (float)(h >> 8) * (1.0f / 16777216.0f) : 0.0f;
  }
  float lam = 0.0f;
  for (int it = 0; it < kPowerIters; ++it) {
    float* buf = lv + (it & 1) * 3 * kMaxPatchRows;
    if (P.q == 0 && P.lr < a.pr)
      for (int c = 0; c < 3; ++c) buf[3 * P.lr + c] = v[c];
    __syncthreads();
    float av[3], w[3];
    local_product(a, patch, P, buf, av);
    mul3(D, av, w);
    if (!P.live) w[0] = w[1] = w[2] = 0.0f;
    double nn[2] = {0.0, 0.0};
    if (P.q == 0) {
      nn[0] = (double)w[0] * w[0] + (double)w[1] * w[1] + (double)w[2] * w[2];
      nn[1] = (double)v[0] * v[0] + (double)v[1] * v[1] + (double)v[2] * v[2];
    }
    wg_sum(nn, ldd);
    const double nw = sqrt(nn[0]), nv = sqrt(nn[1]);
    lam = nv > 0.0 ? (float)(nw / nv) : 0.0f;
    const float sc = nw > 0.0 ? (float)(1.0 / nw) : 0.0f;
    for (int c = 0; c < 3; ++c) v[c] = w[c] * sc;
  }
  const float lmx = (lam > 0.0f && lam < 3.0e38f) ? lam * kLmaxSafety : 1.0f;
  if (threadIdx.x == 0) lmax_out[patch] = lmx;
  float f[3];
  for (int c = 0; c < 3; ++c) f[c] = P.live ? r[3 * ri + c] : 0.0f;
  float y[3];
  patch_chebyshev(a, patch, P, lmx, D, f, lv, y);
  if (P.q == 0 && P.live)
    for (int c = 0; c < 3; ++c) qf[3 * P.row + c] = y[c];
  double s[3] = {0.0, 0.0, 0.0};
  if (P.q == 0)
    for (int c = 0; c < 3; ++c) s[c] = f[c];
  wg_sum(s, ldd);
  if (threadIdx.x == 0)
    for (int c = 0; c < 3; ++c) sr[3 * patch + c] = s[c];
}

// Set-up, per patch (one row per thread): z = qf + (Ac^-1 P^T r) of the patch,
// p = z, partial r.z and max|r|.
__global__ void __launch_bounds__(kMaxPatchRows) k_tl_start(TlArgs a, const float* __restrict__ r,
                                                            const float* __restrict__ qf,
                                                            const double* __restrict__ sr, float* __restrict__ z,
                                                            float* __restrict__ p, double* __restrict__ part_rz,
                                                            float* __restrict__ part_max) {
  __shared__ double ldd[3][kMaxWaves];
  __shared__ double ldsd[kMaxWaves];
  __shared__ float ldsf[kMaxWaves];
  const int patch = xcd_logical_group((int)blockIdx.x, a.npatch);
  const int i = patch * a.pr + threadIdx.x;
  double e[3];
  coarse_rows(a, patch, sr, e);
  wg_sum(e, ldd);
  double rz = 0.0;
  float rm = 0.0f;
  if ((int)threadIdx.x < a.pr && i < a.n) {
    for (int c = 0; c < 3; ++c) {
      const float rv = r[3 * i + c];
      const float zv = qf[3 * i + c] + (float)e[c];
      z[3 * i + c] = zv;
      p[3 * i + c] = zv;
      rz += (double)rv * (double)zv;
      rm = fmaxf(rm, fabsf(rv));
    }
  }
  block_reduce_sum_max(rz, rm, ldsd, ldsf);
  if (threadIdx.x == 0) {
    part_rz[patch] = rz;
    part_max[patch] = rm;
  }
}

// One iteration's product (one workgroup per patch; wave w takes rows
// 32 w .. 32 w + 31 of the patch, lane group q (lanes 0-31 / 32-63) the slots
// of parity q of its row): convergence test and beta from the previous
// update's partials, p = z + beta p on the fly, Ap = A p (mesh slots, then the
// row's self-contact coupling blocks), the partial p.Ap and sa = P^T A p.
__global__ void __launch_bounds__(2 * kMaxPatchRows) k_tl_spmv(TlArgs a, int it, const float* __restrict__ z,
                                                               const float* __restrict__ p_old,
                                                               float* __restrict__ p_new, float* __restrict__ Ap,
                                                               double* __restrict__ sa,
                                                               const double* __restrict__ part_rz,
                                                               const float* __restrict__ part_max,
                                                               double* __restrict__ rz_hist,
                                                               double* __restrict__ part_pAp,
                                                               PcgStatus* __restrict__ st, float tol, int spec) {
  if (spec && st->done) return;  // queued past the predicted convergence: flag first
  __shared__ double ldd[4][kMaxWaves];
  __shared__ double ldsd[kMaxWaves];
  __shared__ float ldsf[kMaxWaves];
  // Patches numbered XCD-contiguously (the same numbering in every patch
  // kernel), so neighbouring patches, which gather each other's rows, share an L2.
  const int patch = xcd_logical_group((int)blockIdx.x, a.npatch);
  const int tid = threadIdx.x;
  const int wv = tid >> 6, lane = tid & 63;
  const int q = lane >> 5;
  const int lr = wv * 32 + (lane & 31);
  const int row = patch * a.pr + lr;
  const bool live = lr < a.pr && row < a.n;
  SliceRow R = slice_row(a, live ? row : a.n - 1);
  if (!live) R.W = 0;
  // The previous update's partials head the longest chain: requested first
  // (the first 256 threads reduce them).
  const bool red = tid < kBlock;
  double pre_rz[kUnroll];
  float pre_max[kUnroll];
  if (it > 0 && red) {
    prefetch_partials(part_rz, a.npatch, pre_rz, tid, kBlock);
    prefetch_partials(part_max, a.npatch, pre_max, tid, kBlock);
  }
  // The coupling pass's row ranges, requested with the first loads.
  int cbr[kCplRounds], cer[kCplRounds];
  if (a.cptr) {
    const int nslot = (int)(blockDim.x >> 3);
#pragma unroll
    for (int k = 0; k < kCplRounds; ++k) {
      const int clr = k * nslot + (tid >> 3), crow = patch * a.pr + clr;
      const bool ok = clr < a.pr && crow < a.n;
      cbr[k] = ok ? a.cpr[2 * crow] : 0;
      cer[k] = ok ? a.cpr[2 * crow + 1] : 0;
    }
  }
  int jc[kSpmvSlots];
  float Ab[kSpmvSlots][9];
#pragma unroll
  for (int m = 0; m < kSpmvSlots; ++m) {
    const int k = kLanes * m + q;
    jc[m] = 0;
    if (k < R.W) {
      jc[m] = slot_col(a, R, k);
#pragma unroll
      for (int e = 0; e < 9; ++e) Ab[m][e] = slot_val(a, R, k, e);
    }
  }
  const int ri = live ? row : a.n - 1;
  float zi[3], pi[3];
  for (int c = 0; c < 3; ++c) {
    zi[c] = z[3 * ri + c];
    pi[c] = p_old[3 * ri + c];
  }
  float zo[kSpmvSlots][3], po[kSpmvSlots][3];
#pragma unroll
  for (int m = 0; m < kSpmvSlots; ++m)
    if (kLanes * m + q < R.W)
      for (int c = 0; c < 3; ++c) {
        zo[m][c] = z[3 * jc[m] + c];
        po[m][c] = p_old[3 * jc[m] + c];
      }
  if (st->done) return;
  float beta = 0.0f;
  if (it > 0) {
    const float bmax = st->bmax;
    const double rz_prev = rz_hist[it - 1];
    double rz = red ? finish_sum_partials(part_rz, a.npatch, pre_rz, tid, kBlock) : 0.0;
    float rmax = red ? finish_max_partials(part_max, a.npatch, pre_max, tid, kBlock) : 0.0f;
    block_reduce_sum_max(rz, rmax, ldsd, ldsf);
    const float rel = ((double)bmax < 1e-16) ? rmax : rmax / bmax;
    if (rel < tol) {
      if (tid == 0) {
        st->rel = rel;
        st->iters = it;
        st->done = 1;
      }
      return;
    }
    beta = (float)(rz / rz_prev);
    if (patch == 0 && tid == 0) {
      rz_hist[it] = rz;
      st->rel = rel;  // progress for the host's chunk sizing
      st->rel_it = it;
    }
  }
  float T[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int m = 0; m < kSpmvSlots; ++m) {
    if (kLanes * m + q < R.W) {
      float pv[3];
      for (int c = 0; c < 3; ++c) pv[c] = it > 0 ? zo[m][c] + beta * po[m][c] : po[m][c];
      T[0] = T[0] + (Ab[m][0] * pv[0] + (Ab[m][1] * pv[1] + Ab[m][2] * pv[2]));
      T[1] = T[1] + (Ab[m][3] * pv[0] + (Ab[m][4] * pv[1] + Ab[m][5] * pv[2]));
      T[2] = T[2] + (Ab[m][6] * pv[0] + (Ab[m][7] * pv[1] + Ab[m][8] * pv[2]));
    }
  }
  for (int k = kSellChunk + q; k < R.W; k += kLanes) {
    const int j = slot_col(a, R, k);
    float B[9];
    for (int e = 0; e < 9; ++e) B[e] = slot_val(a, R, k, e);
    float pv[3];
    for (int c = 0; c < 3; ++c) pv[c] = it > 0 ? z[3 * j + c] + beta * p_old[3 * j + c] : p_old[3 * j + c];
    T[0] = T[0] + (B[0] * pv[0] + (B[1] * pv[1] + B[2] * pv[2]));
    T[1] = T[1] + (B[3] * pv[0] + (B[4] * pv[1] + B[5] * pv[2]));
    T[2] = T[2] + (B[6] * pv[0] + (B[7] * pv[1] + B[8] * pv[2]));
  }
  float ap[3];
  for (int c = 0; c < 3; ++c) ap[c] = T[c] + __shfl_xor(T[c], 32, 64);
  // Self-contact coupling blocks of the patch rows (original numbering),
  // multiplied by the whole workgroup: 8 lanes per row, rows in rounds of
  // blockDim / 8, each lane requesting four blocks (columns, values, p and z
  // entries) before using any, in block order.  A row in a contact cluster can
  // carry dozens of blocks; spreading them over 8 lanes keeps such a row from
  // holding the workgroup.  The row sums reach the row's lanes through LDS.
  if (a.cptr) {
    __shared__ float lcp[3 * kMaxPatchRows];
    const int cl = tid & 7, nslot = (int)(blockDim.x >> 3);
    constexpr int kU = 4;
    for (int k = 0; k < kCplRounds; ++k) {
      const int clr = k * nslot + (tid >> 3);
      const int cb = cbr[k], ce = cer[k];
      float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
      for (int b0 = cb + cl; b0 < ce; b0 += 8 * kU) {
        int j[kU];
        float B[kU][9], pv[kU][3];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
          const int b = b0 + 8 * u;
          if (b < ce) {
            j[u] = a.ccolp[b];
#pragma unroll
            for (int e = 0; e < 9; ++e) B[u][e] = a.cval[9 * (size_t)b + e];
          }
        }
#pragma unroll
        for (int u = 0; u < kU; ++u)
          if (b0 + 8 * u < ce)
            for (int c = 0; c < 3; ++c)
              pv[u][c] = it > 0 ? z[3 * j[u] + c] + beta * p_old[3 * j[u] + c] : p_old[3 * j[u] + c];
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
      if (cl == 0 && clr < a.pr) {
        lcp[3 * clr] = s0;
        lcp[3 * clr + 1] = s1;
        lcp[3 * clr + 2] = s2;
      }
    }
    __syncthreads();
    if (live)
      for (int c = 0; c < 3; ++c) ap[c] = ap[c] + lcp[3 * lr + c];
  }
  // p.Ap partial, and the patch sum of A p for the coarse part.
  double sums[4] = {0.0, 0.0, 0.0, 0.0};
  if (q == 0 && live) {
    float pn[3];
    for (int c = 0; c < 3; ++c) pn[c] = it > 0 ? zi[c] + beta * pi[c] : pi[c];
    if (it > 0)
      for (int c = 0; c < 3; ++c) p_new[3 * row + c] = pn[c];
    for (int c = 0; c < 3; ++c) Ap[3 * row + c] = ap[c];
    sums[0] = (double)pn[0] * ap[0] + (double)pn[1] * ap[1] + (double)pn[2] * ap[2];
    for (int c = 0; c < 3; ++c) sums[1 + c] = ap[c];
  }
  __shared__ float ldsf_end[kMaxWaves];
  float no_max = 0.0f;
  const double tot = wg_sum_max_first<4, true>(sums, no_max, ldd, ldsf_end);
  if (tid == 0)
    part_pAp[patch] = tot;
  else if (tid < 4)
    sa[3 * patch + tid - 1] = tot;
}

// alpha = r.z / p.Ap, x += alpha p, r -= alpha Ap, z = C_p r + (Ac^-1 P^T r)
// with P^T r = sr - alpha sa, partial r.z and max|r|, and the patch sum of the
// new r (sr_next); one workgroup per patch, one row per thread.
__global__ void __launch_bounds__(kPatchLanes * kMaxPatchRows) k_tl_update(TlArgs a, int it, const float* __restrict__ p,
                                                             const float* __restrict__ Ap,
                                                             const double* __restrict__ sr,
                                                             const double* __restrict__ sa,
                                                             double* __restrict__ sr_next, float* __restrict__ x,
                                                             float* __restrict__ r, float* __restrict__ z,
                                                             const double* __restrict__ part_pAp,
                                                             const double* __restrict__ rz_hist,
                                                             double* __restrict__ part_rz,
                                                             float* __restrict__ part_max,
                                                             PcgStatus* __restrict__ st, int spec) {
  if (spec && st->done) return;  // queued past the predicted convergence: flag first
  __shared__ float lv[2 * 3 * kMaxPatchRows];
  __shared__ double ldd[2][kMaxWaves];
  __shared__ double ldf[4][kMaxWaves];
  __shared__ float ldsf[kMaxWaves];
  const int patch = xcd_logical_group((int)blockIdx.x, a.npatch);
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6, nw = (int)(blockDim.x >> 6);
  // Scalars first: they head the longest chains.
  const int done = st->done;
  const double rz_it = rz_hist[it];
  const bool red = tid < kBlock;
  double pre[2 * kUnroll];
  if (red) prefetch_partials(part_pAp, a.npatch, pre, tid, kBlock);
  // Both coarse products do not depend on alpha: entry j of (rows 3 patch ..
  // 3 patch + 2 of Ac^-1 sr, the same rows of Ac^-1 sa) is summed by wave
  // nw - 6 + j (a workgroup has at least 8 waves), every lane requesting its
  // kCoarseU columns before adding.
  constexpr int kCoarseU = 12;
  const int cj = wid - (nw - 6);
  double cv = 0.0;
  if (cj >= 0 && a.nc > 0) {
    const double* sv = cj < 3 ? sr : sa;
    const float* crow = a.Acinv + (size_t)(3 * patch + (cj < 3 ? cj : cj - 3)) * a.nc;
    for (int base = 0; base < a.nc; base += 64 * kCoarseU) {
      float cm[kCoarseU];
      double cs[kCoarseU];
#pragma unroll
      for (int u = 0; u < kCoarseU; ++u) {
        const int col = base + 64 * u + lane;
        const bool in = col < a.nc;
        cm[u] = in ? crow[col] : 0.0f;
        cs[u] = in ? sv[col] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kCoarseU; ++u) cv += (double)cm[u] * cs[u];
    }
  }
  PatchRow P;
  load_patch_row(a, patch, P);
  const int i = P.row;
  const int ii = P.live ? i : a.n - 1;
  float pv[3], av[3], xv[3], rv0[3], D[9];
  for (int c = 0; c < 3; ++c) {
    pv[c] = p[3 * ii + c];
    av[c] = Ap[3 * ii + c];
    xv[c] = x[3 * ii + c];
    rv0[c] = r[3 * ii + c];
  }
  for (int e = 0; e < 9; ++e) D[e] = a.Dinv[(size_t)e * a.n + ii];
  const float lmx = a.lmax[patch];
  if (done) return;
  // One LDS round: the p.Ap partial sums of the first four waves and the six
  // coarse entries; every thread then reads the ten values.
  double pp = red ? finish_sum_partials(part_pAp, a.npatch, pre, tid, kBlock) : 0.0;
  for (int o = 32; o > 0; o >>= 1) {
    pp += __shfl_down(pp, o, 64);
    cv += __shfl_down(cv, o, 64);
  }
  if (lane == 0) {
    if (wid < kBlock / 64) ldd[0][wid] = pp;
    if (cj >= 0) ldd[1][cj] = cv;
  }
  __syncthreads();
  double pAp = ldd[0][0];
  for (int w = 1; w < kBlock / 64; ++w) pAp += ldd[0][w];
  double v[7];
  v[0] = pAp;
  for (int j = 0; j < 6; ++j) v[1 + j] = ldd[1][j];
  if (fabs(pAp) < 1e-16) {
    if (tid == 0) {
      st->failed = 1;
      st->iters = it + 1;
      st->done = 1;
    }
    return;
  }
  const float alpha = (float)(rz_it / pAp);
  float rn[3];
  for (int c = 0; c < 3; ++c) rn[c] = P.live ? rv0[c] - alpha * av[c] : 0.0f;
  if (P.q == 0 && P.live)
    for (int c = 0; c < 3; ++c) {
      x[3 * i + c] = xv[c] + alpha * pv[c];
      r[3 * i + c] = rn[c];
    }
  float y[3];
  patch_chebyshev(a, patch, P, lmx, D, rn, lv, y);
  // w = (r.z, patch sum of the new r), with max|r|: only thread c < 5 needs
  // them, to write the partials.
  double w[4] = {0.0, 0.0, 0.0, 0.0};
  float rm = 0.0f;
  if (P.q == 0 && P.live) {
    for (int c = 0; c < 3; ++c) w[1 + c] = rn[c];
    for (int c = 0; c < 3; ++c) {
      const float zn = y[c] + (float)(v[1 + c] - (double)alpha * v[4 + c]);
      z[3 * i + c] = zn;
      w[0] += (double)rn[c] * (double)zn;
      rm = fmaxf(rm, fabsf(rn[c]));
    }
    if (rm != rm) rm = __int_as_float(0x7f800000);
  }
  const double tot = wg_sum_max_first<4, true>(w, rm, ldf, ldsf);
  if (tid == 0)
    part_rz[patch] = tot;
  else if (tid < 4)
    sr_next[3 * patch + tid - 1] = tot;
  else if (tid == 4)
    part_max[patch] = rm;
}

__global__ void k_tl_scatter(int n, const int32_t* __restrict__ perm, const float* __restrict__ xp,
                             float* __restrict__ xsol) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int o = perm[i];
  for (int c = 0; c < 3; ++c) xsol[3 * o + c] = xp[3 * i + c];
}

// ---- Coarse operator Ac = P^T A P and its inverse.  Ac is dense in FP64,
// padded to a multiple of kGjB rows with an identity block; one wave per
// coarse block (I, J) sums its fine blocks blk[ptr[k] .. ptr[k + 1]) in FP64
// (64 strided lanes and a shuffle tree, a fixed order).
constexpr int kGjB = 32;  // tile of the blocked Gauss-Jordan inversion

__global__ void k_tl_zero(int nc, int ncp, double* __restrict__ out, int* __restrict__ fail) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) *fail = 0;
  if (i >= (size_t)ncp * ncp) return;
  const int r = (int)(i / ncp), c = (int)(i % ncp);
  out[i] = (r == c && r >= nc) ? 1.0 : 0.0;
}

__global__ void __launch_bounds__(kBlock) k_tl_coarse(int ncb, int ncp, const int32_t* __restrict__ cb_ij,
                                                      const int32_t* __restrict__ cb_ptr,
                                                      const int32_t* __restrict__ cb_blk,
                                                      const float* __restrict__ A, double* __restrict__ Ac) {
  const int k = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  if (k >= ncb) return;
  double s[9];
  for (int e = 0; e < 9; ++e) s[e] = 0.0;
  for (int t = cb_ptr[k] + lane; t < cb_ptr[k + 1]; t += 64) {
    const float* B = A + 9 * (size_t)cb_blk[t];
    for (int e = 0; e < 9; ++e) s[e] += (double)B[e];
  }
  for (int e = 0; e < 9; ++e)
    for (int o = 32; o > 0; o >>= 1) s[e] += __shfl_down(s[e], o, 64);
  if (lane == 0) {
    const int I = cb_ij[2 * k], J = cb_ij[2 * k + 1];
    for (int u = 0; u < 3; ++u)
      for (int w = 0; w < 3; ++w) Ac[(size_t)(3 * I + u) * ncp + 3 * J + w] = s[3 * u + w];
  }
}

// Blocked Gauss-Jordan inversion of the symmetric positive definite Ac
// (ncp = nb kGjB, row-major).  Step k: P = X_kk^-1 (k_tl_gj_pivot), then every
// tile from X into Y (k_tl_gj_update):
//   Y_kk = P,  Y_kj = P X_kj,  Y_ik = -X_ik P,  Y_ij = X_ij - X_ik P X_kj;
// after nb steps the last buffer holds Ac^-1.  Positive definiteness makes
// every pivot positive, so no pivoting is needed; a non-positive pivot sets
// *fail, and the coarse correction is then left out.
//
// The pivot block is inverted by one wave with the block in registers: lane l
// holds column l mod 32 of rows 16 (l / 32) .. 16 (l / 32) + 15.  Each of the
// 32 pivot steps broadcasts the pivot row and column through cross-lane
// operations (the pivot by readlane; row q from this lane or from lane l xor 32;
// column q from lane q or q + 32), with no LDS and no barrier; every entry goes
// through the same operations in the same order as an in-LDS elimination (new
// T_rc = T_rc - (T_rq / T_qq) T_qc with the reciprocal formed once, row q
// scaled by it, column q by its negative, the pivot replaced by it).
static_assert(kGjB == 32, "the register layout of k_tl_gj_pivot assumes 32 x 32 pivot blocks");
__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffll), lane);
  const int hi = __builtin_amdgcn_readlane((int)(u >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__global__ void __launch_bounds__(64) k_tl_gj_pivot(int ncp, int k, const double* __restrict__ X,
                                                    double* __restrict__ P, int* __restrict__ fail) {
  const int lane = threadIdx.x;
  const int c = lane & 31, h = lane >> 5;
  double A[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) A[i] = X[(size_t)(k * kGjB + 16 * h + i) * ncp + k * kGjB + c];
#pragma unroll
  for (int q = 0; q < kGjB; ++q) {
    const int hq = q >> 4, iq = q & 15;
    const double piv = readlane_f64(A[iq], q + 32 * hq);  // T_qq
    if (!(piv > 0.0)) {
      if (lane == 0) *fail = 1;
      return;  // uniform: every lane read the same pivot
    }
    const double ip = 1.0 / piv;
    const double other = __shfl_xor(A[iq], 32, 64);
    const double rq = h == hq ? A[iq] : other;  // T_qc before the step
    double cqv[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) cqv[i] = __shfl(A[i], q + 32 * h, 64);  // T_rq before the step
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int r = 16 * h + i;
      const double cq = cqv[i];
      if (r != q && c != q) A[i] = A[i] - cq * ip * rq;
      else if (r == q && c != q) A[i] = rq * ip;
      else if (r != q) A[i] = -cq * ip;
      else A[i] = ip;
    }
  }
#pragma unroll
  for (int i = 0; i < 16; ++i) P[(16 * h + i) * kGjB + c] = A[i];
}

// One tile of step k (one workgroup per tile, 256 threads, 4 entries each).
__global__ void __launch_bounds__(kBlock) k_tl_gj_update(int nb, int ncp, int k, const double* __restrict__ X,
                                                         const double* __restrict__ P, double* __restrict__ Y,
                                                         const int* __restrict__ fail) {
  __shared__ double Ps[kGjB][kGjB + 1], Xk[kGjB][kGjB + 1], Xi[kGjB][kGjB + 1], Tm[kGjB][kGjB + 1];
  if (*fail) return;
  const int ti = blockIdx.x / nb, tj = blockIdx.x % nb;
  const int t = threadIdx.x;
  auto tile = [&](int bi, int bj, int r, int c) { return X[(size_t)(bi * kGjB + r) * ncp + bj * kGjB + c]; };
  for (int e = t; e < kGjB * kGjB; e += blockDim.x) {
    const int r = e / kGjB, c = e % kGjB;
    Ps[r][c] = P[e];
    Xk[r][c] = tile(k, tj, r, c);  // X_kj
    Xi[r][c] = tile(ti, k, r, c);  // X_ik
  }
  __syncthreads();
  // Tm = P X_kj (the new row block k)
  for (int e = t; e < kGjB * kGjB; e += blockDim.x) {
    const int r = e / kGjB, c = e % kGjB;
    double acc = 0.0;
    for (int m = 0; m < kGjB; ++m) acc += Ps[r][m] * Xk[m][c];
    Tm[r][c] = acc;
  }
  __syncthreads();
  for (int e = t; e < kGjB * kGjB; e += blockDim.x) {
    const int r = e / kGjB, c = e % kGjB;
    double v;
    if (ti == k && tj == k) {
      v = Ps[r][c];
    } else if (ti == k) {
      v = Tm[r][c];
    } else if (tj == k) {
      double acc = 0.0;
      for (int m = 0; m < kGjB; ++m) acc += Xi[r][m] * Ps[m][c];
      v = -acc;
    } else {
      double acc = 0.0;
      for (int m = 0; m < kGjB; ++m) acc += Xi[r][m] * Tm[m][c];
      v = tile(ti, tj, r, c) - acc;
    }
    Y[(size_t)(ti * kGjB + r) * ncp + tj * kGjB + c] = v;
  }
}

// The inverse (ncp-wide, row-major FP64) to a symmetric nc x nc row-major FP32
// matrix; after a failed inversion, zeros (no coarse correction, M stays
// positive definite).
__global__ void k_tl_cinv(int nc, int ncp, const double* __restrict__ X, const int* __restrict__ fail,
                          float* __restrict__ out) {
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (size_t)nc * nc) return;
  const int r = (int)(k / nc), c = (int)(k % nc);
  const double v = 0.5 * (X[(size_t)r * ncp + c] + X[(size_t)c * ncp + r]);
  out[k] = *fail ? 0.0f : (float)v;
}

}  // namespace tl
}  // namespace ando
