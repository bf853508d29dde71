// Host driver of the gfx950 cloth step and the C ABI declared in
// include/andosim_hip.h.  One context = one cloth (mesh + state + constraints)
// resident in HBM, one HIP stream.  Integrator::step (integrator.cpp:70-150)
// runs as a sequence of kernel launches; the host only reads back the scalars
// that steer control flow (gradient max-norm, line-search mask, PCG status,
// contact totals).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/andosim_hip.h"
#include "kernels.hip"
#include "model.h"
#include "pcg.hip"
#include "tl.hip"
#include "learned.hip"
#include "lbvh.hip"
#include "frame.hip"
#include "physics.hip"

namespace ando {

static thread_local std::string g_last_error;

struct Fail : std::runtime_error {
  int code;
  Fail(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
// A detection whose candidates would exceed kMaxCandidates.
struct CandidateOverflow : Fail {
  using Fail::Fail;
};

#define HIPCHK(expr)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      throw Fail(ANDO_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));         \
  } while (0)

template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void ensure(size_t count) {
    if (count <= n && p) return;
    release();
    size_t c = std::max<size_t>(count, 1);
    HIPCHK(hipMalloc(&p, c * sizeof(T)));
    n = c;
  }
  void upload(const T* src, size_t count, hipStream_t s) {
    ensure(count);
    if (count) HIPCHK(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T>& v, hipStream_t s) { upload(v.data(), v.size(), s); }
};

// Page-locked host buffer for small read-backs: a copy into it is a direct DMA,
// where a pageable destination costs the runtime a staging copy.  Grows on
// demand; each use ends with a stream synchronisation before the next.
struct PinnedStage {
  void* p = nullptr;
  size_t n = 0;
  PinnedStage() = default;
  PinnedStage(const PinnedStage&) = delete;
  PinnedStage& operator=(const PinnedStage&) = delete;
  ~PinnedStage() {
    if (p) (void)hipHostFree(p);
  }
  void* ensure(size_t bytes) {
    if (bytes <= n && p) return p;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
    const size_t b = std::max<size_t>(bytes, 256);
    HIPCHK(hipHostMalloc(&p, b, hipHostMallocDefault));
    n = b;
    return p;
  }
};

static inline int nblk(long long work, int bs = kBlock) { return (int)((work + bs - 1) / bs); }
// Workgroups of the block-row SpMV with R rows per lane group
// (R * kRowsPerPass block rows per workgroup); R = 1 is the largest count.
static inline int spmv_groups(long long n, int R = 1) { return nblk(n, ando::kRowsPerPass * R); }

struct BodyCache {
  std::vector<float> local;
  std::vector<int32_t> tris;
  DBuf<float> d_local, d_world, d_gbox, d_pose, d_bbox;
  DBuf<int32_t> d_tris;
  // Pose and culling-group size the device world vertices and boxes were made
  // for; an unchanged body is not transformed again.
  bool placed = false;
  float pose[12] = {};
  int group = 0;
};

}  // namespace ando

using namespace ando;

// Sliced-ELL metadata of one block pattern.
struct SpmvMeta {
  // Sliced-ELL layout of the pattern (k_pcg_spmv_sell): per slice of 64
  // rows its first slot position and width, per slot position its BSR block
  // (-1 = padding) and column, and room for the values.
  bool sell_ok = false;
  int slices = 0;
  int uniform_w = 0;  // the common slice width, or 0 when widths differ
  long long slots = 0;        // slot positions over all slices (padding included)
  uint64_t layout_hash = 0;   // FNV-1a of the slice offsets and widths
  DBuf<int32_t> slice_off, slice_w, sell_perm, sell_col;
  DBuf<float> sell_vals;
};

void build_sell_meta(SpmvMeta& sm, int n, const std::vector<int32_t>& row_ptr, const std::vector<int32_t>& cols,
                     hipStream_t s) {
  sm.sell_ok = false;
  if (n <= 0) return;
  const int S = (n + 63) / 64;
  std::vector<int32_t> off(S + 1, 0), w(S, 0);
  for (int sl = 0; sl < S; ++sl) {
    int W = 0;
    for (int i = sl * 64; i < std::min(n, sl * 64 + 64); ++i) W = std::max(W, row_ptr[i + 1] - row_ptr[i]);
    w[sl] = W;
    const long long next = (long long)off[sl] + 64LL * W;
    if (9 * next >= (1LL << 31)) return;
    off[sl + 1] = (int32_t)next;
  }
  const size_t P = (size_t)off[S];
  std::vector<int32_t> perm(P, -1), col(P, 0);
  for (int sl = 0; sl < S; ++sl)
    for (int l = 0; l < 64; ++l) {
      const int i = sl * 64 + l;
      for (int k = 0; k < w[sl]; ++k) {
        const size_t pos = (size_t)off[sl] + 64 * (size_t)k + l;
        if (i < n && row_ptr[i] + k < row_ptr[i + 1]) {
          perm[pos] = row_ptr[i] + k;
          col[pos] = cols[row_ptr[i] + k];
        } else {
          col[pos] = i < n ? i : 0;
        }
      }
    }
  sm.slice_off.upload(off, s);
  sm.slice_w.upload(w, s);
  sm.sell_perm.upload(perm, s);
  sm.sell_col.upload(col, s);
  sm.sell_vals.ensure(9 * std::max<size_t>(P, 1));
  sm.uniform_w = w[0];
  for (int sl = 1; sl < S; ++sl)
    if (w[sl] != w[0]) sm.uniform_w = 0;
  sm.slices = S;
  sm.slots = (long long)P;
  uint64_t h = 1469598103934665603ull;
  auto mix = [&h](int32_t v) {
    for (int b = 0; b < 4; ++b) {
      h ^= (uint64_t)((uint32_t)v >> (8 * b) & 0xffu);
      h *= 1099511628211ull;
    }
  };
  for (int sl = 0; sl < S; ++sl) {
    mix(off[sl]);
    mix(w[sl]);
  }
  sm.layout_hash = h;
  sm.sell_ok = true;
}

void build_spmv_meta(SpmvMeta& sm, int n, const std::vector<int32_t>& row_ptr, const std::vector<int32_t>& cols,
                     hipStream_t s) {
  build_sell_meta(sm, n, row_ptr, cols, s);
}

// Two-level Schwarz PCG (tl.hip): the pattern's rows renumbered (Morton order
// of the rest positions for a mesh), the sliced layout over the renumbered rows,
// the coarse pattern (per patch pair, its fine blocks in ascending order), and
// the solve's buffers.
struct TlMeta {
  bool ok = false;
  int n = 0, pr = 256, npatch = 0, nc = 0, ncb = 0;
  long long nb = 0;
  DBuf<int32_t> perm, iperm;
  SpmvMeta sm;  // sell_col holds renumbered columns
  DBuf<int32_t> cb_ij, cb_ptr, cb_blk;
  int ncp = 0;                // nc padded to a multiple of the inversion tile
  DBuf<double> Ac, Acw, gjP;  // coarse operator, inversion work buffer, pivot block
  DBuf<float> Acinv;
  DBuf<int32_t> info;         // inversion failed
  bool coarse_ready = false;  // Acinv holds the inverse of an earlier solve's coarse operator
  long long solves = 0;       // two-level solves on this layout
  DBuf<float> Dinv, lmax, x, r, z, p0, p1, Ap, qf;
  DBuf<double> sa, sr0, sr1;  // per patch: sums of A p and of r (3 each)
  DBuf<double> part_rz, part_pAp;
  DBuf<float> part_max, part_bmax;
  DBuf<int32_t> cpr, ccolp;   // coupling blocks in the permuted numbering (TlArgs::cpr, ccolp)
};

// Coarse unknowns stay dense: at most this many patches get a coarse level.
constexpr int kTlMaxCoarsePatches = 1024;

// Vertex order by 30-bit Morton code of the positions (10 bits per axis over
// the bounding cube), ties by index.
std::vector<int32_t> morton_order(const float* v, int n) {
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], (double)v[3 * (size_t)i + k]);
      hi[k] = std::max(hi[k], (double)v[3 * (size_t)i + k]);
    }
  const double ext = std::max(std::max(hi[0] - lo[0], hi[1] - lo[1]), std::max(hi[2] - lo[2], 1e-30));
  auto spread = [](uint32_t x) {
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
  };
  std::vector<std::pair<uint32_t, int32_t>> key(n);
  for (int i = 0; i < n; ++i) {
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
      const double t = ((double)v[3 * (size_t)i + k] - lo[k]) / ext * 1023.0;
      q[k] = (uint32_t)std::min(1023.0, std::max(0.0, t));
    }
    key[i] = {spread(q[0]) | (spread(q[1]) << 1) | (spread(q[2]) << 2), i};
  }
  std::sort(key.begin(), key.end());
  std::vector<int32_t> perm(n);
  for (int i = 0; i < n; ++i) perm[i] = key[i].second;
  return perm;
}

// Builds T for the pattern (row_ptr, cols) with the row order perm (new ->
// old): sliced layout of the renumbered rows (slot k of a row is its k-th BSR
// block, columns renumbered), the coarse pattern and the solve's buffers.
void build_tl_meta(TlMeta& T, int n, const std::vector<int32_t>& row_ptr, const std::vector<int32_t>& cols,
                   const std::vector<int32_t>& perm, hipStream_t s) {
  using namespace ando::tl;
  T.ok = false;
  T.coarse_ready = false;
  T.solves = 0;
  std::vector<int32_t> iperm(n);
  for (int i = 0; i < n; ++i) iperm[perm[i]] = i;
  const int S = (n + 63) / 64;
  std::vector<int32_t> off(S + 1, 0), w(S, 0);
  for (int sl = 0; sl < S; ++sl) {
    int W = 0;
    for (int i = sl * 64; i < std::min(n, sl * 64 + 64); ++i) W = std::max(W, row_ptr[perm[i] + 1] - row_ptr[perm[i]]);
    w[sl] = W;
    const long long next = (long long)off[sl] + 64LL * W;
    if (9 * next >= (1LL << 31)) return;
    off[sl + 1] = (int32_t)next;
  }
  const size_t P = (size_t)off[S];
  std::vector<int32_t> sperm(P, -1), scol(P, 0);
  for (int sl = 0; sl < S; ++sl)
    for (int l = 0; l < 64; ++l) {
      const int i = sl * 64 + l;
      for (int k = 0; k < w[sl]; ++k) {
        const size_t pos = (size_t)off[sl] + 64 * (size_t)k + l;
        if (i < n && row_ptr[perm[i]] + k < row_ptr[perm[i] + 1]) {
          const int b = row_ptr[perm[i]] + k;
          sperm[pos] = b;
          scol[pos] = iperm[cols[b]];
        } else {
          scol[pos] = i < n ? i : 0;
        }
      }
    }
  SpmvMeta& sm = T.sm;
  sm.slice_off.upload(off, s);
  sm.slice_w.upload(w, s);
  sm.sell_perm.upload(sperm, s);
  sm.sell_col.upload(scol, s);
  sm.sell_vals.ensure(9 * std::max<size_t>(P, 1));
  sm.uniform_w = w[0];
  for (int sl = 1; sl < S; ++sl)
    if (w[sl] != w[0]) sm.uniform_w = 0;
  sm.slices = S;
  sm.slots = (long long)P;
  sm.sell_ok = true;
  T.n = n;
  T.nb = row_ptr[n];
  T.pr = patch_rows(n);
  // ANDO_TL_PATCH_ROWS: rows per patch (256 to 512; the kernels need at least
  // eight waves per workgroup and at most kMaxPatchRows rows).
  if (const char* w = std::getenv("ANDO_TL_PATCH_ROWS"))
    T.pr = std::min(ando::tl::kMaxPatchRows, std::max(256, std::atoi(w)));
  T.npatch = (n + T.pr - 1) / T.pr;
  T.perm.upload(perm, s);
  T.iperm.upload(iperm, s);
  // Coarse pattern: the fine blocks of each patch pair (I, J), in block order.
  T.nc = 0;
  T.ncb = 0;
  if (T.npatch <= kTlMaxCoarsePatches) {
    const long long nb = row_ptr[n];
    std::vector<std::pair<long long, int32_t>> kb((size_t)nb);
    for (int o = 0; o < n; ++o)
      for (int b = row_ptr[o]; b < row_ptr[o + 1]; ++b)
        kb[(size_t)b] = {(long long)(iperm[o] / T.pr) * T.npatch + iperm[cols[b]] / T.pr, b};
    std::sort(kb.begin(), kb.end());
    std::vector<int32_t> ij, ptr, blk((size_t)nb);
    for (size_t t = 0; t < kb.size(); ++t) {
      if (t == 0 || kb[t].first != kb[t - 1].first) {
        ij.push_back((int32_t)(kb[t].first / T.npatch));
        ij.push_back((int32_t)(kb[t].first % T.npatch));
        ptr.push_back((int32_t)t);
      }
      blk[t] = kb[t].second;
    }
    ptr.push_back((int32_t)nb);
    T.ncb = (int)(ptr.size() - 1);
    T.cb_ij.upload(ij, s);
    T.cb_ptr.upload(ptr, s);
    T.cb_blk.upload(blk, s);
    T.nc = 3 * T.npatch;
    T.ncp = ando::tl::kGjB * ((T.nc + ando::tl::kGjB - 1) / ando::tl::kGjB);
    T.Ac.ensure((size_t)T.ncp * T.ncp);
    T.Acw.ensure((size_t)T.ncp * T.ncp);
    T.gjP.ensure((size_t)ando::tl::kGjB * ando::tl::kGjB);
    T.Acinv.ensure((size_t)T.nc * T.nc);
    T.info.ensure(1);
  }
  const size_t n3 = 3 * (size_t)n;
  T.x.ensure(n3); T.r.ensure(n3); T.z.ensure(n3); T.p0.ensure(n3); T.p1.ensure(n3); T.Ap.ensure(n3);
  T.qf.ensure(n3);
  T.Dinv.ensure(9 * (size_t)n);
  T.sa.ensure(3 * (size_t)T.npatch);
  T.sr0.ensure(3 * (size_t)T.npatch);
  T.sr1.ensure(3 * (size_t)T.npatch);
  T.lmax.ensure((size_t)T.npatch);
  T.part_rz.ensure((size_t)T.npatch + 1);
  T.part_pAp.ensure((size_t)T.npatch + 1);
  T.part_max.ensure((size_t)T.npatch + 1);
  T.ok = true;
}

#include "learned_host.hip"

// One linear BVH (triangles or edges) and its build scratch.
struct LbvhTree {
  int P = 0;
  // The primitives (index array, vertices per primitive) whose sorted order
  // and hierarchy the tree holds, and the refits made on them since.
  const int32_t* topo_prims = nullptr;
  int topo_kind = 0, refits = 0;
  DBuf<float> boxes, table, cbox, bpart, bounds;
  DBuf<uint32_t> keys, keys2;
  DBuf<int32_t> vals, vals2, counts, offsets, left, right, first, last, scan_sums, scan_total;
};

struct ando_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int n = 0, m = 0, nb = 0;
  ando::Material mat;
  float mu = 0.0f;
  bool have_mesh = false, have_state = false;
  // Host copies of topology.
  std::vector<int32_t> tris, row_ptr, cols, diag, block_row;
  // Device: mesh and topology.
  DBuf<float> mverts, dm, area, ec, fe, egrad;
  DBuf<int32_t> d_tris, d_row_ptr, d_cols, d_diag, d_block_row, blk_cptr, blk_contrib, v2f_ptr, v2f, face_blk;
  // Device: state.
  DBuf<float> x, v, masses, xold, xhat, g, d, xprev;
  bool have_prev = false;
  DBuf<BodyState> body_state;
  // Device: system.
  DBuf<float> Hconst, Helast3, H;
  bool const_valid = false;
  float const_dt = 0.0f;
  // Strain scratch.
  DBuf<int32_t> sflags;
  DBuf<float> sgrad, shess;
  // Pins and walls.
  int num_pins = 0;
  std::vector<int32_t> pin_vertex;
  std::vector<float> pin_target;
  DBuf<int32_t> pin_ptr;
  DBuf<float> pin_t;
  int num_walls = 0;
  std::vector<float> wall_n, wall_off, wall_gap;
  DBuf<float> d_wall_n, d_wall_off, d_wall_gap, kw;
  bool kw_valid = false;
  float kw_wall_gap = -1.0f, kw_min_gap = -1.0f;
  // Rigid contacts.
  std::vector<BodyCache> bodies;
  DBuf<int32_t> rc_count, rc_ptr, scan_sums, scan_total;
  DBuf<ando::RigidContact> rc;
  DBuf<int32_t> rc_hits;
  DBuf<double> rw_part, rw_out;  // rigid coupling: per-workgroup and per-body force/torque sums  // per vertex, the first kHitCache (body << 24 | triangle) hits of the count pass
  int rc_total = 0;
  // PCG.
  DBuf<float> r, z, pbuf0, pbuf1, Ap, Pinv;
  DBuf<double> part_rz, part_pAp, rz_hist;
  DBuf<float> part_max, part_bmax;
  DBuf<ando::PcgStatus> pcg_status;
  // Small readbacks.
  DBuf<uint32_t> scalars;  // [0] gradient max bits, [1] line-search mask, [2] strain active count
  PcgStatus* status_pinned = nullptr;  // host copy of the PCG status (pinned)
  uint32_t* sact_pinned = nullptr;     // host copy of the strain-active count (pinned)
  PinnedStage hstage;                  // small read-backs (read_small)
  PinnedStage bstage;                  // rigid body states after the coupled update, read at the step's end
  int bodies_pending = 0;              // bodies in bstage still to be copied into the caller's records
  hipEvent_t status_ev = nullptr;
  int pcg_chunk = 8;
  int pcg_chunk_max = 256;  // ANDO_PCG_CHUNK_MAX: largest chunk (chunks double from pcg_chunk)
  int rigid_group = kGroup;  // rigid triangles per culling box
  ando_stats stats{};
  int fp64_near = 1;
  int kernel_timing = 0;
  int spmv_rowsets = 3;    // block rows per lane group of the row SpMV (1..4)
  int spmv_sell = 1;       // 1: sliced-ELL SpMV (values copied once per solve), 0: BSR kernels
  int sell_lanes = 2;      // ANDO_SELL_LANES: lanes per row of the sliced-ELL SpMV (1, 2 or 4)
  int sell_wg_slices = 1;  // ANDO_SELL_WG_SLICES: slices per workgroup with two lanes per row (0: a wave, 1, 2, 4)
  int pcg_vec_block = kBlock;  // ANDO_PCG_VEC_BLOCK: threads per workgroup of the PCG vector kernels (64, 128, 256)
  int pcg_autotune = 0;    // ANDO_PCG_AUTOTUNE: 1 pick the sliced-ELL launch shape by timing, 2 also print the table
  int tuned_n = -1;        // block rows the launch shape was tuned (or looked up) for
  int autotuned = 0;       // the launch shape in use came from the autotuner
  bool tuning = false;     // inside a trial solve of the autotuner
  int spmv_env = -1;       // ANDO_PCG_SPMV: -1 unset, 0 row, 2 sell (overrides the mode flag)
  SpmvMeta spmv_meta;      // for the mesh's block pattern
  std::mutex call_mutex;   // held by every entry point for the whole call (CtxLock)
  hipEvent_t batch_ev = nullptr;  // a batched solve's set-up is done on the context's stream
  // Corrected contact mode (self point-triangle and edge-edge through the LBVH).
  int corrected = 0;
  int static_rigids = 0;   // rigid bodies are fixed colliders: no coupling, no body update
  int additive_ccd = 0;    // corrected mode: additive CCD instead of the 10-sample test in the line search
  // Corrected physics (ANDO_MODE_CORRECTED_PHYSICS; physics.hip).
  int cp = 0;
  float pin_k = 1e4f;      // attracting pin stiffness in units of m / dt^2
  DBuf<float> fgk, fpi;    // per face: 2k B B^T and I - n n^T (6 floats each)
  DBuf<float> Hiso3, Hbase;  // per block: isotropic elastic stiffness (3), plus mass / dt^2 (9)
  DBuf<float> vbox, ed;    // per vertex: candidate boxes (6), trial motion 1.25 d (3)
  DBuf<float> cfe, cegrad; // membrane gradient at x: per face (9), per vertex (3)
  ando::DetectBodies cpD{};  // the step's rigid bodies (they do not move during a step)
  double mean_edge = 0.0;    // mean rest edge length (FP64, edge order)
  DBuf<unsigned long long> motion2;  // line search: longest |d_i|^2 and |x_hat_i - x_i|^2 (double bits)
  DBuf<uint8_t> hold0, hold1;        // line search: vertices held this iteration (double-buffered passes)
  DBuf<uint32_t> hold_changed;
  float pt_box = 1e-4f, keep_gap = 0.01f;
  std::vector<int32_t> edges_h;
  DBuf<int32_t> d_edges;
  LbvhTree tri_tree, edge_tree;
  DBuf<ando::SelfContact> sc;
  DBuf<int32_t> sc_key, sc_count, sc_ptr, sc_sums, sc_total_d;
  DBuf<int32_t> sc_cache;  // per query, the first kQueryCache hits of the count pass
  DBuf<uint32_t> sc_err;
  DBuf<unsigned long long> cand64;  // 64-bit total of a detection's hit counts
  DBuf<ando::SelfContact> sc_tmp;   // segmented sort of long hit lists: gathered records
  DBuf<int32_t> sc_ktmp;            // and their keys
  int sc_pt = 0, sc_total = 0;
  DBuf<ando::SelfTerm> sterm;
  DBuf<uint32_t> sv_keys, sv_keys2;
  DBuf<int32_t> sv_vals, sv_vals2, sv_counts, sv_offsets, sv_sums, sv_total, sv_hist, sv_ptr;
  DBuf<int32_t> ccol;
  DBuf<float> cval;
  DBuf<int32_t> crow, crow_flag, crow_pos;  // rows with coupling blocks (ncrow of them)
  int ncrow = 0;
  int ncblk = 0;                            // coupling blocks (3 per vertex-view entry)
  DBuf<float> Ap2;                          // their coupling products during a solve
  DBuf<int32_t> pt_of;     // faithful one-triangle mesh: PT contact of each point vertex, or -1
  bool sv_ready = false;
  float barrier_tol = 1e-12f;
  int learned = 0;         // learned preconditioner + warm start in the step's PCG
  nnh::Net nn;
  nnh::Work nnw;
  // Per-frame consumers: partials, results, an uploaded contact list, and the
  // generation of the contact table (bumped by every detection) that the last
  // ando_compute_contacts listed.
  DBuf<ando::EnergyAcc> fe_part;
  DBuf<ando::MetricAcc> fm_part;
  DBuf<ando_energy> fe_out;
  DBuf<ando_collision_metrics> fm_out;
  DBuf<ando_contact> fm_list;
  DBuf<ando::ExtremaAcc> fx_part;
  DBuf<float> fx_out;
  uint64_t det_gen = 0, listed_gen = ~0ull;
  std::vector<hipEvent_t> events;  // pool: start/stop pairs for timed launches
  // Two-level Schwarz PCG (ANDO_MODE_TWO_LEVEL; tl.hip).
  int two_level = 0;
  int tl_cheb = 9;          // ANDO_TL_CHEB: Chebyshev degree of the patch solves (odd)
  float tl_ratio = 5.0f;    // ANDO_TL_RATIO: lmax / lmin of the Chebyshev interval
  int tl_coarse = 1;        // ANDO_TL_COARSE=0: no coarse level (patch polynomials only)
  int tl_coarse_every = 2;  // ANDO_TL_COARSE_EVERY=k: the coarse inverse is rebuilt on every k-th solve
  int tl_warm = 1;          // ANDO_TL_WARM=0: every Newton solve starts from x0 = 0 (else from the scaled last direction)
  int lbvh_refit = 1;       // ANDO_LBVH_REFIT=0: every detection builds its trees in full
  std::vector<int32_t> morton;  // the mesh's vertices in Morton order of the rest positions
  TlMeta tl;                // for the mesh's block pattern
  ~ando_ctx() {
    for (hipEvent_t e : events) (void)hipEventDestroy(e);
    if (status_pinned) (void)hipHostFree(status_pinned);
    if (sact_pinned) (void)hipHostFree(sact_pinned);
    if (status_ev) (void)hipEventDestroy(status_ev);
  }
};

namespace {

void set_error(const std::string& s) { ando::g_last_error = s; }

template <class F>
int guarded(F&& fn) {
  try {
    fn();
    return ANDO_OK;
  } catch (const Fail& f) {
    set_error(f.what());
    return f.code;
  } catch (const std::invalid_argument& e) {
    set_error(e.what());
    return ANDO_EINVAL;
  } catch (const std::exception& e) {
    set_error(e.what());
    return ANDO_EDEVICE;
  }
}

// Calls on one context are serialised: the Python module releases the GIL
// while a step runs, so another thread can reach the same context meanwhile.
// Every entry point that takes a context holds its mutex for the whole call;
// distinct contexts run concurrently (ando_step_batch relies on that).
struct CtxLock {
  std::unique_lock<std::mutex> lock;
  explicit CtxLock(ando_ctx* c) {
    if (c) lock = std::unique_lock<std::mutex>(c->call_mutex);
  }
};

void need(bool cond, int code, const char* msg) {
  if (!cond) throw Fail(code, msg);
}

void sync(ando_ctx* c) { HIPCHK(hipStreamSynchronize(c->stream)); }

// Copies count values from device memory through the context's pinned stage
// and waits for them (everything queued before on the stream finishes first).
template <class T>
void read_small(ando_ctx* c, T* dst, const T* src, size_t count) {
  T* h = static_cast<T*>(c->hstage.ensure(count * sizeof(T)));
  HIPCHK(hipMemcpyAsync(h, src, count * sizeof(T), hipMemcpyDeviceToHost, c->stream));
  sync(c);
  std::memcpy(dst, h, count * sizeof(T));
}

// Block pattern, gather lists and per-face block slots.
void build_topology(ando_ctx* c, const std::vector<int32_t>& edges) {
  const int n = c->n, m = c->m;
  std::vector<std::vector<int32_t>> nbr(n);
  for (int i = 0; i < n; ++i) nbr[i].push_back(i);
  for (size_t e = 0; e + 1 < edges.size(); e += 2) {
    nbr[edges[e]].push_back(edges[e + 1]);
    nbr[edges[e + 1]].push_back(edges[e]);
  }
  c->row_ptr.assign(n + 1, 0);
  c->cols.clear();
  c->diag.assign(n, -1);
  for (int i = 0; i < n; ++i) {
    auto& l = nbr[i];
    std::sort(l.begin(), l.end());
    l.erase(std::unique(l.begin(), l.end()), l.end());
    for (int j : l) {
      if (j == i) c->diag[i] = (int32_t)c->cols.size();
      c->cols.push_back(j);
    }
    c->row_ptr[i + 1] = (int32_t)c->cols.size();
  }
  c->nb = (int)c->cols.size();
  c->block_row.resize(c->nb);
  for (int i = 0; i < n; ++i)
    for (int b = c->row_ptr[i]; b < c->row_ptr[i + 1]; ++b) c->block_row[b] = i;
  auto blk = [&](int i, int j) {
    auto b0 = c->cols.begin() + c->row_ptr[i], b1 = c->cols.begin() + c->row_ptr[i + 1];
    auto it = std::lower_bound(b0, b1, j);
    return (int32_t)(it - c->cols.begin());
  };
  std::vector<int32_t> face_blk(9 * (size_t)m), cnt(c->nb + 1, 0), vcnt(n + 1, 0);
  for (int f = 0; f < m; ++f)
    for (int a = 0; a < 3; ++a) {
      ++vcnt[c->tris[3 * f + a] + 1];
      for (int b = 0; b < 3; ++b) {
        int32_t bi = blk(c->tris[3 * f + a], c->tris[3 * f + b]);
        face_blk[9 * (size_t)f + 3 * a + b] = bi;
        ++cnt[bi + 1];
      }
    }
  for (int b = 0; b < c->nb; ++b) cnt[b + 1] += cnt[b];
  for (int i = 0; i < n; ++i) vcnt[i + 1] += vcnt[i];
  std::vector<int32_t> contrib(cnt[c->nb]), v2f(vcnt[n]);
  std::vector<int32_t> fill(cnt.begin(), cnt.end() - 1), vfill(vcnt.begin(), vcnt.end() - 1);
  for (int f = 0; f < m; ++f)
    for (int a = 0; a < 3; ++a) {
      v2f[vfill[c->tris[3 * f + a]]++] = 3 * f + a;
      for (int b = 0; b < 3; ++b) contrib[fill[face_blk[9 * (size_t)f + 3 * a + b]]++] = 9 * f + 3 * a + b;
    }
  hipStream_t s = c->stream;
  c->d_row_ptr.upload(c->row_ptr, s);
  c->d_cols.upload(c->cols, s);
  c->d_diag.upload(c->diag, s);
  c->d_block_row.upload(c->block_row, s);
  c->blk_cptr.upload(cnt, s);
  c->blk_contrib.upload(contrib, s);
  c->v2f_ptr.upload(vcnt, s);
  c->v2f.upload(v2f, s);
  c->face_blk.upload(face_blk, s);
  build_spmv_meta(c->spmv_meta, n, c->row_ptr, c->cols, s);
  sync(c);
}

void alloc_state(ando_ctx* c) {
  size_t n3 = 3 * (size_t)c->n;
  c->x.ensure(n3); c->v.ensure(n3); c->masses.ensure(c->n); c->xold.ensure(n3); c->xhat.ensure(n3);
  c->g.ensure(n3); c->d.ensure(n3); c->xprev.ensure(n3);
  c->have_prev = false;
  c->Hconst.ensure(9 * (size_t)c->nb); c->Helast3.ensure(3 * (size_t)c->nb); c->H.ensure(9 * (size_t)c->nb);
  c->sflags.ensure(c->m); c->sgrad.ensure(18 * (size_t)c->m); c->shess.ensure(20 * (size_t)c->m);
  c->r.ensure(n3); c->z.ensure(n3); c->pbuf0.ensure(n3); c->pbuf1.ensure(n3); c->Ap.ensure(n3);
  c->Pinv.ensure(9 * (size_t)c->n);
  c->part_rz.ensure(nblk(c->n) + 1); c->part_max.ensure(nblk(c->n) + 1); c->part_pAp.ensure(spmv_groups(c->n) + 1);
  c->rc_count.ensure(c->n + 1); c->rc_ptr.ensure(c->n + 1);
  c->scan_sums.ensure((c->n + 1023) / 1024 + 1); c->scan_total.ensure(1);
}

StepParams step_params(ando_ctx* c, const ando_params& p) {
  StepParams P;
  P.dt = p.dt;
  P.gap_max = p.contact_gap_max;
  P.normal_eps = p.contact_normal_epsilon;
  P.barrier_tol = p.barrier_tolerance;
  P.min_gap = p.min_gap;
  P.num_walls = c->num_walls;
  P.strain = p.enable_strain_limiting ? 1 : 0;
  P.fp64_near = c->fp64_near;
  P.friction = (p.enable_friction && (double)p.friction_mu > 0.0) ? 1 : 0;
  P.mu = p.friction_mu;
  P.friction_eps = p.friction_epsilon;
  P.friction_tangent = p.friction_tangent_threshold;
  P.cp = c->cp;
  P.pin_k = c->pin_k;
  return P;
}

void ensure_const(ando_ctx* c, const ando_params& p) {
  hipStream_t s = c->stream;
  if (!c->const_valid || c->const_dt != p.dt) {
    float dtinv2 = 1.0f / (p.dt * p.dt);
    hipLaunchKernelGGL(k_fold_const, dim3(c->n), dim3(64), 0, s, c->n, c->d_row_ptr.p, c->d_cols.p, c->blk_cptr.p,
                       c->blk_contrib.p, c->ec.p, c->masses.p, dtinv2, c->Hconst.p, c->Helast3.p);
    if (c->cp) {
      c->fgk.ensure(6 * (size_t)std::max(c->m, 1));
      c->fpi.ensure(6 * (size_t)std::max(c->m, 1));
      c->Hiso3.ensure(3 * (size_t)c->nb);
      c->Hbase.ensure(9 * (size_t)c->nb);
      if (c->m)
        hipLaunchKernelGGL(ando::cp::k_face_gk, dim3(nblk(c->m)), dim3(kBlock), 0, s, c->m, (const float*)c->dm.p,
                           (const float*)c->area.p, c->mat.thickness, c->mu, c->fgk.p);
      hipLaunchKernelGGL(ando::cp::k_iso_fold, dim3(c->n), dim3(64), 0, s, c->n, (const int32_t*)c->d_row_ptr.p,
                         (const int32_t*)c->d_cols.p, (const int32_t*)c->blk_cptr.p,
                         (const int32_t*)c->blk_contrib.p, (const float*)c->fgk.p, (const float*)c->masses.p, dtinv2,
                         c->Hbase.p, c->Hiso3.p);
    }
    HIPCHK(hipGetLastError());
    c->const_valid = true;
    c->const_dt = p.dt;
    c->kw_valid = false;
  }
  if (c->num_walls && (!c->kw_valid || c->kw_wall_gap != p.wall_gap || c->kw_min_gap != p.min_gap)) {
    c->kw.ensure((size_t)c->num_walls * c->n);
    hipLaunchKernelGGL(k_wall_stiffness, dim3(nblk(c->n)), dim3(kBlock), 0, s, c->n, c->num_walls, c->d_diag.p,
                       c->cp ? c->Hbase.p : c->Hconst.p, c->masses.p, c->d_wall_n.p, p.wall_gap, p.min_gap, c->kw.p);
    HIPCHK(hipGetLastError());
    c->kw_valid = true;
    c->kw_wall_gap = p.wall_gap;
    c->kw_min_gap = p.min_gap;
  }
}

void upload_bodies(ando_ctx* c, const ando_rigid_body* bodies, int nbod) {
  need(nbod >= 0 && nbod <= kMaxBodies, ANDO_EINVAL, "too many rigid bodies (limit 16)");
  for (int b = 0; b < nbod; ++b)
    need(bodies[b].num_triangles >= 0 && bodies[b].num_triangles < (1 << 24), ANDO_EINVAL,
         "a rigid body has too many triangles (limit 2^24 - 1)");
  c->bodies.resize(nbod);
  hipStream_t s = c->stream;
  for (int b = 0; b < nbod; ++b) {
    const ando_rigid_body& B = bodies[b];
    BodyCache& C = c->bodies[b];
    need(B.num_vertices >= 0 && B.num_triangles >= 0, ANDO_EINVAL, "negative rigid body size");
    size_t nv3 = 3 * (size_t)B.num_vertices, nt3 = 3 * (size_t)B.num_triangles;
    for (size_t t = 0; t < nt3; ++t)
      need(B.triangles[t] >= 0 && B.triangles[t] < B.num_vertices, ANDO_EINVAL, "rigid triangle index out of range");
    bool same = C.local.size() == nv3 && C.tris.size() == nt3 &&
                (nv3 == 0 || std::memcmp(C.local.data(), B.local_vertices, nv3 * sizeof(float)) == 0) &&
                (nt3 == 0 || std::memcmp(C.tris.data(), B.triangles, nt3 * sizeof(int32_t)) == 0);
    if (!same) {
      C.local.assign(B.local_vertices, B.local_vertices + nv3);
      C.tris.assign(B.triangles, B.triangles + nt3);
      C.d_local.upload(C.local, s);
      C.d_tris.upload(C.tris, s);
      C.d_world.ensure(nv3);
      C.placed = false;
    }
    float pose[12];
    std::memcpy(pose, B.position, 3 * sizeof(float));
    std::memcpy(pose + 3, B.rotation, 9 * sizeof(float));
    if (C.placed && C.group == c->rigid_group && std::memcmp(pose, C.pose, sizeof(pose)) == 0) continue;
    C.d_gbox.ensure(6 * (size_t)((B.num_triangles + c->rigid_group - 1) / c->rigid_group + 1));
    C.d_pose.upload(pose, 12, s);
    std::memcpy(C.pose, pose, sizeof(pose));
    C.group = c->rigid_group;
    C.placed = true;
    if (B.num_vertices)
      hipLaunchKernelGGL(k_rigid_world, dim3(nblk(B.num_vertices)), dim3(kBlock), 0, s, B.num_vertices,
                         C.d_local.p, C.d_pose.p, C.d_world.p);
    int ng = (B.num_triangles + c->rigid_group - 1) / c->rigid_group;
    if (ng)
      hipLaunchKernelGGL(k_rigid_groups, dim3(nblk(ng)), dim3(kBlock), 0, s, B.num_triangles, c->rigid_group,
                         C.d_tris.p, C.d_world.p, C.d_gbox.p);
    C.d_bbox.ensure(6);
    hipLaunchKernelGGL(k_rigid_bbox, dim3(1), dim3(kBlock), 0, s, ng, C.d_gbox.p, C.d_bbox.p);
    HIPCHK(hipGetLastError());
  }
}

DetectBodies detect_bodies(ando_ctx* c, const ando_rigid_body* bodies) {
  DetectBodies D{};
  D.count = (int)c->bodies.size();
  for (int b = 0; b < D.count; ++b) {
    D.b[b].nv = bodies[b].num_vertices;
    D.b[b].nt = bodies[b].num_triangles;
    D.b[b].grp = c->rigid_group;
    D.b[b].world = c->bodies[b].d_world.p;
    D.b[b].tris = c->bodies[b].d_tris.p;
    D.b[b].gbox = c->bodies[b].d_gbox.p;
    D.b[b].bbox = c->bodies[b].d_bbox.p;
  }
  return D;
}

// Rigid contacts at the current positions, ordered (vertex, body, triangle).
void detect_rigid(ando_ctx* c, const ando_rigid_body* bodies, int nbod, float dt) {
  ++c->det_gen;
  c->rc_total = 0;
  if (nbod == 0) return;
  hipStream_t s = c->stream;
  DetectBodies D = detect_bodies(c, bodies);
  c->rc_hits.ensure((size_t)c->n * kHitCache);
  hipLaunchKernelGGL(k_rigid_detect, dim3(nblk(c->n)), dim3(kBlock), 0, s, c->n, 0, c->x.p, D, c->rc_ptr.p,
                     c->rc_count.p, c->rc.p, c->masses.p, c->d_diag.p, c->Helast3.p, dt, c->rc_hits.p);
  int nbs = (c->n + 1023) / 1024;
  hipLaunchKernelGGL(k_scan_local, dim3(nbs), dim3(kBlock), 0, s, c->n, c->rc_count.p, c->rc_ptr.p,
                     c->scan_sums.p);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kBlock), 0, s, nbs, c->scan_sums.p, c->scan_total.p);
  hipLaunchKernelGGL(k_scan_add, dim3(nblk(c->n + 1)), dim3(kBlock), 0, s, c->n, c->rc_ptr.p, c->scan_sums.p,
                     c->scan_total.p);
  HIPCHK(hipGetLastError());
  int32_t total = 0;
  read_small(c, &total, (const int32_t*)c->scan_total.p, 1);
  c->rc_total = total;
  c->rc.ensure((size_t)total + 1);
  if (total)
    hipLaunchKernelGGL(k_rigid_detect, dim3(nblk(c->n)), dim3(kBlock), 0, s, c->n, 1, c->x.p, D, c->rc_ptr.p,
                       c->rc_count.p, c->rc.p, c->masses.p, c->d_diag.p, c->Helast3.p, dt, c->rc_hits.p);
  HIPCHK(hipGetLastError());
}

// Exclusive scan of n int32 (out has n + 1 entries); returns the total, read
// back with a stream synchronisation unless read_total is false (then -1).
int32_t exclusive_scan(ando_ctx* c, int n, const int32_t* in, int32_t* out, DBuf<int32_t>& sums,
                       DBuf<int32_t>& total, bool read_total = true) {
  hipStream_t s = c->stream;
  const int nbs = std::max(1, (n + 1023) / 1024);
  sums.ensure((size_t)nbs + 1);
  total.ensure(1);
  hipLaunchKernelGGL(k_scan_local, dim3(nbs), dim3(kBlock), 0, s, n, in, out, sums.p);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kBlock), 0, s, nbs, sums.p, total.p);
  hipLaunchKernelGGL(k_scan_add, dim3(nblk(n + 1)), dim3(kBlock), 0, s, n, out, (const int32_t*)sums.p,
                     (const int32_t*)total.p);
  HIPCHK(hipGetLastError());
  if (!read_total) return -1;
  int32_t t = 0;
  read_small(c, &t, (const int32_t*)total.p, 1);
  return t;
}

// Stable LSD radix sort of P (key, value) pairs, 8-bit digits; the result is
// left in keys / vals (the second pair of buffers is scratch).
void radix_sort(ando_ctx* c, int P, DBuf<uint32_t>& keys, DBuf<int32_t>& vals, DBuf<uint32_t>& keys2,
                DBuf<int32_t>& vals2, DBuf<int32_t>& counts, DBuf<int32_t>& offsets, DBuf<int32_t>& sums,
                DBuf<int32_t>& total) {
  using namespace ando::lbvh;
  hipStream_t s = c->stream;
  const int G = nblk(P, 256), nc = 256 * G;
  keys2.ensure(P);
  vals2.ensure(P);
  counts.ensure(nc);
  offsets.ensure((size_t)nc + 1);
  for (int shift = 0; shift < 32; shift += 8) {
    hipLaunchKernelGGL(k_radix_count, dim3(G), dim3(256), 0, s, P, (const uint32_t*)keys.p, shift, counts.p);
    exclusive_scan(c, nc, counts.p, offsets.p, sums, total, false);
    hipLaunchKernelGGL(k_radix_scatter, dim3(G), dim3(256), 0, s, P, (const uint32_t*)keys.p,
                       (const int32_t*)vals.p, shift, (const int32_t*)offsets.p, keys2.p, vals2.p);
    std::swap(keys, keys2);
    std::swap(vals, vals2);
  }
}

// Candidates one detection may hold: beyond this a trial motion spans so
// much of the mesh that the lists would take tens of GB.
constexpr unsigned long long kMaxCandidates = 1ull << 27;

// Exclusive scan of a detection pass's hit counts, the total checked in 64
// bits against kMaxCandidates (the int32 scan would wrap).
int32_t candidate_scan(ando_ctx* c, int n, const int32_t* in, int32_t* out, DBuf<int32_t>& sums,
                       DBuf<int32_t>& total) {
  hipStream_t s = c->stream;
  c->cand64.ensure(1);
  HIPCHK(hipMemsetAsync(c->cand64.p, 0, sizeof(unsigned long long), s));
  hipLaunchKernelGGL(ando::lbvh::k_sum_counts64, dim3(std::min(nblk(n), 1024)), dim3(256), 0, s, n, in,
                     c->cand64.p);
  exclusive_scan(c, n, in, out, sums, total, false);
  unsigned long long t = 0;
  read_small(c, &t, (const unsigned long long*)c->cand64.p, 1);
  if (t > kMaxCandidates)
    throw CandidateOverflow(ANDO_EDEVICE, "contact candidates of one detection exceed 2^27 (" + std::to_string(t) +
                                              "): a trial motion spans too much of the mesh");
  c->stats.max_candidates = std::max(c->stats.max_candidates, (int32_t)t);
  return (int32_t)t;
}

// The `count` hits of one pass at sc[base..], grouped by query (ptr: the
// queries' exclusive scan, nq + 1 entries), sorted by (query, key) when a
// query had more hits than it sorts itself.
void sort_long_hits(ando_ctx* c, int base, int count, int nq, const int32_t* ptr) {
  using namespace ando::lbvh;
  if (count <= 1) return;
  hipStream_t s = c->stream;
  ++c->stats.long_hit_sorts;
  c->sv_keys.ensure(count);
  c->sv_vals.ensure(count);
  hipLaunchKernelGGL(k_hit_keys, dim3(nblk(count)), dim3(kBlock), 0, s, count, (const int32_t*)(c->sc_key.p + base),
                     c->sv_keys.p, c->sv_vals.p);
  radix_sort(c, count, c->sv_keys, c->sv_vals, c->sv_keys2, c->sv_vals2, c->sv_counts, c->sv_offsets, c->sv_sums,
             c->sv_total);
  hipLaunchKernelGGL(k_hit_queries, dim3(nblk(count)), dim3(kBlock), 0, s, count, nq, ptr,
                     (const int32_t*)c->sv_vals.p, c->sv_keys.p);
  radix_sort(c, count, c->sv_keys, c->sv_vals, c->sv_keys2, c->sv_vals2, c->sv_counts, c->sv_offsets, c->sv_sums,
             c->sv_total);
  c->sc_tmp.ensure(count);
  c->sc_ktmp.ensure(count);
  hipLaunchKernelGGL(k_gather_hits, dim3(nblk(count)), dim3(kBlock), 0, s, count, (const int32_t*)c->sv_vals.p,
                     (const ando::SelfContact*)(c->sc.p + base), (const int32_t*)(c->sc_key.p + base), c->sc_tmp.p,
                     c->sc_ktmp.p);
  HIPCHK(hipMemcpyAsync(c->sc.p + base, c->sc_tmp.p, sizeof(ando::SelfContact) * count, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync(c->sc_key.p + base, c->sc_ktmp.p, sizeof(int32_t) * count, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipGetLastError());
}

// After a fill pass: bit 0 of the error flags is a traversal stack overflow,
// bit 1 a query that left its hits unsorted (cleared here once sorted).
void finish_pass(ando_ctx* c, int base, int count, int nq) {
  uint32_t err = 0;
  read_small(c, &err, (const uint32_t*)c->sc_err.p, 1);
  if (err & 1u) throw Fail(ANDO_EDEVICE, "BVH traversal stack exceeded its depth");
  if (err & 2u) {
    sort_long_hits(c, base, count, nq, c->sc_ptr.p);
    HIPCHK(hipMemsetAsync(c->sc_err.p, 0, sizeof(uint32_t), c->stream));
  }
}

// LBVH over P primitives of `kind` vertices each (device index array prims):
// boxes, Morton codes, stable radix sort, Karras hierarchy, node boxes.
// vbox (optional): per-vertex boxes whose unions are the primitive boxes;
// otherwise the primitives' own boxes at the current positions.
// refit: when the tree already holds a hierarchy over the same primitives,
// keep its sorted order and hierarchy and recompute only the boxes (leaf,
// sparse table, node and child boxes).  A query finds every primitive whose
// box overlaps its own whatever the hierarchy, and sorts its hits, so the
// candidate lists are the same as after a full build; only the traversal
// work depends on how well the old order fits the new boxes.  At most
// kMaxRefits refits follow one build.
constexpr int kMaxRefits = 64;
void lbvh_build(ando_ctx* c, LbvhTree& T, int P, int kind, const int32_t* prims, const float* vbox = nullptr,
                bool refit = false) {
  using namespace ando::lbvh;
  hipStream_t s = c->stream;
  refit = refit && c->lbvh_refit && P >= 2 && T.P == P && T.topo_prims == prims && T.topo_kind == kind &&
          T.refits < kMaxRefits;
  T.P = P;
  if (P == 0) {
    T.topo_prims = nullptr;
    return;
  }
  const int G = nblk(P);
  T.boxes.ensure(6 * (size_t)P);
  int levels = 1;
  while ((1 << levels) <= P) ++levels;
  T.table.ensure(6 * (size_t)P * levels);
  const bool fused_leaves = refit && vbox;  // boxes and table level 0 in one pass
  if (fused_leaves)
    hipLaunchKernelGGL(k_refit_leaves, dim3(G), dim3(256), 0, s, P, kind, prims, (const int32_t*)T.vals.p, vbox,
                       T.boxes.p, T.table.p);
  else if (vbox)
    hipLaunchKernelGGL(k_prim_boxes_union, dim3(G), dim3(256), 0, s, P, kind, prims, vbox, T.boxes.p);
  else
    hipLaunchKernelGGL(k_prim_boxes, dim3(G), dim3(256), 0, s, P, kind, prims, (const float*)c->x.p, T.boxes.p);
  if (!refit) {
    T.bpart.ensure(6 * (size_t)G);
    T.bounds.ensure(6);
    hipLaunchKernelGGL(k_centroid_bounds, dim3(G), dim3(256), 0, s, P, (const float*)T.boxes.p, T.bpart.p);
    hipLaunchKernelGGL(k_centroid_bounds_final, dim3(1), dim3(256), 0, s, G, (const float*)T.bpart.p, T.bounds.p);
    T.keys.ensure(P); T.keys2.ensure(P); T.vals.ensure(P); T.vals2.ensure(P);
    hipLaunchKernelGGL(k_morton, dim3(G), dim3(256), 0, s, P, (const float*)T.boxes.p, (const float*)T.bounds.p,
                       T.keys.p, T.vals.p);
    radix_sort(c, P, T.keys, T.vals, T.keys2, T.vals2, T.counts, T.offsets, T.scan_sums, T.scan_total);
  }
  if (!fused_leaves)
    hipLaunchKernelGGL(k_table_level0, dim3(G), dim3(256), 0, s, P, (const float*)T.boxes.p,
                       (const int32_t*)T.vals.p, T.table.p);
  for (int k0 = 0; k0 + 1 < levels; k0 += 3)
    hipLaunchKernelGGL(k_table_levels, dim3(nblk(P - (2 << k0) + 1)), dim3(256), 0, s, P, k0,
                       std::min(3, levels - 1 - k0), T.table.p);
  if (P >= 2) {
    T.left.ensure(P - 1); T.right.ensure(P - 1); T.first.ensure(P - 1); T.last.ensure(P - 1);
    if (!refit)
      hipLaunchKernelGGL(k_karras, dim3(nblk(P - 1)), dim3(256), 0, s, P, (const uint32_t*)T.keys.p, T.left.p,
                         T.right.p, T.first.p, T.last.p);
    T.cbox.ensure(12 * (size_t)(P - 1));
    hipLaunchKernelGGL(k_child_boxes, dim3(nblk(P - 1)), dim3(256), 0, s, P, (const int32_t*)T.left.p,
                       (const int32_t*)T.right.p, (const int32_t*)T.first.p, (const int32_t*)T.last.p,
                       (const float*)T.table.p, T.cbox.p);
  }
  HIPCHK(hipGetLastError());
  if (refit) {
    ++T.refits;
  } else {
    T.topo_prims = prims;
    T.topo_kind = kind;
    T.refits = 0;
  }
}

// Self contacts exist in corrected mode, and in faithful mode on a mesh of
// exactly one triangle: the reference's BVH then has a single primitive-bearing
// leaf, and its broad and narrow phases find the PT contacts of the vertices
// not on the triangle (collision.cpp:78-86,198-244,279-387) -- the corrected
// predicate restricted to one triangle; its three edges share vertices, so
// there are no EE contacts.
bool self_active(const ando_ctx* c) { return c->corrected || c->cp || c->m == 1; }

// Self detection at the current positions: c->sc holds the PT contacts
// (vertex, triangle order) followed by the EE contacts (e1, e2).
void detect_self(ando_ctx* c) {
  using namespace ando::lbvh;
  hipStream_t s = c->stream;
  ++c->det_gen;
  c->sc_pt = c->sc_total = 0;
  c->sv_ready = false;
  if (!self_active(c)) return;
  const int n = c->n, m = c->m, E = (int)c->edges_h.size() / 2;
  lbvh_build(c, c->tri_tree, m, 3, c->d_tris.p);
  lbvh_build(c, c->edge_tree, E, 2, c->d_edges.p);
  c->sc_err.ensure(1);
  HIPCHK(hipMemsetAsync(c->sc_err.p, 0, sizeof(uint32_t), s));
  const int q = std::max(n, E);
  c->sc_count.ensure((size_t)q + 1);
  c->sc_ptr.ensure((size_t)q + 2);
  c->sc_cache.ensure((size_t)q * ando::lbvh::kQueryCache + 1);
  const LbvhTree& T = c->tri_tree;
  const LbvhTree& Q = c->edge_tree;
  auto pt = [&](int pass, ando::SelfContact* out, int32_t* key) {
    // Traversals: one thread per query, 64-thread workgroups so that even
    // ~30k queries spread over every CU.
    hipLaunchKernelGGL(k_query_pt, dim3(nblk(n, 64)), dim3(64), 0, s, pass, n, (const float*)c->x.p, c->pt_box,
                       c->keep_gap, (const int32_t*)c->d_tris.p, T.P, (const int32_t*)T.left.p,
                       (const int32_t*)T.right.p, (const float*)T.cbox.p, (const float*)T.table.p,
                       (const int32_t*)T.vals.p, c->sc_count.p, (const int32_t*)c->sc_ptr.p, out, key, c->sc_err.p,
                       c->sc_cache.p);
  };
  auto ee = [&](int pass, ando::SelfContact* out, int32_t* key) {
    hipLaunchKernelGGL(k_query_ee, dim3(nblk(E, 64)), dim3(64), 0, s, pass, E, (const float*)c->x.p, c->keep_gap,
                       (const int32_t*)c->d_edges.p, Q.P, (const int32_t*)Q.left.p, (const int32_t*)Q.right.p,
                       (const float*)Q.cbox.p, (const float*)Q.table.p, (const int32_t*)Q.vals.p, c->sc_count.p,
                       (const int32_t*)c->sc_ptr.p, out, key, c->sc_err.p, c->sc_cache.p);
  };
  int npt = 0, nee = 0;
  if (m > 0) {
    pt(0, nullptr, nullptr);
    npt = candidate_scan(c, n, c->sc_count.p, c->sc_ptr.p, c->sc_sums, c->sc_total_d);
  }
  c->sc.ensure((size_t)npt + 1);
  c->sc_key.ensure((size_t)npt + 1);
  if (npt) {
    pt(1, c->sc.p, c->sc_key.p);
    finish_pass(c, 0, npt, n);
  }
  if (E > 1) {
    ee(0, nullptr, nullptr);
    nee = candidate_scan(c, E, c->sc_count.p, c->sc_ptr.p, c->sc_sums, c->sc_total_d);
  }
  if (nee) {
    // Append the EE part, growing the list (keeping the PT part, with a
    // quarter of headroom so later detections reuse it) only when it is full.
    if (c->sc.n < (size_t)npt + nee) {
      DBuf<ando::SelfContact> all;
      all.ensure(((size_t)npt + nee) * 5 / 4 + 1);
      if (npt)
        HIPCHK(hipMemcpyAsync(all.p, c->sc.p, sizeof(ando::SelfContact) * npt, hipMemcpyDeviceToDevice, s));
      c->sc = std::move(all);
    }
    c->sc_key.ensure((size_t)npt + nee);
    ee(1, c->sc.p + npt, c->sc_key.p + npt);
    finish_pass(c, npt, nee, E);
  }
  HIPCHK(hipGetLastError());
  if (!npt && !nee) finish_pass(c, 0, 0, 0);
  c->sc_pt = npt;
  c->sc_total = npt + nee;
}

// Corrected physics: the candidate pairs of the per-vertex boxes of the
// motion x -> x + ed (ed null: the points x), grown by gm on the query side.
// Self candidates (PT by vertex then triangle, then EE by e1 then e2) replace
// c->sc, rigid ones (per vertex by body then triangle) replace c->rc; both
// carry indices only until refresh_cp.
void detect_cp_rigid(ando_ctx* c, const ando::DetectBodies& D, float gm);
void detect_cp(ando_ctx* c, const ando::DetectBodies& D, float gm, const float* ed) {
  using namespace ando::lbvh;
  hipStream_t s = c->stream;
  ++c->det_gen;
  c->sv_ready = false;
  c->sc_pt = c->sc_total = 0;
  c->rc_total = 0;
  const int n = c->n, m = c->m, E = (int)c->edges_h.size() / 2;
  c->vbox.ensure(6 * (size_t)n);
  hipLaunchKernelGGL(ando::cp::k_vertex_boxes, dim3(nblk(n)), dim3(kBlock), 0, s, n, (const float*)c->x.p, ed,
                     c->vbox.p);
  HIPCHK(hipGetLastError());
  // The swept boxes of a step's line searches refit the trees the step's
  // first detection (ed null: the points x) built.
  lbvh_build(c, c->tri_tree, m, 3, c->d_tris.p, c->vbox.p, ed != nullptr);
  lbvh_build(c, c->edge_tree, E, 2, c->d_edges.p, c->vbox.p, ed != nullptr);
  c->sc_err.ensure(1);
  HIPCHK(hipMemsetAsync(c->sc_err.p, 0, sizeof(uint32_t), s));
  const int q = std::max(n, E);
  c->sc_count.ensure((size_t)q + 1);
  c->sc_ptr.ensure((size_t)q + 2);
  c->sc_cache.ensure((size_t)q * ando::lbvh::kQueryCache + 1);
  const LbvhTree& T = c->tri_tree;
  const LbvhTree& Q = c->edge_tree;
  auto pt = [&](int pass, ando::SelfContact* out, int32_t* key) {
    hipLaunchKernelGGL(k_query_pt_box, dim3(nblk(n, 64)), dim3(64), 0, s, pass, n, (const float*)c->vbox.p, gm,
                       (const int32_t*)c->d_tris.p, T.P, (const int32_t*)T.left.p, (const int32_t*)T.right.p,
                       (const float*)T.cbox.p, (const float*)T.table.p, (const int32_t*)T.vals.p, c->sc_count.p,
                       (const int32_t*)c->sc_ptr.p, out, key, c->sc_err.p, c->sc_cache.p);
  };
  auto ee = [&](int pass, ando::SelfContact* out, int32_t* key) {
    hipLaunchKernelGGL(k_query_ee_box, dim3(nblk(E, 64)), dim3(64), 0, s, pass, E, (const float*)Q.boxes.p, gm,
                       (const int32_t*)c->d_edges.p, Q.P, (const int32_t*)Q.left.p, (const int32_t*)Q.right.p,
                       (const float*)Q.cbox.p, (const float*)Q.table.p, (const int32_t*)Q.vals.p, c->sc_count.p,
                       (const int32_t*)c->sc_ptr.p, out, key, c->sc_err.p, c->sc_cache.p);
  };
  int npt = 0, nee = 0;
  if (m > 0) {
    pt(0, nullptr, nullptr);
    npt = candidate_scan(c, n, c->sc_count.p, c->sc_ptr.p, c->sc_sums, c->sc_total_d);
  }
  c->sc.ensure((size_t)npt + 1);
  c->sc_key.ensure((size_t)npt + 1);
  if (npt) {
    pt(1, c->sc.p, c->sc_key.p);
    finish_pass(c, 0, npt, n);
  }
  if (E > 1) {
    ee(0, nullptr, nullptr);
    nee = candidate_scan(c, E, c->sc_count.p, c->sc_ptr.p, c->sc_sums, c->sc_total_d);
  }
  if (nee) {
    if (c->sc.n < (size_t)npt + nee) {
      DBuf<ando::SelfContact> all;
      all.ensure(((size_t)npt + nee) * 5 / 4 + 1);
      if (npt)
        HIPCHK(hipMemcpyAsync(all.p, c->sc.p, sizeof(ando::SelfContact) * npt, hipMemcpyDeviceToDevice, s));
      c->sc = std::move(all);
    }
    c->sc_key.ensure((size_t)npt + nee);
    ee(1, c->sc.p + npt, c->sc_key.p + npt);
    finish_pass(c, npt, nee, E);
  }
  HIPCHK(hipGetLastError());
  if (!npt && !nee) finish_pass(c, 0, 0, 0);
  c->sc_pt = npt;
  c->sc_total = npt + nee;
  detect_cp_rigid(c, D, gm);
}

// The rigid candidates of the vertex boxes in c->vbox (see detect_cp).
void detect_cp_rigid(ando_ctx* c, const ando::DetectBodies& D, float gm) {
  hipStream_t s = c->stream;
  const int n = c->n;
  c->rc_total = 0;
  if (D.count > 0) {
    c->rc_hits.ensure((size_t)n * kHitCache);
    hipLaunchKernelGGL(ando::cp::k_rigid_cand, dim3(nblk(n)), dim3(kBlock), 0, s, n, 0, (const float*)c->vbox.p, gm, D,
                       (const int32_t*)c->rc_ptr.p, c->rc_count.p, c->rc.p, c->rc_hits.p);
    const int32_t total = candidate_scan(c, n, c->rc_count.p, c->rc_ptr.p, c->scan_sums, c->scan_total);
    c->rc_total = total;
    c->rc.ensure((size_t)total + 1);
    if (total)
      hipLaunchKernelGGL(ando::cp::k_rigid_cand, dim3(nblk(n)), dim3(kBlock), 0, s, n, 1, (const float*)c->vbox.p, gm,
                         D, (const int32_t*)c->rc_ptr.p, c->rc_count.p, c->rc.p, c->rc_hits.p);
    HIPCHK(hipGetLastError());
  }
}

ando::SelfView self_view(ando_ctx* c) {
  ando::SelfView v{nullptr, nullptr, nullptr, nullptr};
  if (c->sv_ready)
    v = ando::SelfView{c->sv_ptr.p, c->sv_vals.p, c->sc.p, c->sterm.p, c->corrected ? nullptr : c->pt_of.p};
  return v;
}

// Step-constant terms of the self contacts and their vertex view (after
// detect_self; needs the elastic-only blocks from ensure_const).
void prepare_self(ando_ctx* c, const StepParams& P) {
  using namespace ando::lbvh;
  hipStream_t s = c->stream;
  c->sv_ready = false;
  const int nc = c->sc_total;
  if (nc == 0) return;
  c->sterm.ensure(nc);
  hipLaunchKernelGGL(k_self_terms, dim3(nblk(nc)), dim3(kBlock), 0, s, nc, c->n, P,
                     (const ando::SelfContact*)c->sc.p, (const float*)c->masses.p, (const int32_t*)c->d_row_ptr.p,
                     (const int32_t*)c->d_cols.p, (const float*)(c->cp ? c->Hiso3.p : c->Helast3.p), c->sterm.p);
  const int cnt = 4 * nc;
  c->sv_keys.ensure(cnt);
  c->sv_vals.ensure(cnt);
  hipLaunchKernelGGL(k_self_slots, dim3(nblk(cnt)), dim3(kBlock), 0, s, nc, (const ando::SelfContact*)c->sc.p,
                     (const ando::SelfTerm*)c->sterm.p, c->sv_keys.p, c->sv_vals.p);
  radix_sort(c, cnt, c->sv_keys, c->sv_vals, c->sv_keys2, c->sv_vals2, c->sv_counts, c->sv_offsets, c->sv_sums,
             c->sv_total);
  c->sv_hist.ensure((size_t)c->n + 1);
  c->sv_ptr.ensure((size_t)c->n + 1);
  HIPCHK(hipMemsetAsync(c->sv_hist.p, 0, sizeof(int32_t) * c->n, s));
  hipLaunchKernelGGL(k_count_keys, dim3(nblk(cnt)), dim3(kBlock), 0, s, cnt, c->n, (const uint32_t*)c->sv_keys.p,
                     c->sv_hist.p);
  const int entries = exclusive_scan(c, c->n, c->sv_hist.p, c->sv_ptr.p, c->sv_sums, c->sv_total);
  if (!c->corrected) {
    c->pt_of.ensure((size_t)c->n);
    HIPCHK(hipMemsetAsync(c->pt_of.p, 0xff, sizeof(int32_t) * c->n, s));
    if (c->sc_pt > 0)
      hipLaunchKernelGGL(k_pt_of, dim3(nblk(c->sc_pt)), dim3(kBlock), 0, s, c->sc_pt, c->n,
                         (const ando::SelfContact*)c->sc.p, c->pt_of.p);
  }
  c->sv_ready = true;
  c->ccol.ensure(3 * (size_t)entries + 1);
  c->cval.ensure(27 * (size_t)entries + 1);
  c->ncrow = 0;
  c->ncblk = 3 * entries;
  if (entries) {
    hipLaunchKernelGGL(k_self_blocks, dim3(nblk(3LL * entries)), dim3(kBlock), 0, s, entries, self_view(c),
                       c->barrier_tol, c->ccol.p, c->cval.p);
    c->crow_flag.ensure((size_t)c->n + 1);
    c->crow_pos.ensure((size_t)c->n + 2);
    hipLaunchKernelGGL(k_couple_row_flags, dim3(nblk(c->n)), dim3(kBlock), 0, s, c->n, (const int32_t*)c->sv_ptr.p,
                       c->crow_flag.p);
    c->ncrow = exclusive_scan(c, c->n, c->crow_flag.p, c->crow_pos.p, c->sv_sums, c->sv_total);
    c->crow.ensure((size_t)c->ncrow + 1);
    hipLaunchKernelGGL(k_couple_row_fill, dim3(nblk(c->n)), dim3(kBlock), 0, s, c->n, (const int32_t*)c->sv_ptr.p,
                       (const int32_t*)c->crow_pos.p, c->crow.p);
  }
  HIPCHK(hipGetLastError());
}

// Corrected physics: the candidates' gaps, normals, weights and (rigid)
// stiffness at the current positions, then the self terms and their vertex
// view (the set of contacts that add something changes with the gaps).
void refresh_cp(ando_ctx* c, const ando_params& p) {
  hipStream_t s = c->stream;
  if (c->sc_total > 0)
    hipLaunchKernelGGL(ando::cp::k_self_refresh, dim3(nblk(c->sc_total)), dim3(kBlock), 0, s, c->sc_total,
                       (const float*)c->x.p, c->sc.p);
  if (c->rc_total > 0) {
    const float ks = std::max(std::max(p.contact_gap_max, p.min_gap), 1e-12f);
    hipLaunchKernelGGL(ando::cp::k_rigid_refresh, dim3(nblk(c->rc_total)), dim3(kBlock), 0, s, c->rc_total,
                       (const float*)c->x.p, c->cpD, c->rc.p, (const float*)c->masses.p, (const int32_t*)c->d_diag.p,
                       (const float*)c->Hiso3.p, p.dt, ks);
  }
  HIPCHK(hipGetLastError());
  prepare_self(c, step_params(c, p));
}

// PCG on the matrix in c->H (or a caller's buffers) with b = -g; result in c->d.
struct PcgOut {
  int iters;
  bool converged;
};

// Batched PCG for ando_step_batch.  The worker threads step their contexts as
// usual; a solve that qualifies (sliced-ELL SpMV on the default grid, no
// coupling blocks, no learned preconditioner, no kernel timing) is set up on
// the context's stream and handed over here.  The first waiting thread leads:
// once every active worker waits (or 100 us after it arrived) it takes all
// waiting solves with its own iteration cap, runs their iterations together
// (k_pcg_spmv_sell_batch + k_pcg_update_batch, one status read-back per chunk
// for all of them) on the batch stream, and releases them; the contexts'
// streams wait for the batch stream before going on.  Each system's
// arithmetic is that of its own launch, so results do not depend on batching.
struct BatchJob {
  ando_ctx* c = nullptr;
  SpmvMeta* sm = nullptr;
  float* xsol = nullptr;
  int n = 0, nwg_v = 0, nwg_r = 0, maxit = 0;
  float tol = 0.0f;
  const float* A = nullptr;
  const int32_t* diag = nullptr;
  const float* rhs = nullptr;
  hipEvent_t ready = nullptr;
  PcgOut out{0, false};
  bool done = false;
  std::string error;
};

// The batch the calling thread steps its contexts for (ando_step_batch), or null.
struct BatchPcg;
thread_local BatchPcg* t_batch = nullptr;

struct BatchPcg {
  std::mutex mu;
  std::condition_variable cv;
  int active = 0;       // workers that still have contexts to step
  bool leading = false;
  std::vector<BatchJob*> waiting;
  int device = 0;
  hipStream_t stream = nullptr;
  DBuf<ando::PcgSys> sys;
  DBuf<int32_t> wg_r, wg_v, wg_f;
  DBuf<PcgStatus> st_all;
  PcgStatus* st_host = nullptr;  // pinned
  int st_cap = 0;
  int chunk0 = 8, chunk_max = 256;
  int gather_us = 100;  // how long a solve waits for others before leading a round
  long rounds = 0, solves = 0;

  ~BatchPcg() {
    if (st_host) (void)hipHostFree(st_host);
    if (stream) (void)hipStreamDestroy(stream);
  }

  void worker_done() {
    std::lock_guard<std::mutex> lk(mu);
    --active;
    cv.notify_all();
  }

  PcgOut solve(BatchJob& job) {
    HIPCHK(hipEventRecord(job.ready, job.c->stream));
    const auto deadline = std::chrono::steady_clock::now() + std::chrono::microseconds(gather_us);
    std::unique_lock<std::mutex> lk(mu);
    waiting.push_back(&job);
    cv.notify_all();
    while (!job.done) {
      if (leading) {
        cv.wait(lk);
        continue;
      }
      if ((int)waiting.size() < active && std::chrono::steady_clock::now() < deadline) {
        cv.wait_until(lk, deadline);
        continue;
      }
      // Lead a round: every waiting solve with this solve's iteration cap.
      leading = true;
      std::vector<BatchJob*> round, rest;
      for (BatchJob* j : waiting) (j->maxit == job.maxit ? round : rest).push_back(j);
      waiting.swap(rest);
      lk.unlock();
      std::string err;
      try {
        run_round(round);
      } catch (const std::exception& e) {
        err = e.what();
      }
      lk.lock();
      for (BatchJob* j : round) {
        j->error = err;
        j->done = true;
      }
      leading = false;
      cv.notify_all();
    }
    lk.unlock();
    if (!job.error.empty()) throw Fail(ANDO_EDEVICE, "batched PCG: " + job.error);
    // The round ended with a synchronisation of the batch stream, so the
    // solution is in place before the context's stream goes on.
    return job.out;
  }

  void run_round(std::vector<BatchJob*>& jobs) {
    HIPCHK(hipSetDevice(device));
    const int ns = (int)jobs.size();
    std::vector<ando::PcgSys> h(ns);
    std::vector<int32_t> hr, hv, hf;
    for (int k = 0; k < ns; ++k) {
      BatchJob& j = *jobs[k];
      ando_ctx* c = j.c;
      HIPCHK(hipStreamWaitEvent(stream, j.ready, 0));
      ando::PcgSys& S = h[k];
      S.n = j.n;
      S.slices = j.sm->slices;
      S.uniform_w = j.sm->uniform_w;
      S.nwg_v = j.nwg_v;
      S.nwg_r = j.nwg_r;
      S.wg0_r = (int)hr.size();
      S.wg0_v = (int)hv.size();
      S.wg0_f = (int)hf.size();
      S.A = j.A;
      S.diag = j.diag;
      S.rhs = j.rhs;
      S.perm = j.sm->sell_perm.p;
      S.vals = j.sm->sell_vals.p;
      S.scol = j.sm->sell_col.p;
      S.slice_off = j.sm->slice_off.p;
      S.slice_w = j.sm->slice_w.p;
      S.Pinv = c->Pinv.p;
      S.pe = j.n;
      S.pi = 1;
      S.z = c->z.p;
      S.p0 = c->pbuf0.p;
      S.p1 = c->pbuf1.p;
      S.Ap = c->Ap.p;
      S.x = j.xsol;
      S.r = c->r.p;
      S.part_rz = c->part_rz.p;
      S.part_max = c->part_max.p;
      S.rz_hist = c->rz_hist.p;
      S.part_pAp = c->part_pAp.p;
      S.st = c->pcg_status.p;
      S.tol = j.tol;
      hr.insert(hr.end(), j.nwg_r, k);
      hv.insert(hv.end(), j.nwg_v, k);
      hf.insert(hf.end(), nblk(j.sm->slices, kBlock / 64), k);
    }
    sys.upload(h, stream);
    wg_r.upload(hr, stream);
    wg_v.upload(hv, stream);
    wg_f.upload(hf, stream);
    st_all.ensure((size_t)ns);
    if (st_cap < ns) {
      if (st_host) HIPCHK(hipHostFree(st_host));
      st_host = nullptr;
      HIPCHK(hipHostMalloc((void**)&st_host, sizeof(PcgStatus) * ns));
      st_cap = ns;
    }
    // The tables are read by the kernels below; the uploads finish first.
    HIPCHK(hipStreamSynchronize(stream));
    const int maxit = jobs[0]->maxit;
    // Set-up of all systems (what run_pcg does per system on its own stream).
    hipLaunchKernelGGL(ando::k_sell_fill_batch, dim3((int)hf.size()), dim3(kBlock), 0, stream,
                       (const ando::PcgSys*)sys.p, (const int32_t*)wg_f.p);
    hipLaunchKernelGGL(ando::k_pcg_init_batch, dim3((int)hv.size()), dim3(kBlock), 0, stream,
                       (const ando::PcgSys*)sys.p, (const int32_t*)wg_v.p);
    hipLaunchKernelGGL(ando::k_pcg_start_batch, dim3(ns), dim3(kBlock), 0, stream, (const ando::PcgSys*)sys.p);
    auto read_status = [&] {
      hipLaunchKernelGGL(ando::k_pcg_status_gather, dim3((ns + 63) / 64), dim3(64), 0, stream, ns,
                         (const ando::PcgSys*)sys.p, st_all.p);
      HIPCHK(hipMemcpyAsync(st_host, st_all.p, sizeof(PcgStatus) * ns, hipMemcpyDeviceToHost, stream));
      HIPCHK(hipStreamSynchronize(stream));
    };
    int it = 0, chunk = chunk0;
    for (;;) {
      const int end = std::min(maxit, it + chunk);
      for (; it < end; ++it) {
        hipLaunchKernelGGL(ando::k_pcg_spmv_sell_batch, dim3((int)hr.size()), dim3(128), 0, stream,
                           (const ando::PcgSys*)sys.p, (const int32_t*)wg_r.p, it);
        hipLaunchKernelGGL(ando::k_pcg_update_batch, dim3((int)hv.size()), dim3(kBlock), 0, stream,
                           (const ando::PcgSys*)sys.p, (const int32_t*)wg_v.p, it);
      }
      HIPCHK(hipGetLastError());
      read_status();
      bool all = true;
      for (int k = 0; k < ns; ++k) all = all && st_host[k].done;
      if (all) break;
      if (it >= maxit) {
        for (int k = 0; k < ns; ++k)
          if (!st_host[k].done)
            hipLaunchKernelGGL(ando::k_pcg_final, dim3(1), dim3(kBlock), 0, stream, it, h[k].part_max, h[k].nwg_v,
                               h[k].st, h[k].tol);
        read_status();
        break;
      }
      chunk = std::min(chunk * 2, chunk_max);
    }
    for (int k = 0; k < ns; ++k) jobs[k]->out = PcgOut{st_host[k].iters, !st_host[k].failed};
    ++rounds;
    solves += ns;
  }
};

void autotune_pcg(ando_ctx* c, int n, const float* A, const int32_t* rp, const int32_t* cols, const int32_t* diag,
                  const float* rhs, int negate, float* xsol, int maxit, SpmvMeta* sm, const ando::SelfView* sv);

// Iterations to queue in a solve's next chunk.  After each chunk the host
// reads the status, which carries the max-norm relative residual of the
// latest convergence test (rel at iteration rel_it).  From the rate at which
// its logarithm fell since the previous read (from 1 at iteration 0 on the
// first read) the iterations still needed are estimated and queued (plus one;
// at least 8, at most the largest chunk); without a usable rate (a residual
// that rose, a tolerance of 0) the chunk doubles.  Launches queued past
// convergence cost a few microseconds each when they issue their loads first
// and about one when they read the done flag first (spec), while a status
// read costs one host round trip.  The chunking changes no arithmetic.
struct ChunkPlan {
  float last_rel = 1.0f;
  int last_it = 0;
  // Iterations from here on lie past the predicted convergence: their launches
  // read the done flag first (spec), so those that find the solve over cost
  // little; the others pay one flag read before their loads.
  int spec_from = 1 << 30;
  int next(int chunk, int issued, const PcgStatus& st, float tol, int chunk_max) {
    int out = std::min(chunk * 2, chunk_max);
    if (tol > 0.0f && st.rel_it > last_it && st.rel > 0.0f && st.rel < last_rel) {
      const double rate =
          (std::log((double)st.rel) - std::log((double)last_rel)) / (double)(st.rel_it - last_it);
      const double need = std::log((double)tol / (double)st.rel) / rate;  // iterations past rel_it
      if (std::isfinite(need) && need > 0.0) {
        const double ahead = need + 1.0 - (double)(issued - st.rel_it);
        out = (int)std::min<double>((double)chunk_max, std::max(8.0, std::ceil(ahead)));
        // CG usually converges faster than the rate so far predicts, so the
        // last quarter of the prediction is already queued as speculative.
        spec_from = st.rel_it + (int)std::min(std::floor(0.75 * need), 1e9);
      }
    }
    if (st.rel_it > last_it) {
      last_rel = st.rel;
      last_it = st.rel_it;
    }
    return out;
  }
};

// P_given / x0_given (device, both or neither): a preconditioner of 9 floats
// per vertex and a starting guess instead of block-Jacobi from x0 = 0
// (x0_given may be xsol itself).  guard_x0: a starting guess whose residual is
// no smaller than b's (relative residual >= 1, that of x0 = 0) is replaced by
// x0 = 0 (one status read-back more).  With sv the starting residual
// includes the coupling blocks' share of A x0.
// sv (nullable): self-contact blocks that couple two vertices, multiplied
// inside every SpMV (corrected modes).
PcgOut run_pcg(ando_ctx* c, int n, const float* A, const int32_t* rp, const int32_t* cols, const int32_t* diag,
               const float* rhs, int negate, float* xsol, float tol, int maxit, SpmvMeta* sm,
               const float* P_given = nullptr, const float* x0_given = nullptr,
               const ando::SelfView* sv = nullptr, bool guard_x0 = false) {
  hipStream_t s = c->stream;
  const bool couple = sv != nullptr && sv->ptr != nullptr;
  const int R = c->spmv_rowsets;
  const bool sell = c->spmv_sell && sm != nullptr && sm->sell_ok;
  if (c->pcg_autotune && !c->tuning && sell && c->tuned_n != n && P_given == nullptr && x0_given == nullptr &&
      !(t_batch && t_batch->device == c->device))
    autotune_pcg(c, n, A, rp, cols, diag, rhs, negate, xsol, maxit, sm, sv);
  const int sell_slices = (c->sell_lanes == 2) ? c->sell_wg_slices : 4;  // slices per SpMV workgroup (0: a wave)
  const int vb = c->pcg_vec_block;  // threads per workgroup of the vector kernels (init, update)
  int nwg_v = nblk(n, vb);
  int nwg_r = sell ? (sell_slices == 0 ? 2 * sm->slices : nblk(sm->slices, sell_slices)) : spmv_groups(n, R);
  c->part_rz.ensure((size_t)nwg_v + 1);
  c->part_max.ensure((size_t)nwg_v + 1);
  c->stats.pcg_kernel = sell ? 3 : 0;
  c->stats.spmv_lanes = sell ? c->sell_lanes : 0;
  c->stats.spmv_wg_slices = sell ? sell_slices : 0;
  c->stats.pcg_vec_block = vb;
  c->stats.pcg_autotuned = sell ? c->autotuned : 0;
  c->rz_hist.ensure((size_t)maxit + 2);
  c->pcg_status.ensure(1);
  c->part_pAp.ensure((size_t)nwg_r + 1);
  // Batched solves (ando_step_batch) are set up by the batch: the Newton
  // system (b = -g) on the default sliced-ELL grid, without coupling blocks
  // (a self-contact view with no coupling rows adds nothing to the solve, so
  // it may join), learned preconditioner or kernel timing.  A shape chosen
  // through the environment keeps the solve out of the batch; one chosen by
  // the autotuner gives way to the batch's grid.
  const bool default_shape = c->sell_lanes == 2 && sell_slices == 1 && vb == kBlock;
  const bool coupling_rows = couple && c->ncrow > 0;
  if (t_batch && t_batch->device == c->device && sell && !coupling_rows && P_given == nullptr && x0_given == nullptr &&
      negate == 1 && c->kernel_timing == 0 && (default_shape || c->autotuned)) {
    c->stats.spmv_lanes = 2;
    c->stats.spmv_wg_slices = 1;
    c->stats.pcg_vec_block = kBlock;
    c->stats.pcg_autotuned = 0;
    // The batch's grid: one slice per SpMV workgroup, 256-thread vector
    // workgroups (the partial buffers above were sized for the context's shape).
    const int bv = nblk(n, kBlock), br = sm->slices;
    c->part_rz.ensure((size_t)bv + 1);
    c->part_max.ensure((size_t)bv + 1);
    c->part_pAp.ensure((size_t)br + 1);
    BatchJob job;
    job.c = c;
    job.sm = sm;
    job.xsol = xsol;
    job.n = n;
    job.nwg_v = bv;
    job.nwg_r = br;
    job.maxit = maxit;
    job.tol = tol;
    job.A = A;
    job.diag = diag;
    job.rhs = rhs;
    if (!c->batch_ev) HIPCHK(hipEventCreateWithFlags(&c->batch_ev, hipEventDisableTiming));
    job.ready = c->batch_ev;
    c->stats.batched_solves++;
    return t_batch->solve(job);
  }
  // Coupling blocks of the self contacts: workgroups appended to the SpMV
  // grid (8 lanes per row with blocks) write them to Ap2; their partials
  // follow the mesh workgroups'.
  const int spmv_threads = sell ? (c->sell_lanes == 2 && sell_slices == 0   ? 64
                                   : c->sell_lanes == 2 && sell_slices == 1 ? 128
                                   : c->sell_lanes == 2 && sell_slices == 2 ? 256
                                   : c->sell_lanes == 2                      ? 2 * kBlock
                                   : c->sell_lanes == 4                      ? 4 * kBlock
                                                                             : kBlock)
                                : kBlock;
  ando::Coupling cpl{0, nwg_r, 0, nullptr, nullptr, nullptr, nullptr, nullptr};
  int nwg_c = 0;
  if (couple && c->ncrow > 0) {
    c->Ap2.ensure(3 * (size_t)n);
    HIPCHK(hipMemsetAsync(c->Ap2.p, 0, 3 * sizeof(float) * (size_t)n, s));
    nwg_c = nblk(c->ncrow, spmv_threads / 8);
    cpl = ando::Coupling{c->ncrow, nwg_r, nwg_c, c->crow.p, sv->ptr, c->ccol.p, c->cval.p, c->Ap2.p};
    c->part_pAp.ensure((size_t)nwg_r + nwg_c + 1);
  }
  const float* Pc = c->Pinv.p;
  int pe = n, pi = 1;  // block-Jacobi blocks as one plane per entry
  if (sell)
    hipLaunchKernelGGL(k_sell_fill, dim3(nblk(sm->slices, kBlock / 64)), dim3(kBlock), 0, s, sm->slices,
                       (const int32_t*)sm->slice_off.p, (const int32_t*)sm->slice_w.p, (const int32_t*)sm->sell_perm.p,
                       A, sm->sell_vals.p);
  if (P_given) {
    pe = 1;
    pi = 9;
    c->part_bmax.ensure((size_t)nwg_v + 1);
    // The coupling blocks' share of A x0 (in Ap2, zero off the coupling rows).
    if (x0_given && cpl.ncrow > 0)
      hipLaunchKernelGGL(ando::k_couple_mul, dim3(nblk(cpl.ncrow, 32)), dim3(256), 0, s, cpl, x0_given, cpl.Ap2);
    hipLaunchKernelGGL(k_pcg_init_given, dim3(nwg_v), dim3(vb), 0, s, n, A, rp, cols, P_given, rhs, negate,
                       x0_given, xsol, c->r.p, c->z.p, c->pbuf0.p, c->part_rz.p, c->part_max.p, c->part_bmax.p,
                       (const float*)cpl.Ap2);
    Pc = P_given;
  } else {
    hipLaunchKernelGGL(k_pcg_init, dim3(nwg_v), dim3(vb), 0, s, n, A, diag, rhs, xsol, c->r.p, c->z.p,
                       c->pbuf0.p, c->Pinv.p, pe, pi, c->part_rz.p, c->part_max.p, negate);
  }
  hipLaunchKernelGGL(k_pcg_start, dim3(1), dim3(kBlock), 0, s, nwg_v, (const double*)c->part_rz.p,
                     (const float*)c->part_max.p, (const float*)(P_given ? c->part_bmax.p : nullptr), c->rz_hist.p,
                     c->pcg_status.p, tol);
  HIPCHK(hipGetLastError());
  if (P_given && x0_given && guard_x0) {
    PcgStatus s0{};
    read_small(c, &s0, (const PcgStatus*)c->pcg_status.p, 1);
    if (!(s0.rel < 1.0f)) {
      hipLaunchKernelGGL(k_pcg_init_given, dim3(nwg_v), dim3(vb), 0, s, n, A, rp, cols, P_given, rhs, negate,
                         (const float*)nullptr, xsol, c->r.p, c->z.p, c->pbuf0.p, c->part_rz.p, c->part_max.p,
                         c->part_bmax.p, (const float*)nullptr);
      hipLaunchKernelGGL(k_pcg_start, dim3(1), dim3(kBlock), 0, s, nwg_v, (const double*)c->part_rz.p,
                         (const float*)c->part_max.p, (const float*)c->part_bmax.p, c->rz_hist.p, c->pcg_status.p,
                         tol);
      HIPCHK(hipGetLastError());
    }
  }
  PcgStatus st{};
  int it = 0;
  // The iterations go out in chunks: 8 first, then as many as the residual's
  // convergence rate predicts (ChunkPlan; doubling up to 256 without a rate);
  // after each chunk the status is copied to pinned host memory and read.
  // Launches queued after convergence return at the done flag (after issuing
  // their first loads, which lets a live launch overlap the flag's latency).
  // Tried and slower on this path: queueing the next chunk before waiting, and
  // sizing the first chunk from the previous step's solve.
  int chunk = std::max(1, c->pcg_chunk);
  ChunkPlan plan;
  const bool timing = c->kernel_timing != 0;
  size_t ev_used = 0;
  // Timed launches record their start and end with the dispatch itself
  // (hipExtLaunchKernelGGL), so the events add no packets between kernels.
  auto ev = [&](size_t k) -> hipEvent_t {
    while (c->events.size() <= k) {
      hipEvent_t e;
      HIPCHK(hipEventCreate(&e));
      c->events.push_back(e);
    }
    return c->events[k];
  };
  auto launch_chunk = [&](int count) {
    const int end = std::min(maxit, it + count);
    for (; it < end; ++it) {
      float* pold = (it == 0) ? c->pbuf0.p : ((it - 1) & 1 ? c->pbuf1.p : c->pbuf0.p);
      float* pcur = (it & 1) ? c->pbuf1.p : c->pbuf0.p;
      hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr, e3 = nullptr;
      const int spec = it >= plan.spec_from ? 1 : 0;
      if (timing) {
        e0 = ev(ev_used); e1 = ev(ev_used + 1); e2 = ev(ev_used + 2); e3 = ev(ev_used + 3);
        ev_used += 4;
      }
      auto spmv = [&](auto kernel) {
        hipExtLaunchKernelGGL(kernel, dim3(nwg_r + nwg_c), dim3(kBlock), 0, s, e0, e1, 0, n, it, A, rp, cols,
                              (const float*)c->z.p, (const float*)pold, pcur, c->Ap.p, (const double*)c->part_rz.p,
                              (const float*)c->part_max.p, nwg_v, c->rz_hist.p, c->part_pAp.p, c->pcg_status.p, tol,
                              cpl, spec);
      };
      if (sell) {
        auto sell_launch = [&](auto kernel, int threads) {
          hipExtLaunchKernelGGL(kernel, dim3(nwg_r + nwg_c), dim3(threads), 0, s, e0, e1, 0, n, sm->slices, it,
                                (const float*)sm->sell_vals.p, (const int32_t*)sm->sell_col.p,
                                (const int32_t*)sm->slice_off.p, (const int32_t*)sm->slice_w.p, sm->uniform_w,
                                (const float*)c->z.p, (const float*)pold, pcur, c->Ap.p,
                                (const double*)c->part_rz.p, (const float*)c->part_max.p, nwg_v, c->rz_hist.p,
                                c->part_pAp.p, c->pcg_status.p, tol, cpl, spec);
        };
        if (c->sell_lanes == 2 && sell_slices == 0) sell_launch(k_pcg_spmv_sell_split<2, 0>, 64);
        else if (c->sell_lanes == 2 && sell_slices == 1) sell_launch(k_pcg_spmv_sell_split<2, 1>, 128);
        else if (c->sell_lanes == 2 && sell_slices == 2) sell_launch(k_pcg_spmv_sell_split<2, 2>, 256);
        else if (c->sell_lanes == 2) sell_launch(k_pcg_spmv_sell_split<2>, 2 * kBlock);
        else if (c->sell_lanes == 4) sell_launch(k_pcg_spmv_sell_split<4>, 4 * kBlock);
        else sell_launch(k_pcg_spmv_sell, kBlock);
      } else {
        switch (R) {
          case 1: spmv(k_pcg_spmv<1>); break;
          case 2: spmv(k_pcg_spmv<2>); break;
          case 4: spmv(k_pcg_spmv<4>); break;
          default: spmv(k_pcg_spmv<3>); break;
        }
      }
      hipExtLaunchKernelGGL(k_pcg_update, dim3(nwg_v), dim3(vb), 0, s, e2, e3, 0, n, it, (const float*)pcur,
                            (const float*)c->Ap.p, (const float*)cpl.Ap2, Pc, pe, pi, xsol, c->r.p, c->z.p,
                            (const double*)c->part_pAp.p, nwg_r + nwg_c, (const double*)c->rz_hist.p, c->part_rz.p,
                            c->part_max.p, c->pcg_status.p, spec);
    }
    HIPCHK(hipGetLastError());
  };
  for (;;) {
    launch_chunk(chunk);
    HIPCHK(hipMemcpyAsync(c->status_pinned, c->pcg_status.p, sizeof(PcgStatus), hipMemcpyDeviceToHost, s));
    HIPCHK(hipEventRecord(c->status_ev, s));
    HIPCHK(hipEventSynchronize(c->status_ev));
    st = *c->status_pinned;
    if (st.done) break;
    if (it >= maxit) {
      hipLaunchKernelGGL(k_pcg_final, dim3(1), dim3(kBlock), 0, s, it, c->part_max.p, nwg_v, c->pcg_status.p, tol);
      HIPCHK(hipMemcpyAsync(c->status_pinned, c->pcg_status.p, sizeof(PcgStatus), hipMemcpyDeviceToHost, s));
      sync(c);
      st = *c->status_pinned;
      break;
    }
    chunk = plan.next(chunk, it, st, tol, c->pcg_chunk_max);
  }
  if (timing) sync(c);
  if (timing) {
    // Launch k of the solve ran in full iff k < iterations.
    for (size_t k = 0; 4 * k + 3 < ev_used && (int)k < st.iters; ++k) {
      float a = 0.0f, b = 0.0f;
      HIPCHK(hipEventElapsedTime(&a, c->events[4 * k], c->events[4 * k + 1]));
      HIPCHK(hipEventElapsedTime(&b, c->events[4 * k + 2], c->events[4 * k + 3]));
      c->stats.timed_launches[0]++;
      c->stats.timed_launches[1]++;
      c->stats.timed_ms[0] += a;
      c->stats.timed_ms[1] += b;
      // The coupling workgroups' algorithmic bytes in that SpMV launch: 40 B
      // per block; per row its index, range, Ap2 and z, p_old entries.
      if (cpl.ncrow > 0) c->stats.timed_coupling_bytes += 40.0 * c->ncblk + 48.0 * cpl.ncrow;
    }
  }
  return PcgOut{st.iters, !st.failed};
}

// Two-level Schwarz PCG (tl.hip) on the matrix A with pattern T; same
// contract as run_pcg with x0 = 0 (b = -g when negate), the result in xsol.
PcgOut run_pcg_tl(ando_ctx* c, TlMeta& T, const float* A, const int32_t* diag, const float* rhs, int negate,
                  float* xsol, float tol, int maxit, const ando::SelfView* sv, const float* x0 = nullptr,
                  float x0_scale = 0.0f) {
  using namespace ando::tl;
  hipStream_t s = c->stream;
  const int n = T.n;
  SpmvMeta& sm = T.sm;
  c->stats.pcg_kernel = 4;
  c->stats.spmv_lanes = kLanes;
  c->stats.spmv_wg_slices = 0;
  c->stats.pcg_vec_block = T.pr;
  c->stats.pcg_autotuned = 0;
  c->rz_hist.ensure((size_t)maxit + 2);
  c->pcg_status.ensure(1);
  hipLaunchKernelGGL(k_sell_fill, dim3(nblk(sm.slices, kBlock / 64)), dim3(kBlock), 0, s, sm.slices,
                     (const int32_t*)sm.slice_off.p, (const int32_t*)sm.slice_w.p, (const int32_t*)sm.sell_perm.p, A,
                     sm.sell_vals.p);
  // Coarse operator Ac = P^T A P and its inverse (blocked Gauss-Jordan, FP64),
  // made before the solve's iterations start: for every solve, or with
  // tl_coarse_every = k on every k-th solve of this layout, the solves between
  // keeping the last inverse (any fixed symmetric positive definite coarse
  // term leaves the preconditioner symmetric positive definite and constant
  // within a solve, which is all PCG needs).
  const int nc = c->tl_coarse ? T.nc : 0;
  const bool rebuild =
      nc > 0 && (!T.coarse_ready || c->tl_coarse_every <= 1 || T.solves % c->tl_coarse_every == 0);
  ++T.solves;
  if (rebuild) {
    const size_t nn = (size_t)nc * nc;
    const int ncp = T.ncp, nbt = ncp / kGjB;
    const size_t nnp = (size_t)ncp * ncp;
    hipLaunchKernelGGL(k_tl_zero, dim3(nblk((long long)nnp)), dim3(kBlock), 0, s, nc, ncp, T.Ac.p, T.info.p);
    hipLaunchKernelGGL(k_tl_coarse, dim3(nblk(64LL * T.ncb)), dim3(kBlock), 0, s, T.ncb, ncp,
                       (const int32_t*)T.cb_ij.p, (const int32_t*)T.cb_ptr.p, (const int32_t*)T.cb_blk.p, A, T.Ac.p);
    // Blocked Gauss-Jordan (tl.hip): nbt steps of a pivot-block inverse and a
    // tile update, ping-ponging between Ac and Acw; then the FP32 copy.  The
    // failure flag was cleared by k_tl_zero.
    double* X = T.Ac.p;
    double* Y = T.Acw.p;
    for (int q = 0; q < nbt; ++q) {
      hipLaunchKernelGGL(k_tl_gj_pivot, dim3(1), dim3(64), 0, s, ncp, q, (const double*)X, T.gjP.p, T.info.p);
      hipLaunchKernelGGL(k_tl_gj_update, dim3(nbt * nbt), dim3(kBlock), 0, s, nbt, ncp, q, (const double*)X,
                         (const double*)T.gjP.p, Y, (const int*)T.info.p);
      std::swap(X, Y);
    }
    hipLaunchKernelGGL(k_tl_cinv, dim3(nblk((long long)nn)), dim3(kBlock), 0, s, nc, ncp, (const double*)X,
                       (const int*)T.info.p, T.Acinv.p);
    HIPCHK(hipGetLastError());
    T.coarse_ready = true;
  }
  const bool couple = sv != nullptr && sv->ptr != nullptr && c->ncblk > 0;
  TlArgs a{};
  a.n = n;
  a.pr = T.pr;
  a.npatch = T.npatch;
  a.slices = sm.slices;
  a.uniform_w = sm.uniform_w;
  a.nc = nc;
  a.cheb = c->tl_cheb;
  a.lmin_ratio = c->tl_ratio;
  a.perm = T.perm.p;
  a.iperm = T.iperm.p;
  a.vals = sm.sell_vals.p;
  a.scol = sm.sell_col.p;
  a.slice_off = sm.slice_off.p;
  a.slice_w = sm.slice_w.p;
  a.Dinv = T.Dinv.p;
  a.lmax = T.lmax.p;
  a.Acinv = T.Acinv.p;
  a.cptr = couple ? sv->ptr : nullptr;
  a.ccol = couple ? c->ccol.p : nullptr;
  a.cval = couple ? c->cval.p : nullptr;
  if (couple) {
    T.cpr.ensure(2 * (size_t)n);
    T.ccolp.ensure((size_t)c->ncblk);
    hipLaunchKernelGGL(k_tl_couple_perm, dim3(nblk(std::max(n, c->ncblk))), dim3(kBlock), 0, s, n, c->ncblk,
                       (const int32_t*)T.perm.p, (const int32_t*)T.iperm.p, sv->ptr, (const int32_t*)c->ccol.p,
                       T.cpr.p, T.ccolp.p);
    HIPCHK(hipGetLastError());
    a.cpr = T.cpr.p;
    a.ccolp = T.ccolp.p;
  }
  // Start: x0 = 0, or a warm start x0 = x0_scale * x0 (r0 = b - A x0, the
  // residual still measured against max|b|), dropped for x0 = 0 when its
  // residual is no smaller than b's.
  auto start = [&](bool warm) {
    if (warm) {
      hipLaunchKernelGGL(k_tl_gather_x0, dim3(nblk(n)), dim3(kBlock), 0, s, n, (const int32_t*)T.perm.p, x0, x0_scale,
                         T.x.p);
      hipLaunchKernelGGL(k_tl_ax0, dim3(nblk(n)), dim3(kBlock), 0, s, a, (const float*)T.x.p, T.Ap.p);
      T.part_bmax.ensure((size_t)nblk(n) + 1);
    }
    hipLaunchKernelGGL(k_tl_init, dim3(nblk(n)), dim3(kBlock), 0, s, a, A, diag, rhs, negate, T.x.p, T.r.p, T.Dinv.p,
                       warm ? (const float*)T.Ap.p : nullptr, warm ? T.part_bmax.p : nullptr);
    hipLaunchKernelGGL(k_tl_prep, dim3(T.npatch), dim3(patch_threads(T.pr)), 0, s, a, T.lmax.p, (const float*)T.r.p,
                       T.qf.p, T.sr0.p);
    hipLaunchKernelGGL(k_tl_start, dim3(T.npatch), dim3(64 * ((T.pr + 63) / 64)), 0, s, a, (const float*)T.r.p,
                       (const float*)T.qf.p, (const double*)T.sr0.p, T.z.p, T.p0.p, T.part_rz.p, T.part_max.p);
    hipLaunchKernelGGL(k_pcg_start, dim3(1), dim3(kBlock), 0, s, T.npatch, (const double*)T.part_rz.p,
                       (const float*)T.part_max.p, (const float*)nullptr, c->rz_hist.p, c->pcg_status.p, tol);
    if (warm)
      hipLaunchKernelGGL(k_tl_fix_bmax, dim3(1), dim3(kBlock), 0, s, nblk(n), (const float*)T.part_bmax.p,
                         (const float*)T.part_max.p, T.npatch, c->pcg_status.p, tol);
    HIPCHK(hipGetLastError());
  };
  start(x0 != nullptr);
  if (x0 != nullptr) {
    PcgStatus s0{};
    read_small(c, &s0, (const PcgStatus*)c->pcg_status.p, 1);
    if (!(s0.rel < 1.0f)) start(false);
  }
  PcgStatus st{};
  int it = 0;
  int chunk = std::max(1, c->pcg_chunk);
  ChunkPlan plan;
  const bool timing = c->kernel_timing != 0;
  size_t ev_used = 0;
  auto ev = [&](size_t k) -> hipEvent_t {
    while (c->events.size() <= k) {
      hipEvent_t e;
      HIPCHK(hipEventCreate(&e));
      c->events.push_back(e);
    }
    return c->events[k];
  };
  auto launch_chunk = [&](int count) {
    const int end = std::min(maxit, it + count);
    for (; it < end; ++it) {
      float* pold = (it == 0) ? T.p0.p : ((it - 1) & 1 ? T.p1.p : T.p0.p);
      float* pcur = (it & 1) ? T.p1.p : T.p0.p;
      hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr, e3 = nullptr;
      const int spec = it >= plan.spec_from ? 1 : 0;
      if (timing) {
        e0 = ev(ev_used); e1 = ev(ev_used + 1); e2 = ev(ev_used + 2); e3 = ev(ev_used + 3);
        ev_used += 4;
      }
      double* sr_cur = (it & 1) ? T.sr1.p : T.sr0.p;
      double* sr_next = (it & 1) ? T.sr0.p : T.sr1.p;
      hipExtLaunchKernelGGL(k_tl_spmv, dim3(T.npatch), dim3(spmv_threads(T.pr)), 0, s, e0, e1, 0, a, it,
                            (const float*)T.z.p, (const float*)pold, pcur, T.Ap.p, T.sa.p,
                            (const double*)T.part_rz.p, (const float*)T.part_max.p, c->rz_hist.p, T.part_pAp.p,
                            c->pcg_status.p, tol, spec);
      hipExtLaunchKernelGGL(k_tl_update, dim3(T.npatch), dim3(patch_threads(T.pr)), 0, s, e2, e3, 0, a, it, (const float*)pcur,
                            (const float*)T.Ap.p, (const double*)sr_cur, (const double*)T.sa.p, sr_next, T.x.p,
                            T.r.p, T.z.p, (const double*)T.part_pAp.p, (const double*)c->rz_hist.p, T.part_rz.p,
                            T.part_max.p, c->pcg_status.p, spec);
    }
    HIPCHK(hipGetLastError());
  };
  for (;;) {
    launch_chunk(chunk);
    HIPCHK(hipMemcpyAsync(c->status_pinned, c->pcg_status.p, sizeof(PcgStatus), hipMemcpyDeviceToHost, s));
    HIPCHK(hipEventRecord(c->status_ev, s));
    HIPCHK(hipEventSynchronize(c->status_ev));
    st = *c->status_pinned;
    if (st.done) break;
    if (it >= maxit) {
      hipLaunchKernelGGL(k_pcg_final, dim3(1), dim3(kBlock), 0, s, it, T.part_max.p, T.npatch, c->pcg_status.p, tol);
      HIPCHK(hipMemcpyAsync(c->status_pinned, c->pcg_status.p, sizeof(PcgStatus), hipMemcpyDeviceToHost, s));
      sync(c);
      st = *c->status_pinned;
      break;
    }
    chunk = plan.next(chunk, it, st, tol, c->pcg_chunk_max);
  }
  hipLaunchKernelGGL(k_tl_scatter, dim3(nblk(n)), dim3(kBlock), 0, s, n, (const int32_t*)T.perm.p,
                     (const float*)T.x.p, xsol);
  HIPCHK(hipGetLastError());
  if (timing) {
    sync(c);
    for (size_t k = 0; 4 * k + 3 < ev_used && (int)k < st.iters; ++k) {
      float ta = 0.0f, tb = 0.0f;
      HIPCHK(hipEventElapsedTime(&ta, c->events[4 * k], c->events[4 * k + 1]));
      HIPCHK(hipEventElapsedTime(&tb, c->events[4 * k + 2], c->events[4 * k + 3]));
      c->stats.timed_launches[0]++;
      c->stats.timed_launches[1]++;
      c->stats.timed_ms[0] += ta;
      c->stats.timed_ms[1] += tb;
      if (couple) c->stats.timed_coupling_bytes += 40.0 * c->ncblk;
    }
  }
  return PcgOut{st.iters, !st.failed};
}

// The mesh pattern's two-level layout, built on first use.
TlMeta& mesh_tl(ando_ctx* c) {
  if (!c->tl.ok) build_tl_meta(c->tl, c->n, c->row_ptr, c->cols, c->morton, c->stream);
  if (!c->tl.ok) throw Fail(ANDO_EDEVICE, "two-level PCG: sliced layout too large");
  return c->tl;
}

// Launch-shape autotuner (ANDO_PCG_AUTOTUNE).  On the first sliced-ELL solve of
// a system size, a short trial solve (32 iterations, tolerance 0) runs with each
// candidate shape -- lanes per row, slices of 64 rows per SpMV workgroup, threads
// per vector workgroup -- on the system at hand, and every SpMV and update launch
// is timed with events recorded by the dispatch.  The shape with the least kernel
// time per iteration (best of two trials) is kept for the context and, through a
// process-wide table keyed by the device and the sliced-ELL structure, for every
// later context with the same structure.  The default shape stays unless another is at least 3% faster, so
// timing noise between near-equal shapes does not move it.  The lanes alone
// change no result; the grid and the vector block change the association of the
// p.Ap and r.z sums, so results agree with the default shape to rounding.  The
// trial solves write only the solve's own work vectors and x, which the real
// solve then initialises again; their iterations are not counted in the stats.
struct PcgShape {
  int lanes, slices, vb;
};
constexpr PcgShape kPcgShapes[] = {{2, 1, kBlock}, {2, 0, kBlock}, {2, 2, kBlock}, {2, 4, kBlock},
                                   {4, 4, kBlock}, {1, 4, kBlock}, {2, 1, 128},    {2, 1, 64}};
// A tuned shape is reused only for the same device and the same sliced-ELL
// structure (rows, slices, slot count and slice layout), not merely the same size.
struct TuneKey {
  int device, n, slices;
  long long slots;
  uint64_t layout;
  bool operator==(const TuneKey& o) const {
    return device == o.device && n == o.n && slices == o.slices && slots == o.slots && layout == o.layout;
  }
};
std::mutex g_tune_mu;
std::vector<std::pair<TuneKey, PcgShape>> g_tune_table;

void set_shape(ando_ctx* c, const PcgShape& sh) {
  c->sell_lanes = sh.lanes;
  c->sell_wg_slices = sh.slices;
  c->pcg_vec_block = sh.vb;
}

void autotune_pcg(ando_ctx* c, int n, const float* A, const int32_t* rp, const int32_t* cols, const int32_t* diag,
                  const float* rhs, int negate, float* xsol, int maxit, SpmvMeta* sm, const ando::SelfView* sv) {
  const TuneKey key{c->device, n, sm->slices, sm->slots, sm->layout_hash};
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    for (const auto& e : g_tune_table)
      if (e.first == key) {
        set_shape(c, e.second);
        c->tuned_n = n;
        c->autotuned = 1;
        return;
      }
  }
  const int K = std::min(maxit, 32);
  if (K < 4) return;
  constexpr int NS = (int)(sizeof(kPcgShapes) / sizeof(kPcgShapes[0]));
  // The trials switch the context into timing mode; this guard puts the
  // context back (mode, chunk, timed statistics, shape) on every exit,
  // including a trial that throws.
  struct Restore {
    ando_ctx* c;
    int timing, chunk;
    int32_t tl[2];
    double tm[2];
    PcgShape shape;
    ~Restore() {
      c->tuning = false;
      c->kernel_timing = timing;
      c->pcg_chunk = chunk;
      c->stats.timed_launches[0] = tl[0];
      c->stats.timed_launches[1] = tl[1];
      c->stats.timed_ms[0] = tm[0];
      c->stats.timed_ms[1] = tm[1];
      set_shape(c, shape);
    }
  } restore{c, c->kernel_timing, c->pcg_chunk, {c->stats.timed_launches[0], c->stats.timed_launches[1]},
            {c->stats.timed_ms[0], c->stats.timed_ms[1]}, PcgShape{c->sell_lanes, c->sell_wg_slices, c->pcg_vec_block}};
  c->tuning = true;
  c->kernel_timing = 1;
  c->pcg_chunk = K;  // one chunk: no status read-back inside the trial
  double us[NS];
  for (int k = 0; k < NS; ++k) {
    set_shape(c, kPcgShapes[k]);
    us[k] = INFINITY;
    for (int rep = 0; rep < 2; ++rep) {
      c->stats.timed_launches[0] = c->stats.timed_launches[1] = 0;
      c->stats.timed_ms[0] = c->stats.timed_ms[1] = 0.0;
      run_pcg(c, n, A, rp, cols, diag, rhs, negate, xsol, 0.0f, K, sm, nullptr, nullptr, sv);
      if (c->stats.timed_launches[0] > 0)
        us[k] = std::min(us[k], 1e3 * (c->stats.timed_ms[0] + c->stats.timed_ms[1]) / c->stats.timed_launches[0]);
    }
  }
  int best = 0;
  for (int k = 1; k < NS; ++k)
    if (us[k] < us[best]) best = k;
  if (!(us[best] < 0.97 * us[0])) best = 0;
  restore.shape = kPcgShapes[best];
  if (!std::isfinite(us[0])) return;  // nothing ran (b = 0): try again on the next solve
  c->tuned_n = n;
  c->autotuned = 1;
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune_table.push_back({key, kPcgShapes[best]});
  }
  if (c->pcg_autotune == 2) {
    std::fprintf(stderr, "ando autotune n=%d:", n);
    for (int k = 0; k < NS; ++k)
      std::fprintf(stderr, " (%d,%d,%d)=%.2fus", kPcgShapes[k].lanes, kPcgShapes[k].slices, kPcgShapes[k].vb, us[k]);
    std::fprintf(stderr, " -> (%d,%d,%d)\n", kPcgShapes[best].lanes, kPcgShapes[best].slices, kPcgShapes[best].vb);
  }
}

float read_gmax(ando_ctx* c) {
  uint32_t bits = 0;
  read_small(c, &bits, (const uint32_t*)c->scalars.p, 1);
  float f;
  std::memcpy(&f, &bits, sizeof(float));
  return f;
}

// Gradient at the current x (compute_gradient), result in c->g; returns max |g|.
void launch_gradient(ando_ctx* c, const ando_params& p, bool with_rigid) {
  hipStream_t s = c->stream;
  StepParams P = step_params(c, p);
  HIPCHK(hipMemsetAsync(c->scalars.p, 0, 3 * sizeof(uint32_t), s));
  if (P.strain) {
    float tau = p.strain_tau > 1.0f ? p.strain_tau * 0.01f : p.strain_tau;
    float eps = p.strain_epsilon > 0.0f ? (p.strain_epsilon > 1.0f ? p.strain_epsilon * 0.01f : p.strain_epsilon)
                                        : tau;
    if (eps <= 0.0f) {
      P.strain = 0;
    } else {
      StrainParams sp;
      sp.tau = tau;
      sp.eps = eps;
      sp.min_gap = std::max(p.min_gap, 1e-8f);
      sp.svd_eps = std::max(p.strain_svd_epsilon, 1e-8f);
      sp.thick = c->mat.thickness;
      sp.rho = c->mat.density;
      sp.fp64_near = c->fp64_near;
      sp.cp = c->cp;
      hipLaunchKernelGGL(k_strain_face, dim3(nblk(c->m)), dim3(kBlock), 0, s, c->m, sp, c->d_tris.p, c->x.p,
                         c->dm.p, c->area.p, c->face_blk.p, c->cp ? c->Hiso3.p : c->Helast3.p, c->sflags.p,
                         c->sgrad.p, c->shess.p, c->scalars.p + 2);
    }
  }
  const float* egrad = c->egrad.p;
  if (c->cp) {
    // Membrane gradient and face projectors at the current positions.
    c->cfe.ensure(9 * (size_t)std::max(c->m, 1));
    c->cegrad.ensure(3 * (size_t)c->n);
    if (c->m)
      hipLaunchKernelGGL(ando::cp::k_face_x, dim3(nblk(c->m)), dim3(kBlock), 0, s, c->m, (const int32_t*)c->d_tris.p,
                         (const float*)c->x.p, (const float*)c->dm.p, (const float*)c->area.p, c->mat.thickness, c->mu,
                         c->cfe.p, c->fpi.p);
    hipLaunchKernelGGL(k_egrad_gather, dim3(nblk(c->n)), dim3(kBlock), 0, s, c->n, c->v2f_ptr.p, c->v2f.p, c->cfe.p,
                       c->cegrad.p);
    egrad = c->cegrad.p;
  }
  const int32_t* rcp = (with_rigid && c->rc_total > 0) ? c->rc_ptr.p : nullptr;
  const int32_t* pinp = c->num_pins ? c->pin_ptr.p : nullptr;
  hipLaunchKernelGGL(k_gradient, dim3(nblk(c->n)), dim3(kBlock), 0, s, c->n, P, c->x.p, c->xhat.p, c->masses.p,
                     egrad, c->v2f_ptr.p, c->v2f.p, c->sflags.p, c->sgrad.p, rcp, c->rc.p, pinp, c->pin_t.p,
                     c->d_diag.p, c->Hconst.p, c->d_wall_n.p, c->d_wall_off.p, c->kw.p, c->xprev.p, c->g.p,
                     c->scalars.p, self_view(c));
  HIPCHK(hipGetLastError());
}

void launch_assemble(ando_ctx* c, const ando_params& p, bool with_rigid) {
  StepParams P = step_params(c, p);
  if (P.strain) {
    float tau = p.strain_tau > 1.0f ? p.strain_tau * 0.01f : p.strain_tau;
    float eps = p.strain_epsilon > 0.0f ? (p.strain_epsilon > 1.0f ? p.strain_epsilon * 0.01f : p.strain_epsilon)
                                        : tau;
    if (eps <= 0.0f) P.strain = 0;
  }
  const int32_t* rcp = (with_rigid && c->rc_total > 0) ? c->rc_ptr.p : nullptr;
  const int32_t* pinp = c->num_pins ? c->pin_ptr.p : nullptr;
  hipLaunchKernelGGL(k_assemble, dim3(nblk(c->nb)), dim3(kBlock), 0, c->stream, c->n, P, c->d_block_row.p,
                     c->d_cols.p, c->blk_cptr.p, c->blk_contrib.p, c->Hconst.p, c->sflags.p, c->shess.p, c->x.p,
                     c->masses.p, rcp, c->rc.p, pinp, c->pin_t.p, c->d_wall_n.p, c->d_wall_off.p, c->kw.p, c->xprev.p, c->H.p,
                     c->nb, self_view(c), (const float*)c->fgk.p, (const float*)c->fpi.p);
  HIPCHK(hipGetLastError());
}

// LineSearch::search: the first active wall, and (corrected mode) the CCD of
// the self point-triangle and edge-edge contacts.
float line_search(ando_ctx* c) {
  hipStream_t s = c->stream;
  const bool self = c->sc_total > 0;
  bool wall = c->num_walls > 0;
  f3 w = mk(0.0f, 0.0f, 0.0f);
  if (wall) {
    const float* wn = c->wall_n.data();
    w = mk(wn[0], wn[1], wn[2]);
    double sq = (double)(w.x * w.x + (w.y * w.y + w.z * w.z));
    wall = sq > 0.5;
  }
  if (!wall && !self) return 1.0f;
  HIPCHK(hipMemsetD32Async(c->scalars.p + 1, 0xFFFFF, 1, s));  // bit t: alpha = 2^-t still feasible
  if (wall)
    hipLaunchKernelGGL(k_ls_mask, dim3(nblk(c->n)), dim3(kBlock), 0, s, c->n, w, c->wall_off[0], 1, c->x.p, c->d.p,
                       c->scalars.p + 1);
  if (self)
    hipLaunchKernelGGL(ando::lbvh::k_ls_self, dim3(nblk(c->sc_total)), dim3(kBlock), 0, s, c->sc_total,
                       (const ando::SelfContact*)c->sc.p, (const float*)c->x.p, (const float*)c->d.p, 1.25f,
                       c->additive_ccd, c->scalars.p + 1);
  HIPCHK(hipGetLastError());
  uint32_t mask = 0;
  read_small(c, &mask, (const uint32_t*)(c->scalars.p + 1), 1);
  for (int t = 0; t < 20; ++t)
    if (mask & (1u << t)) return std::ldexp(1.0f, -t);
  return 0.0f;
}

// Corrected physics: the first line-search trial, the smallest t0 (at most 19)
// with 1.25 * 2^-t0 * max |d_i| <= L = max(4 * mean rest edge, 2 * max
// |x_hat_i - x_i|); arguments are the squared maxima (FP64).
int cp_first_trial(double mean_edge, double md2, double mh2) {
  const double L = std::max(4.0 * mean_edge, 2.0 * std::sqrt(mh2));
  const double len = 1.25 * std::sqrt(md2);
  int t0 = 0;
  while (t0 < 19 && std::ldexp(len, -t0) > L) ++t0;
  return t0;
}

// Corrected physics line search: trials start at t0 (cp_first_trial); the
// candidates of the trial motion x -> x + ed, ed = 1.25 * 2^-t0 d, replace
// the step's candidates; alpha_t = 2^-t (t >= t0) is feasible when 2^-(t - t0)
// is at most every candidate's time of impact over ed and no vertex ends up
// more violated at the first wall than it is (g1 >= min(0, g0)).
float line_search_cp(ando_ctx* c, const ando_params& p) {
  hipStream_t s = c->stream;
  const int n3 = 3 * c->n;
  c->ed.ensure((size_t)n3);
  c->motion2.ensure(2);
  HIPCHK(hipMemsetAsync(c->motion2.p, 0, 2 * sizeof(unsigned long long), s));
  hipLaunchKernelGGL(ando::cp::k_motion_max, dim3(nblk(c->n)), dim3(kBlock), 0, s, c->n, (const float*)c->d.p,
                     (const float*)c->x.p, (const float*)c->xhat.p, c->motion2.p);
  unsigned long long mb[2] = {0, 0};
  read_small(c, mb, (const unsigned long long*)c->motion2.p, 2);
  double md2, mh2;
  std::memcpy(&md2, &mb[0], sizeof(double));
  std::memcpy(&mh2, &mb[1], sizeof(double));
  int t0 = cp_first_trial(c->mean_edge, md2, mh2);
  // A trial motion whose swept boxes would give a detection more than
  // kMaxCandidates candidates (some vertex moving centimetres makes the bound
  // above large, and a garbage direction moves every vertex that far) starts
  // the trials one halving later instead; the oracle, which has no such
  // capacity limit, takes the first trial as it is.
  for (;;) {
    hipLaunchKernelGGL(ando::cp::k_scale, dim3(nblk(n3)), dim3(kBlock), 0, s, n3, std::ldexp(1.25f, -t0),
                       (const float*)c->d.p, c->ed.p);
    HIPCHK(hipGetLastError());
    try {
      detect_cp(c, c->cpD, p.contact_gap_max, c->ed.p);
      break;
    } catch (const CandidateOverflow&) {
      if (t0 >= 19) throw;
      ++t0;
      ++c->stats.capped_trials;
    }
  }
  // Pairs within their separation that would not move apart hold their cloth
  // vertices: up to four passes (a hold changes its neighbours' motion), then
  // every such pair's vertices.
  if (c->sc_total > 0 || c->rc_total > 0) {
    const int n = c->n;
    c->hold0.ensure((size_t)n);
    c->hold1.ensure((size_t)n);
    c->hold_changed.ensure(1);
    HIPCHK(hipMemsetAsync(c->hold0.p, 0, (size_t)n, s));
    uint8_t* fin = c->hold0.p;
    uint8_t* fout = c->hold1.p;
    bool any = false;
    for (int pass = 0; pass < 5; ++pass) {
      HIPCHK(hipMemcpyAsync(fout, fin, (size_t)n, hipMemcpyDeviceToDevice, s));
      HIPCHK(hipMemsetAsync(c->hold_changed.p, 0, sizeof(uint32_t), s));
      const int all = pass == 4 ? 1 : 0;
      if (c->sc_total > 0)
        hipLaunchKernelGGL(ando::cp::k_band_self, dim3(nblk(c->sc_total)), dim3(kBlock), 0, s, c->sc_total,
                           (const ando::SelfContact*)c->sc.p, (const float*)c->x.p, (const float*)c->ed.p,
                           (const uint8_t*)fin, fout, c->hold_changed.p, all);
      if (c->rc_total > 0)
        hipLaunchKernelGGL(ando::cp::k_band_rigid, dim3(nblk(c->rc_total)), dim3(kBlock), 0, s, c->rc_total,
                           (const RigidContact*)c->rc.p, c->cpD, (const float*)c->x.p, (const float*)c->ed.p,
                           (const uint8_t*)fin, fout, c->hold_changed.p, all);
      HIPCHK(hipGetLastError());
      uint32_t changed = 0;
      read_small(c, &changed, (const uint32_t*)c->hold_changed.p, 1);
      std::swap(fin, fout);
      if (!changed) break;
      any = true;
    }
    if (any) {
      hipLaunchKernelGGL(ando::cp::k_apply_hold, dim3(nblk(n)), dim3(kBlock), 0, s, n, (const uint8_t*)fin, c->d.p,
                         c->ed.p);
      HIPCHK(hipGetLastError());
      ++c->stats.held_iterations;
    }
  }
  bool wall = c->num_walls > 0;
  f3 w = mk(0.0f, 0.0f, 0.0f);
  if (wall) {
    const float* wn = c->wall_n.data();
    w = mk(wn[0], wn[1], wn[2]);
    double sq = (double)(w.x * w.x + (w.y * w.y + w.z * w.z));
    wall = sq > 0.5;
  }
  HIPCHK(hipMemsetD32Async(c->scalars.p + 1, 0xFFFFFu & ~((1u << t0) - 1u), 1, s));
  if (c->sc_total > 0)
    hipLaunchKernelGGL(ando::cp::k_ls_toi_self, dim3(nblk(c->sc_total)), dim3(kBlock), 0, s, c->sc_total, t0,
                       (const ando::SelfContact*)c->sc.p, (const float*)c->x.p, (const float*)c->ed.p,
                       c->scalars.p + 1);
  if (c->rc_total > 0)
    hipLaunchKernelGGL(ando::cp::k_ls_toi_rigid, dim3(nblk(c->rc_total)), dim3(kBlock), 0, s, c->rc_total, t0,
                       (const RigidContact*)c->rc.p, c->cpD, (const float*)c->x.p, (const float*)c->ed.p,
                       c->scalars.p + 1);
  if (wall)
    hipLaunchKernelGGL(ando::cp::k_ls_wall, dim3(nblk(c->n)), dim3(kBlock), 0, s, c->n, t0, w, c->wall_off[0],
                       (const float*)c->x.p, (const float*)c->ed.p, c->scalars.p + 1);
  HIPCHK(hipGetLastError());
  uint32_t mask = 0;
  read_small(c, &mask, (const uint32_t*)(c->scalars.p + 1), 1);
  for (int t = 0; t < 20; ++t)
    if (mask & (1u << t)) return std::ldexp(1.0f, -t);
  return 0.0f;
}

// Integrator::inner_newton_step (integrator.cpp:152-236).
float inner_newton(ando_ctx* c, const ando_params& p) {
  int maxit = p.max_newton_steps;
  if (p.enable_friction) maxit = std::max(maxit, p.friction_min_newton_steps);
  hipStream_t s = c->stream;
  // The last iteration's accepted step: x moved by 1.25 alpha d, so the next
  // Newton direction is close to (1 - 1.25 alpha) d where the system changed
  // little (the two-level solver's optional warm start).
  float prev_alpha = -1.0f;
  for (int it = 0; it < maxit; ++it) {
    c->stats.newton_iterations++;
    if (c->cp) refresh_cp(c, p);
    launch_gradient(c, p, true);
    float gmax = read_gmax(c);
    if (gmax < p.pcg_tol) return 1.0f;
    launch_assemble(c, p, true);
    PcgOut po;
    const ando::SelfView sv = self_view(c);
    if (c->learned && c->nn.ready) {
      // Learned preconditioner and warm start, within the plain solve's
      // iteration cap: the warm start is dropped when its residual is no
      // smaller than b's; the learned preconditioner gets half the cap, and a
      // solve that misses the tolerance continues with block-Jacobi from its
      // iterate for the rest.  Warm-start-only mode is block-Jacobi PCG from
      // the warm start with the whole cap.
      nnh::forward(c->nn, c->nnw, c->n, c->H.p, c->d_row_ptr.p, c->d_cols.p, c->d_diag.p, c->g.p, 1, false,
                   c->stream);
      nnh::warm_start(c->nnw, c->n, c->H.p, c->d_row_ptr.p, c->d_cols.p, c->g.p, 1, c->stream);
      const int cap = p.pcg_max_iters;
      const bool precond = c->learned == 1;
      const int first_cap = precond ? std::max(1, cap / 2) : cap;
      po = run_pcg(c, c->n, c->H.p, c->d_row_ptr.p, c->d_cols.p, c->d_diag.p, c->g.p, 1, c->d.p, p.pcg_tol,
                   first_cap, &c->spmv_meta, precond ? c->nnw.P.p : c->nnw.Dinv.p, c->nnw.y.p, &sv, true);
      c->stats.learned_solves++;
      if (!po.converged && precond && cap - po.iters > 0) {
        c->stats.learned_fallbacks++;
        const int first = po.iters;
        po = run_pcg(c, c->n, c->H.p, c->d_row_ptr.p, c->d_cols.p, c->d_diag.p, c->g.p, 1, c->d.p, p.pcg_tol,
                     cap - first, &c->spmv_meta, c->nnw.Dinv.p, c->d.p, &sv, true);
        c->stats.fallback_iterations += po.iters;
        po.iters += first;
      }
    } else if (c->two_level) {
      const float ws = 1.0f - 1.25f * prev_alpha;
      const bool warm = c->tl_warm && prev_alpha >= 0.0f && ws > 0.0f;
      po = run_pcg_tl(c, mesh_tl(c), c->H.p, c->d_diag.p, c->g.p, 1, c->d.p, p.pcg_tol, p.pcg_max_iters, &sv,
                      warm ? (const float*)c->d.p : nullptr, ws);
    } else {
      po = run_pcg(c, c->n, c->H.p, c->d_row_ptr.p, c->d_cols.p, c->d_diag.p, c->g.p, 1, c->d.p, p.pcg_tol,
                   p.pcg_max_iters, &c->spmv_meta, nullptr, nullptr, &sv);
    }
    if (c->stats.pcg_solves < ANDO_MAX_RECORDED_SOLVES) c->stats.pcg_per_solve[c->stats.pcg_solves] = po.iters;
    c->stats.pcg_solves++;
    c->stats.pcg_iterations += po.iters;
    if (!po.converged) c->stats.pcg_failures++;
    float alpha = c->cp ? line_search_cp(c, p) : line_search(c);
    c->stats.last_alpha = alpha;
    if ((double)alpha < 1e-8) {
      c->stats.line_search_failures++;
      return 0.0f;
    }
    float sc = (float)((double)alpha * 1.25);
    hipLaunchKernelGGL(k_axpy, dim3(nblk(3LL * c->n)), dim3(kBlock), 0, s, 3 * c->n, sc, c->d.p, c->x.p);
    HIPCHK(hipGetLastError());
    if ((double)alpha > 0.99) return 1.0f;
    prev_alpha = alpha;
  }
  return 0.5f;
}

// Body records of the caller <-> device body states (rigid_body.h:72-85).
void pack_bodies(const ando_rigid_body* bodies, int nbod, std::vector<BodyState>& bs) {
  bs.resize(nbod);
  for (int b = 0; b < nbod; ++b) {
    const ando_rigid_body& B = bodies[b];
    std::memcpy(bs[b].pos, B.position, sizeof(bs[b].pos));
    std::memcpy(bs[b].rot, B.rotation, sizeof(bs[b].rot));
    std::memcpy(bs[b].lin, B.linear_velocity, sizeof(bs[b].lin));
    std::memcpy(bs[b].ang, B.angular_velocity, sizeof(bs[b].ang));
    std::memcpy(bs[b].Ibinv, B.inertia_body_inv, sizeof(bs[b].Ibinv));
    bs[b].mass = B.mass;
  }
}

// After the step's final synchronisation: the coupled bodies' new pose and
// velocities (k_rigid_integrate) into the caller's records.
void finish_bodies(ando_ctx* c, ando_rigid_body* bodies) {
  const int nbod = c->bodies_pending;
  c->bodies_pending = 0;
  const BodyState* bs = static_cast<const BodyState*>(c->bstage.p);
  for (int b = 0; b < nbod; ++b) {
    std::memcpy(bodies[b].position, bs[b].pos, sizeof(bs[b].pos));
    std::memcpy(bodies[b].rotation, bs[b].rot, sizeof(bs[b].rot));
    std::memcpy(bodies[b].linear_velocity, bs[b].lin, sizeof(bs[b].lin));
    std::memcpy(bodies[b].angular_velocity, bs[b].ang, sizeof(bs[b].ang));
  }
}

// Integrator::apply_rigid_coupling (integrator.cpp:676-754): contacts at the
// final positions push the bodies with -dV/dg n at the witness point.
// The per-body sums (k_rigid_forces) and RigidBody::integrate
// (k_rigid_integrate) run on the device; the new body states are copied to
// page-locked memory with the step's last synchronisation and written into
// the caller's records by finish_bodies, so the coupling adds no host round
// trip.
// Corrected physics: the contacts within gap_max of the final positions (the
// point boxes grown by gap_max) with the corrected stiffness, and the force
// V' n, which pushes the body away from the cloth.
void rigid_coupling(ando_ctx* c, const ando_params& p, ando_rigid_body* bodies, int nbod) {
  upload_bodies(c, bodies, nbod);
  hipStream_t s = c->stream;
  if (c->cp) {
    c->cpD = detect_bodies(c, bodies);
    c->vbox.ensure(6 * (size_t)c->n);
    hipLaunchKernelGGL(ando::cp::k_vertex_boxes, dim3(nblk(c->n)), dim3(kBlock), 0, s, c->n, (const float*)c->x.p,
                       (const float*)nullptr, c->vbox.p);
    ++c->det_gen;
    detect_cp_rigid(c, c->cpD, p.contact_gap_max);
    if (c->rc_total > 0) {
      const float ks = std::max(std::max(p.contact_gap_max, p.min_gap), 1e-12f);
      hipLaunchKernelGGL(ando::cp::k_rigid_refresh, dim3(nblk(c->rc_total)), dim3(kBlock), 0, s, c->rc_total,
                         (const float*)c->x.p, c->cpD, c->rc.p, (const float*)c->masses.p,
                         (const int32_t*)c->d_diag.p, (const float*)c->Hiso3.p, p.dt, ks);
    }
    HIPCHK(hipGetLastError());
  } else {
    detect_rigid(c, bodies, nbod, p.dt);
  }
  BodyFrames bf{};
  for (int b = 0; b < nbod; ++b) {
    for (int k = 0; k < 3; ++k) bf.pos[b][k] = bodies[b].position[k];
    bf.mass[b] = bodies[b].mass;
  }
  const int G = std::max(1, std::min(64, nblk(c->rc_total)));
  c->rw_part.ensure((size_t)G * nbod * 6);
  c->rw_out.ensure((size_t)nbod * 6);
  hipLaunchKernelGGL(k_rigid_forces, dim3(G), dim3(kBlock), 0, s, c->rc_total, (const RigidContact*)c->rc.p, nbod,
                     bf, p.contact_gap_max, c->cp ? 1.0f : -1.0f, c->rw_part.p);
  hipLaunchKernelGGL(k_rigid_forces_final, dim3(1), dim3(kBlock), 0, s, G, nbod, (const double*)c->rw_part.p,
                     c->rw_out.p);
  HIPCHK(hipGetLastError());
  std::vector<BodyState> bs;
  pack_bodies(bodies, nbod, bs);
  c->body_state.upload(bs, s);
  hipLaunchKernelGGL(k_rigid_integrate, dim3(1), dim3(64), 0, s, nbod, (const double*)c->rw_out.p, c->body_state.p,
                     p.dt);
  HIPCHK(hipGetLastError());
  BodyState* h = static_cast<BodyState*>(c->bstage.ensure((size_t)nbod * sizeof(BodyState)));
  HIPCHK(hipMemcpyAsync(h, c->body_state.p, (size_t)nbod * sizeof(BodyState), hipMemcpyDeviceToHost, s));
  c->bodies_pending = nbod;
}

// Integrator::apply_contact_restitution (integrator.cpp:565-674).
void restitution(ando_ctx* c, const ando_params& p, ando_rigid_body* bodies, int nbod) {
  float e = std::min(std::max(p.contact_restitution, 0.0f), 1.0f);
  if (e <= 0.0f) return;
  hipStream_t s = c->stream;
  // Self contacts first (they head the contact list), detected again at the
  // end-of-step positions.
  if (self_active(c)) {
    detect_self(c);
    const int nc = c->sc_total, cnt = 4 * nc;
    if (nc > 0) {
      using namespace ando::lbvh;
      c->sv_keys.ensure(cnt);
      c->sv_vals.ensure(cnt);
      hipLaunchKernelGGL(k_rest_slots, dim3(nblk(cnt)), dim3(kBlock), 0, s, nc, c->n, (const ando::SelfContact*)c->sc.p,
                         std::max(p.contact_gap_max, 1e-5f), c->sv_keys.p, c->sv_vals.p);
      radix_sort(c, cnt, c->sv_keys, c->sv_vals, c->sv_keys2, c->sv_vals2, c->sv_counts, c->sv_offsets, c->sv_sums,
                 c->sv_total);
      c->sv_hist.ensure((size_t)c->n + 1);
      c->sv_ptr.ensure((size_t)c->n + 1);
      HIPCHK(hipMemsetAsync(c->sv_hist.p, 0, sizeof(int32_t) * c->n, s));
      hipLaunchKernelGGL(k_count_keys, dim3(nblk(cnt)), dim3(kBlock), 0, s, cnt, c->n, (const uint32_t*)c->sv_keys.p,
                         c->sv_hist.p);
      exclusive_scan(c, c->n, c->sv_hist.p, c->sv_ptr.p, c->sv_sums, c->sv_total, false);
      hipLaunchKernelGGL(k_rest_self, dim3(nblk(c->n)), dim3(kBlock), 0, s, c->n, (const int32_t*)c->sv_ptr.p,
                         (const int32_t*)c->sv_vals.p, (const ando::SelfContact*)c->sc.p, e, c->v.p);
      HIPCHK(hipGetLastError());
    }
  }
  if (nbod > 0) {
    upload_bodies(c, bodies, nbod);
    detect_rigid(c, bodies, nbod, p.dt);
    std::vector<BodyState> bs;
    pack_bodies(bodies, nbod, bs);
    c->body_state.upload(bs, s);
    if (c->rc_total)
      hipLaunchKernelGGL(k_restitution_rigid, dim3(1), dim3(64), 0, s, c->rc_total, c->rc.p, c->v.p, c->masses.p,
                         c->body_state.p, nbod, e, std::max(p.contact_gap_max, 1e-5f));
    read_small(c, bs.data(), (const BodyState*)c->body_state.p, (size_t)nbod);
    for (int b = 0; b < nbod; ++b) {
      std::memcpy(bodies[b].linear_velocity, bs[b].lin, sizeof(bs[b].lin));
      std::memcpy(bodies[b].angular_velocity, bs[b].ang, sizeof(bs[b].ang));
    }
  }
  if (c->num_walls)
    hipLaunchKernelGGL(k_restitution_walls, dim3(nblk(c->n)), dim3(kBlock), 0, s, c->n, c->num_walls,
                       c->d_wall_n.p, c->d_wall_off.p, c->d_wall_gap.p, p.contact_gap_max, e, c->x.p, c->v.p);
  HIPCHK(hipGetLastError());
}

void check_params(const ando_params* p) {
  need(p != nullptr, ANDO_EINVAL, "params is null");
  need(std::isfinite(p->dt), ANDO_EINVAL, "dt must be finite");
  need(p->pcg_max_iters >= 0 && p->pcg_max_iters <= 1000000, ANDO_EINVAL, "pcg_max_iters out of range");
  need(p->max_newton_steps >= 0 && p->max_newton_steps <= 100000, ANDO_EINVAL, "max_newton_steps out of range");
}

}  // namespace

extern "C" {

int32_t ando_abi_version(void) { return 1; }
const char* ando_last_error(void) { return ando::g_last_error.c_str(); }

int ando_create(int32_t device, ando_ctx** out) {
  return guarded([&] {
    need(out != nullptr, ANDO_EINVAL, "out is null");
    int count = 0;
    HIPCHK(hipGetDeviceCount(&count));
    need(device >= 0 && device < count, ANDO_EINVAL, "device index out of range");
    HIPCHK(hipSetDevice(device));
    auto* c = new ando_ctx;
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
      delete c;
      throw Fail(ANDO_EDEVICE, "hipStreamCreate failed");
    }
    c->scalars.ensure(4);
    if (hipHostMalloc((void**)&c->status_pinned, sizeof(PcgStatus), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&c->sact_pinned, sizeof(uint32_t), hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&c->status_ev, hipEventDisableTiming) != hipSuccess) {
      delete c;
      throw Fail(ANDO_EDEVICE, "pinned status or event allocation failed");
    }
    if (const char* w = std::getenv("ANDO_PCG_CHUNK")) c->pcg_chunk = std::min(256, std::max(1, std::atoi(w)));
    if (const char* w = std::getenv("ANDO_PCG_CHUNK_MAX"))
      c->pcg_chunk_max = std::min(1024, std::max(1, std::atoi(w)));
    if (const char* w = std::getenv("ANDO_RIGID_GROUP")) c->rigid_group = std::min(256, std::max(1, std::atoi(w)));
    if (const char* w = std::getenv("ANDO_PCG_VEC_BLOCK")) {
      const int k = std::atoi(w);
      c->pcg_vec_block = (k == 64 || k == 128) ? k : kBlock;
    }
    if (const char* w = std::getenv("ANDO_SELL_WG_SLICES")) {
      const int k = std::atoi(w);
      c->sell_wg_slices = (k == 0 || k == 2 || k == 4) ? k : 1;
    }
    if (const char* w = std::getenv("ANDO_SELL_LANES")) {
      const int l = std::atoi(w);
      c->sell_lanes = (l == 1 || l == 4) ? l : 2;
    }
    if (const char* w = std::getenv("ANDO_PCG_AUTOTUNE")) c->pcg_autotune = std::min(2, std::max(0, std::atoi(w)));
    if (const char* w = std::getenv("ANDO_TL_CHEB")) {
      const int k = std::min(ando::tl::kMaxCheb, std::max(1, std::atoi(w)));
      c->tl_cheb = (k % 2) ? k : k + 1;  // odd degrees keep the patch polynomial positive
    }
    if (const char* w = std::getenv("ANDO_TL_RATIO")) c->tl_ratio = std::max(1.5f, (float)std::atof(w));
    if (const char* w = std::getenv("ANDO_TL_COARSE")) c->tl_coarse = std::atoi(w) != 0 ? 1 : 0;
    if (const char* w = std::getenv("ANDO_TL_COARSE_EVERY")) c->tl_coarse_every = std::max(1, std::atoi(w));
    if (const char* w = std::getenv("ANDO_TL_WARM")) c->tl_warm = std::atoi(w) != 0 ? 1 : 0;
    if (const char* w = std::getenv("ANDO_LBVH_REFIT")) c->lbvh_refit = std::atoi(w) != 0 ? 1 : 0;
    if (const char* w = std::getenv("ANDO_PCG_ROWSETS")) c->spmv_rowsets = std::min(4, std::max(1, std::atoi(w)));
    if (const char* w = std::getenv("ANDO_PCG_SPMV")) {
      c->spmv_env = std::strcmp(w, "sell") == 0 ? 2 : 0;
      c->spmv_sell = c->spmv_env == 2 ? 1 : 0;
    }
    *out = c;
  });
}

int ando_destroy(ando_ctx* c) {
  if (!c) return ANDO_OK;
  return guarded([&] {
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    (void)hipStreamDestroy(c->stream);
    delete c;
  });
}

int ando_set_mesh(ando_ctx* c, const float* vertices, int32_t n, const int32_t* triangles, int32_t m,
                  const ando_material* material) {
  return guarded([&] {
    CtxLock lock(c);
    need(c && vertices && material && (m == 0 || triangles), ANDO_EINVAL, "null argument");
    need(n > 0 && m >= 0, ANDO_EINVAL, "mesh needs at least one vertex");
    HIPCHK(hipSetDevice(c->device));
    c->n = n;
    c->m = m;
    c->mat.youngs_modulus = material->youngs_modulus;
    c->mat.poisson_ratio = material->poisson_ratio;
    c->mat.density = material->density;
    c->mat.thickness = material->thickness;
    c->mat.bending_stiffness = material->bending_stiffness;
    c->mu = ando::shear_modulus(c->mat);
    c->tris.assign(triangles, triangles + 3 * (size_t)m);
    RestState rs = compute_rest_state(vertices, n, c->tris.data(), m);
    build_topology(c, rs.edges);
    c->edges_h = rs.edges;
    c->d_edges.upload(c->edges_h, c->stream);
    c->morton = morton_order(vertices, n);
    c->tl.ok = false;
    {
      double sum = 0.0;
      const size_t E = c->edges_h.size() / 2;
      for (size_t e = 0; e < E; ++e) {
        const float* a = vertices + 3 * (size_t)c->edges_h[2 * e];
        const float* b = vertices + 3 * (size_t)c->edges_h[2 * e + 1];
        const double dx = (double)b[0] - (double)a[0], dy = (double)b[1] - (double)a[1],
                     dz = (double)b[2] - (double)a[2];
        sum += std::sqrt(dx * dx + dy * dy + dz * dz);
      }
      c->mean_edge = E ? sum / (double)E : 0.0;
    }
    alloc_state(c);
    hipStream_t s = c->stream;
    c->mverts.upload(vertices, 3 * (size_t)n, s);
    c->d_tris.upload(c->tris, s);
    c->dm.upload(rs.dm_inv, s);
    c->area.upload(rs.area, s);
    c->ec.ensure(9 * (size_t)std::max(m, 1));
    c->fe.ensure(9 * (size_t)std::max(m, 1));
    c->egrad.ensure(3 * (size_t)n);
    if (m)
      hipLaunchKernelGGL(k_elastic_face, dim3(nblk(m)), dim3(kBlock), 0, s, m, c->d_tris.p, c->mverts.p, c->dm.p,
                         c->area.p, c->mat.thickness, c->mu, c->ec.p, c->fe.p);
    hipLaunchKernelGGL(k_egrad_gather, dim3(nblk(n)), dim3(kBlock), 0, s, n, c->v2f_ptr.p, c->v2f.p, c->fe.p,
                       c->egrad.p);
    HIPCHK(hipGetLastError());
    sync(c);
    c->have_mesh = true;
    c->have_state = false;
    c->const_valid = false;
    c->kw_valid = false;
    c->num_pins = 0;
    c->num_walls = 0;
    c->pin_vertex.clear();
    c->pin_target.clear();
    c->wall_n.clear();
    c->wall_off.clear();
    c->wall_gap.clear();
  });
}

int ando_set_mesh_vertices(ando_ctx* c, const float* vertices, int32_t n) {
  return guarded([&] {
    CtxLock lock(c);
    need(c && vertices, ANDO_EINVAL, "null argument");
    need(c->have_mesh && n == c->n, ANDO_EINVAL, "vertex count does not match the mesh");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    c->mverts.upload(vertices, 3 * (size_t)n, s);
    // The rest state (Dm_inv, areas) is not recomputed by the reference's
    // vertex setter either; only the elastic gradient reads mesh.vertices.
    if (c->m)
      hipLaunchKernelGGL(k_elastic_face, dim3(nblk(c->m)), dim3(kBlock), 0, s, c->m, c->d_tris.p, c->mverts.p,
                         c->dm.p, c->area.p, c->mat.thickness, c->mu, c->ec.p, c->fe.p);
    hipLaunchKernelGGL(k_egrad_gather, dim3(nblk(n)), dim3(kBlock), 0, s, n, c->v2f_ptr.p, c->v2f.p, c->fe.p,
                       c->egrad.p);
    HIPCHK(hipGetLastError());
    sync(c);
  });
}

int ando_lumped_masses(const float* vertices, int32_t n, const int32_t* triangles, int32_t m,
                       const ando_material* material, float* masses_out) {
  return guarded([&] {
    need(vertices && material && masses_out && (m == 0 || triangles), ANDO_EINVAL, "null argument");
    need(n > 0 && m >= 0, ANDO_EINVAL, "Mesh must be initialised before creating a state");
    ando::Material mat;
    mat.youngs_modulus = material->youngs_modulus;
    mat.poisson_ratio = material->poisson_ratio;
    mat.density = material->density;
    mat.thickness = material->thickness;
    mat.bending_stiffness = material->bending_stiffness;
    RestState rs = compute_rest_state(vertices, n, triangles, m);
    std::vector<float> ms = lumped_masses(n, triangles, m, rs.area, mat);
    std::memcpy(masses_out, ms.data(), sizeof(float) * n);
  });
}

int ando_set_state(ando_ctx* c, const float* positions, const float* velocities, const float* masses, int32_t n) {
  return guarded([&] {
    CtxLock lock(c);
    need(c && c->have_mesh, ANDO_ESTATE, "mesh must be set before the state");
    need(n == c->n, ANDO_EINVAL, "vertex count does not match the mesh");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    if (positions) c->x.upload(positions, 3 * (size_t)n, s);
    if (velocities) c->v.upload(velocities, 3 * (size_t)n, s);
    if (masses) {
      c->masses.upload(masses, n, s);
      c->const_valid = false;
      c->kw_valid = false;
    }
    if (positions && velocities && masses && !c->have_prev) {
      HIPCHK(hipMemcpyAsync(c->xprev.p, c->x.p, 3 * sizeof(float) * (size_t)n, hipMemcpyDeviceToDevice, s));
      c->have_prev = true;
    }
    // A new state starts the coarse-inverse schedule afresh: the first solve
    // from it rebuilds the coarse inverse, so the steps that follow depend on
    // the state alone and not on the solves made before it was set.
    c->tl.coarse_ready = false;
    c->tl.solves = 0;
    sync(c);
    if (positions && velocities && masses) c->have_state = true;
  });
}

int ando_set_previous_positions(ando_ctx* c, const float* prev, int32_t n) {
  return guarded([&] {
    CtxLock lock(c);
    need(c && c->have_mesh && prev, ANDO_ESTATE, "mesh must be set before the state");
    need(n == c->n, ANDO_EINVAL, "vertex count does not match the mesh");
    HIPCHK(hipSetDevice(c->device));
    c->xprev.upload(prev, 3 * (size_t)n, c->stream);
    c->have_prev = true;
    sync(c);
  });
}

int ando_get_state(ando_ctx* c, float* positions, float* velocities, int32_t n) {
  return guarded([&] {
    CtxLock lock(c);
    need(c && c->have_state, ANDO_ESTATE, "state has not been set");
    need(n == c->n, ANDO_EINVAL, "vertex count does not match the state");
    HIPCHK(hipSetDevice(c->device));
    if (positions)
      HIPCHK(hipMemcpyAsync(positions, c->x.p, 3 * sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
    if (velocities)
      HIPCHK(hipMemcpyAsync(velocities, c->v.p, 3 * sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
    sync(c);
  });
}

int ando_host_alloc(size_t bytes, void** out) {
  return guarded([&] {
    need(out != nullptr && bytes > 0, ANDO_EINVAL, "bad pinned allocation request");
    *out = nullptr;
    if (hipHostMalloc(out, bytes, hipHostMallocDefault) != hipSuccess) throw Fail(ANDO_EDEVICE, "hipHostMalloc failed");
  });
}

int ando_host_free(void* ptr) {
  return guarded([&] {
    if (ptr) HIPCHK(hipHostFree(ptr));
  });
}

int ando_set_pins(ando_ctx* c, const int32_t* vertices, const float* targets, int32_t count) {
  return guarded([&] {
    CtxLock lock(c);
    need(c && c->have_mesh, ANDO_ESTATE, "mesh must be set before pins");
    need(count >= 0 && (count == 0 || (vertices && targets)), ANDO_EINVAL, "bad pin arrays");
    for (int k = 0; k < count; ++k)
      need(vertices[k] >= 0 && vertices[k] < c->n, ANDO_EINVAL, "pin vertex index out of range");
    HIPCHK(hipSetDevice(c->device));
    c->num_pins = count;
    c->pin_vertex.assign(vertices, vertices + count);
    c->pin_target.assign(targets, targets + 3 * (size_t)count);
    // Per-vertex lists in pin order.
    std::vector<int32_t> ptr(c->n + 1, 0);
    for (int k = 0; k < count; ++k) ++ptr[vertices[k] + 1];
    for (int i = 0; i < c->n; ++i) ptr[i + 1] += ptr[i];
    std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
    std::vector<float> tgt(3 * (size_t)std::max(count, 1));
    for (int k = 0; k < count; ++k) {
      int slot = fill[vertices[k]]++;
      for (int q = 0; q < 3; ++q) tgt[3 * slot + q] = targets[3 * k + q];
    }
    c->pin_ptr.upload(ptr, c->stream);
    c->pin_t.upload(tgt, c->stream);
    sync(c);
  });
}

int ando_set_walls(ando_ctx* c, const float* normals, const float* offsets, const float* gaps, int32_t count) {
  return guarded([&] {
    CtxLock lock(c);
    need(c && c->have_mesh, ANDO_ESTATE, "mesh must be set before walls");
    need(count >= 0 && (count == 0 || (normals && offsets && gaps)), ANDO_EINVAL, "bad wall arrays");
    HIPCHK(hipSetDevice(c->device));
    c->num_walls = count;
    c->wall_n.assign(normals, normals + 3 * (size_t)count);
    c->wall_off.assign(offsets, offsets + count);
    c->wall_gap.assign(gaps, gaps + count);
    if (count) {
      c->d_wall_n.upload(c->wall_n, c->stream);
      c->d_wall_off.upload(c->wall_off, c->stream);
      c->d_wall_gap.upload(c->wall_gap, c->stream);
    }
    c->kw_valid = false;
    sync(c);
  });
}

int ando_set_mode(ando_ctx* c, int32_t flags) {
  return guarded([&] {
    CtxLock lock(c);
    need(c != nullptr, ANDO_EINVAL, "null context");
    need((flags & ~(ANDO_MODE_FP64_NEAR_CONTACT | ANDO_MODE_KERNEL_TIMING |
                    ANDO_MODE_LEARNED_PRECOND | ANDO_MODE_LEARNED_WARM_START_ONLY |
                    ANDO_MODE_CORRECTED_CONTACT | ANDO_MODE_PCG_ROW_SPMV | ANDO_MODE_STATIC_RIGIDS |
                    ANDO_MODE_ADDITIVE_CCD | ANDO_MODE_CORRECTED_PHYSICS | ANDO_MODE_TWO_LEVEL)) == 0,
         ANDO_EINVAL, "unknown mode flag");
    need(!(flags & ANDO_MODE_LEARNED_PRECOND) || c->nn.ready, ANDO_ESTATE,
         "learned preconditioner mode needs ando_set_precond_weights first");
    c->fp64_near = (flags & ANDO_MODE_FP64_NEAR_CONTACT) ? 1 : 0;
    c->kernel_timing = (flags & ANDO_MODE_KERNEL_TIMING) ? 1 : 0;
    c->learned = (flags & ANDO_MODE_LEARNED_PRECOND) ? ((flags & ANDO_MODE_LEARNED_WARM_START_ONLY) ? 2 : 1) : 0;
    const int cp = (flags & ANDO_MODE_CORRECTED_PHYSICS) ? 1 : 0;
    if (cp != c->cp) {
      c->const_valid = false;
      c->kw_valid = false;
    }
    c->cp = cp;
    c->corrected = ((flags & ANDO_MODE_CORRECTED_CONTACT) || cp) ? 1 : 0;
    if (c->spmv_env < 0) c->spmv_sell = (flags & ANDO_MODE_PCG_ROW_SPMV) ? 0 : 1;
    c->static_rigids = (flags & ANDO_MODE_STATIC_RIGIDS) ? 1 : 0;
    c->additive_ccd = (flags & ANDO_MODE_ADDITIVE_CCD) ? 1 : 0;
    c->two_level = (flags & ANDO_MODE_TWO_LEVEL) ? 1 : 0;
  });
}

int ando_apply_gravity(ando_ctx* c, const float* g3, float dt) {
  return guarded([&] {
    CtxLock lock(c);
    need(c && c->have_state, ANDO_ESTATE, "State has not been initialised");
    need(g3 != nullptr, ANDO_EINVAL, "gravity is null");
    HIPCHK(hipSetDevice(c->device));
    // Corrected physics: velocity only (the step's x_hat = x + dt v carries it).
    hipLaunchKernelGGL(k_gravity, dim3(nblk(c->n)), dim3(kBlock), 0, c->stream, c->n, g3[0], g3[1], g3[2], dt,
                       c->x.p, c->v.p, c->cp ? 0 : 1);
    HIPCHK(hipGetLastError());
  });
}

int ando_step(ando_ctx* c, const ando_params* params, ando_rigid_body* bodies, int32_t num_bodies) {
  return guarded([&] {
    CtxLock lock(c);
    need(c && c->have_state, ANDO_ESTATE, "state has not been set");
    check_params(params);
    need(num_bodies == 0 || bodies != nullptr, ANDO_EINVAL, "bodies is null");
    HIPCHK(hipSetDevice(c->device));
    auto t0 = std::chrono::steady_clock::now();
    const ando_params& p = *params;
    hipStream_t s = c->stream;
    c->stats = ando_stats{};
    ensure_const(c, p);
    hipLaunchKernelGGL(k_predict, dim3(nblk(3LL * c->n)), dim3(kBlock), 0, s, 3 * c->n, p.dt, c->x.p, c->v.p,
                       c->xold.p, c->xhat.p);
    HIPCHK(hipGetLastError());
    upload_bodies(c, bodies, num_bodies);
    c->barrier_tol = p.barrier_tolerance;
    if (c->cp) {
      c->cpD = detect_bodies(c, bodies);
      detect_cp(c, c->cpD, p.contact_gap_max, nullptr);
    } else {
      detect_rigid(c, bodies, num_bodies, p.dt);
      detect_self(c);
      prepare_self(c, step_params(c, p));
    }
    c->stats.contacts = c->rc_total + c->sc_total;
    c->stats.rigid_contacts = c->rc_total;
    float beta = 0.0f;
    int bi = 0;
    while (beta < p.beta_max && bi < 20) {
      float a = inner_newton(c, p);
      beta = (float)((double)beta + (1.0 - (double)beta) * (double)a);
      ++bi;
      if ((double)a < 1e-6) break;
    }
    if ((double)beta > 1e-6) inner_newton(c, p);
    c->stats.beta = beta;
    if (c->cp) {
      // The implicit-Euler velocity of the positions reached.
      float damp = p.velocity_damping;
      int do_damp = (double)damp > 0.0 ? 1 : 0;
      float scale = 1.0f - std::min(std::max(damp, 0.0f), 1.0f);
      hipLaunchKernelGGL(ando::cp::k_velocity_dt, dim3(nblk(3LL * c->n)), dim3(kBlock), 0, s, 3 * c->n, p.dt, scale,
                         do_damp, (const float*)c->x.p, (const float*)c->xold.p, c->v.p);
      HIPCHK(hipGetLastError());
      if ((double)p.contact_restitution > 0.0) restitution(c, p, bodies, num_bodies);
    } else if ((double)beta > 1e-6) {
      float bdt = beta * p.dt;
      float damp = p.velocity_damping;
      int do_damp = (double)damp > 0.0 ? 1 : 0;
      float scale = 1.0f - std::min(std::max(damp, 0.0f), 1.0f);
      hipLaunchKernelGGL(k_velocity, dim3(nblk(3LL * c->n)), dim3(kBlock), 0, s, 3 * c->n, bdt, scale, do_damp,
                         c->x.p, c->xold.p, c->v.p);
      HIPCHK(hipGetLastError());
      if ((double)p.contact_restitution > 0.0) restitution(c, p, bodies, num_bodies);
    }
    // Read with the step's last synchronisation.
    HIPCHK(hipMemcpyAsync(c->sact_pinned, c->scalars.p + 2, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    c->bodies_pending = 0;
    if (num_bodies > 0 && !c->static_rigids) rigid_coupling(c, p, bodies, num_bodies);
    sync(c);
    finish_bodies(c, bodies);
    c->stats.strain_active = (int32_t)*c->sact_pinned;
    c->stats.step_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  });
}

int ando_get_stats(ando_ctx* c, ando_stats* out) {
  return guarded([&] {
    CtxLock lock(c);
    need(c && out, ANDO_EINVAL, "null argument");
    *out = c->stats;
  });
}

int ando_compute_contacts(ando_ctx* c, const ando_rigid_body* bodies, int32_t num_bodies, ando_contact* out,
                          int32_t capacity, int32_t* count) {
  return guarded([&] {
    CtxLock lock(c);
    need(c && c->have_state && count, ANDO_ESTATE, "state has not been set");
    need(num_bodies == 0 || bodies != nullptr, ANDO_EINVAL, "bodies is null");
    HIPCHK(hipSetDevice(c->device));
    upload_bodies(c, bodies, num_bodies);
    float dt = c->const_valid ? c->const_dt : 0.0f;
    if (!c->const_valid) {
      // Stiffness is not part of the contact record; Helast3 may be unset.
      HIPCHK(hipMemsetAsync(c->Helast3.p, 0, 3 * sizeof(float) * (size_t)c->nb, c->stream));
    }
    detect_rigid(c, bodies, num_bodies, dt);
    detect_self(c);
    c->listed_gen = c->det_gen;
    *count = c->sc_total + c->rc_total;
    if (!out || capacity <= 0) return;
    std::vector<RigidContact> hc(c->rc_total);
    std::vector<int32_t> ptr(c->n + 1);
    if (c->rc_total) {
      HIPCHK(hipMemcpyAsync(hc.data(), c->rc.p, sizeof(RigidContact) * c->rc_total, hipMemcpyDeviceToHost,
                            c->stream));
      HIPCHK(hipMemcpyAsync(ptr.data(), c->rc_ptr.p, sizeof(int32_t) * (c->n + 1), hipMemcpyDeviceToHost,
                            c->stream));
    }
    std::vector<float> xh(3 * (size_t)c->n);
    HIPCHK(hipMemcpyAsync(xh.data(), c->x.p, 3 * sizeof(float) * (size_t)c->n, hipMemcpyDeviceToHost, c->stream));
    std::vector<ando::SelfContact> hs(c->sc_total);
    if (c->sc_total)
      HIPCHK(hipMemcpyAsync(hs.data(), c->sc.p, sizeof(ando::SelfContact) * c->sc_total, hipMemcpyDeviceToHost,
                            c->stream));
    sync(c);
    int w = 0;
    // Self contacts first (detect_all_collisions order), then rigid ones.
    for (int e = 0; e < c->sc_total && w < capacity; ++e, ++w) {
      const ando::SelfContact& r = hs[e];
      ando_contact& o = out[w];
      std::memset(&o, 0, sizeof(o));
      o.type = r.type;
      for (int k = 0; k < 4; ++k) o.idx[k] = r.idx[k];
      o.rigid_body_index = -1;
      o.vertex_count = 4;
      o.gap = r.gap;
      o.normal[0] = r.nx; o.normal[1] = r.ny; o.normal[2] = r.nz;
      for (int k = 0; k < 3; ++k) o.witness_p[k] = xh[3 * r.idx[0] + k];
      o.witness_q[0] = r.qx; o.witness_q[1] = r.qy; o.witness_q[2] = r.qz;
      for (int k = 0; k < 4; ++k) o.weights[k] = r.w[k];
    }
    for (int i = 0; i < c->n && w < capacity; ++i)
      for (int e = ptr[i]; e < ptr[i + 1] && w < capacity; ++e, ++w) {
        const RigidContact& r = hc[e];
        ando_contact& o = out[w];
        std::memset(&o, 0, sizeof(o));
        o.type = 3;
        o.idx[0] = i;
        const int32_t* t = bodies[r.body].triangles + 3 * r.tri;
        o.idx[1] = t[0]; o.idx[2] = t[1]; o.idx[3] = t[2];
        o.rigid_body_index = r.body;
        o.vertex_count = 1;
        o.gap = r.gap;
        o.normal[0] = r.nx; o.normal[1] = r.ny; o.normal[2] = r.nz;
        o.witness_p[0] = xh[3 * i]; o.witness_p[1] = xh[3 * i + 1]; o.witness_p[2] = xh[3 * i + 2];
        o.witness_q[0] = r.qx; o.witness_q[1] = r.qy; o.witness_q[2] = r.qz;
        o.weights[0] = 1.0f;
      }
  });
}

int ando_get_pattern(ando_ctx* c, int32_t* row_ptr, int32_t* cols, int32_t* num_blocks) {
  return guarded([&] {
    CtxLock lock(c);
    need(c && c->have_mesh && num_blocks, ANDO_ESTATE, "mesh has not been set");
    *num_blocks = c->nb;
    if (row_ptr) std::memcpy(row_ptr, c->row_ptr.data(), sizeof(int32_t) * (c->n + 1));
    if (cols) std::memcpy(cols, c->cols.data(), sizeof(int32_t) * c->nb);
  });
}

int ando_linearize(ando_ctx* c, const ando_params* params, const float* x_target, const ando_rigid_body* bodies,
                   int32_t num_bodies, float* gradient, float* values) {
  return guarded([&] {
    CtxLock lock(c);
    need(c && c->have_state, ANDO_ESTATE, "state has not been set");
    check_params(params);
    need(x_target != nullptr, ANDO_EINVAL, "x_target is null");
    HIPCHK(hipSetDevice(c->device));
    const ando_params& p = *params;
    ensure_const(c, p);
    c->xhat.upload(x_target, 3 * (size_t)c->n, c->stream);
    upload_bodies(c, bodies, num_bodies);
    c->barrier_tol = p.barrier_tolerance;
    if (c->cp) {
      c->cpD = detect_bodies(c, bodies);
      detect_cp(c, c->cpD, p.contact_gap_max, nullptr);
      refresh_cp(c, p);
    } else {
      detect_rigid(c, bodies, num_bodies, p.dt);
      detect_self(c);
      prepare_self(c, step_params(c, p));
    }
    launch_gradient(c, p, true);
    launch_assemble(c, p, true);
    if (gradient)
      HIPCHK(hipMemcpyAsync(gradient, c->g.p, 3 * sizeof(float) * (size_t)c->n, hipMemcpyDeviceToHost, c->stream));
    if (values)
      HIPCHK(hipMemcpyAsync(values, c->H.p, 9 * sizeof(float) * (size_t)c->nb, hipMemcpyDeviceToHost, c->stream));
    sync(c);
  });
}

int ando_apply_restitution(ando_ctx* c, const ando_params* params, ando_rigid_body* bodies, int32_t num_bodies) {
  return guarded([&] {
    CtxLock lock(c);
    need(c && c->have_state, ANDO_ESTATE, "state has not been set");
    check_params(params);
    need(num_bodies >= 0 && (num_bodies == 0 || bodies != nullptr), ANDO_EINVAL, "bodies is null");
    HIPCHK(hipSetDevice(c->device));
    ensure_const(c, *params);
    restitution(c, *params, bodies, num_bodies);
    sync(c);
  });
}

int ando_pcg(ando_ctx* c, int32_t num_rows, const int32_t* row_ptr, const int32_t* cols, const float* values,
             const float* b, float* x, float tol, int32_t max_iters, int32_t* iterations, int32_t* converged) {
  return guarded([&] {
    CtxLock lock(c);
    need(c && row_ptr && cols && values && b && x, ANDO_EINVAL, "null argument");
    need(num_rows > 0 && max_iters >= 0, ANDO_EINVAL, "bad sizes");
    HIPCHK(hipSetDevice(c->device));
    int n = num_rows;
    int nbk = row_ptr[n];
    need(nbk >= n, ANDO_EINVAL, "pattern needs a diagonal block per row");
    std::vector<int32_t> dg(n, -1);
    for (int i = 0; i < n; ++i) {
      need(row_ptr[i + 1] >= row_ptr[i], ANDO_EINVAL, "row_ptr not monotone");
      for (int k = row_ptr[i]; k < row_ptr[i + 1]; ++k) {
        need(cols[k] >= 0 && cols[k] < n, ANDO_EINVAL, "column index out of range");
        if (cols[k] == i) dg[i] = k;
      }
      need(dg[i] >= 0, ANDO_EINVAL, "every row needs its diagonal block");
    }
    DBuf<int32_t> drp, dci, ddg;
    DBuf<float> dv, db, dx;
    hipStream_t s = c->stream;
    drp.upload(row_ptr, n + 1, s);
    dci.upload(cols, nbk, s);
    ddg.upload(dg, s);
    dv.upload(values, 9 * (size_t)nbk, s);
    db.upload(b, 3 * (size_t)n, s);
    dx.ensure(3 * (size_t)n);
    if (n > c->n || !c->r.p) {
      size_t n3 = 3 * (size_t)n;
      c->r.ensure(n3); c->z.ensure(n3); c->pbuf0.ensure(n3); c->pbuf1.ensure(n3); c->Ap.ensure(n3);
      c->Pinv.ensure(9 * (size_t)n);
      c->part_rz.ensure(nblk(n) + 1); c->part_max.ensure(nblk(n) + 1); c->part_pAp.ensure(spmv_groups(n) + 1);
    }
    PcgOut po{0, false};
    if (c->two_level) {
      // Rows in the Morton order of the context's mesh when the system has its
      // size, otherwise in the given order.
      std::vector<int32_t> order(n);
      if (c->have_mesh && n == c->n) order = c->morton;
      else
        for (int i = 0; i < n; ++i) order[i] = i;
      TlMeta T;
      build_tl_meta(T, n, std::vector<int32_t>(row_ptr, row_ptr + n + 1), std::vector<int32_t>(cols, cols + nbk),
                    order, s);
      need(T.ok, ANDO_EINVAL, "two-level PCG: sliced layout too large");
      po = run_pcg_tl(c, T, dv.p, ddg.p, db.p, 0, dx.p, tol, max_iters, nullptr);
      sync(c);  // T's buffers are released on return
    } else {
      SpmvMeta sm;
      build_spmv_meta(sm, n, std::vector<int32_t>(row_ptr, row_ptr + n + 1), std::vector<int32_t>(cols, cols + nbk),
                      s);
      po = run_pcg(c, n, dv.p, drp.p, dci.p, ddg.p, db.p, 0, dx.p, tol, max_iters, &sm);
    }
    HIPCHK(hipMemcpyAsync(x, dx.p, 3 * sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, s));
    sync(c);
    if (iterations) *iterations = po.iters;
    if (converged) *converged = po.converged ? 1 : 0;
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free PCG (ando_pcg_noalloc).  Its work vectors, preconditioner,
// partial sums, history, status and error word live in one file-scope device
// array, `scratch`, sized for kNoAllocMaxRows block rows and kNoAllocMaxIters
// iterations.  Each GPU holds its own copy for the life of the process; host
// code never names the array in a launch or copy, only its per-device address
// from hipGetSymbolAddress.  Calls on one GPU share the array: the library
// serialises them with a per-device mutex, and every buffer that must start at
// zero is cleared by a kernel at the start of each call.
// ---------------------------------------------------------------------------
namespace {

constexpr int kNoAllocMaxRows = 1 << 17;
constexpr int kNoAllocMaxIters = 8192;
constexpr size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

struct NoAllocLayout {
  size_t r, z, p0, p1, Ap, Pinv, diag, part_rz, part_max, part_pAp, rz_hist, status, err, total;
};
constexpr NoAllocLayout noalloc_layout() {
  NoAllocLayout L{};
  const size_t N = kNoAllocMaxRows;
  const size_t vec = align256(12 * N);
  const size_t wg_v = (N + kBlock - 1) / kBlock + 1;
  const size_t wg_r = (N + ando::kRowsPerPass - 1) / ando::kRowsPerPass + 1;
  size_t o = 0;
  L.r = o; o += vec;
  L.z = o; o += vec;
  L.p0 = o; o += vec;
  L.p1 = o; o += vec;
  L.Ap = o; o += vec;
  L.Pinv = o; o += align256(36 * N);
  L.diag = o; o += align256(4 * N);
  L.part_rz = o; o += align256(8 * wg_v);
  L.part_max = o; o += align256(4 * wg_v);
  L.part_pAp = o; o += align256(8 * wg_r);
  L.rz_hist = o; o += align256(8 * (size_t)(kNoAllocMaxIters + 2));
  L.status = o; o += align256(sizeof(ando::PcgStatus));
  L.err = o; o += align256(sizeof(uint32_t));
  L.total = o;
  return L;
}
constexpr NoAllocLayout kNoAlloc = noalloc_layout();

static __device__ __align__(256) unsigned char scratch[kNoAlloc.total];

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu[kMaxDevices];
unsigned char* g_noalloc_ptr[kMaxDevices] = {};  // device address of `scratch`, per device

// Checks a BSR pattern before any kernel follows it: row_ptr starts at 0, is
// monotone and ends at num_blocks, every column is a row index, and every row
// has its diagonal block (whose position goes to diag).  Violations set bits
// of *err: 1 row pointers, 2 columns, 4 a missing diagonal.
__global__ void k_check_pattern(int n, int nb, const int32_t* __restrict__ rp, const int32_t* __restrict__ cols,
                                int32_t* __restrict__ diag, uint32_t* __restrict__ err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int a = rp[i], b = rp[i + 1];
  if ((i == 0 && a != 0) || (i == n - 1 && b != nb) || a < 0 || b < a || b > nb) {
    atomicOr(err, 1u);
    diag[i] = -1;
    return;
  }
  int d = -1;
  for (int k = a; k < b; ++k) {
    const int c = cols[k];
    if (c < 0 || c >= n) atomicOr(err, 2u);
    if (c == i) d = k;
  }
  if (d < 0) atomicOr(err, 4u);
  diag[i] = d;
}

__global__ void k_zero_u32(uint32_t* __restrict__ p, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0u;
}

}  // namespace

extern "C" {

int ando_pcg_noalloc(int32_t device, int32_t num_rows, int32_t num_blocks, const int32_t* row_ptr, const int32_t* cols,
                     const float* values, const float* b, float* x, float tol, int32_t max_iters, void* stream,
                     int32_t* iterations, int32_t* converged) {
  return guarded([&] {
    need(row_ptr && cols && values && b && x, ANDO_EINVAL, "null argument");
    need(device >= 0 && device < kMaxDevices, ANDO_EINVAL, "device index out of range");
    need(num_rows > 0 && num_blocks >= num_rows, ANDO_EINVAL, "bad sizes");
    need(num_rows <= kNoAllocMaxRows, ANDO_EINVAL,
         "ando_pcg_noalloc: more than 131072 block rows (its fixed scratch holds at most 131072)");
    need(max_iters >= 0 && max_iters <= kNoAllocMaxIters, ANDO_EINVAL,
         "ando_pcg_noalloc: max_iters must be within 0..8192 (its fixed scratch holds that history)");
    std::lock_guard<std::mutex> lk(g_noalloc_mu[device]);
    HIPCHK(hipSetDevice(device));
    if (!g_noalloc_ptr[device]) {
      void* ptr = nullptr;
      HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
      need(ptr != nullptr, ANDO_EDEVICE, "scratch has no device address");
      g_noalloc_ptr[device] = static_cast<unsigned char*>(ptr);
    }
    unsigned char* base = g_noalloc_ptr[device];
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int n = num_rows;
    float* r = reinterpret_cast<float*>(base + kNoAlloc.r);
    float* z = reinterpret_cast<float*>(base + kNoAlloc.z);
    float* p0 = reinterpret_cast<float*>(base + kNoAlloc.p0);
    float* p1 = reinterpret_cast<float*>(base + kNoAlloc.p1);
    float* Ap = reinterpret_cast<float*>(base + kNoAlloc.Ap);
    float* Pinv = reinterpret_cast<float*>(base + kNoAlloc.Pinv);
    int32_t* diag = reinterpret_cast<int32_t*>(base + kNoAlloc.diag);
    double* part_rz = reinterpret_cast<double*>(base + kNoAlloc.part_rz);
    float* part_max = reinterpret_cast<float*>(base + kNoAlloc.part_max);
    double* part_pAp = reinterpret_cast<double*>(base + kNoAlloc.part_pAp);
    double* rz_hist = reinterpret_cast<double*>(base + kNoAlloc.rz_hist);
    PcgStatus* st = reinterpret_cast<PcgStatus*>(base + kNoAlloc.status);
    uint32_t* err = reinterpret_cast<uint32_t*>(base + kNoAlloc.err);
    // The pattern first: nothing below reads a column or row pointer it has not vetted.
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, err, 1);
    hipLaunchKernelGGL(k_check_pattern, dim3(nblk(n)), dim3(kBlock), 0, s, n, num_blocks, row_ptr, cols, diag, err);
    HIPCHK(hipGetLastError());
    uint32_t e = 0;
    HIPCHK(hipMemcpyAsync(&e, err, sizeof(e), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    need(!(e & 1u), ANDO_EINVAL, "row_ptr must start at 0, be monotone and end at num_blocks");
    need(!(e & 2u), ANDO_EINVAL, "column index out of range");
    need(!(e & 4u), ANDO_EINVAL, "every row needs its diagonal block");
    // PCGSolver::solve (pcg_solver.cpp:8-146) with the row SpMV, as ando_pcg
    // runs it in ANDO_MODE_PCG_ROW_SPMV: same kernels, grid and chunk schedule.
    constexpr int R = 3;
    const int nwg_v = nblk(n), nwg_r = spmv_groups(n, R);
    hipLaunchKernelGGL(k_pcg_init, dim3(nwg_v), dim3(kBlock), 0, s, n, values, (const int32_t*)diag, b, x, r, z, p0,
                       Pinv, n, 1, part_rz, part_max, 0);
    hipLaunchKernelGGL(k_pcg_start, dim3(1), dim3(kBlock), 0, s, nwg_v, (const double*)part_rz,
                       (const float*)part_max, (const float*)nullptr, rz_hist, st, tol);
    HIPCHK(hipGetLastError());
    PcgStatus hs{};
    int it = 0, chunk = 8;
    for (;;) {
      const int end = std::min(max_iters, it + chunk);
      for (; it < end; ++it) {
        float* pold = (it == 0) ? p0 : ((it - 1) & 1 ? p1 : p0);
        float* pcur = (it & 1) ? p1 : p0;
        hipLaunchKernelGGL(k_pcg_spmv<R>, dim3(nwg_r), dim3(kBlock), 0, s, n, it, values, row_ptr, cols,
                           (const float*)z, (const float*)pold, pcur, Ap, (const double*)part_rz,
                           (const float*)part_max, nwg_v, rz_hist, part_pAp, st, tol,
                           ando::Coupling{0, nwg_r, 0, nullptr, nullptr, nullptr, nullptr, nullptr}, 0);
        hipLaunchKernelGGL(k_pcg_update, dim3(nwg_v), dim3(kBlock), 0, s, n, it, (const float*)pcur,
                           (const float*)Ap, (const float*)nullptr, (const float*)Pinv, n, 1, x, r, z, (const double*)part_pAp, nwg_r,
                           (const double*)rz_hist, part_rz, part_max, st, 0);
      }
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(&hs, st, sizeof(hs), hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      if (hs.done) break;
      if (it >= max_iters) {
        hipLaunchKernelGGL(k_pcg_final, dim3(1), dim3(kBlock), 0, s, it, part_max, nwg_v, st, tol);
        HIPCHK(hipMemcpyAsync(&hs, st, sizeof(hs), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        break;
      }
      chunk = std::min(chunk * 2, 256);
    }
    if (iterations) *iterations = hs.iters;
    if (converged) *converged = hs.failed ? 0 : 1;
  });
}

int ando_count_intersections(ando_ctx* c, const ando_rigid_body* bodies, int32_t num_bodies, int32_t* edge_triangle,
                              int32_t* inside_rigid) {
  return guarded([&] {
    CtxLock lock(c);
    need(c && c->have_state, ANDO_ESTATE, "state has not been set");
    need(num_bodies == 0 || bodies != nullptr, ANDO_EINVAL, "bodies is null");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int E = (int)c->edges_h.size() / 2;
    int32_t et = 0, in = 0;
    if (E > 0 && c->m > 0) {
      lbvh_build(c, c->tri_tree, c->m, 3, c->d_tris.p);
      c->sc_err.ensure(1);
      HIPCHK(hipMemsetAsync(c->sc_err.p, 0, sizeof(uint32_t), s));
      c->sc_count.ensure((size_t)std::max(c->n, E) + 1);
      c->sc_ptr.ensure((size_t)std::max(c->n, E) + 2);
      const LbvhTree& T = c->tri_tree;
      hipLaunchKernelGGL(ando::lbvh::k_count_edge_tri, dim3(nblk(E, 64)), dim3(64), 0, s, E, (const float*)c->x.p,
                         (const int32_t*)c->d_edges.p, (const int32_t*)c->d_tris.p, T.P, (const int32_t*)T.left.p,
                         (const int32_t*)T.right.p, (const float*)T.cbox.p, (const float*)T.table.p,
                         (const int32_t*)T.vals.p, c->sc_count.p, c->sc_err.p);
      et = exclusive_scan(c, E, c->sc_count.p, c->sc_ptr.p, c->sc_sums, c->sc_total_d);
      uint32_t err = 0;
      HIPCHK(hipMemcpyAsync(&err, c->sc_err.p, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
      sync(c);
      if (err) throw Fail(ANDO_EDEVICE, "BVH traversal stack exceeded its depth");
    }
    if (num_bodies > 0) {
      upload_bodies(c, bodies, num_bodies);
      DetectBodies D = detect_bodies(c, bodies);
      c->sc_count.ensure((size_t)std::max(c->n, E) + 1);
      c->sc_ptr.ensure((size_t)std::max(c->n, E) + 2);
      hipLaunchKernelGGL(ando::lbvh::k_count_inside, dim3(nblk(c->n)), dim3(kBlock), 0, s, c->n,
                         (const float*)c->x.p, num_bodies, D, c->sc_count.p);
      in = exclusive_scan(c, c->n, c->sc_count.p, c->sc_ptr.p, c->sc_sums, c->sc_total_d);
    }
    if (edge_triangle) *edge_triangle = et;
    if (inside_rigid) *inside_rigid = in;
  });
}

int ando_step_batch(ando_ctx* const* ctxs, const ando_params* params, ando_rigid_body* const* bodies,
                    const int32_t* num_bodies, int32_t count, int32_t threads) {
  std::vector<int> rc;
  std::vector<std::string> err;
  const int status = guarded([&] {
    need(count >= 0 && (count == 0 || (ctxs && params)), ANDO_EINVAL, "null argument");
    std::vector<const ando_ctx*> seen(ctxs, ctxs + count);
    std::sort(seen.begin(), seen.end());
    need(std::adjacent_find(seen.begin(), seen.end()) == seen.end(), ANDO_EINVAL,
         "a context appears twice in the batch");
    for (int k = 0; k < count; ++k) need(ctxs[k] != nullptr, ANDO_EINVAL, "null context in the batch");
    rc.assign(count, ANDO_OK);
    err.assign(count, std::string());
    int T = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
    T = std::max(1, std::min({T, count, 256}));
    // Batched PCG when every context is on one device (ANDO_BATCH_PCG=0 turns it off).
    std::unique_ptr<BatchPcg> batch;
    bool same_device = count > 0;
    for (int k = 1; k < count && same_device; ++k) same_device = ctxs[k]->device == ctxs[0]->device;
    const char* bp = std::getenv("ANDO_BATCH_PCG");
    if (same_device && T > 1 && !(bp && std::atoi(bp) == 0)) {
      batch.reset(new BatchPcg);
      batch->device = ctxs[0]->device;
      batch->active = T;
      batch->chunk0 = std::max(1, ctxs[0]->pcg_chunk);
      batch->chunk_max = ctxs[0]->pcg_chunk_max;
      if (const char* w = std::getenv("ANDO_BATCH_GATHER_US")) batch->gather_us = std::max(0, std::atoi(w));
      HIPCHK(hipSetDevice(batch->device));
      HIPCHK(hipStreamCreateWithFlags(&batch->stream, hipStreamNonBlocking));
    }
    std::atomic<int> next{0};
    auto work = [&] {
      t_batch = batch.get();
      for (int k = next++; k < count; k = next++) {
        rc[k] = ando_step(ctxs[k], &params[k], bodies ? bodies[k] : nullptr, num_bodies ? num_bodies[k] : 0);
        if (rc[k] != ANDO_OK) err[k] = ando_last_error();
      }
      t_batch = nullptr;
      if (batch) batch->worker_done();
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < T; ++t) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
    if (batch && std::getenv("ANDO_BATCH_REPORT"))
      std::fprintf(stderr, "ando_step_batch: %d contexts on %d threads, %ld solves in %ld batched rounds\n", count, T,
                   batch->solves, batch->rounds);
  });
  if (status != ANDO_OK) return status;
  for (int k = 0; k < (int)rc.size(); ++k)
    if (rc[k] != ANDO_OK) {
      set_error("batch entry " + std::to_string(k) + ": " + err[k]);
      return rc[k];
    }
  return ANDO_OK;
}

int ando_energy_diagnostics(ando_ctx* c, ando_energy* out) {
  return guarded([&] {
    CtxLock lock(c);
    need(c && out, ANDO_EINVAL, "null argument");
    need(c->have_state, ANDO_ESTATE, "state has not been set");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int G = std::max(1, std::min(ando::kFrameMaxBlocks, nblk(std::max(c->n, c->m), ando::kFrameBlock)));
    c->fe_part.ensure((size_t)G);
    c->fe_out.ensure(1);
    hipLaunchKernelGGL(ando::k_frame_energy, dim3(G), dim3(ando::kFrameBlock), 0, s, c->n, c->m,
                       (const float*)c->x.p, (const float*)c->v.p, (const float*)c->masses.p,
                       (const int32_t*)c->d_tris.p, (const float*)c->mverts.p, (const float*)c->dm.p,
                       (const float*)c->area.p, c->mat.thickness, c->mu, c->fe_part.p);
    hipLaunchKernelGGL(ando::k_frame_energy_final, dim3(1), dim3(ando::kFrameBlock), 0, s, G,
                       (const ando::EnergyAcc*)c->fe_part.p, c->fe_out.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, c->fe_out.p, sizeof(ando_energy), hipMemcpyDeviceToHost, s));
    sync(c);
  });
}

int ando_contact_metrics(ando_ctx* c, const ando_contact* contacts, int32_t count, float gap_max, int32_t ccd,
                         ando_collision_metrics* out) {
  return guarded([&] {
    CtxLock lock(c);
    need(c && out, ANDO_EINVAL, "null argument");
    need(c->have_state, ANDO_ESTATE, "state has not been set");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    int total = 0;
    if (contacts) {
      need(count >= 0, ANDO_EINVAL, "negative contact count");
      // Indices the relative-velocity terms read (the reference would read out
      // of bounds; here a bad index is an argument error).
      for (int e = 0; e < count; ++e) {
        const ando_contact& k = contacts[e];
        const int used = k.type == 0 || k.type == 1 ? 4 : (k.type == 2 ? 1 : 0);
        for (int q = 0; q < used; ++q)
          need(k.idx[q] >= 0 && k.idx[q] < c->n, ANDO_EINVAL, "contact vertex index out of range");
      }
      total = count;
      if (total) c->fm_list.upload(contacts, (size_t)total, s);
    } else {
      need(c->listed_gen == c->det_gen, ANDO_ESTATE,
           "the contacts listed by ando_compute_contacts have been replaced by a later detection");
      total = c->sc_total + c->rc_total;
    }
    const int G = std::max(1, std::min(ando::kFrameMaxBlocks, nblk(total, ando::kFrameBlock)));
    c->fm_part.ensure((size_t)G);
    c->fm_out.ensure(1);
    const int flag = ccd ? 1 : 0;
    if (contacts) {
      ando::ListSource src{c->fm_list.p};
      hipLaunchKernelGGL(ando::k_frame_metrics<ando::ListSource>, dim3(G), dim3(ando::kFrameBlock), 0, s, total, src,
                         (const float*)c->v.p, gap_max, flag, c->fm_part.p);
    } else {
      ando::TableSource src{c->sc.p, c->sc_total, c->rc.p};
      hipLaunchKernelGGL(ando::k_frame_metrics<ando::TableSource>, dim3(G), dim3(ando::kFrameBlock), 0, s, total,
                         src, (const float*)c->v.p, gap_max, flag, c->fm_part.p);
    }
    hipLaunchKernelGGL(ando::k_frame_metrics_final, dim3(1), dim3(ando::kFrameBlock), 0, s, G,
                       (const ando::MetricAcc*)c->fm_part.p, flag, c->fm_out.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, c->fm_out.p, sizeof(ando_collision_metrics), hipMemcpyDeviceToHost, s));
    sync(c);
  });
}

int ando_cfl_inputs(ando_ctx* c, float* max_velocity, float* min_edge_length) {
  return guarded([&] {
    CtxLock lock(c);
    need(c && max_velocity && min_edge_length, ANDO_EINVAL, "null argument");
    need(c->have_state, ANDO_ESTATE, "state has not been set");
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    const int E = (int)c->edges_h.size() / 2;
    const int G = std::max(1, std::min(ando::kFrameMaxBlocks, nblk(std::max(c->n, E), ando::kFrameBlock)));
    c->fx_part.ensure((size_t)G);
    c->fx_out.ensure(2);
    hipLaunchKernelGGL(ando::k_frame_extrema, dim3(G), dim3(ando::kFrameBlock), 0, s, c->n, E, (const float*)c->v.p,
                       (const int32_t*)c->d_edges.p, (const float*)c->mverts.p, c->fx_part.p);
    hipLaunchKernelGGL(ando::k_frame_extrema_final, dim3(1), dim3(ando::kFrameBlock), 0, s, G,
                       (const ando::ExtremaAcc*)c->fx_part.p, c->fx_out.p);
    HIPCHK(hipGetLastError());
    float out[2];
    HIPCHK(hipMemcpyAsync(out, c->fx_out.p, sizeof(out), hipMemcpyDeviceToHost, s));
    sync(c);
    *max_velocity = out[0];
    *min_edge_length = out[1];
  });
}

int32_t ando_nn_num_weights(void) { return ando::nn::kNumWeights; }

int ando_set_precond_weights(ando_ctx* c, const float* weights, int32_t count) {
  return guarded([&] {
    CtxLock lock(c);
    need(c && weights, ANDO_EINVAL, "null argument");
    need(count == ando::nn::kNumWeights, ANDO_EINVAL, "weight count differs from ando_nn_num_weights()");
    HIPCHK(hipSetDevice(c->device));
    nnh::upload(c->nn, weights, c->stream);
    sync(c);
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Trainer for the learned preconditioner.
// ---------------------------------------------------------------------------
struct ando_trainer {
  ando_ctx* ctx = nullptr;
  nnh::Net net;
  nnh::Work wk;
  struct System {
    int n = 0, nb = 0;
    DBuf<int32_t> rp, cols, diag;
    DBuf<float> A, b;
    SpmvMeta sm;
  };
  std::vector<System> systems;
  DBuf<double> loss_hist;
};

namespace {

void ensure_pcg_buffers(ando_ctx* c, int n) {
  const size_t n3 = 3 * (size_t)n;
  c->r.ensure(n3); c->z.ensure(n3); c->pbuf0.ensure(n3); c->pbuf1.ensure(n3); c->Ap.ensure(n3);
  c->Pinv.ensure(9 * (size_t)n);
  c->part_rz.ensure(nblk(n) + 1); c->part_max.ensure(nblk(n) + 1); c->part_pAp.ensure(spmv_groups(n) + 1);
}

ando_trainer::System& trainer_system(ando_trainer* tr, int32_t i) {
  need(tr != nullptr, ANDO_EINVAL, "null trainer");
  need(i >= 0 && i < (int32_t)tr->systems.size(), ANDO_EINVAL, "system index out of range");
  need(tr->net.ready, ANDO_ESTATE, "trainer has no weights (ando_trainer_init_weights or _set_weights)");
  return tr->systems[i];
}

}  // namespace

extern "C" {

int ando_trainer_create(int32_t device, ando_trainer** out) {
  return guarded([&] {
    need(out != nullptr, ANDO_EINVAL, "out is null");
    ando_ctx* c = nullptr;
    int rc = ando_create(device, &c);
    if (rc != ANDO_OK) throw Fail(rc, ando::g_last_error);
    auto* tr = new ando_trainer;
    tr->ctx = c;
    *out = tr;
  });
}

int ando_trainer_destroy(ando_trainer* tr) {
  if (!tr) return ANDO_OK;
  ando_ctx* c = tr->ctx;
  (void)hipSetDevice(c->device);
  delete tr;
  return ando_destroy(c);
}

int ando_trainer_add_system(ando_trainer* tr, int32_t num_rows, const int32_t* row_ptr, const int32_t* cols,
                            const float* values, const float* b, int32_t* index) {
  return guarded([&] {
    need(tr && row_ptr && cols && values && b, ANDO_EINVAL, "null argument");
    need(num_rows > 0, ANDO_EINVAL, "bad sizes");
    const int n = num_rows, nbk = row_ptr[n];
    std::vector<int32_t> dg(n, -1);
    for (int i = 0; i < n; ++i) {
      need(row_ptr[i + 1] >= row_ptr[i], ANDO_EINVAL, "row_ptr not monotone");
      for (int k = row_ptr[i]; k < row_ptr[i + 1]; ++k) {
        need(cols[k] >= 0 && cols[k] < n, ANDO_EINVAL, "column index out of range");
        if (cols[k] == i) dg[i] = k;
      }
      need(dg[i] >= 0, ANDO_EINVAL, "every row needs its diagonal block");
    }
    ando_ctx* c = tr->ctx;
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    tr->systems.emplace_back();
    auto& S = tr->systems.back();
    S.n = n;
    S.nb = nbk;
    S.rp.upload(row_ptr, n + 1, s);
    S.cols.upload(cols, nbk, s);
    S.diag.upload(dg, s);
    S.A.upload(values, 9 * (size_t)nbk, s);
    S.b.upload(b, 3 * (size_t)n, s);
    build_spmv_meta(S.sm, n, std::vector<int32_t>(row_ptr, row_ptr + n + 1), std::vector<int32_t>(cols, cols + nbk),
                    s);
    sync(c);
    if (index) *index = (int32_t)tr->systems.size() - 1;
  });
}

int ando_trainer_init_weights(ando_trainer* tr, uint64_t seed) {
  return guarded([&] {
    need(tr != nullptr, ANDO_EINVAL, "null trainer");
    HIPCHK(hipSetDevice(tr->ctx->device));
    auto w = nnh::init_weights(seed);
    nnh::upload(tr->net, w.data(), tr->ctx->stream);
    sync(tr->ctx);
  });
}

int ando_trainer_set_weights(ando_trainer* tr, const float* weights, int32_t count) {
  return guarded([&] {
    need(tr && weights, ANDO_EINVAL, "null argument");
    need(count == ando::nn::kNumWeights, ANDO_EINVAL, "weight count differs from ando_nn_num_weights()");
    HIPCHK(hipSetDevice(tr->ctx->device));
    nnh::upload(tr->net, weights, tr->ctx->stream);
    sync(tr->ctx);
  });
}

int ando_trainer_get_weights(ando_trainer* tr, float* weights, int32_t count) {
  return guarded([&] {
    need(tr && weights, ANDO_EINVAL, "null argument");
    need(count == ando::nn::kNumWeights, ANDO_EINVAL, "weight count differs from ando_nn_num_weights()");
    need(tr->net.ready, ANDO_ESTATE, "trainer has no weights");
    HIPCHK(hipSetDevice(tr->ctx->device));
    HIPCHK(hipMemcpyAsync(weights, tr->net.w.p, sizeof(float) * count, hipMemcpyDeviceToHost, tr->ctx->stream));
    sync(tr->ctx);
  });
}

int ando_trainer_forward(ando_trainer* tr, int32_t system, float* P, float* y) {
  return guarded([&] {
    auto& S = trainer_system(tr, system);
    ando_ctx* c = tr->ctx;
    HIPCHK(hipSetDevice(c->device));
    nnh::forward(tr->net, tr->wk, S.n, S.A.p, S.rp.p, S.cols.p, S.diag.p, S.b.p, 0, false, c->stream);
    if (P) HIPCHK(hipMemcpyAsync(P, tr->wk.P.p, sizeof(float) * 9 * S.n, hipMemcpyDeviceToHost, c->stream));
    if (y) HIPCHK(hipMemcpyAsync(y, tr->wk.y.p, sizeof(float) * 3 * S.n, hipMemcpyDeviceToHost, c->stream));
    sync(c);
  });
}

int ando_trainer_loss_grad(ando_trainer* tr, int32_t system, double* loss, float* grad) {
  return guarded([&] {
    auto& S = trainer_system(tr, system);
    ando_ctx* c = tr->ctx;
    HIPCHK(hipSetDevice(c->device));
    nnh::forward(tr->net, tr->wk, S.n, S.A.p, S.rp.p, S.cols.p, S.diag.p, S.b.p, 0, true, c->stream);
    nnh::backward(tr->net, tr->wk, S.n, S.A.p, S.rp.p, S.cols.p, S.b.p, 0, c->stream);
    double fin[3] = {0, 0, 0};
    HIPCHK(hipMemcpyAsync(fin, tr->wk.lfin.p, sizeof(fin), hipMemcpyDeviceToHost, c->stream));
    if (grad)
      HIPCHK(hipMemcpyAsync(grad, tr->net.grad.p, sizeof(float) * ando::nn::kNumWeights, hipMemcpyDeviceToHost,
                            c->stream));
    sync(c);
    if (loss) *loss = fin[1];
  });
}

int ando_trainer_train(ando_trainer* tr, int32_t steps, float lr, uint64_t seed, double* losses) {
  return guarded([&] {
    need(tr != nullptr, ANDO_EINVAL, "null trainer");
    need(steps >= 0 && lr > 0.0f, ANDO_EINVAL, "bad steps or learning rate");
    need(!tr->systems.empty(), ANDO_ESTATE, "trainer has no systems");
    need(tr->net.ready, ANDO_ESTATE, "trainer has no weights");
    ando_ctx* c = tr->ctx;
    HIPCHK(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    tr->loss_hist.ensure((size_t)steps + 1);
    uint64_t st = seed;
    for (int t = 0; t < steps; ++t) {
      uint64_t z = (st += 0x9E3779B97F4A7C15ull);
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      z ^= z >> 31;
      auto& S = tr->systems[z % tr->systems.size()];
      nnh::forward(tr->net, tr->wk, S.n, S.A.p, S.rp.p, S.cols.p, S.diag.p, S.b.p, 0, true, s);
      nnh::backward(tr->net, tr->wk, S.n, S.A.p, S.rp.p, S.cols.p, S.b.p, 0, s);
      HIPCHK(hipMemcpyAsync(tr->loss_hist.p + t, tr->wk.lfin.p + 1, sizeof(double), hipMemcpyDeviceToDevice, s));
      nnh::adam(tr->net, lr, s);
    }
    if (losses && steps > 0)
      HIPCHK(hipMemcpyAsync(losses, tr->loss_hist.p, sizeof(double) * steps, hipMemcpyDeviceToHost, s));
    sync(c);
  });
}

int ando_trainer_solve(ando_trainer* tr, int32_t system, int32_t learned, float tol, int32_t max_iters,
                       int32_t* iterations, int32_t* converged) {
  return guarded([&] {
    auto& S = trainer_system(tr, system);
    need(max_iters >= 0, ANDO_EINVAL, "bad max_iters");
    ando_ctx* c = tr->ctx;
    HIPCHK(hipSetDevice(c->device));
    ensure_pcg_buffers(c, S.n);
    DBuf<float> x;
    x.ensure(3 * (size_t)S.n);
    PcgOut po;
    if (learned) {
      nnh::forward(tr->net, tr->wk, S.n, S.A.p, S.rp.p, S.cols.p, S.diag.p, S.b.p, 0, false, c->stream);
      nnh::warm_start(tr->wk, S.n, S.A.p, S.rp.p, S.cols.p, S.b.p, 0, c->stream);
      po = run_pcg(c, S.n, S.A.p, S.rp.p, S.cols.p, S.diag.p, S.b.p, 0, x.p, tol, max_iters, &S.sm,
                   learned == 2 ? tr->wk.Dinv.p : tr->wk.P.p, tr->wk.y.p);
    } else {
      po = run_pcg(c, S.n, S.A.p, S.rp.p, S.cols.p, S.diag.p, S.b.p, 0, x.p, tol, max_iters, &S.sm);
    }
    sync(c);
    if (iterations) *iterations = po.iters;
    if (converged) *converged = po.converged ? 1 : 0;
  });
}

}  // extern "C"
