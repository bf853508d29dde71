// Per-frame consumers of the add-on on the device: EnergyTracker::compute
// (energy_tracker.cpp:7-76) and CollisionValidator::compute_metrics
// (collision_validator.cpp:7-157,178-230).  Each per-item term is the
// reference's float expression; sums are taken in double (the reference
// accumulates in float), in a fixed tree order so a result does not depend on
// scheduling.  Included by ctx.hip after kernels.hip.
#include "../../include/andosim_hip.h"

namespace ando {

constexpr int kFrameBlock = 256;
constexpr int kFrameMaxBlocks = 512;

// Tree reduction of one value per thread over a kFrameBlock workgroup; the
// result is valid in thread 0.
template <class T>
__device__ inline void block_tree(T& v, T* lds) {
  const int t = threadIdx.x;
  lds[t] = v;
  __syncthreads();
  for (int s = kFrameBlock / 2; s > 0; s >>= 1) {
    if (t < s) lds[t].add(lds[t + s]);
    __syncthreads();
  }
  v = lds[0];
}

struct EnergyAcc {
  double ke, el, lm[3], am[3];
  float vmax;
  __device__ void zero() {
    ke = el = 0.0;
    for (int k = 0; k < 3; ++k) lm[k] = am[k] = 0.0;
    vmax = 0.0f;
  }
  __device__ void add(const EnergyAcc& o) {
    ke += o.ke;
    el += o.el;
    for (int k = 0; k < 3; ++k) {
      lm[k] += o.lm[k];
      am[k] += o.am[k];
    }
    vmax = fmaxf(vmax, o.vmax);
  }
};

// Elastic energy of one face, Elasticity::compute_energy (elasticity.cpp:5-14,
// 115-126) with F = Ds Dm^-1 from the mesh vertices (mesh.cpp:126-150).
__device__ inline float face_energy(int f, const int32_t* tris, const float* mverts, const float* dm,
                                    const float* area, float thick, float mu) {
  const float M00 = dm[4 * f], M01 = dm[4 * f + 1], M10 = dm[4 * f + 2], M11 = dm[4 * f + 3];
  const f3 v0 = ld3(mverts, tris[3 * f]), v1 = ld3(mverts, tris[3 * f + 1]), v2 = ld3(mverts, tris[3 * f + 2]);
  const f3 e1 = v1 - v0, e2 = v2 - v0;
  const f3 nn = normalized(cross(e1, e2));
  const f3 t1 = normalized(e1);
  const f3 t2 = cross(nn, t1);
  const float D00 = dot(e1, t1), D01 = dot(e2, t1), D10 = dot(e1, t2), D11 = dot(e2, t2);
  const float F00 = D00 * M00 + D01 * M10, F01 = D00 * M01 + D01 * M11;
  const float F10 = D10 * M00 + D11 * M10, F11 = D10 * M01 + D11 * M11;
  const float k = area[f] * thick * mu;
  const float d00 = F00 - 1.0f, d01 = F01, d10 = F10, d11 = F11 - 1.0f;
  const float sq = (d00 * d00 + d10 * d10) + (d01 * d01 + d11 * d11);
  return k * sq;
}

// One partial per workgroup over vertices [0, n) and faces [0, m).
__global__ void __launch_bounds__(kFrameBlock)
k_frame_energy(int n, int m, const float* __restrict__ x, const float* __restrict__ v,
               const float* __restrict__ masses, const int32_t* __restrict__ tris, const float* __restrict__ mverts,
               const float* __restrict__ dm, const float* __restrict__ area, float thick, float mu,
               EnergyAcc* __restrict__ part) {
  __shared__ EnergyAcc lds[kFrameBlock];
  EnergyAcc a;
  a.zero();
  const int stride = gridDim.x * kFrameBlock;
  const int top = n > m ? n : m;
  for (int i = blockIdx.x * kFrameBlock + threadIdx.x; i < top; i += stride) {
    if (i < n) {
      const f3 vi = ld3(v, i), ri = ld3(x, i);
      const float mi = masses[i];
      const float vsq = dot(vi, vi);
      a.ke += 0.5 * (double)mi * (double)vsq;                 // energy_tracker.cpp:7-15
      const f3 p = mi * vi;                                   // :17-23
      a.lm[0] += (double)p.x;
      a.lm[1] += (double)p.y;
      a.lm[2] += (double)p.z;
      const f3 c = cross(ri, p);                              // :25-33
      a.am[0] += (double)c.x;
      a.am[1] += (double)c.y;
      a.am[2] += (double)c.z;
      a.vmax = fmaxf(a.vmax, sqrtf(vsq));                     // :35-42
    }
    if (i < m) a.el += (double)face_energy(i, tris, mverts, dm, area, thick, mu);
  }
  block_tree(a, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = a;
}

__global__ void __launch_bounds__(kFrameBlock)
k_frame_energy_final(int G, const EnergyAcc* __restrict__ part, ando_energy* __restrict__ out) {
  __shared__ EnergyAcc lds[kFrameBlock];
  EnergyAcc a;
  a.zero();
  for (int g = threadIdx.x; g < G; g += kFrameBlock) a.add(part[g]);
  block_tree(a, lds);
  if (threadIdx.x == 0) {
    out->kinetic_energy = a.ke;
    out->elastic_energy = a.el;
    for (int k = 0; k < 3; ++k) {
      out->linear_momentum[k] = a.lm[k];
      out->angular_momentum[k] = a.am[k];
    }
    out->max_velocity = (double)a.vmax;
  }
}

struct MetricAcc {
  double sgap, spen, srel;
  float gmin, gmax, pmax, rmax;
  int npt, nee, nwall, ntot, npen, nccd, nrel, major, tunnel;
  __device__ void zero() {
    sgap = spen = srel = 0.0;
    gmin = INFINITY;
    gmax = -INFINITY;
    pmax = rmax = 0.0f;
    npt = nee = nwall = ntot = npen = nccd = nrel = major = tunnel = 0;
  }
  __device__ void add(const MetricAcc& o) {
    sgap += o.sgap;
    spen += o.spen;
    srel += o.srel;
    gmin = fminf(gmin, o.gmin);
    gmax = fmaxf(gmax, o.gmax);
    pmax = fmaxf(pmax, o.pmax);
    rmax = fmaxf(rmax, o.rmax);
    npt += o.npt;
    nee += o.nee;
    nwall += o.nwall;
    ntot += o.ntot;
    npen += o.npen;
    nccd += o.nccd;
    nrel += o.nrel;
    major |= o.major;
    tunnel |= o.tunnel;
  }
};

// What compute_metrics reads of a contact.
struct MetricItem {
  int type;
  int idx[4];
  float gap;
  f3 n;
};

// The contacts ando_compute_contacts lists: self contacts, then rigid ones.
struct TableSource {
  const SelfContact* sc;
  int nsc;
  const RigidContact* rc;
  __device__ MetricItem get(int e) const {
    MetricItem it;
    if (e < nsc) {
      const SelfContact& s = sc[e];
      it.type = s.type;
      for (int k = 0; k < 4; ++k) it.idx[k] = s.idx[k];
      it.gap = s.gap;
      it.n = f3{s.nx, s.ny, s.nz};
    } else {
      const RigidContact& r = rc[e - nsc];
      it.type = 3;
      it.idx[0] = r.vertex;
      it.idx[1] = it.idx[2] = it.idx[3] = -1;
      it.gap = r.gap;
      it.n = f3{r.nx, r.ny, r.nz};
    }
    return it;
  }
};

// A caller-supplied contact list (indices checked on the host).
struct ListSource {
  const ando_contact* c;
  __device__ MetricItem get(int e) const {
    const ando_contact& s = c[e];
    MetricItem it;
    it.type = s.type;
    for (int k = 0; k < 4; ++k) it.idx[k] = s.idx[k];
    it.gap = s.gap;
    it.n = f3{s.normal[0], s.normal[1], s.normal[2]};
    return it;
  }
};

// CollisionMetrics::compute_from_contacts (collision_validator.cpp:7-111) and
// compute_relative_velocities (:178-230) for one contact.
__device__ inline void metric_item(const MetricItem& c, const float* v, float gap_max, int ccd, MetricAcc& a) {
  ++a.ntot;
  if (c.type == 0) ++a.npt;
  else if (c.type == 1) ++a.nee;
  else if (c.type == 2) ++a.nwall;
  const float g = c.gap;
  a.gmin = fminf(a.gmin, g);
  a.gmax = fmaxf(a.gmax, g);
  a.sgap += (double)g;
  if (g < 0.0f) {
    ++a.npen;
    const float pen = -g;
    a.pmax = fmaxf(a.pmax, pen);
    a.spen += (double)pen;
    if ((double)pen > 0.001) a.major = 1;
    if ((double)pen > 0.1 * (double)gap_max) a.tunnel = 1;
  }
  if (ccd && (double)g < (double)gap_max * 0.5) ++a.nccd;
  float rel;
  if (c.type == 0) {
    const f3 vp = ld3(v, c.idx[0]), va = ld3(v, c.idx[1]), vb = ld3(v, c.idx[2]), vd = ld3(v, c.idx[3]);
    const f3 vt{(float)((double)((va.x + vb.x) + vd.x) / 3.0), (float)((double)((va.y + vb.y) + vd.y) / 3.0),
                (float)((double)((va.z + vb.z) + vd.z) / 3.0)};
    rel = fabsf(dot(vp - vt, c.n));
  } else if (c.type == 1) {
    const f3 va = ld3(v, c.idx[0]), vb = ld3(v, c.idx[1]), vd = ld3(v, c.idx[2]), ve = ld3(v, c.idx[3]);
    const f3 e0{(float)((double)(va.x + vb.x) / 2.0), (float)((double)(va.y + vb.y) / 2.0),
                (float)((double)(va.z + vb.z) / 2.0)};
    const f3 e1{(float)((double)(vd.x + ve.x) / 2.0), (float)((double)(vd.y + ve.y) / 2.0),
                (float)((double)(vd.z + ve.z) / 2.0)};
    rel = fabsf(dot(e0 - e1, c.n));
  } else if (c.type == 2) {
    rel = fabsf(dot(ld3(v, c.idx[0]), c.n));
  } else {
    return;  // rigid contacts carry no relative velocity in the reference
  }
  ++a.nrel;
  a.rmax = fmaxf(a.rmax, rel);
  a.srel += (double)rel;
}

template <class Src>
__global__ void __launch_bounds__(kFrameBlock)
k_frame_metrics(int count, Src src, const float* __restrict__ v, float gap_max, int ccd, MetricAcc* __restrict__ part) {
  __shared__ MetricAcc lds[kFrameBlock];
  MetricAcc a;
  a.zero();
  for (int e = blockIdx.x * kFrameBlock + threadIdx.x; e < count; e += gridDim.x * kFrameBlock)
    metric_item(src.get(e), v, gap_max, ccd, a);
  block_tree(a, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = a;
}

// Sums the partials and forms the reported metrics (averages as float
// quotients of the rounded sums, as the reference divides its float sums).
__global__ void __launch_bounds__(kFrameBlock)
k_frame_metrics_final(int G, const MetricAcc* __restrict__ part, int ccd, ando_collision_metrics* __restrict__ out) {
  __shared__ MetricAcc lds[kFrameBlock];
  MetricAcc a;
  a.zero();
  for (int g = threadIdx.x; g < G; g += kFrameBlock) a.add(part[g]);
  block_tree(a, lds);
  if (threadIdx.x != 0) return;
  ando_collision_metrics r;
  r.num_point_triangle = a.npt;
  r.num_edge_edge = a.nee;
  r.num_wall = a.nwall;
  r.num_total_contacts = a.ntot;
  r.num_penetrations = a.npen;
  r.num_ccd_contacts = ccd ? a.nccd : 0;
  r.num_broad_phase_contacts = a.ntot - r.num_ccd_contacts;
  r.num_relative_velocities = a.nrel;
  r.has_major_penetration = a.major;
  r.has_tunneling = a.tunnel;
  r.sum_gap = a.sgap;
  r.sum_penetration = a.spen;
  r.sum_relative_velocity = a.srel;
  if (a.ntot == 0) {
    r.min_gap = r.max_gap = r.avg_gap = 0.0f;
    r.max_penetration = r.avg_penetration = 0.0f;
    r.ccd_effectiveness = 0.0f;
  } else {
    r.min_gap = a.gmin;
    r.max_gap = a.gmax;
    r.avg_gap = (float)a.sgap / (float)a.ntot;
    r.max_penetration = a.pmax;
    r.avg_penetration = a.npen > 0 ? (float)a.spen / (float)a.npen : 0.0f;
    r.ccd_effectiveness =
        ccd ? (float)((double)((float)r.num_ccd_contacts / (float)a.ntot) * 100.0) : 0.0f;
  }
  r.max_relative_velocity = a.nrel ? a.rmax : 0.0f;
  r.avg_relative_velocity = a.nrel ? (float)a.srel / (float)a.nrel : 0.0f;
  *out = r;
}

// AdaptiveTimestep's inputs (adaptive_timestep.cpp:82-150): the largest
// squared velocity and the smallest finite squared edge length.
struct ExtremaAcc {
  float v2max, e2min;
  int found;
  __device__ void zero() {
    v2max = 0.0f;
    e2min = INFINITY;
    found = 0;
  }
  __device__ void add(const ExtremaAcc& o) {
    v2max = fmaxf(v2max, o.v2max);
    e2min = fminf(e2min, o.e2min);
    found |= o.found;
  }
};

__global__ void __launch_bounds__(kFrameBlock)
k_frame_extrema(int n, int E, const float* __restrict__ v, const int32_t* __restrict__ edges,
                const float* __restrict__ mverts, ExtremaAcc* __restrict__ part) {
  __shared__ ExtremaAcc lds[kFrameBlock];
  ExtremaAcc a;
  a.zero();
  const int top = n > E ? n : E;
  for (int i = blockIdx.x * kFrameBlock + threadIdx.x; i < top; i += gridDim.x * kFrameBlock) {
    if (i < n) {
      const f3 q = ld3(v, i);
      a.v2max = fmaxf(a.v2max, dot(q, q));
    }
    if (i < E) {
      const f3 d = ld3(mverts, edges[2 * i + 1]) - ld3(mverts, edges[2 * i]);
      const float l2 = dot(d, d);
      if (isfinite(l2)) {
        a.e2min = fminf(a.e2min, l2);
        a.found = 1;
      }
    }
  }
  block_tree(a, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = a;
}

// out[0] = max |v| (sqrt of the largest square), out[1] = min edge length
// (0 without a finite edge), as compute_max_velocity / compute_min_edge_length.
__global__ void __launch_bounds__(kFrameBlock)
k_frame_extrema_final(int G, const ExtremaAcc* __restrict__ part, float* __restrict__ out) {
  __shared__ ExtremaAcc lds[kFrameBlock];
  ExtremaAcc a;
  a.zero();
  for (int g = threadIdx.x; g < G; g += kFrameBlock) a.add(part[g]);
  block_tree(a, lds);
  if (threadIdx.x == 0) {
    out[0] = sqrtf(a.v2max);
    float r = a.found ? sqrtf(fmaxf(a.e2min, 0.0f)) : 0.0f;
    out[1] = isfinite(r) ? r : 0.0f;
  }
}

}  // namespace ando
