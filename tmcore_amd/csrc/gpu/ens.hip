// Batched ensemble of shallow-water members on meshes of uneven size,
// advanced by one persistent kernel per run.
//
// Layout: the members of a level group are cut into slices of consecutive
// members (tm_ens_set_slices); every state and intermediate array of a slice
// is [element][member] (member fastest).  The slices are the kernel's
// sub-levels: each XCD takes its contiguous eighth of every level's chunk
// list (slice after slice), so with eight slices an XCD advances one slice
// over the whole mesh and its gathered member rows stay in that XCD's L2.
// A workgroup takes a chunk of elements;
// each thread carries W consecutive members of one element as one vector
// value (W = 4 for FP32 and 2 for FP64 when M allows), so an element's
// constants (indices, TRiSK and fit weights, geometry) are loaded once per W
// members and each neighbour gather is one W-wide load of a contiguous row.
// The arithmetic per element is the single-model code (edge_prep_e, cell_e,
// vertex_e, edge_pv_e, edge_tend_e) with the Members<W> layout.
//
// Schedule: the run is T = max(nsteps) ticks.  Level l takes a step in tick t
// when floor((t+1) S_l / T) > floor(t S_l / T), so its S_l steps are spread
// evenly.  A tick is an edge pass, 4 stages x 3 phases + reduce + finalize;
// phases are separated by a grid barrier.  The barrier waits are bounded by a wall-clock
// limit: a wave that waits longer raises the abort flag, every other wave
// sees it at its next wait, and the kernel returns with an error instead of
// hanging.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "swm_kernels.h"
#include "tm_mesh_dev.h"
#include "tmcore_gpu.h"

namespace tmk {

constexpr int kEnsBlock = 256;
constexpr int kMaxLevels = 4;   // level groups
constexpr int kMaxSlices = 8;   // member slices per level group
constexpr int kMaxSub = kMaxLevels * kMaxSlices;

template <typename R>
struct EnsLevelDev {
  DMesh m;
  int M;      // members
  int nT;     // tracers per member
  double dt;
  int nsteps;
  int pv_order, conserve_energy;
  double inv_total_area, inv_sum_area_tri;
  R *gd_old, *iap_old, *q_old, *ghs;
  R *gdS[2], *iap_s, *qS[2];
  R *acc_gd, *acc_iap, *acc_q;
  R *t_gd[2], *t_iap[2];
  R *gd_e, *u, *fk, *udc, *E, *tend_gd, *pv_v, *pv_e;  // fk: [2 nE][M] (StageBufs::fk)
  R *gd_e2, *u2;  // second gd_edge/u set: stages 2 and 4 (the tendency phase forms the next stage's)
  double2 *part;  // [kNumSums][M][part_stride]
  int part_stride;
  double *sums;   // [M][kNumSums]
  double *beta;   // [M] beta of the last step
  double *inv;    // [M][4]
};

struct EnsCtl {
  unsigned *bar_count;
  unsigned *bar_gen;
  int *abort;
  unsigned long long wait_limit;  // wall-clock ticks (100 MHz) a barrier may wait
  int T;                          // ticks of this run
  int chunk;                      // elements per chunk
  int nlev;                       // sub-levels (slices of all level groups)
  int ngroups;                    // level groups
  int first[kMaxLevels];          // first sub-level of each group
  int nsl[kMaxLevels];            // slices of each group
};

// Grid barrier with a bounded wait.  Returns false when the run is aborted.
__device__ __forceinline__ bool grid_sync(const EnsCtl &c) {
  __shared__ int ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    int good = __hip_atomic_load(c.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
    const unsigned gen = __hip_atomic_load(c.bar_gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    __atomic_thread_fence(__ATOMIC_RELEASE);
    const unsigned arrived =
        __hip_atomic_fetch_add(c.bar_count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) + 1u;
    if (arrived == gridDim.x) {
      __hip_atomic_store(c.bar_count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(c.bar_gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      const unsigned long long t0 = wall_clock64();
      while (__hip_atomic_load(c.bar_gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
        if (__hip_atomic_load(c.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
          good = 0;
          break;
        }
        if (wall_clock64() - t0 > c.wait_limit) {
          __hip_atomic_store(c.abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          good = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
    }
    __atomic_thread_fence(__ATOMIC_ACQUIRE);
    ok = good;
  }
  __syncthreads();
  return ok != 0;
}

__device__ __forceinline__ bool level_active(int S, int T, int t) {
  return (long long)(t + 1) * S / T > (long long)t * S / T;
}

template <typename R>
__device__ __forceinline__ StageBufs<R> ens_bufs(const EnsLevelDev<R> &L, int k, bool want_inv) {
  StageBufs<R> b{};
  const int in_slot = (k == 2 || k == 4) ? 0 : 1;
  const int out_slot = (k == 1 || k == 3) ? 0 : 1;
  b.gd = (k <= 1) ? L.gd_old : L.gdS[in_slot];
  b.iap = (k <= 1) ? L.iap_old : L.iap_s;
  b.q = (k <= 1) ? L.q_old : L.qS[in_slot];
  b.gd_base = L.gd_old;
  b.iap_base = L.iap_old;
  b.q_base = L.q_old;
  b.gd_next = L.gdS[out_slot];
  b.iap_next = L.iap_s;
  b.q_next = L.qS[out_slot];
  b.acc_gd = L.acc_gd;
  b.acc_iap = L.acc_iap;
  b.acc_q = L.acc_q;
  b.tprev_gd = L.t_gd[(k + 1) & 1];
  b.tprev_iap = L.t_iap[(k + 1) & 1];
  b.tcur_gd = L.t_gd[k & 1];
  b.tcur_iap = L.t_iap[k & 1];
  // the tendency phase of stages 1-3 forms the next stage's edge values
  // (kXTendPrep) into the other gd_edge/u set: stages 1 and 3 read set 0,
  // stages 2 and 4 set 1
  const int es = (k == 2 || k == 4) ? 1 : 0;
  b.gd_e = es ? L.gd_e2 : L.gd_e;
  b.u = es ? L.u2 : L.u;
  b.gd_e_n = es ? L.gd_e : L.gd_e2;
  b.u_n = es ? L.u : L.u2;
  b.prep_next = k >= 1 && k <= 3 ? 1 : 0;
  b.fk = L.fk;
  b.udc = L.udc;
  b.E = L.E;
  b.tend_gd = L.tend_gd;
  b.pv_v = L.pv_v;
  b.pv_e = L.pv_e;
  b.ghs = L.ghs;
  b.nT = L.nT;
  b.k = k;
  const double cf[5] = {0.0, 0.5, 0.5, 1.0, 0.0};
  b.cdt = cf[k] * L.dt;
  b.want_inv = want_inv ? 1 : 0;
  b.pv_order = L.pv_order;
  return b;
}

enum EnsKind { kEdge = 0, kCell = 1, kVert = 2 };

template <typename R>
__device__ __forceinline__ int n_elems(const EnsLevelDev<R> &L, int kind) {
  return kind == kEdge ? L.m.nE : (kind == kCell ? L.m.nC : L.m.nV);
}

// Folds the per-thread partials (W members each) of one chunk and stores
// them at part[id][member][chunk].  lds holds 2 * kEnsBlock * W doubles.
template <int N, int W>
__device__ __forceinline__ void chunk_sums(const val_t<double, W> (&v)[N], const int (&ids)[N], double2 *part, int M,
                                           int stride, int chunk, int lanes, double *lds) {
  const int G = M / W;
  __syncthreads();
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int w = 0; w < W; ++w) lds[(i * kEnsBlock + threadIdx.x) * W + w] = vget<double, W>(v[i], w);
  __syncthreads();
  if ((int)threadIdx.x < M) {
    const int g = threadIdx.x / W, w = threadIdx.x % W;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      double2 a = make_double2(0.0, 0.0);
      for (int l = 0; l < lanes; ++l) a = dd_add(a, make_double2(lds[(i * kEnsBlock + l * G + g) * W + w], 0.0));
      part[((size_t)ids[i] * M + threadIdx.x) * stride + chunk] = a;
    }
  }
}

// One chunk of one phase of level L, W members per thread.
template <typename R, int W>
__device__ void ens_chunk(const EnsLevelDev<R> &L, int phase, int k, bool inv, int ch, int chunk, double *lds) {
  typedef val_t<double, W> D;
  typedef Members<W> Lay;
  const int M = L.M;
  const int G = M / W;  // threads per element
  const int lanes = kEnsBlock / G;
  const int lane = threadIdx.x / G;
  const Lay at{M, (int)(threadIdx.x % G) * W};
  const bool live = lane < lanes;
  const int kind = phase == 1 ? kCell : (phase == 4 ? kVert : kEdge);
  const int n = phase == 5 ? L.m.nE + L.m.nC : n_elems(L, kind);
  const int e0 = ch * chunk, e1 = min(e0 + chunk, n);
  const StageBufs<R> b = ens_bufs(L, k, inv);
  if (phase == 0) {  // edge prep
    D en = D(0.0);
    if (live)
      for (int e = e0 + lane; e < e1; e += lanes) en += edge_prep_e<R, Lay>(L.m, b, at, e);
    if (inv) {
      const D v[1] = {en};
      const int ids[1] = {kEnergyE};
      chunk_sums<1, W>(v, ids, L.part, M, L.part_stride, ch, lanes, lds);
    }
  } else if (phase == 1) {  // cells
    D s[4] = {D(0.0), D(0.0), D(0.0), D(0.0)};
    if (live)
      for (int c = e0 + lane; c < e1; c += lanes) {
        const CellSums<D> r = cell_e<R, Lay>(L.m, b, at, c);
        s[0] += r.mass;
        s[1] += r.enc;
        s[2] += r.rp;
        s[3] += r.r44;
      }
    const int rid = k == 2 ? kR12c : (k == 3 ? kR23c : kR34c);
    if (inv) {
      const D v[2] = {s[0], s[1]};
      const int ids[2] = {kMass, kEnergyC};
      chunk_sums<2, W>(v, ids, L.part, M, L.part_stride, ch, lanes, lds);
    } else if (k == 4) {
      const D v[2] = {s[2], s[3]};
      const int ids[2] = {rid, kR44c};
      chunk_sums<2, W>(v, ids, L.part, M, L.part_stride, ch, lanes, lds);
    } else if (k >= 2) {
      const D v[1] = {s[2]};
      const int ids[1] = {rid};
      chunk_sums<1, W>(v, ids, L.part, M, L.part_stride, ch, lanes, lds);
    }
  } else if (phase == 4) {  // vertices
    D s[2] = {D(0.0), D(0.0)};
    if (live)
      for (int v = e0 + lane; v < e1; v += lanes) {
        const VertSums<D> r = vertex_e<R, Lay>(L.m, b, at, v);
        s[0] += r.ens;
        s[1] += r.av;
      }
    if (inv) {
      const D v[2] = {s[0], s[1]};
      const int ids[2] = {kEnstrophy, kAbsVor};
      chunk_sums<2, W>(v, ids, L.part, M, L.part_stride, ch, lanes, lds);
    }
  } else if (phase == 2) {  // PV on edges
    if (live)
      for (int e = e0 + lane; e < e1; e += lanes) edge_pv_e<R, Lay>(L.m, b, at, e);
  } else if (phase == 3) {  // PV flux, tendency, RK update
    D s[2] = {D(0.0), D(0.0)};
    if (live)
      for (int e = e0 + lane; e < e1; e += lanes) {
        const EdgeSums<D> r = edge_tend_e<R, Lay, kXTendPrep>(L.m, b, at, e);
        s[0] += r.rp;
        s[1] += r.r44;
      }
    const int rid = k == 2 ? kR12e : (k == 3 ? kR23e : kR34e);
    if (k == 4) {
      const D v[2] = {s[0], s[1]};
      const int ids[2] = {rid, kR44e};
      chunk_sums<2, W>(v, ids, L.part, M, L.part_stride, ch, lanes, lds);
    } else if (k >= 2) {
      const D v[1] = {s[0]};
      const int ids[1] = {rid};
      chunk_sums<1, W>(v, ids, L.part, M, L.part_stride, ch, lanes, lds);
    }
  } else if (phase == 5) {  // finalize: update_state(beta*dt, ...) per member
    if (live) {
      typedef val_t<R, W> V;
      double beta[W];
      V bdt;
#pragma unroll
      for (int w = 0; w < W; ++w) {
        const double *r = L.sums + (size_t)(at.m0 + w) * kNumSums;
        beta[w] = 1.0;
        if (L.conserve_energy) {
          const double R44 = (r[kR44e] + r[kR44c]) * L.inv_total_area;
          const double R12 = (r[kR12e] + r[kR12c]) * L.inv_total_area;
          const double R23 = (r[kR23e] + r[kR23c]) * L.inv_total_area;
          const double R34 = (r[kR34e] + r[kR34c]) * L.inv_total_area;
          beta[w] = (R12 + R23 + R34) / (3.0 * R44);
        }
        vset<R, W>(bdt, w, (R)(beta[w] * L.dt));
      }
      // edges first, then cells (chunk covers [e0,e1) of a combined range)
      for (int i = e0 + lane; i < e1; i += lanes) {
        if (i < L.m.nE) {
          at.st(L.iap_old, i, at.ld(L.iap_old, i) + bdt * at.ld(L.acc_iap, i));
        } else {
          const int c = i - L.m.nE;
          at.st(L.gd_old, c, at.ld(L.gd_old, c) + bdt * at.ld(L.acc_gd, c));
          for (int t = 0; t < L.nT; ++t) {
            const size_t q = (size_t)t * L.m.nC + c;
            at.st(L.q_old, q, at.ld(L.q_old, q) + (R)L.dt * at.ld(L.acc_q, q));
          }
        }
      }
      if (ch == 0 && lane == 0)
#pragma unroll
        for (int w = 0; w < W; ++w) {
          const double *r = L.sums + (size_t)(at.m0 + w) * kNumSums;
          L.beta[at.m0 + w] = beta[w];
          L.inv[(size_t)(at.m0 + w) * 4 + 1] = r[kEnergyE] + r[kEnergyC];
        }
    }
  }
}


// Number of chunks of a phase for one level.
template <typename R>
__device__ __forceinline__ int phase_chunks(const EnsLevelDev<R> &L, int phase, int chunk) {
  int n;
  if (phase == 1)
    n = L.m.nC;
  else if (phase == 4)
    n = L.m.nV;
  else if (phase == 5)
    n = L.m.nE + L.m.nC;
  else
    n = L.m.nE;
  return (n + chunk - 1) / chunk;
}

// Deals the chunks of all active sub-levels to the workgroups.  Each level
// group's chunk list (its slices one after another) is cut into kXcd
// contiguous parts, XCD x taking part x of every group; the XCD's workgroups
// stride over its parts.  f(sub-level, chunk).
template <typename R, typename F>
__device__ __forceinline__ void for_chunks(const EnsLevelDev<R> *lv, const EnsCtl &c, const bool *act, int phase,
                                           F f) {
  const int G = gridDim.x, xcd = blockIdx.x % kXcd, G8 = G / kXcd, bi = blockIdx.x / kXcd;
  int lo[kMaxLevels], cnt[kMaxLevels], per[kMaxLevels], total = 0;
  for (int g = 0; g < c.ngroups; ++g) {
    const int s0 = c.first[g];
    per[g] = act[s0] ? phase_chunks(lv[s0], phase, c.chunk) : 0;
    const long long n = (long long)per[g] * c.nsl[g];
    lo[g] = (int)(n * xcd / kXcd);
    cnt[g] = (int)(n * (xcd + 1) / kXcd) - lo[g];
    total += cnt[g];
  }
  for (int i = bi; i < total; i += G8) {
    int g = 0, r = i;
    while (r >= cnt[g]) {
      r -= cnt[g];
      ++g;
    }
    r += lo[g];
    f(c.first[g] + r / per[g], r % per[g]);
  }
}

// Per (level, member, sum) folding of the chunk partials into sums[].
template <typename R>
__device__ void ens_reduce(const EnsLevelDev<R> *lv, const EnsCtl &c, const bool *act, bool inv_only) {
  __shared__ double2 sh[kEnsBlock / 64];
  int total = 0, cnt[kMaxSub];
  for (int l = 0; l < c.nlev; ++l) {
    cnt[l] = act[l] ? lv[l].M * kNumSums : 0;
    total += cnt[l];
  }
  for (int g = blockIdx.x; g < total; g += gridDim.x) {
    int l = 0, r = g;
    while (r >= cnt[l]) {
      r -= cnt[l];
      ++l;
    }
    const EnsLevelDev<R> &L = lv[l];
    const int mem = r / kNumSums, id = r % kNumSums;
    const bool edge = id == kR12e || id == kR23e || id == kR34e || id == kR44e || id == kEnergyE;
    const bool vert = id == kEnstrophy || id == kAbsVor;
    const int n = edge ? L.m.nE : (vert ? L.m.nV : L.m.nC);
    const int nch = (n + c.chunk - 1) / c.chunk;
    const bool needed = id >= kMass || !inv_only;
    double2 a = make_double2(0.0, 0.0);
    if (needed)
      for (int j = threadIdx.x; j < nch; j += kEnsBlock)
        a = dd_add(a, L.part[((size_t)id * L.M + mem) * L.part_stride + j]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) a = dd_add(a, make_double2(__shfl_xor(a.x, off, 64), __shfl_xor(a.y, off, 64)));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
      double2 t = sh[0];
      for (int q = 1; q < kEnsBlock / 64; ++q) t = dd_add(t, sh[q]);
      const double v = t.x + t.y;
      L.sums[(size_t)mem * kNumSums + id] = v;
      // the invariants of the base state of this step (or of the final state
      // in the refresh pass), tmcore_swm_mod.F90:215-255
      double *iv = L.inv + (size_t)mem * 4;
      if (id == kMass) iv[0] = v;
      if (id == kEnstrophy) iv[2] = v;
      if (id == kAbsVor) iv[3] = v * L.inv_sum_area_tri;
    }
  }
}

template <typename R>
__device__ void ens_energy_out(const EnsLevelDev<R> *lv, const EnsCtl &c) {
  for (int l = 0; l < c.nlev; ++l)
    for (int mem = blockIdx.x * kEnsBlock + threadIdx.x; mem < lv[l].M; mem += gridDim.x * kEnsBlock) {
      const double *r = lv[l].sums + (size_t)mem * kNumSums;
      lv[l].inv[(size_t)mem * 4 + 1] = r[kEnergyE] + r[kEnergyC];
    }
}

template <typename R, int W>
__global__ __launch_bounds__(kEnsBlock) void k_ens_run(const EnsLevelDev<R> *__restrict__ lv, EnsCtl c) {
  __shared__ double lds[2 * kEnsBlock * W];
  bool act[kMaxSub];
  for (int t = 0; t < c.T; ++t) {
    for (int l = 0; l < c.nlev; ++l) act[l] = level_active(lv[l].nsteps, c.T, t);
    for (int k = 1; k <= 4; ++k) {
      // phases: 0 edge prep (stage 1 only: the tendency phase of stages 1-3
      // forms the next stage's edge values), 1 cells (+ 4 vertices in the
      // same pass), 2 PV, 3 tendency
      if (k == 1) {
        for_chunks<R>(lv, c, act, 0, [&](int l, int ch) { ens_chunk<R, W>(lv[l], 0, k, true, ch, c.chunk, lds); });
        if (!grid_sync(c)) return;
      }
      for_chunks<R>(lv, c, act, 1, [&](int l, int ch) { ens_chunk<R, W>(lv[l], 1, k, k == 1, ch, c.chunk, lds); });
      for_chunks<R>(lv, c, act, 4, [&](int l, int ch) { ens_chunk<R, W>(lv[l], 4, k, k == 1, ch, c.chunk, lds); });
      if (!grid_sync(c)) return;
      for_chunks<R>(lv, c, act, 2, [&](int l, int ch) { ens_chunk<R, W>(lv[l], 2, k, k == 1, ch, c.chunk, lds); });
      if (!grid_sync(c)) return;
      for_chunks<R>(lv, c, act, 3, [&](int l, int ch) { ens_chunk<R, W>(lv[l], 3, k, k == 1, ch, c.chunk, lds); });
      if (!grid_sync(c)) return;
    }
    ens_reduce<R>(lv, c, act, false);
    if (!grid_sync(c)) return;
    for_chunks<R>(lv, c, act, 5, [&](int l, int ch) { ens_chunk<R, W>(lv[l], 5, 0, false, ch, c.chunk, lds); });
    if (!grid_sync(c)) return;
  }
  // refresh: invariants of the final state of every member
  bool all[kMaxSub];
  for (int l = 0; l < c.nlev; ++l) all[l] = true;
  for_chunks<R>(lv, c, all, 0, [&](int l, int ch) { ens_chunk<R, W>(lv[l], 0, 0, true, ch, c.chunk, lds); });
  if (!grid_sync(c)) return;
  for_chunks<R>(lv, c, all, 1, [&](int l, int ch) { ens_chunk<R, W>(lv[l], 1, 0, true, ch, c.chunk, lds); });
  for_chunks<R>(lv, c, all, 4, [&](int l, int ch) { ens_chunk<R, W>(lv[l], 4, 0, true, ch, c.chunk, lds); });
  if (!grid_sync(c)) return;
  ens_reduce<R>(lv, c, all, true);
  if (!grid_sync(c)) return;
  ens_energy_out<R>(lv, c);
}

// iap_u = sqrt(c2e(gd)) * u for every member (tmcore_swm_mod.F90:62-64)
template <typename R>
__global__ void k_ens_iap(DMesh m, int M, const R *gd, const R *u, R *iap) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)m.nE * M) return;
  const int e = (int)(i / M), mem = (int)(i % M);
  const Member at{M, mem};
  const R gde = R(0.5) * (gd[at(m.coe[e])] + gd[at(m.coe[m.nE + e])]);
  iap[i] = tm_sqrt(gde) * u[i];
}

}  // namespace tmk

namespace tmk {

// Column m of a [n][M] array <-> a contiguous vector.
template <typename R>
__global__ void k_ens_put(const R *src, R *dst, int n, int M, int m) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < (size_t)n) dst[i * M + m] = src[i];
}
template <typename R>
__global__ void k_ens_get(const R *src, R *dst, int n, int M, int m) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < (size_t)n) dst[i] = src[i * M + m];
}
// iap_u of member m from its u (contiguous) and the gd column.
template <typename R>
__global__ void k_ens_iap_member(DMesh d, int M, int m, const R *gd, const R *u, R *iap) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= d.nE) return;
  const Member at{M, m};
  const R gde = R(0.5) * (gd[at(d.coe[e])] + gd[at(d.coe[d.nE + e])]);
  iap[at(e)] = tm_sqrt(gde) * u[e];
}

}  // namespace tmk

using namespace tmk;

struct tm_ens {
  tm_opts o{};
  std::string err;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int nlev = 0;  // level groups
  size_t rsz = 8;
  std::vector<MeshDev> md;
  std::vector<int> M;
  std::vector<double> dt;
  std::vector<int> S;      // member slices per level group
  std::vector<int> first;  // first sub-level (slice) of each group
  int nsub = 0;
  bool uploaded = false;   // slices are fixed once state is on the device
  std::vector<void *> allocs;
  std::vector<void *> level_allocs;  // the sub-levels' storage (replaced by tm_ens_set_slices)
  std::vector<EnsLevelDev<double>> lv64;
  std::vector<EnsLevelDev<float>> lv32;
  void *d_lv = nullptr;
  void *scratch = nullptr;  // staging for one member's column
  std::vector<void *> backup;  // autotune copies of gd_old, iap_old, q_old per level
  unsigned *bar = nullptr;     // [count, gen]
  int *abort_flag = nullptr;
  int chunk = 256, bpc = 2, cus = 1;
  int W = 1;                 // members per thread, divides every level's M
  int max_bpc[3] = {1, 1, 1};  // occupancy of the kernel for W = 1, 2, 4
  unsigned long long wait_limit = 0;

  int fail(int code, const std::string &m) {
    err = m;
    return code;
  }
  void *alloc(size_t bytes, bool level = false) {
    void *p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(bytes, 16)) != hipSuccess) return nullptr;
    (level ? level_allocs : allocs).push_back(p);
    return p;
  }
  // sub-level and member within it of member m of group l
  int sub_of(int l, int m, int *j) const {
    const int ms = M[l] / S[l];
    *j = m % ms;
    return first[l] + m / ms;
  }
};

#define ENSCHK(e, call)                                                                                   \
  do {                                                                                                   \
    hipError_t e_ = (call);                                                                              \
    if (e_ != hipSuccess) return (e)->fail(TM_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace {

// Default slices of a level group of M members: the most (up to eight, one
// per XCD) that leave at least 16 members in a slice.  Each slice re-reads
// the mesh constants, so slices pay only where the member rows are wide
// (measured on the 128/32/8-member level 6/7/8 mix: 8 slices of the 128
// members +13%, slices of 8 or 32 members neutral or slower).
int default_slices(int M) {
  for (int S : {8, 4, 2})
    if (M % S == 0 && M / S >= 16) return S;
  return 1;
}

template <typename R>
int ens_alloc_level(tm_ens *E, int l, EnsLevelDev<R> &L) {
  const MeshDev &md = E->md[l];
  const int M = E->M[l] / E->S[l], nT = E->o.ntracers;
  const size_t nC = md.dm.nC, nE = md.dm.nE, nV = md.dm.nV, rs = sizeof(R);
  L.m = md.dm;
  L.M = M;
  L.nT = nT;
  L.dt = E->dt[l];
  L.nsteps = 0;
  L.pv_order = E->o.interp_pv_order;
  L.conserve_energy = E->o.conserve_energy;
  L.inv_total_area = 1.0 / md.totalArea;
  L.inv_sum_area_tri = 1.0 / md.sumAreaTri;
  auto A = [&](size_t n) { return (R *)E->alloc(n * M * rs, true); };
  R **cells[] = {&L.gd_old, &L.ghs, &L.gdS[0], &L.gdS[1], &L.acc_gd, &L.t_gd[0], &L.t_gd[1], &L.E, &L.tend_gd};
  for (R **p : cells)
    if (!(*p = A(nC))) return E->fail(TM_ENOMEM, "hipMalloc ensemble cell state");
  R **edges[] = {&L.iap_old, &L.iap_s, &L.acc_iap, &L.t_iap[0], &L.t_iap[1], &L.gd_e, &L.u, &L.udc, &L.pv_e,
                  &L.gd_e2, &L.u2};
  for (R **p : edges)
    if (!(*p = A(nE))) return E->fail(TM_ENOMEM, "hipMalloc ensemble edge state");
  if (!(L.fk = A(2 * nE))) return E->fail(TM_ENOMEM, "hipMalloc ensemble edge state");
  if (!(L.pv_v = A(nV))) return E->fail(TM_ENOMEM, "hipMalloc ensemble vertex state");
  const size_t nq = std::max<size_t>((size_t)nT * nC, 1);
  R **tr[] = {&L.q_old, &L.qS[0], &L.qS[1], &L.acc_q};
  for (R **p : tr)
    if (!(*p = A(nq))) return E->fail(TM_ENOMEM, "hipMalloc ensemble tracer state");
  // partials sized for the smallest chunk the tuner may pick (32 elements)
  L.part_stride = (int)((std::max(nE, std::max(nC, nV)) + 31) / 32);
  if (!(L.part = (double2 *)E->alloc((size_t)kNumSums * M * L.part_stride * sizeof(double2), true)) ||
      !(L.sums = (double *)E->alloc((size_t)M * kNumSums * sizeof(double), true)) ||
      !(L.beta = (double *)E->alloc((size_t)M * sizeof(double), true)) ||
      !(L.inv = (double *)E->alloc((size_t)M * 4 * sizeof(double), true)))
    return E->fail(TM_ENOMEM, "hipMalloc ensemble sums");
  ENSCHK(E, hipMemsetAsync(L.gd_old, 0, nC * M * rs, E->stream));
  ENSCHK(E, hipMemsetAsync(L.iap_old, 0, nE * M * rs, E->stream));
  ENSCHK(E, hipMemsetAsync(L.ghs, 0, nC * M * rs, E->stream));
  ENSCHK(E, hipMemsetAsync(L.q_old, 0, nq * M * rs, E->stream));
  ENSCHK(E, hipMemsetAsync(L.inv, 0, (size_t)M * 4 * sizeof(double), E->stream));
  return TM_OK;
}

template <typename R>
std::vector<EnsLevelDev<R>> &levels(tm_ens *E);
template <>
std::vector<EnsLevelDev<double>> &levels<double>(tm_ens *E) {
  return E->lv64;
}
template <>
std::vector<EnsLevelDev<float>> &levels<float>(tm_ens *E) {
  return E->lv32;
}

// Copies a reference-order FP64 host array into column m of dst ([n][M]).
template <typename R>
int put_column(tm_ens *E, int l, R *dst, const double *src, const std::vector<int> &perm, int n, int copies) {
  std::vector<R> h((size_t)n * copies);
  for (int t = 0; t < copies; ++t)
    for (int i = 0; i < n; ++i) h[(size_t)t * n + i] = (R)src[(size_t)t * n + perm[i]];
  ENSCHK(E, hipMemcpyAsync(E->scratch, h.data(), h.size() * sizeof(R), hipMemcpyHostToDevice, E->stream));
  const int N = n * copies;
  (void)l;
  return N;  // caller launches the scatter (needs M, m)
}

template <typename R>
int ens_upload(tm_ens *E, int l, int m0, const double *gd, const double *u, const double *ghs, const double *q) {
  int m = 0;
  EnsLevelDev<R> &L = levels<R>(E)[E->sub_of(l, m0, &m)];
  const MeshDev &md = E->md[l];
  const int M = L.M, nC = md.dm.nC, nE = md.dm.nE;
  auto scatter = [&](R *dst, const double *src, const std::vector<int> &perm, int n, int copies) -> int {
    std::vector<R> h((size_t)n * copies);
    for (int t = 0; t < copies; ++t)
      for (int i = 0; i < n; ++i) h[(size_t)t * n + i] = (R)src[(size_t)t * n + perm[i]];
    ENSCHK(E, hipMemcpyAsync(E->scratch, h.data(), h.size() * sizeof(R), hipMemcpyHostToDevice, E->stream));
    const int N = n * copies;
    hipLaunchKernelGGL(k_ens_put<R>, dim3((N + 255) / 256), dim3(256), 0, E->stream, (const R *)E->scratch, dst, N,
                       M, m);
    ENSCHK(E, hipGetLastError());
    ENSCHK(E, hipStreamSynchronize(E->stream));
    return TM_OK;
  };
  int rc;
  if ((rc = scatter(L.gd_old, gd, md.pC, nC, 1))) return rc;
  if ((rc = scatter(L.ghs, ghs, md.pC, nC, 1))) return rc;
  if (L.nT > 0 && (rc = scatter(L.q_old, q, md.pC, nC, L.nT))) return rc;
  // u (contiguous in scratch) -> iap column
  std::vector<R> h(nE);
  for (int i = 0; i < nE; ++i) h[i] = (R)u[md.pE[i]];
  ENSCHK(E, hipMemcpyAsync(E->scratch, h.data(), h.size() * sizeof(R), hipMemcpyHostToDevice, E->stream));
  hipLaunchKernelGGL(k_ens_iap_member<R>, dim3((nE + 255) / 256), dim3(256), 0, E->stream, md.dm, M, m,
                     (const R *)L.gd_old, (const R *)E->scratch, L.iap_old);
  ENSCHK(E, hipGetLastError());
  ENSCHK(E, hipStreamSynchronize(E->stream));
  return TM_OK;
}

template <typename R>
int ens_download(tm_ens *E, int l, int m0, double *gd, double *iap, double *q) {
  int m = 0;
  EnsLevelDev<R> &L = levels<R>(E)[E->sub_of(l, m0, &m)];
  const MeshDev &md = E->md[l];
  const int M = L.M;
  auto gather = [&](const R *src, double *dst, const std::vector<int> &perm, int n, int copies) -> int {
    const int N = n * copies;
    hipLaunchKernelGGL(k_ens_get<R>, dim3((N + 255) / 256), dim3(256), 0, E->stream, src, (R *)E->scratch, N, M, m);
    ENSCHK(E, hipGetLastError());
    std::vector<R> h(N);
    ENSCHK(E, hipMemcpyAsync(h.data(), E->scratch, (size_t)N * sizeof(R), hipMemcpyDeviceToHost, E->stream));
    ENSCHK(E, hipStreamSynchronize(E->stream));
    for (int t = 0; t < copies; ++t)
      for (int i = 0; i < n; ++i) dst[(size_t)t * n + perm[i]] = (double)h[(size_t)t * n + i];
    return TM_OK;
  };
  int rc;
  if (gd && (rc = gather(L.gd_old, gd, md.pC, md.dm.nC, 1))) return rc;
  if (iap && (rc = gather(L.iap_old, iap, md.pE, md.dm.nE, 1))) return rc;
  if (q && L.nT > 0 && (rc = gather(L.q_old, q, md.pC, md.dm.nC, L.nT))) return rc;
  return TM_OK;
}

template <typename R>
int ens_launch(tm_ens *E, const int *nsteps, double *ms) {
  auto &lv = levels<R>(E);
  int T = 0;
  for (int l = 0; l < E->nlev; ++l)
    if (nsteps[l] < 0) return E->fail(TM_EINVAL, "negative step count");
  EnsCtl c{E->bar, E->bar + 1, E->abort_flag, E->wait_limit, 0, E->chunk, E->nsub, E->nlev, {}, {}};
  for (int l = 0; l < E->nlev; ++l) {
    for (int s = 0; s < E->S[l]; ++s) lv[E->first[l] + s].nsteps = nsteps[l];
    c.first[l] = E->first[l];
    c.nsl[l] = E->S[l];
    T = std::max(T, nsteps[l]);
  }
  c.T = T;
  ENSCHK(E, hipMemcpyAsync(E->d_lv, lv.data(), lv.size() * sizeof(EnsLevelDev<R>), hipMemcpyHostToDevice,
                           E->stream));
  ENSCHK(E, hipMemsetAsync(E->abort_flag, 0, sizeof(int), E->stream));
  const int G = std::max(kXcd, (E->cus * E->bpc / kXcd) * kXcd);
  // The persistent launch is captured in a hipGraph and replayed once.
  hipGraph_t g = nullptr;
  hipGraphExec_t ge = nullptr;
  ENSCHK(E, hipStreamBeginCapture(E->stream, hipStreamCaptureModeThreadLocal));
  const EnsLevelDev<R> *d_lv = (const EnsLevelDev<R> *)E->d_lv;
  if (E->W == 4)
    hipLaunchKernelGGL((k_ens_run<R, 4>), dim3(G), dim3(kEnsBlock), 0, E->stream, d_lv, c);
  else if (E->W == 2)
    hipLaunchKernelGGL((k_ens_run<R, 2>), dim3(G), dim3(kEnsBlock), 0, E->stream, d_lv, c);
  else
    hipLaunchKernelGGL((k_ens_run<R, 1>), dim3(G), dim3(kEnsBlock), 0, E->stream, d_lv, c);
  hipError_t ce = hipStreamEndCapture(E->stream, &g);
  if (ce != hipSuccess) return E->fail(TM_EDEVICE, std::string("capture: ") + hipGetErrorString(ce));
  ce = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  if (ce != hipSuccess) {
    (void)hipGraphDestroy(g);
    return E->fail(TM_EDEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ce));
  }
  ENSCHK(E, hipEventRecord(E->ev0, E->stream));
  ce = hipGraphLaunch(ge, E->stream);
  ENSCHK(E, hipEventRecord(E->ev1, E->stream));
  hipError_t se = hipStreamSynchronize(E->stream);
  (void)hipGraphExecDestroy(ge);
  (void)hipGraphDestroy(g);
  if (ce != hipSuccess || se != hipSuccess)
    return E->fail(TM_EDEVICE, std::string("ensemble launch: ") + hipGetErrorString(ce != hipSuccess ? ce : se));
  int ab = 0;
  ENSCHK(E, hipMemcpy(&ab, E->abort_flag, sizeof(int), hipMemcpyDeviceToHost));
  if (ab) {
    // reset the barrier so the next run starts clean
    ENSCHK(E, hipMemset(E->bar, 0, 2 * sizeof(unsigned)));
    return E->fail(TM_EDEVICE, "ensemble grid barrier timed out (persistent grid not co-resident?)");
  }
  float f = 0.f;
  ENSCHK(E, hipEventElapsedTime(&f, E->ev0, E->ev1));
  if (ms) *ms = f;
  return TM_OK;
}

// Storage of every slice of every level group (sub-levels in group order).
template <typename R>
int ens_create_levels(tm_ens *E) {
  auto &lv = levels<R>(E);
  E->first.assign(E->nlev, 0);
  E->nsub = 0;
  for (int l = 0; l < E->nlev; ++l) {
    E->first[l] = E->nsub;
    E->nsub += E->S[l];
  }
  lv.assign(E->nsub, EnsLevelDev<R>{});
  for (int l = 0; l < E->nlev; ++l)
    for (int s = 0; s < E->S[l]; ++s) {
      int rc = ens_alloc_level<R>(E, l, lv[E->first[l] + s]);
      if (rc) return rc;
    }
  if (!E->d_lv && !(E->d_lv = E->alloc(sizeof(EnsLevelDev<R>) * kMaxSub))) return E->fail(TM_ENOMEM, "hipMalloc levels");
  int occ[3] = {0, 0, 0};
  ENSCHK(E, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ[0], k_ens_run<R, 1>, kEnsBlock, 0));
  ENSCHK(E, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ[1], k_ens_run<R, 2>, kEnsBlock, 0));
  ENSCHK(E, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ[2], k_ens_run<R, 4>, kEnsBlock, 0));
  for (int i = 0; i < 3; ++i) E->max_bpc[i] = std::max(1, occ[i]);
  // widest vector of members that divides every slice's member count:
  // 16-byte rows (4 x FP32, 2 x FP64)
  E->W = 1;
  for (int w : {2, 4}) {
    if (w * sizeof(R) > 16) continue;
    bool ok = true;
    for (int l = 0; l < E->nlev; ++l) ok = ok && ((E->M[l] / E->S[l]) % w == 0);
    if (ok) E->W = w;
  }
  E->bpc = std::min(E->bpc, E->max_bpc[E->W == 4 ? 2 : E->W - 1]);
  return TM_OK;
}

}  // namespace

extern "C" int tm_ens_create(tm_ens **out, int nlevels, const tm_ens_level *levels_in, const tm_opts *o) {
  if (!out || !levels_in || !o || nlevels < 1 || nlevels > kMaxLevels) return TM_EINVAL;
  if (o->adv_order != 2 || o->pv_scheme != 3 || o->interp_pv_order > 4 || o->time_scheme != 2) {
    std::fprintf(stderr, "tm_ens_create: the ensemble runs CRK4, adv_order 2, pv_scheme 3, interp_pv_order <= 4\n");
    return TM_EINVAL;
  }
  *out = nullptr;
  tm_ens *E = new tm_ens();
  E->o = *o;
  E->nlev = nlevels;
  E->rsz = o->precision == 64 ? 8 : 4;
  auto bail = [&](int rc) {
    std::fprintf(stderr, "tm_ens_create: %s\n", E->err.c_str());
    tm_ens_destroy(E);
    return rc;
  };
  for (int l = 0; l < nlevels; ++l) {
    const tm_ens_level &L = levels_in[l];
    if (L.members < 1 || L.members > kEnsBlock || !(L.dt > 0.0))
      return bail(E->fail(TM_EINVAL, "members must be 1..256 and dt > 0"));
    std::string e;
    tm_opts lo = *o;
    lo.dt = L.dt;
    int rc = validate_mesh(L.mesh, L.fit, lo, e);
    if (rc) return bail(E->fail(rc, e));
  }
  if (hipSetDevice(o->device) != hipSuccess || hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&E->ev0) != hipSuccess || hipEventCreate(&E->ev1) != hipSuccess)
    return bail(E->fail(TM_EDEVICE, "HIP device/stream setup failed"));
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, o->device) != hipSuccess) return bail(E->fail(TM_EDEVICE, "hipGetDeviceProperties"));
  E->cus = prop.multiProcessorCount;
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, o->device) != hipSuccess || khz <= 0) khz = 100000;
  E->wait_limit = (unsigned long long)khz * 1000ull * 2ull;  // 2 s
  E->md.resize(nlevels);
  size_t maxn = 1;
  for (int l = 0; l < nlevels; ++l) {
    tm_opts lo = *o;
    lo.dt = levels_in[l].dt;
    std::string e;
    int rc = build_mesh_dev(E->md[l], levels_in[l].mesh, levels_in[l].fit, lo, E->stream, E->allocs, e);
    if (rc) return bail(E->fail(rc, e));
    E->M.push_back(levels_in[l].members);
    E->S.push_back(default_slices(levels_in[l].members));
    E->dt.push_back(levels_in[l].dt);
    maxn = std::max(maxn, (size_t)std::max(levels_in[l].mesh->nEdges,
                                           levels_in[l].mesh->nCells * std::max(o->ntracers, 1)));
  }
  if (!(E->scratch = E->alloc(maxn * 8))) return bail(E->fail(TM_ENOMEM, "hipMalloc scratch"));
  if (!(E->bar = (unsigned *)E->alloc(2 * sizeof(unsigned))) || !(E->abort_flag = (int *)E->alloc(sizeof(int))))
    return bail(E->fail(TM_ENOMEM, "hipMalloc barrier"));
  if (hipMemset(E->bar, 0, 2 * sizeof(unsigned)) != hipSuccess) return bail(E->fail(TM_EDEVICE, "hipMemset"));
  int rc = E->rsz == 8 ? ens_create_levels<double>(E) : ens_create_levels<float>(E);
  if (rc) return bail(rc);
  if (hipStreamSynchronize(E->stream) != hipSuccess) return bail(E->fail(TM_EDEVICE, "sync"));
  *out = E;
  return TM_OK;
}

extern "C" int tm_ens_destroy(tm_ens *E) {
  if (!E) return TM_EINVAL;
  for (void *p : E->allocs) (void)hipFree(p);
  for (void *p : E->level_allocs) (void)hipFree(p);
  for (void *p : E->backup) (void)hipFree(p);
  if (E->ev0) (void)hipEventDestroy(E->ev0);
  if (E->ev1) (void)hipEventDestroy(E->ev1);
  if (E->stream) (void)hipStreamDestroy(E->stream);
  delete E;
  return TM_OK;
}

extern "C" const char *tm_ens_last_error(const tm_ens *E) { return E ? E->err.c_str() : "null ensemble"; }

extern "C" int tm_ens_upload(tm_ens *E, int l, int m, const double *gd, const double *u, const double *ghs,
                             const double *q) {
  if (!E) return TM_EINVAL;
  if (l < 0 || l >= E->nlev || m < 0 || m >= E->M[l]) return E->fail(TM_EINVAL, "level/member out of range");
  if (!gd || !u || !ghs || (E->o.ntracers > 0 && !q)) return E->fail(TM_EINVAL, "missing state array");
  E->uploaded = true;
  const MeshDev &md = E->md[l];
  for (int i = 0; i < md.dm.nC; ++i)
    if (!(gd[i] > 0.0) || !std::isfinite(gd[i]) || !std::isfinite(ghs[i]))
      return E->fail(TM_ENONFINITE, "gd must be finite and positive, ghs finite");
  for (int i = 0; i < md.dm.nE; ++i)
    if (!std::isfinite(u[i])) return E->fail(TM_ENONFINITE, "non-finite u");
  return E->rsz == 8 ? ens_upload<double>(E, l, m, gd, u, ghs, q) : ens_upload<float>(E, l, m, gd, u, ghs, q);
}

extern "C" int tm_ens_download(tm_ens *E, int l, int m, double *gd, double *iap, double *q) {
  if (!E) return TM_EINVAL;
  if (l < 0 || l >= E->nlev || m < 0 || m >= E->M[l]) return E->fail(TM_EINVAL, "level/member out of range");
  return E->rsz == 8 ? ens_download<double>(E, l, m, gd, iap, q) : ens_download<float>(E, l, m, gd, iap, q);
}

extern "C" int tm_ens_set_slices(tm_ens *E, int l, int slices) {
  if (!E) return TM_EINVAL;
  if (l < 0 || l >= E->nlev) return E->fail(TM_EINVAL, "level out of range");
  if (slices < 1 || slices > kMaxSlices || E->M[l] % slices)
    return E->fail(TM_EINVAL, "slices must be 1..8 and divide the level's members");
  if (E->uploaded) return E->fail(TM_EINVAL, "slices are set before the first upload");
  ENSCHK(E, hipStreamSynchronize(E->stream));
  for (void *p : E->level_allocs) (void)hipFree(p);
  E->level_allocs.clear();
  for (void *p : E->backup) (void)hipFree(p);
  E->backup.clear();
  E->S[l] = slices;
  int rc = E->rsz == 8 ? ens_create_levels<double>(E) : ens_create_levels<float>(E);
  if (rc) return rc;
  ENSCHK(E, hipStreamSynchronize(E->stream));
  return TM_OK;
}

extern "C" int tm_ens_get_slices(tm_ens *E, int l, int *slices) {
  if (!E || !slices) return TM_EINVAL;
  if (l < 0 || l >= E->nlev) return E->fail(TM_EINVAL, "level out of range");
  *slices = E->S[l];
  return TM_OK;
}

extern "C" int tm_ens_run(tm_ens *E, const int *nsteps, double *ms) {
  if (!E || !nsteps) return TM_EINVAL;
  return E->rsz == 8 ? ens_launch<double>(E, nsteps, ms) : ens_launch<float>(E, nsteps, ms);
}

extern "C" int tm_ens_invariants(tm_ens *E, int l, int m, double out[4]) {
  if (!E || !out) return TM_EINVAL;
  if (l < 0 || l >= E->nlev || m < 0 || m >= E->M[l]) return E->fail(TM_EINVAL, "level/member out of range");
  int j = 0;
  const int sl = E->sub_of(l, m, &j);
  const double *src = E->rsz == 8 ? E->lv64[sl].inv : E->lv32[sl].inv;
  ENSCHK(E, hipMemcpy(out, src + (size_t)j * 4, 4 * sizeof(double), hipMemcpyDeviceToHost));
  return TM_OK;
}

namespace {
int wslot(int w) { return w == 4 ? 2 : w - 1; }
bool w_valid(const tm_ens *E, int w) {
  if (w != 1 && w != 2 && w != 4) return false;
  if ((size_t)w * E->rsz > 16) return false;
  for (int l = 0; l < E->nlev; ++l)
    if ((E->M[l] / E->S[l]) % w) return false;
  return true;
}
}  // namespace

extern "C" int tm_ens_set_tuning(tm_ens *E, int chunk, int bpc, int members_per_thread) {
  if (!E) return TM_EINVAL;
  if (chunk < 32 || chunk > 4096 || bpc < 1) return E->fail(TM_EINVAL, "chunk must be 32..4096, blocks_per_cu >= 1");
  if (!w_valid(E, members_per_thread))
    return E->fail(TM_EINVAL,
                   "members_per_thread must be 1, 2 or 4, divide every slice's members and span <= 16 bytes");
  if (bpc > E->max_bpc[wslot(members_per_thread)])
    return E->fail(TM_EINVAL, "blocks_per_cu exceeds the occupancy of the persistent kernel");
  E->chunk = chunk;
  E->bpc = bpc;
  E->W = members_per_thread;
  return TM_OK;
}

namespace {
template <typename R>
int ens_state_copy(tm_ens *E, bool save) {
  auto &lv = levels<R>(E);
  if (save && E->backup.empty()) {
    for (int l = 0; l < E->nsub; ++l) {
      const size_t M = lv[l].M, nC = lv[l].m.nC, nE = lv[l].m.nE, nq = std::max<size_t>((size_t)lv[l].nT * nC, 1);
      for (size_t n : {nC, nE, nq}) {
        void *p = nullptr;
        if (hipMalloc(&p, n * M * sizeof(R)) != hipSuccess) return E->fail(TM_ENOMEM, "hipMalloc autotune backup");
        E->backup.push_back(p);
      }
    }
  }
  for (int l = 0; l < E->nsub; ++l) {
    const size_t M = lv[l].M, nC = lv[l].m.nC, nE = lv[l].m.nE, nq = std::max<size_t>((size_t)lv[l].nT * nC, 1);
    void *dev[3] = {lv[l].gd_old, lv[l].iap_old, lv[l].q_old};
    const size_t n[3] = {nC, nE, nq};
    for (int i = 0; i < 3; ++i) {
      void *b = E->backup[3 * l + i];
      ENSCHK(E, hipMemcpyAsync(save ? b : dev[i], save ? dev[i] : b, n[i] * M * sizeof(R), hipMemcpyDeviceToDevice,
                               E->stream));
    }
  }
  ENSCHK(E, hipStreamSynchronize(E->stream));
  return TM_OK;
}
}  // namespace

extern "C" int tm_ens_autotune(tm_ens *E, int ticks, int best[3], double *best_ms) {
  if (!E || ticks < 1) return TM_EINVAL;
  int rc = E->rsz == 8 ? ens_state_copy<double>(E, true) : ens_state_copy<float>(E, true);
  if (rc) return rc;
  std::vector<int> steps(E->nlev, ticks);
  double bms = 1e300;
  int bc = E->chunk, bb = E->bpc, bw = E->W;
  for (int w : {1, 2, 4}) {
    if (!w_valid(E, w)) continue;
    for (int chunk : {64, 128, 256, 512, 1024})
      for (int bpc : {1, 2, 4, 8}) {
        if (bpc > E->max_bpc[wslot(w)]) continue;
        E->chunk = chunk;
        E->bpc = bpc;
        E->W = w;
        double ms = 0.0;
        rc = tm_ens_run(E, steps.data(), &ms);
        int rc2 = E->rsz == 8 ? ens_state_copy<double>(E, false) : ens_state_copy<float>(E, false);
        if (rc) return rc;
        if (rc2) return rc2;
        if (ms < bms) {
          bms = ms;
          bc = chunk;
          bb = bpc;
          bw = w;
        }
      }
  }
  E->chunk = bc;
  E->bpc = bb;
  E->W = bw;
  if (best) {
    best[0] = bc;
    best[1] = bb;
    best[2] = bw;
  }
  if (best_ms) *best_ms = bms / ticks;
  return TM_OK;
}
