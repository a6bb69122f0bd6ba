// libtmgpu: MI355X (gfx950) HIP path of the TMCORE TRiSK shallow-water RK
// stage behind the C-ABI of include/tmcore_gpu.h.
//
// Host side: validates the reference-layout mesh, renumbers cells, edges and
// vertices along a 3-D Hilbert curve (so stencil gathers of neighbouring
// threads hit neighbouring cache lines), stores connectivity slot-major (SoA,
// padded to the widest stencil), owns every device buffer, and drives the
// CRK4 step as a captured hipGraph on one stream.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <random>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "closure_kernels.h"
#include "swm_kernels.h"
#include "tm_mesh_dev.h"
#include "tmcore_gpu.h"

using namespace tmk;

// ------------------------------------------ device scratch of tm_invariants_noalloc
// The block partials (double-double, kNumSums x at most kNoAllocMaxBlocks),
// the kNumSums global sums and the four invariants of tm_invariants_noalloc
// live in this array, so that call allocates nothing.  Each GPU holds its own
// copy for the life of the process; host code reaches it only through the
// per-device address that hipGetSymbolAddress returns (noalloc_scratch_for).
// Calls on one GPU share it: calls that overlap in time race.
constexpr int kNoAllocMaxBlocks = 16384;  // workgroups per element kind: <= 4,194,304 elements
constexpr size_t kNoAllocPartBytes = (size_t)kNumSums * kNoAllocMaxBlocks * sizeof(double2);
constexpr size_t kNoAllocBytes = kNoAllocPartBytes + 256;  // partials, then sums[13] and inv[4]
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Address of `scratch` on `device` (which must be current), resolved once per
// device; nullptr with err set on failure.
void *noalloc_scratch_for(int device, std::string &err) {
  if (device < 0 || device >= kMaxDevices) {
    err = "device ordinal outside the scratch table";
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return nullptr;
    }
    g_scratch_ptr[device] = ptr;
  }
  return g_scratch_ptr[device];
}

// ------------------------------------------------------------- Hilbert order
// Skilling, "Programming the Hilbert curve", AIP Conf. Proc. 707 (2004):
// axes -> transposed Hilbert index, then bit interleave.
uint64_t hilbert3(uint32_t x, uint32_t y, uint32_t z, int bits) {
  uint32_t X[3] = {x, y, z};
  const uint32_t M = 1u << (bits - 1);
  for (uint32_t Q = M; Q > 1; Q >>= 1) {
    const uint32_t P = Q - 1;
    for (int i = 0; i < 3; ++i) {
      if (X[i] & Q) {
        X[0] ^= P;
      } else {
        uint32_t t = (X[0] ^ X[i]) & P;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
  for (int i = 1; i < 3; ++i) X[i] ^= X[i - 1];
  uint32_t t = 0;
  for (uint32_t Q = M; Q > 1; Q >>= 1)
    if (X[2] & Q) t ^= Q - 1;
  for (int i = 0; i < 3; ++i) X[i] ^= t;
  uint64_t key = 0;
  for (int b = bits - 1; b >= 0; --b)
    for (int i = 0; i < 3; ++i) key = (key << 1) | ((X[i] >> b) & 1u);
  return key;
}

// new -> old permutation sorting points on the unit sphere by Hilbert key.
std::vector<int> hilbert_order(int n, const double *x, const double *y, const double *z, bool enabled,
                               double pos_scale) {
  std::vector<int> p(n);
  std::iota(p.begin(), p.end(), 0);
  if (!enabled) return p;
  const int bits = 20;
  const double s = (double)((1u << bits) - 1) / 2.0;
  std::vector<uint64_t> key(n);
  for (int i = 0; i < n; ++i) {
    auto q = [&](double v) {
      return (uint32_t)std::lround((std::min(1.0, std::max(-1.0, v * pos_scale)) + 1.0) * s);
    };
    key[i] = hilbert3(q(x[i]), q(y[i]), q(z[i]), bits);
  }
  std::stable_sort(p.begin(), p.end(), [&](int a, int b) { return key[a] < key[b]; });
  return p;
}

inline int unsign_host(int idx) { return idx < 0 ? ~idx : idx; }

// Which edge kernels derive edgesOnEdge from the cells' rows by default
// (kXEoePv, kXEoeTend), per state precision: measured at level 8 with four
// tracers, deriving them speeds k_edge_pv up in both precisions, while the
// FP64 k_edge_tend is faster streaming the table.
int cell_rows_default(size_t rsz) { return rsz == 8 ? kXEoePv : kXEoePv | kXEoeTend; }

int round_blocks(int n) {
  int nb = (n + kBlock - 1) / kBlock;
  return ((nb + kXcd - 1) / kXcd) * kXcd;
}

}  // namespace

struct tm_ctx {
  tm_opts o{};
  std::string err;
  hipStream_t stream = nullptr;
  int nC = 0, nE = 0, nV = 0, nT = 0;
  int maxEdges = 6;  // the caller's edgesOnCell stride
  size_t rsz = 8;  // bytes per state value
  std::vector<int> pC, pE, pV;   // new -> old
  DMesh dm{};
  MeshDev md;
  std::vector<void *> allocs;
  double totalArea = 0.0, sumAreaTri = 0.0;
  double step_bytes[3] = {0, 0, 0};
  // state (element type R = double or float, stored untyped)
  void *gd_old = nullptr, *iap_old = nullptr, *q_old = nullptr, *ghs = nullptr;
  void *gdS[2] = {nullptr, nullptr}, *iap_s = nullptr, *qS[2] = {nullptr, nullptr};
  void *acc_gd = nullptr, *acc_iap = nullptr, *acc_q = nullptr;
  void *qe = nullptr;  // [nT][nE] corrected tracer edge values (adv_order >= 3)
  // pv_scheme 1/2 work arrays (null at scheme 3)
  void *pv_c = nullptr, *tgd_v = nullptr, *dpv = nullptr, *pvf = nullptr, *dpvf = nullptr;
  double2 *cv_part = nullptr;
  double *cv_eps = nullptr;
  void *t_gd[2] = {nullptr, nullptr}, *t_iap[2] = {nullptr, nullptr};
  void *gd_e = nullptr, *u = nullptr, *fk = nullptr, *udc = nullptr;  // fk: [nE][2] (StageBufs::fk)
  void *gd_e2 = nullptr, *u2 = nullptr;  // second gd_edge/u set (kXTendPrep, stages 2 and 4)
  void *E = nullptr, *tend_gd = nullptr, *pv_v = nullptr, *pv_e = nullptr;
  void *scratch_e[4] = {}, *scratch_c[2] = {}, *scratch_v[2] = {};
  double2 *part = nullptr;
  int part_stride = 0;
  double *sums = nullptr, *beta_hist = nullptr, *inv_hist = nullptr;
  int *hist_idx = nullptr;
  int hist_cap = 4096;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  // closure over several steps: step graphs without the MLP (kind 1: apply
  // the held correction; 2: MLP only; 3: nothing; 4: apply interval x dt)
  hipGraph_t graph_k[5] = {};
  hipGraphExec_t gexec_k[5] = {};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  void *noalloc_scratch = nullptr;  // this device's `scratch` (tm_invariants_noalloc)
  // pinned host buffer the state transfers stage through (tm_state_upload /
  // tm_state_download), allocated at the first transfer, grown on demand
  void *host_stage = nullptr;
  size_t host_stage_bytes = 0;
  bool have_state = false;
  bool timed = false;
  int serpentine = 1;  // alternate the tile sweep direction between kernels
  int xmask = 0;       // compile-time variant of the stage kernels (kX* bits of swm_kernels.h)
  int nstages = 4;     // stages per step: 4 (CRK4), 3 (CRK3, PC2)
  // learned closure: attached MLP, its per-step buffers and the training set
  tm_mlp *cl = nullptr;
  int cl_interval = 1;  // MLP every cl_interval steps
  int cl_lumped = 0;    // 0: correction added every step; 1: once per interval, times the interval
  double cl_gain = 1.0; // factor on the applied correction
  // history output: reconstruction coefficients [3][6][nC] and cos/sin of
  // latCell, lonCell [4][nC] (new numbering), uploaded by tm_history_setup
  double *hist_rc = nullptr, *hist_trig = nullptr;
  double *hist_out = nullptr;  // [4][nC] zonal, meridional, h, div
  void *hist_vor = nullptr;    // [nV] vorticity
  long long cl_step = 0;  // steps since the state or the closure changed
  float *cl_X = nullptr, *cl_Y = nullptr, *cl_r = nullptr, *cl_dY = nullptr;
  void *cl_sq = nullptr, *cl_vor = nullptr, *cl_div = nullptr;
  double *cl_loss = nullptr;
  double *cl_lpart = nullptr;  // [nbE + nbC] per-block loss partials
  int cl_loss_cap = 0;
  struct ClSample {
    float *X, *tu, *tg;
    float tgs;
  };
  std::vector<ClSample> ds;

  int fail(int code, const std::string &m) {
    err = m;
    return code;
  }
  template <typename T>
  T *alloc(size_t n) {
    void *p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return nullptr;
    allocs.push_back(p);
    return (T *)p;
  }
  void *alloc_r(size_t n) {
    void *p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(n, 1) * rsz) != hipSuccess) return nullptr;
    allocs.push_back(p);
    return p;
  }
  // frees one allocation made by alloc/alloc_r before the context ends
  void release(void *p) {
    if (!p) return;
    (void)hipFree(p);
    auto it = std::find(allocs.begin(), allocs.end(), p);
    if (it != allocs.end()) allocs.erase(it);
  }
};

#define HIPCHK(ctx, call)                                                                    \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess) return (ctx)->fail(TM_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace {

template <typename T>
int upload(tm_ctx *c, T *dst, const std::vector<T> &h) {
  HIPCHK(c, hipMemcpyAsync(dst, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TM_OK;
}

// f(lo, hi) over [0, n) in contiguous chunks on up to 16 host threads (one
// thread below 64k elements).
template <typename F>
void par_for(size_t n, F f) {
  const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const size_t nt = n < (size_t(1) << 16) ? 1 : std::min<size_t>(hw, n >> 15);
  if (nt <= 1) {
    f(size_t(0), n);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(nt - 1);
  size_t k = 1;
  try {
    for (; k < nt; ++k) th.emplace_back(f, n * k / nt, n * (k + 1) / nt);
  } catch (const std::system_error &) {
    // no more threads: the chunks not handed out run on this thread
  }
  f(size_t(0), n / nt);
  for (size_t r = k; r < nt; ++r) f(n * r / nt, n * (r + 1) / nt);
  for (auto &t : th) t.join();
}

// The context's pinned staging buffer, at least `bytes` long (null if the
// pinned allocation fails).
void *host_stage(tm_ctx *c, size_t bytes) {
  if (c->host_stage_bytes >= bytes) return c->host_stage;
  if (c->host_stage) (void)hipHostFree(c->host_stage);
  c->host_stage = nullptr;
  c->host_stage_bytes = 0;
  if (hipHostMalloc(&c->host_stage, bytes, hipHostMallocDefault) != hipSuccess) {
    c->host_stage = nullptr;
    return nullptr;
  }
  c->host_stage_bytes = bytes;
  return c->host_stage;
}

// Convert a reference-order FP64 host array to new order and the state type,
// through the pinned staging buffer.
int upload_state(tm_ctx *c, void *dst, const double *src, const std::vector<int> &perm, int ncopies, int n) {
  const size_t len = (size_t)n * ncopies;
  void *h = host_stage(c, len * c->rsz);
  if (!h) return c->fail(TM_EDEVICE, "hipHostMalloc of the transfer staging buffer failed");
  const int *pm = perm.data();
  if (c->rsz == 8) {
    double *hd = (double *)h;
    par_for(len, [=](size_t lo, size_t hi) {
      for (size_t j = lo; j < hi; ++j) hd[j] = src[j - j % n + pm[j % n]];
    });
  } else {
    float *hf = (float *)h;
    par_for(len, [=](size_t lo, size_t hi) {
      for (size_t j = lo; j < hi; ++j) hf[j] = (float)src[j - j % n + pm[j % n]];
    });
  }
  HIPCHK(c, hipMemcpyAsync(dst, h, len * c->rsz, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TM_OK;
}

int download_state(tm_ctx *c, double *dst, const void *src, const std::vector<int> &perm, int ncopies, int n) {
  const size_t len = (size_t)n * ncopies;
  void *h = host_stage(c, len * c->rsz);
  if (!h) return c->fail(TM_EDEVICE, "hipHostMalloc of the transfer staging buffer failed");
  HIPCHK(c, hipMemcpyAsync(h, src, len * c->rsz, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const int *pm = perm.data();
  if (c->rsz == 8) {
    const double *hd = (const double *)h;
    par_for(len, [=](size_t lo, size_t hi) {
      for (size_t j = lo; j < hi; ++j) dst[j - j % n + pm[j % n]] = hd[j];
    });
  } else {
    const float *hf = (const float *)h;
    par_for(len, [=](size_t lo, size_t hi) {
      for (size_t j = lo; j < hi; ++j) dst[j - j % n + pm[j % n]] = hf[j];
    });
  }
  return TM_OK;
}

// add_upwind_correction_on_cell coefficients (operators_mod.F90:436-441;
// coef4 keeps its initial 0.25 unless adv_order is 6)
template <typename R>
void set_adv(StageBufs<R> &b, const tm_opts &o) {
  b.adv_order = o.adv_order;
  b.adv_coef2 = o.adv_order == 3 ? (o.adv_mono ? 0.25 : 1.0) : 0.0;
  b.adv_coef4 = o.adv_order == 6 ? 0.0 : 0.25;
}

template <typename R>
StageBufs<R> stage_bufs(tm_ctx *c, int k, bool want_inv) {
  StageBufs<R> b{};
  auto P = [](void *p) { return (R *)p; };
  // gd and the tracers ping-pong: k_cell_vertex reads neighbour cells of the
  // stage input while it writes the next stage state.
  const int in_slot = (k == 2 || k == 4) ? 0 : 1;
  const int out_slot = (k == 1 || k == 3) ? 0 : 1;
  b.gd = (k <= 1) ? P(c->gd_old) : P(c->gdS[in_slot]);
  b.iap = (k <= 1) ? P(c->iap_old) : P(c->iap_s);
  b.q = (k <= 1) ? P(c->q_old) : P(c->qS[in_slot]);
  b.gd_base = P(c->gd_old);
  b.iap_base = P(c->iap_old);
  b.q_base = P(c->q_old);
  b.gd_next = P(c->gdS[out_slot]);
  b.iap_next = P(c->iap_s);
  b.q_next = P(c->qS[out_slot]);
  b.acc_gd = P(c->acc_gd);
  b.acc_iap = P(c->acc_iap);
  b.acc_q = P(c->acc_q);
  b.tprev_gd = P(c->t_gd[(k + 1) & 1]);
  b.tprev_iap = P(c->t_iap[(k + 1) & 1]);
  b.tcur_gd = P(c->t_gd[k & 1]);
  b.tcur_iap = P(c->t_iap[k & 1]);
  // gd_edge and u: set 0 for stages 0, 1 and 3, set 1 for stages 2 and 4 when
  // k_edge_tend forms the next stage's edge values (kXTendPrep); else set 0
  const bool tp = (c->xmask & kXTendPrep) != 0;
  const int es = tp && (k == 2 || k == 4) ? 1 : 0;
  b.gd_e = P(es ? c->gd_e2 : c->gd_e);
  b.u = P(es ? c->u2 : c->u);
  b.gd_e_n = P(es ? c->gd_e : c->gd_e2);
  b.u_n = P(es ? c->u : c->u2);
  b.prep_next = tp && k >= 1 && k <= 3 ? 1 : 0;
  b.fk = P(c->fk);
  b.udc = P(c->udc);
  b.E = P(c->E);
  b.tend_gd = P(c->tend_gd);
  b.pv_v = P(c->pv_v);
  b.pv_e = P(c->pv_e);
  b.ghs = P(c->ghs);
  b.nT = c->nT;
  b.k = k;
  const double cf[5] = {0.0, 0.5, 0.5, 1.0, 0.0};
  b.cdt = cf[k] * c->o.dt;
  // stage descriptor (kXExt kernels): time_integrate_CRK4/CRK3/PC2,
  // time_scheme_mod.F90:109-154, 65-107, 156-184
  const int ts = c->o.time_scheme;
  b.ts_last = k == c->nstages;
  b.rid_prev = b.rid_first = -1;
  b.cprev_dt = 0.0;
  b.ts_w = 0.0;
  b.tfirst_gd = P(c->t_gd[1]);  // stage 1 writes slot 1 & 1
  b.tfirst_iap = P(c->t_iap[1]);
  if (k >= 1 && ts == 2) {
    b.ts_acc = k == 1 ? 1 : (k < 4 ? 2 : 3);
    b.ts_w = 2.0;
    b.rid_prev = k == 2 ? kR12e : (k == 3 ? kR23e : (k == 4 ? kR34e : -1));
  } else if (k >= 1 && ts == 1) {  // X2 = X1 + dt/2 T1; X3 = (X1 - dt T1) + 2 dt T2; (T1 + 4 T2 + T3) / 6
    b.cdt = (k == 1 ? 0.5 : 2.0) * c->o.dt;
    b.cprev_dt = k == 2 ? -c->o.dt : 0.0;
    b.ts_acc = k == 1 ? 1 : (k == 2 ? 2 : 3);
    b.ts_w = 4.0;
    b.rid_prev = k == 2 ? kR12e : (k == 3 ? kR23e : -1);
    b.rid_first = k == 3 ? kR34e : -1;  // R13
  } else if (k >= 1 && ts == 3) {  // X_new = X_old + dt/2 T twice, then T_new; ip1 = T_b.T_new
    b.cdt = 0.5 * c->o.dt;
    b.ts_acc = k == 3 ? 1 : 0;
    b.rid_prev = k == 3 ? kR34e : -1;
  }
  b.want_inv = want_inv ? 1 : 0;
  b.pv_order = c->o.pv_scheme == 3 ? c->o.interp_pv_order : 2;
  b.pv_scheme = c->o.pv_scheme;
  b.apvm_wdt = c->o.apvm_weight * c->o.dt;
  b.dt = c->o.dt;
  b.pv_c = P(c->pv_c);
  b.tgd_v = P(c->tgd_v);
  b.dpv = P(c->dpv);
  b.pvf = P(c->pvf);
  b.dpvf = P(c->dpvf);
  b.cv_part = c->cv_part;
  b.cv_eps = c->cv_eps;
  set_adv(b, c->o);
  b.qe = P(c->qe);
  b.part = c->part;
  b.part_stride = c->part_stride;
  return b;
}

// Stage part `which` (0 edge prep, 1 cells + vertices, 2 PV on edges, 3 edge
// tendencies) with the kernels' compile-time variant X (kXAdv, kXExt).  Part 2
// is k_edge_pv plus, at pv_scheme 1/2, k_pv_cell before it and, at scheme 2,
// the conservative-APVM kernels after it.
template <typename R, int X>
void launch_part_x(tm_ctx *c, int which, const StageBufs<R> &b) {
  const DMesh &m = c->dm;
  const int ps = c->o.pv_scheme;
  switch (which) {
    case 0:
      hipLaunchKernelGGL((k_edge_prep<R, X>), dim3(m.nbE), dim3(kBlock), 0, c->stream, m, b);
      break;
    case 1:
      hipLaunchKernelGGL((k_cell_vertex<R, X>), dim3(m.nbC + m.nbV), dim3(kBlock), 0, c->stream, m, b);
      break;
    case 2:
      if (ps != 3)  // pv_cell (and gd_tend_vertex for scheme 2)
        hipLaunchKernelGGL(k_pv_cell<R>, dim3(m.nbC + (ps == 2 ? m.nbV : 0)), dim3(kBlock), 0, c->stream, m, b);
      hipLaunchKernelGGL((k_edge_pv<R, X>), dim3(m.nbE), dim3(kBlock), 0, c->stream, m, b);
      if (ps == 2) {
        hipLaunchKernelGGL(k_cvor_flux<R>, dim3(m.nbE), dim3(kBlock), 0, c->stream, m, b);
        hipLaunchKernelGGL(k_cvor_sums<R>, dim3(m.nbV), dim3(kBlock), 0, c->stream, m, b);
        hipLaunchKernelGGL(k_cvor_eps<R>, dim3(1), dim3(kBlock), 0, c->stream, m, b);
        hipLaunchKernelGGL(k_cvor_apply<R>, dim3(m.nbE), dim3(kBlock), 0, c->stream, m, b);
      }
      break;
    default:
      hipLaunchKernelGGL((k_edge_tend<R, X>), dim3(m.nbE), dim3(kBlock), 0, c->stream, m, b);
  }
}

template <typename R>
void launch_part(tm_ctx *c, int which, const StageBufs<R> &b) {
  switch (c->xmask) {
    case 0: launch_part_x<R, 0>(c, which, b); break;
    case kXAdv: launch_part_x<R, kXAdv>(c, which, b); break;
    case kXExt: launch_part_x<R, kXExt>(c, which, b); break;
    case kXAdv | kXExt: launch_part_x<R, kXAdv | kXExt>(c, which, b); break;
    case kXAdv | kXAdv56: launch_part_x<R, kXAdv | kXAdv56>(c, which, b); break;
    case kXEoePv: launch_part_x<R, kXEoePv>(c, which, b); break;
    case kXTendPrep: launch_part_x<R, kXTendPrep>(c, which, b); break;
    case kXTendPrep | kXEoePv: launch_part_x<R, kXTendPrep | kXEoePv>(c, which, b); break;
    case kXTendPrep | kXEoeTend: launch_part_x<R, kXTendPrep | kXEoeTend>(c, which, b); break;
    case kXTendPrep | kXEoePv | kXEoeTend: launch_part_x<R, kXTendPrep | kXEoePv | kXEoeTend>(c, which, b); break;
    case kXEoeTend: launch_part_x<R, kXEoeTend>(c, which, b); break;
    case kXTendLds: launch_part_x<R, kXTendLds>(c, which, b); break;
    case kXTendLds | kXEoePv: launch_part_x<R, kXTendLds | kXEoePv>(c, which, b); break;
    case kXTendLds | kXTendPrep: launch_part_x<R, kXTendLds | kXTendPrep>(c, which, b); break;
    case kXTendLds | kXTendPrep | kXEoePv: launch_part_x<R, kXTendLds | kXTendPrep | kXEoePv>(c, which, b); break;
    case kXEoePv | kXEoeTend: launch_part_x<R, kXEoePv | kXEoeTend>(c, which, b); break;
    case kXPvLds: launch_part_x<R, kXPvLds>(c, which, b); break;
    case kXPvLds | kXTendPrep: launch_part_x<R, kXPvLds | kXTendPrep>(c, which, b); break;
    case kXPvLds | kXEoeTend: launch_part_x<R, kXPvLds | kXEoeTend>(c, which, b); break;
    case kXPvLds | kXEoeTend | kXTendPrep: launch_part_x<R, kXPvLds | kXEoeTend | kXTendPrep>(c, which, b); break;
    case kXPvLds | kXTendLds: launch_part_x<R, kXPvLds | kXTendLds>(c, which, b); break;
    case kXPvLds | kXTendLds | kXTendPrep: launch_part_x<R, kXPvLds | kXTendLds | kXTendPrep>(c, which, b); break;
    case kXExt | kXEoePv: launch_part_x<R, kXExt | kXEoePv>(c, which, b); break;
    case kXExt | kXEoeTend: launch_part_x<R, kXExt | kXEoeTend>(c, which, b); break;
    case kXExt | kXEoePv | kXEoeTend: launch_part_x<R, kXExt | kXEoePv | kXEoeTend>(c, which, b); break;
    default: launch_part_x<R, kXAdv | kXAdv56 | kXExt>(c, which, b);
  }
}

// The four kernels sweep forward, backward, forward, backward (see
// xcd_tile): the tiles each starts on were the last the previous one touched.
template <typename R>
int launch_stage(tm_ctx *c, const StageBufs<R> &b0, bool tail = true) {
  StageBufs<R> b = b0;
  // with kXTendPrep the previous stage's k_edge_tend formed this stage's edge values
  const int first = (c->xmask & kXTendPrep) && b.k >= 2 ? 1 : 0;
  for (int which = first; which < (tail ? 4 : 2); ++which) {
    b.rev = (which & 1) ? c->serpentine : 0;
    launch_part<R>(c, which, b);
  }
  HIPCHK(c, hipGetLastError());
  return TM_OK;
}

ReduceArgs reduce_args(tm_ctx *c, int mode, bool have_inv) {
  ReduceArgs a{};
  a.part = c->part;
  a.stride = c->part_stride;
  a.nbE = c->dm.nbE;
  a.nbC = c->dm.nbC;
  a.nbV = c->dm.nbV;
  a.mode = mode;
  a.have_inv = have_inv ? 1 : 0;
  a.sums = c->sums;
  return a;
}

StepOut step_out(tm_ctx *c, bool have_inv) {
  StepOut o{};
  o.sums = c->sums;
  o.inv_total_area = 1.0 / c->totalArea;  // inner_product divides by totalArea
  o.inv_sum_area_tri = 1.0 / c->sumAreaTri;
  o.conserve_energy = c->o.conserve_energy;
  o.time_scheme = c->o.time_scheme;
  o.have_inv = have_inv ? 1 : 0;
  o.hist_idx = c->hist_idx;
  o.beta_hist = c->beta_hist;
  o.inv_hist = c->inv_hist;
  o.hist_cap = c->hist_cap;
  return o;
}

template <typename R>
FinalBufs<R> final_bufs(tm_ctx *c) {
  FinalBufs<R> f{};
  f.gd = (R *)c->gd_old;
  f.iap = (R *)c->iap_old;
  f.q = (R *)c->q_old;
  f.acc_gd = (const R *)c->acc_gd;
  f.acc_iap = (const R *)c->acc_iap;
  f.acc_q = (const R *)c->acc_q;
  f.dt = c->o.dt;
  f.nT = c->nT;
  f.out = step_out(c, true);
  return f;
}

// Closure inputs of the stage-1 fields (the start-of-step state) into X,
// and sqrt(gd_edge) into sq when sq != nullptr.
// (the stage's vertex pass wrote the vorticity into cl_vor: see cl_stage_bufs)
template <typename R>
int enqueue_cl_features(tm_ctx *c, float *X, void *sq) {
  hipLaunchKernelGGL(k_cl_div<R>, dim3(c->dm.nbC), dim3(kBlock), 0, c->stream, c->dm, (const R *)c->u,
                     (R *)c->cl_div);
  ClFeatArgs<R> a{};
  a.u = (const R *)c->u;
  a.gd_e = (const R *)c->gd_e;
  a.vor = (const R *)c->cl_vor;
  a.div = (const R *)c->cl_div;
  a.norm = mlp_state(c->cl).norm;
  a.X = X;
  a.sq = (R *)sq;
  hipLaunchKernelGGL(k_cl_features<R>, dim3(c->dm.nbE), dim3(kBlock), 0, c->stream, c->dm, a);
  HIPCHK(c, hipGetLastError());
  return TM_OK;
}

template <typename R>
int enqueue_step(tm_ctx *c, bool infer, int apply_factor) {
  for (int k = 1; k <= c->nstages; ++k) {
    StageBufs<R> b = stage_bufs<R>(c, k, k == 1);
    if (k == 1 && c->cl && infer) b.d_vor = (R *)c->cl_vor;  // closure input
    int rc = launch_stage<R>(c, b);
    if (rc) return rc;
    if (k == 1 && c->cl && infer) {
      // MLP of the start-of-step state, between stage 1 and stage 2 (which
      // overwrites u, gd_edge, pv_edge and tend_gd)
      if ((rc = enqueue_cl_features<R>(c, c->cl_X, c->cl_sq))) return rc;
      MlpState &ms = mlp_state(c->cl);
      if ((rc = mlp_forward(ms, c->cl_X, c->nE, c->cl_Y, false, c->stream))) return c->fail(rc, ms.err);
    }
  }
  hipLaunchKernelGGL(k_reduce, dim3(kNumSums), dim3(kReduceBlock), 0, c->stream, reduce_args(c, 1, true));
  hipLaunchKernelGGL(k_finalize<R>, dim3(c->dm.nbE + c->dm.nbC), dim3(kBlock), 0, c->stream, c->dm, final_bufs<R>(c));
  if (c->cl && apply_factor > 0) {
    ClApplyArgs<R> a{};
    a.Y = c->cl_Y;
    a.sq = (const R *)c->cl_sq;
    a.norm = mlp_state(c->cl).norm;
    a.dt = c->o.dt * apply_factor * c->cl_gain;
    a.iap = (R *)c->iap_old;
    a.gd = (R *)c->gd_old;
    hipLaunchKernelGGL(k_cl_apply<R>, dim3(c->dm.nbE + c->dm.nbC), dim3(kBlock), 0, c->stream, c->dm, a);
  }
  HIPCHK(c, hipGetLastError());
  return TM_OK;
}

// Invariants of state(old) (the post-step refresh of tmcore_swm_mod.F90:78-85):
// the refresh stage, with its block partials in `part` (kNumSums x
// c->part_stride), and the reduction into `sums`.
template <typename R>
int enqueue_refresh_sums(tm_ctx *c, double2 *part, double *sums) {
  StageBufs<R> b = stage_bufs<R>(c, 0, true);
  b.part = part;
  int rc = launch_stage<R>(c, b, false);
  if (rc) return rc;
  ReduceArgs a = reduce_args(c, 0, true);
  a.part = part;
  a.sums = sums;
  hipLaunchKernelGGL(k_reduce, dim3(kNumSums), dim3(kReduceBlock), 0, c->stream, a);
  HIPCHK(c, hipGetLastError());
  return TM_OK;
}

template <typename R>
int enqueue_refresh(tm_ctx *c) {
  int rc = enqueue_refresh_sums<R>(c, c->part, c->sums);
  if (rc) return rc;
  hipLaunchKernelGGL(k_inv_out, dim3(1), dim3(64), 0, c->stream, step_out(c, true));
  HIPCHK(c, hipGetLastError());
  return TM_OK;
}

int capture_step(tm_ctx *c, bool infer, int apply_factor, hipGraph_t *graph, hipGraphExec_t *exec) {
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int rc = c->rsz == 8 ? enqueue_step<double>(c, infer, apply_factor) : enqueue_step<float>(c, infer, apply_factor);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (rc) return rc;
  if (e != hipSuccess) return c->fail(TM_EDEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  *graph = g;
  HIPCHK(c, hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
  return TM_OK;
}

void drop_graphs(tm_ctx *c) {
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->gexec = nullptr;
  c->graph = nullptr;
  for (int k = 1; k < 5; ++k) {
    if (c->gexec_k[k]) (void)hipGraphExecDestroy(c->gexec_k[k]);
    if (c->graph_k[k]) (void)hipGraphDestroy(c->graph_k[k]);
    c->gexec_k[k] = nullptr;
    c->graph_k[k] = nullptr;
  }
}

// Step graph of kind k: 0 = MLP + correction (every step of a closure with
// interval 1, or the plain step without closure), 1 = held correction only,
// 2 = MLP only, 3 = plain step, 4 = correction times the interval.
int step_graph(tm_ctx *c, int k, hipGraphExec_t *out) {
  int rc = TM_OK;
  if (k == 0) {
    if (!c->gexec && (rc = capture_step(c, true, 1, &c->graph, &c->gexec))) return rc;
    *out = c->gexec;
    return TM_OK;
  }
  if (!c->gexec_k[k]) {
    const bool infer = k == 2;
    const int factor = k == 1 ? 1 : (k == 4 ? c->cl_interval : 0);
    if ((rc = capture_step(c, infer, factor, &c->graph_k[k], &c->gexec_k[k]))) return rc;
  }
  *out = c->gexec_k[k];
  return TM_OK;
}

// Which step graph the closure schedule runs at step s of the interval.
int schedule_kind(const tm_ctx *c) {
  if (!c->cl || c->cl_interval == 1) return 0;
  const int r = (int)(c->cl_step % c->cl_interval);
  if (!c->cl_lumped) return r == 0 ? 0 : 1;
  if (r == 0) return 2;
  return r == c->cl_interval - 1 ? 4 : 3;
}

int build_graph(tm_ctx *c) {
  hipGraphExec_t g = nullptr;
  return step_graph(c, 0, &g);
}

int run_steps(tm_ctx *c, int nsteps, bool refresh) {
  int rc = build_graph(c);
  if (rc) return rc;
  hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, c->stream, c->hist_idx);
  HIPCHK(c, hipEventRecord(c->ev0, c->stream));
  for (int s = 0; s < nsteps; ++s) {
    hipGraphExec_t g = nullptr;
    if ((rc = step_graph(c, schedule_kind(c), &g))) return rc;
    HIPCHK(c, hipGraphLaunch(g, c->stream));
    ++c->cl_step;
  }
  HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  c->timed = true;
  if (refresh) {
    rc = c->rsz == 8 ? enqueue_refresh<double>(c) : enqueue_refresh<float>(c);
    if (rc) return rc;
  }
  return TM_OK;
}

bool finite_all(const double *a, size_t n) {
  std::atomic<int> bad{0};
  par_for(n, [&](size_t lo, size_t hi) {
    int b = 0;
    for (size_t i = lo; i < hi; ++i) b |= !std::isfinite(a[i]);
    if (b) bad.store(1, std::memory_order_relaxed);
  });
  return bad.load() == 0;
}

}  // namespace

// =================================================================== C-ABI
extern "C" void tm_opts_default(tm_opts *o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->time_scheme = 2;
  o->pv_scheme = 3;
  o->interp_pv_order = 3;
  o->adv_order = 2;
  o->adv_mono = 1;
  o->conserve_energy = 1;
  o->apvm_weight = 0.5;
  o->dt = 600.0;
  o->radius = 6371220.0;
  o->ntracers = 0;
  o->precision = 64;
  o->device = 0;
  o->renumber = 1;
}

extern "C" const char *tm_last_error(const tm_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

extern "C" int tm_ctx_destroy(tm_ctx *c) {
  if (!c) return TM_EINVAL;
  drop_graphs(c);
  for (auto &smp : c->ds) {
    (void)hipFree(smp.X);
    (void)hipFree(smp.tu);
    (void)hipFree(smp.tg);
  }
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  for (void *p : c->allocs) (void)hipFree(p);
  if (c->host_stage) (void)hipHostFree(c->host_stage);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return TM_OK;
}

namespace tmk {

int validate_mesh(const tm_mesh_desc *m, const tm_fit_desc *f, const tm_opts &o, std::string &err) {
  auto failv = [&](int code, const char *msg) {
    err = msg;
    return code;
  };
  if (o.time_scheme < 1 || o.time_scheme > 3) return failv(TM_EINVAL, "time_scheme must be 1 (CRK3), 2 (CRK4) or 3 (PC2)");
  if (o.time_scheme != 2 && o.ntracers > 0) return failv(TM_EINVAL, "tracers are advanced with time_scheme 2 (CRK4) only");
  if (o.pv_scheme < 1 || o.pv_scheme > 3) return failv(TM_EINVAL, "pv_scheme must be 1, 2 or 3");
  if (o.interp_pv_order < 2 || o.interp_pv_order > 6) return failv(TM_EINVAL, "interp_pv_order must be 2..6");
  if (o.adv_order < 2 || o.adv_order > 6) return failv(TM_EINVAL, "adv_order must be 2..6");
  if (o.precision != 64 && o.precision != 32) return failv(TM_EINVAL, "precision must be 64 or 32");
  if (o.ntracers < 0 || o.ntracers > 64) return failv(TM_EINVAL, "ntracers out of range");
  if (!(o.dt > 0.0) || !(o.radius > 0.0)) return failv(TM_EINVAL, "dt and radius must be positive");
  if (!m || m->nCells <= 0 || m->nEdges <= 0 || m->nVertices <= 0 || m->vertexDegree != 3 || m->maxEdges < 3 ||
      m->maxEdges2 < 1)
    return failv(TM_EINVAL, "bad mesh dimensions");
  const bool need_fit = o.pv_scheme == 3 && o.interp_pv_order >= 3;
  if (need_fit && (!f || !f->nFit2Vertices || !f->fit2Vertices || !f->deriv2OnVertex || f->maxFit < 2))
    return failv(TM_EINVAL, "interp_pv_order >= 3 needs nFit2Vertices, fit2Vertices and deriv2OnVertex");
  const bool need_fit4v = o.pv_scheme == 3 && o.interp_pv_order >= 5;
  if (need_fit4v && (!f->nFit4Vertices || !f->fit4Vertices || !f->deriv4OnVertex))
    return failv(TM_EINVAL, "interp_pv_order >= 5 needs nFit4Vertices, fit4Vertices and deriv4OnVertex");
  if (o.adv_order >= 3 && (!f || !f->nFit2Cells || !f->fit2Cells || !f->deriv2OnCell || f->maxFit < 2))
    return failv(TM_EINVAL, "adv_order >= 3 needs nFit2Cells, fit2Cells and deriv2OnCell");
  if (o.adv_order >= 5 && (!f->nFit4Cells || !f->fit4Cells || !f->deriv4OnCell))
    return failv(TM_EINVAL, "adv_order >= 5 needs nFit4Cells, fit4Cells and deriv4OnCell");
  const int nC = m->nCells, nE = m->nEdges, nV = m->nVertices, ME = m->maxEdges, ME2 = m->maxEdges2;
  auto inr = [](int v, int n) { return v >= 1 && v <= n; };
  for (int i = 0; i < nC; ++i) {
    int n = m->nEdgesOnCell[i];
    if (n < 3 || n > ME) return failv(TM_ERANGE, "nEdgesOnCell out of range");
    if (n > kMaxEoC) return failv(TM_ESTENCIL, "a cell has more than 6 edges");
    for (int s = 0; s < n; ++s)
      if (!inr(m->edgesOnCell[(size_t)i * ME + s], nE) || !inr(m->verticesOnCell[(size_t)i * ME + s], nV))
        return failv(TM_ERANGE, "edgesOnCell/verticesOnCell index out of range");
  }
  for (int e = 0; e < nE; ++e) {
    if (!inr(m->cellsOnEdge[2 * (size_t)e], nC) || !inr(m->cellsOnEdge[2 * (size_t)e + 1], nC) ||
        !inr(m->verticesOnEdge[2 * (size_t)e], nV) || !inr(m->verticesOnEdge[2 * (size_t)e + 1], nV))
      return failv(TM_ERANGE, "cellsOnEdge/verticesOnEdge index out of range");
    int n = m->nEdgesOnEdge[e];
    if (n < 0 || n > kMaxEoE || n > ME2) return failv(TM_ESTENCIL, "nEdgesOnEdge exceeds 10");
    for (int s = 0; s < n; ++s)
      if (!inr(m->edgesOnEdge[(size_t)e * ME2 + s], nE)) return failv(TM_ERANGE, "edgesOnEdge index out of range");
  }
  for (int v = 0; v < nV; ++v)
    for (int j = 0; j < 3; ++j)
      if (!inr(m->edgesOnVertex[(size_t)v * 3 + j], nE) || !inr(m->cellsOnVertex[(size_t)v * 3 + j], nC))
        return failv(TM_ERANGE, "edgesOnVertex/cellsOnVertex index out of range");
  if (need_fit)
    for (int v = 0; v < nV; ++v) {
      int n = f->nFit2Vertices[v];
      if (n < 1 || n - 1 > kFit2 || n > f->maxFit) return failv(TM_ESTENCIL, "nFit2Vertices exceeds 13");
      for (int j = 1; j < n; ++j)
        if (!inr(f->fit2Vertices[(size_t)v * f->maxFit + j], nV))
          return failv(TM_ERANGE, "fit2Vertices index out of range");
    }
  if (need_fit4v)
    for (int v = 0; v < nV; ++v) {
      int n = f->nFit4Vertices[v];
      if (n < 1 || n - 1 > kFit4V || n > f->maxFit) return failv(TM_ESTENCIL, "nFit4Vertices exceeds 22");
      for (int j = 1; j < n; ++j)
        if (!inr(f->fit4Vertices[(size_t)v * f->maxFit + j], nV))
          return failv(TM_ERANGE, "fit4Vertices index out of range");
    }
  for (int order = 2; order <= 4; order += 2) {
    if (o.adv_order < order + 1) break;
    const int *nf = order == 2 ? f->nFit2Cells : f->nFit4Cells, *fc = order == 2 ? f->fit2Cells : f->fit4Cells;
    const int ns = order == 2 ? kFit2C : kFit4C;
    for (int c = 0; c < nC; ++c) {
      int n = nf[c];
      if (n < 1 || n - 1 > ns || n > f->maxFit) return failv(TM_ESTENCIL, "cell fit stencil exceeds the device layout");
      for (int j = 1; j < n; ++j)
        if (!inr(fc[(size_t)c * f->maxFit + j], nC)) return failv(TM_ERANGE, "cell fit index out of range");
    }
  }

  return TM_OK;
}

int build_mesh_dev(MeshDev &md, const tm_mesh_desc *m, const tm_fit_desc *f, const tm_opts &o,
                   hipStream_t stream, std::vector<void *> &allocs, std::string &err) {
  int rc = validate_mesh(m, f, o, err);
  if (rc) return rc;
  // the vertex fits are read by pv_scheme 3 at interp_pv_order >= 3 only (the
  // predicate validate_mesh checks them under)
  const bool need_fit = o.pv_scheme == 3 && o.interp_pv_order >= 3;
  const int nC = m->nCells, nE = m->nEdges, nV = m->nVertices, ME = m->maxEdges, ME2 = m->maxEdges2;
  // Hilbert renumbering: new -> old and old -> new.  Keys are taken on the
  // unit sphere whether or not the caller's coordinates are radius-scaled.
  const bool ren = o.renumber != 0;
  const double pos_scale = o.geometry_scaled ? 1.0 / o.radius : 1.0;
  md.pC = hilbert_order(nC, m->xCell, m->yCell, m->zCell, ren, pos_scale);
  md.pE = hilbert_order(nE, m->xEdge, m->yEdge, m->zEdge, ren, pos_scale);
  md.pV = hilbert_order(nV, m->xVertex, m->yVertex, m->zVertex, ren, pos_scale);
  md.iC.resize(nC);
  md.iE.resize(nE);
  md.iV.resize(nV);
  for (int i = 0; i < nC; ++i) md.iC[md.pC[i]] = i;
  for (int i = 0; i < nE; ++i) md.iE[md.pE[i]] = i;
  for (int i = 0; i < nV; ++i) md.iV[md.pV[i]] = i;
  // Lengths and areas: scale unit-sphere values as mesh_mod.F90:205-224 does,
  // or take them as they are when the caller has already scaled them.
  const double a = o.geometry_scaled ? 1.0 : o.radius, a2 = a * a;

  // ---- cells
  std::vector<int8_t> nEoC(nC);
  std::vector<int> eocS((size_t)kMaxEoC * nC, 0), vocS((size_t)kMaxEoC * nC, 0), cocN((size_t)kMaxEoC * nC, 0);
  std::vector<double> areaCell(nC), kiteC((size_t)kMaxEoC * nC, 0.0);
  double tot = 0.0;
  for (int i = 0; i < nC; ++i) tot += m->areaCell[i] * a2;  // mesh_mod.F90:224 order
  md.totalArea = tot;
  for (int ni = 0; ni < nC; ++ni) {
    const int oi = md.pC[ni];
    const int n = m->nEdgesOnCell[oi];
    nEoC[ni] = (int8_t)n;
    areaCell[ni] = m->areaCell[oi] * a2;
    for (int s = n; s < kMaxEoC; ++s) cocN[(size_t)s * nC + ni] = ni;
    for (int s = 0; s < n; ++s) {
      const int oe = m->edgesOnCell[(size_t)oi * ME + s] - 1;
      const int ne = md.iE[oe];
      const bool plus = m->cellsOnEdge[2 * (size_t)oe] - 1 == oi;  // nSignEdge = +1
      eocS[(size_t)s * nC + ni] = plus ? ne : ~ne;
      // the cell across the edge (cellsOnCell(s) lies across edgesOnCell(s))
      const int other = plus ? m->cellsOnEdge[2 * (size_t)oe + 1] - 1 : m->cellsOnEdge[2 * (size_t)oe] - 1;
      cocN[(size_t)s * nC + ni] = md.iC[other];
      const int ov = m->verticesOnCell[(size_t)oi * ME + s] - 1;
      vocS[(size_t)s * nC + ni] = md.iV[ov];
      double k = 0.0;
      for (int j = 0; j < 3; ++j)
        if (m->cellsOnVertex[(size_t)ov * 3 + j] - 1 == oi) k = m->kiteAreasOnVertex[(size_t)ov * 3 + j] * a2;
      kiteC[(size_t)s * nC + ni] = k;
    }
  }
  // ---- edges
  std::vector<int> coe(2 * (size_t)nE), voe(2 * (size_t)nE), eoe((size_t)kMaxEoE * nE, 0);
  std::vector<int8_t> nEoE(nE), kapE(nE);
  std::vector<double> dv(nE), dc(nE), aE(nE), woe((size_t)kMaxEoE * nE, 0.0), d2v((size_t)2 * kFit2 * nE, 0.0);
  for (int ne = 0; ne < nE; ++ne) {
    const int oe = md.pE[ne];
    coe[ne] = md.iC[m->cellsOnEdge[2 * (size_t)oe] - 1];
    coe[(size_t)nE + ne] = md.iC[m->cellsOnEdge[2 * (size_t)oe + 1] - 1];
    {
      // +1 when the normal cellsOnEdge(1) -> (2) points east (north if it is
      // exactly meridional), -1 otherwise
      const int c1 = m->cellsOnEdge[2 * (size_t)oe] - 1, c2 = m->cellsOnEdge[2 * (size_t)oe + 1] - 1;
      const double nx = m->xCell[c2] - m->xCell[c1], ny = m->yCell[c2] - m->yCell[c1], nz = m->zCell[c2] - m->zCell[c1];
      const double de = -nx * m->yEdge[oe] + ny * m->xEdge[oe];
      kapE[ne] = de > 0.0 ? 1 : (de < 0.0 ? -1 : (nz >= 0.0 ? 1 : -1));
    }
    voe[ne] = md.iV[m->verticesOnEdge[2 * (size_t)oe] - 1];
    voe[(size_t)nE + ne] = md.iV[m->verticesOnEdge[2 * (size_t)oe + 1] - 1];
    dv[ne] = m->dvEdge[oe] * a;
    dc[ne] = m->dcEdge[oe] * a;
    aE[ne] = (m->dvEdge[oe] * m->dcEdge[oe]) * a2;  // mesh_mod.F90:178,221
    const int n = m->nEdgesOnEdge[oe];
    nEoE[ne] = (int8_t)n;
    for (int s = 0; s < kMaxEoE; ++s) {
      // padding slots point at the edge itself with weight 0 (never read)
      eoe[pair_at(s, ne, nE)] = s < n ? md.iE[m->edgesOnEdge[(size_t)oe * ME2 + s] - 1] : ne;
      woe[pair_at(s, ne, nE)] = s < n ? m->weightsOnEdge[(size_t)oe * ME2 + s] : 0.0;
    }
    if (need_fit)
      for (int k = 0; k < 2; ++k) {
        // slots past the vertex's stencil are zeroed whatever the caller's
        // array holds there (the reference leaves them unset)
        const int nfit = f->nFit2Vertices[m->verticesOnEdge[2 * (size_t)oe + k] - 1];
        for (int j = 0; j < kFit2; ++j) {
          const int slot = j + 1;
          d2v[pair_at(k * kFit2 + j, ne, nE)] =
              slot < nfit ? f->deriv2OnVertex[((size_t)oe * 2 + k) * f->maxFit + slot] : 0.0;
        }
      }
  }
  // ---- vertices
  std::vector<int> eovS((size_t)3 * nV), cov((size_t)3 * nV), fit2V((size_t)kFit2 * nV, 0);
  std::vector<int8_t> nFit2V(nV, 1);
  std::vector<double> kite((size_t)3 * nV), areaTri(nV), fV(nV);
  double sat = 0.0;
  for (int v = 0; v < nV; ++v) sat += m->areaTriangle[v] * a2;
  md.sumAreaTri = sat;
  for (int nv = 0; nv < nV; ++nv) {
    const int ov = md.pV[nv];
    areaTri[nv] = m->areaTriangle[ov] * a2;
    fV[nv] = 2.0 * 7.292e-5 * std::sin(m->latVertex[ov]);  // mesh_mod.F90:180, const_mod.F90:13
    for (int j = 0; j < 3; ++j) {
      const int oe = m->edgesOnVertex[(size_t)ov * 3 + j] - 1;
      const bool plus = m->verticesOnEdge[2 * (size_t)oe] - 1 == ov;  // tSignEdge = +1
      const int ne = md.iE[oe];
      eovS[(size_t)j * nV + nv] = plus ? ne : ~ne;
      cov[(size_t)j * nV + nv] = md.iC[m->cellsOnVertex[(size_t)ov * 3 + j] - 1];
      kite[(size_t)j * nV + nv] = m->kiteAreasOnVertex[(size_t)ov * 3 + j] * a2;
    }
    if (need_fit) {
      const int n = f->nFit2Vertices[ov];
      nFit2V[nv] = (int8_t)n;
      for (int j = 0; j < kFit2; ++j)
        fit2V[(size_t)nv * kFit2 + j] = j + 1 < n ? md.iV[f->fit2Vertices[(size_t)ov * f->maxFit + j + 1] - 1] : nv;
    }
  }

  // ---- cell-major edge rows.  In the MPAS order, edgesOnEdge(:, e) lists
  // the edges of cellsOnEdge(1, e) that follow e counter-clockwise, then those
  // of cellsOnEdge(2, e), so the edge kernels can derive it from the two
  // cells' rows instead of streaming the table.  Checked for every edge; a
  // mesh whose rows differ keeps the table (md.cell_rows stays false).
  std::vector<int> eocR;
  {
    bool ok = true;
    for (int ne = 0; ne < nE && ok; ++ne) {
      int der[kMaxEoE], nd = 0;
      for (int side = 0; side < 2 && ok; ++side) {
        const int c = coe[(size_t)side * nE + ne], n = nEoC[c];
        int k = -1;
        for (int q = n - 1; q >= 0; --q)  // the first slot holding e, as the kernels search
          if (unsign_host(eocS[(size_t)q * nC + c]) == ne) k = q;
        ok = k >= 0 && nd + n - 1 <= kMaxEoE;
        for (int j = 1; j < n && ok; ++j) der[nd++] = unsign_host(eocS[(size_t)((k + j) % n) * nC + c]);
      }
      for (int q = 0; q < kMaxEoE && ok; ++q) ok = eoe[pair_at(q, ne, nE)] == (q < nd ? der[q] : ne);
    }
    md.cell_rows = ok;
    if (ok) {
      eocR.assign((size_t)8 * nC, 0);
      for (int c = 0; c < nC; ++c) {
        for (int q = 0; q < kMaxEoC; ++q) eocR[(size_t)c * 8 + q] = q < nEoC[c] ? eocS[(size_t)q * nC + c] : 0;
        eocR[(size_t)c * 8 + 6] = nEoC[c];
      }
    }
  }

  // ---- LDS tiles of the tendency kernel: tile t = edges [kTileE t,
  // kTileE (t+1)) of the Hilbert order; its halo is the set of other edges
  // its edgesOnEdge slots name, ascending.  eoeL gives every slot as a local
  // row: the tile's edges at 0..kTileE-1, the halo's after them.  A mesh with
  // a halo beyond kTendHaloMax keeps the table (md.tend_lds stays false).
  std::vector<uint32_t> eoeL;
  std::vector<int> haloE, haloOff;
  {
    static_assert(kTileE + kTendHaloMax <= 65536, "local rows fit 16 bits");
    const int nb = round_blocks(nE);
    haloOff.assign((size_t)nb + 1, 0);
    eoeL.assign((size_t)(kMaxEoE / 2) * nE, 0u);
    std::vector<int> halo;
    bool ok = true;
    for (int t = 0; t < nb && ok; ++t) {
      const int e0 = std::min(nE, t * kTileE), e1 = std::min(nE, e0 + kTileE);
      halo.clear();
      for (int ne = e0; ne < e1; ++ne)
        for (int q = 0; q < kMaxEoE; ++q) {
          const int g = eoe[pair_at(q, ne, nE)];
          if (g < e0 || g >= e1) halo.push_back(g);
        }
      std::sort(halo.begin(), halo.end());
      halo.erase(std::unique(halo.begin(), halo.end()), halo.end());
      ok = (int)halo.size() <= kTendHaloMax;
      for (int ne = e0; ne < e1 && ok; ++ne)
        for (int q = 0; q < kMaxEoE; ++q) {
          const int g = eoe[pair_at(q, ne, nE)];
          const int l = (g >= e0 && g < e1)
                            ? g - e0
                            : kTileE + (int)(std::lower_bound(halo.begin(), halo.end(), g) - halo.begin());
          eoeL[(size_t)(q / 2) * nE + ne] |= (uint32_t)l << (16 * (q % 2));
        }
      haloE.insert(haloE.end(), halo.begin(), halo.end());
      haloOff[(size_t)t + 1] = (int)haloE.size();
    }
    md.tend_lds = ok;
    md.haloTotal = ok ? (double)haloE.size() : 0.0;
  }

  DMesh &d = md.dm;
  d.nC = nC;
  d.nE = nE;
  d.nV = nV;
  d.nbC = round_blocks(nC);
  d.nbE = round_blocks(nE);
  d.nbV = round_blocks(nV);
#define UP(field, vec, T)                                  \
  do {                                                     \
    void *p_ = nullptr;                                    \
    if (hipMalloc(&p_, std::max<size_t>(vec.size(), 1) * sizeof(T)) != hipSuccess) { \
      err = "hipMalloc " #vec;                             \
      return TM_ENOMEM;                                    \
    }                                                      \
    allocs.push_back(p_);                                  \
    if (hipMemcpyAsync(p_, vec.data(), vec.size() * sizeof(T), hipMemcpyHostToDevice, stream) != hipSuccess || \
        hipStreamSynchronize(stream) != hipSuccess) {       \
      err = "upload " #vec;                                \
      return TM_EDEVICE;                                   \
    }                                                      \
    d.field = (T *)p_;                                     \
  } while (0)
  UP(nEoC, nEoC, int8_t);
  UP(eocS, eocS, int);
  UP(areaCell, areaCell, double);
  UP(coe, coe, int);
  UP(voe, voe, int);
  UP(dv, dv, double);
  UP(dc, dc, double);
  UP(aE, aE, double);
  UP(nEoE, nEoE, int8_t);
  UP(kapE, kapE, int8_t);
  UP(eoe, eoe, int);
  UP(woe, woe, double);
  UP(d2v, d2v, double);
  if (o.precision == 32) {
    // FP32 copies slot-major ([slot][n]): the FP64 tables are slot pairs
    std::vector<float> d2v32(d2v.size()), woe32(woe.size());
    for (int ne = 0; ne < nE; ++ne) {
      for (int s = 0; s < kMaxEoE; ++s) woe32[(size_t)s * nE + ne] = (float)woe[pair_at(s, ne, nE)];
      for (int s = 0; s < 2 * kFit2; ++s) d2v32[(size_t)s * nE + ne] = (float)d2v[pair_at(s, ne, nE)];
    }
    std::vector<float> dv32(dv.begin(), dv.end()), dc32(dc.begin(), dc.end()), aE32(aE.begin(), aE.end());
    UP(d2v32, d2v32, float);
    UP(woe32, woe32, float);
    UP(dv32, dv32, float);
    UP(dc32, dc32, float);
    UP(aE32, aE32, float);
  } else {
    d.d2v32 = d.woe32 = d.dv32 = d.dc32 = d.aE32 = nullptr;
  }
  // cell fits of the upwind correction, new numbering, padded to the device slots
  for (int order = 2; order <= 4; order += 2) {
    if (o.adv_order < order + 1) break;
    const int ns = order == 2 ? kFit2C : kFit4C;
    const int *nf = order == 2 ? f->nFit2Cells : f->nFit4Cells, *fc = order == 2 ? f->fit2Cells : f->fit4Cells;
    const double *dw = order == 2 ? f->deriv2OnCell : f->deriv4OnCell;
    const int rw = fit_row(ns);
    std::vector<int> fitc((size_t)rw * nC);
    std::vector<double> dc((size_t)2 * ns * nE, 0.0);
    for (int ncell = 0; ncell < nC; ++ncell) {
      const int oc = md.pC[ncell], n = nf[oc];
      for (int j = 0; j < rw; ++j)
        fitc[(size_t)ncell * rw + j] = j < ns && j + 1 < n ? md.iC[fc[(size_t)oc * f->maxFit + j + 1] - 1] : ncell;
    }
    for (int ne = 0; ne < nE; ++ne) {
      const int oe = md.pE[ne];
      for (int k = 0; k < 2; ++k) {
        const int n = nf[m->cellsOnEdge[2 * (size_t)oe + k] - 1];
        for (int j = 0; j < ns; ++j)
          dc[((size_t)k * ns + j) * nE + ne] = j + 1 < n ? dw[((size_t)oe * 2 + k) * f->maxFit + j + 1] : 0.0;
      }
    }
    std::vector<float> w32;
    if (o.precision == 32) w32.assign(dc.begin(), dc.end());
    if (order == 2) {
      UP(fit2C, fitc, int);
      UP(d2c, dc, double);
      if (o.precision == 32) UP(d2c32, w32, float);
    } else {
      UP(fit4C, fitc, int);
      UP(d4c, dc, double);
      if (o.precision == 32) UP(d4c32, w32, float);
    }
  }
  // vertex fourth-order fits of the scheme-3 PV (interp_pv_order 5, 6)
  if (o.pv_scheme == 3 && o.interp_pv_order >= 5) {
    const int rw = fit_row(kFit4V);
    std::vector<int> fitv((size_t)rw * nV);
    std::vector<double> dv4((size_t)2 * kFit4V * nE, 0.0);
    for (int nv = 0; nv < nV; ++nv) {
      const int ov = md.pV[nv], n = f->nFit4Vertices[ov];
      for (int j = 0; j < rw; ++j)
        fitv[(size_t)nv * rw + j] =
            j < kFit4V && j + 1 < n ? md.iV[f->fit4Vertices[(size_t)ov * f->maxFit + j + 1] - 1] : nv;
    }
    for (int ne = 0; ne < nE; ++ne) {
      const int oe = md.pE[ne];
      for (int k = 0; k < 2; ++k) {
        const int n = f->nFit4Vertices[m->verticesOnEdge[2 * (size_t)oe + k] - 1];
        for (int j = 0; j < kFit4V; ++j)
          dv4[((size_t)k * kFit4V + j) * nE + ne] =
              j + 1 < n ? f->deriv4OnVertex[((size_t)oe * 2 + k) * f->maxFit + j + 1] : 0.0;
      }
    }
    UP(fit4V, fitv, int);
    UP(d4v, dv4, double);
    if (o.precision == 32) {
      std::vector<float> w32(dv4.begin(), dv4.end());
      UP(d4v32, w32, float);
    }
  }
  UP(eovS, eovS, int);
  UP(cov, cov, int);
  UP(kite, kite, double);
  UP(areaTri, areaTri, double);
  UP(fV, fV, double);
  UP(nFit2V, nFit2V, int8_t);
  UP(fit2V, fit2V, int);
  UP(vocS, vocS, int);
  UP(cocN, cocN, int);
  UP(kiteC, kiteC, double);
  if (md.cell_rows)
    UP(eocR, eocR, int);
  else
    d.eocR = nullptr;
  if (md.tend_lds) {
    UP(eoeL, eoeL, uint32_t);
    UP(haloE, haloE, int);
    UP(haloOff, haloOff, int);
  } else {
    d.eoeL = nullptr;
    d.haloE = d.haloOff = nullptr;
  }
#undef UP

  md.sumEoE = md.sumFitE = md.sumFitV = md.sumEoC = 0.0;
  for (int e = 0; e < nE; ++e) md.sumEoE += m->nEdgesOnEdge[e];
  for (int i = 0; i < nC; ++i) md.sumEoC += m->nEdgesOnCell[i];
  if (need_fit) {
    for (int e = 0; e < nE; ++e)
      md.sumFitE += (f->nFit2Vertices[m->verticesOnEdge[2 * (size_t)e] - 1] - 1) +
                    (f->nFit2Vertices[m->verticesOnEdge[2 * (size_t)e + 1] - 1] - 1);
    for (int v = 0; v < nV; ++v) md.sumFitV += f->nFit2Vertices[v] - 1;
  }
  return TM_OK;
}

}  // namespace tmk

static int create_impl(tm_ctx *c, const tm_mesh_desc *m, const tm_fit_desc *f) {
  const tm_opts &o = c->o;
  {
    std::string e;
    int rc = validate_mesh(m, f, o, e);
    if (rc) return c->fail(rc, e);
  }
  const int nC = m->nCells, nE = m->nEdges, nV = m->nVertices;
  c->nC = nC;
  c->nE = nE;
  c->maxEdges = m->maxEdges;
  c->nV = nV;
  c->nT = o.ntracers;
  c->rsz = o.precision == 64 ? 8 : 4;
  HIPCHK(c, hipSetDevice(o.device));
  {
    std::string e;
    if (!(c->noalloc_scratch = noalloc_scratch_for(o.device, e))) return c->fail(TM_EDEVICE, e);
  }
  HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  HIPCHK(c, hipEventCreate(&c->ev0));
  HIPCHK(c, hipEventCreate(&c->ev1));
  {
    std::string e;
    int rc = build_mesh_dev(c->md, m, f, o, c->stream, c->allocs, e);
    if (rc) return c->fail(rc, e);
  }
  c->dm = c->md.dm;
  c->pC = c->md.pC;
  c->pE = c->md.pE;
  c->pV = c->md.pV;
  c->totalArea = c->md.totalArea;
  c->sumAreaTri = c->md.sumAreaTri;
  const DMesh &d = c->dm;
  const size_t nq = (size_t)c->nT * nC;
  void **rbufs_c[] = {&c->gd_old, &c->ghs, &c->gdS[0], &c->gdS[1], &c->acc_gd, &c->t_gd[0], &c->t_gd[1],
                      &c->E, &c->tend_gd, &c->scratch_c[0], &c->scratch_c[1]};
  for (void **p : rbufs_c)
    if (!(*p = c->alloc_r(nC))) return c->fail(TM_ENOMEM, "hipMalloc cell state");
  void **rbufs_e[] = {&c->iap_old, &c->iap_s, &c->acc_iap, &c->t_iap[0], &c->t_iap[1], &c->gd_e, &c->u,
                      &c->udc, &c->pv_e, &c->scratch_e[0], &c->scratch_e[1], &c->scratch_e[2], &c->scratch_e[3]};
  for (void **p : rbufs_e)
    if (!(*p = c->alloc_r(nE))) return c->fail(TM_ENOMEM, "hipMalloc edge state");
  if (!(c->fk = c->alloc_r(2 * (size_t)nE))) return c->fail(TM_ENOMEM, "hipMalloc edge state");
  if (!(c->gd_e2 = c->alloc_r(nE)) || !(c->u2 = c->alloc_r(nE))) return c->fail(TM_ENOMEM, "hipMalloc edge state");
  void **rbufs_v[] = {&c->pv_v, &c->scratch_v[0], &c->scratch_v[1]};
  for (void **p : rbufs_v)
    if (!(*p = c->alloc_r(nV))) return c->fail(TM_ENOMEM, "hipMalloc vertex state");
  void **rbufs_q[] = {&c->q_old, &c->qS[0], &c->qS[1], &c->acc_q};
  for (void **p : rbufs_q)
    if (!(*p = c->alloc_r(std::max<size_t>(nq, 1)))) return c->fail(TM_ENOMEM, "hipMalloc tracer state");
  if (o.adv_order >= 3 && c->nT > 0 && !(c->qe = c->alloc_r((size_t)c->nT * nE)))
    return c->fail(TM_ENOMEM, "hipMalloc tracer edge values");
  c->xmask = (o.adv_order >= 3 ? kXAdv : 0) | (o.adv_order >= 5 ? kXAdv56 : 0) |
             (o.pv_scheme != 3 || o.interp_pv_order >= 5 || o.time_scheme != 2 ? kXExt : 0);
  // edgesOnEdge derived from the cells' edge rows where the mesh keeps the
  // MPAS order (not in the upwind-correction variants)
  if (c->md.cell_rows && o.adv_order < 3) c->xmask |= cell_rows_default(c->rsz);
  // next-stage edge values from the tendency kernel, and the tendency and PV
  // kernels' edgesOnEdge gathers from LDS tiles (plain CRK4, no upwind
  // correction); the tiles replace the cell-row derivation
  if ((c->xmask & (kXAdv | kXExt)) == 0) c->xmask |= kXTendPrep;
  if (c->md.tend_lds && (c->xmask & (kXAdv | kXExt)) == 0)
    c->xmask = (c->xmask & ~(kXEoeTend | kXEoePv)) | kXTendLds | kXPvLds;
  c->nstages = o.time_scheme == 2 ? 4 : 3;
  if (o.pv_scheme != 3 && !(c->pv_c = c->alloc_r(nC))) return c->fail(TM_ENOMEM, "hipMalloc pv_cell");
  if (o.pv_scheme == 2 &&
      (!(c->tgd_v = c->alloc_r(nV)) || !(c->dpv = c->alloc_r(nE)) || !(c->pvf = c->alloc_r(nE)) ||
       !(c->dpvf = c->alloc_r(nE)) || !(c->cv_part = c->alloc<double2>((size_t)3 * d.nbV)) ||
       !(c->cv_eps = c->alloc<double>(1))))
    return c->fail(TM_ENOMEM, "hipMalloc conservative APVM work arrays");
  c->part_stride = std::max(d.nbE, std::max(d.nbC, d.nbV));
  if (!(c->part = c->alloc<double2>((size_t)kNumSums * c->part_stride))) return c->fail(TM_ENOMEM, "hipMalloc part");
  HIPCHK(c, hipMemsetAsync(c->part, 0, sizeof(double2) * kNumSums * c->part_stride, c->stream));
  if (!(c->sums = c->alloc<double>(kNumSums)) || !(c->hist_idx = c->alloc<int>(1)) ||
      !(c->beta_hist = c->alloc<double>(c->hist_cap)) || !(c->inv_hist = c->alloc<double>(4 * (size_t)c->hist_cap)))
    return c->fail(TM_ENOMEM, "hipMalloc history");
  HIPCHK(c, hipStreamSynchronize(c->stream));

  // Algorithmic bytes (compulsory traffic, actual stencil nnz): constants of
  // one stage + 56 B (FP64) per prognostic element, finalize 24 B per element.
  const double sumEoE = c->md.sumEoE, sumFitE = c->md.sumFitE, sumFitV = c->md.sumFitV, sumEoC = c->md.sumEoC;
  // bytes of an edge weight or length read by the stencils (deriv2OnVertex,
  // weightsOnEdge, dvEdge, dcEdge): their FP32 copies in FP32 contexts
  const double fsz = c->rsz == 4 ? 4.0 : 8.0;
  const double cst = nE * (8.0 + 8.0 + 2.0 * fsz) + sumEoE * (4.0 + fsz) + sumFitE * fsz +
                     nV * (12.0 + 12.0 + 24.0 + 8.0 + 8.0) + sumFitV * 4.0 + sumEoC * 4.0 + nC * (8.0 + 8.0);
  const double rs = (double)c->rsz;
  const double prog = (double)nE + (double)nC * (1 + c->nT);
  c->step_bytes[0] = cst + 7.0 * rs * prog;
  c->step_bytes[1] = 3.0 * rs * prog;
  c->step_bytes[2] = 4.0 * c->step_bytes[0] + c->step_bytes[1];
  return TM_OK;
}

// Compulsory bytes of one launch of each kernel (stage 2/3 pattern; the
// tendency kernel's averaged over the step) in the context's current kernel
// layout: each array the kernel touches counted once, gathers at the size of
// the gathered array.
static void kernel_bytes(const tm_ctx *c, double *kb) {
  const double nC = c->nC, nE = c->nE, nV = c->nV, T = c->nT;
  const double sumEoE = c->md.sumEoE, sumFitE = c->md.sumFitE, sumFitV = c->md.sumFitV, sumEoC = c->md.sumEoC;
  const double fsz = c->rsz == 4 ? 4.0 : 8.0, rs = (double)c->rsz;
  const double prog = nE + nC * (1 + T);
  // edgesOnEdge: the table, or (cellsOnEdge plus) the cells' 32-byte edge rows
  const double tiles = nE * 20.0 + 4.0 * c->md.haloTotal + 4.0 * (c->dm.nbE + 1);  // 16-bit local rows + halo lists
  const double eoe_pv = (c->xmask & kXPvLds) ? tiles : (c->xmask & kXEoePv) ? nE * 8.0 + nC * 32.0 : 4.0 * sumEoE;
  // the tendency kernel's edgesOnEdge: the table, the cells' rows, or the LDS
  // tile's local rows (16-bit) with the halo lists
  const double eoe_tend = (c->xmask & kXTendLds) ? tiles : (c->xmask & kXEoeTend) ? nC * 32.0 : 4.0 * sumEoE;
  kb[TM_K_EDGE_PREP] = nE * (8.0 + rs + 24.0 + 5.0 * rs) + nC * rs;
  // cells, vertices, the edge products gathered (fk, u dcEdge) and the
  // tracers' cellsOnCell, u and dvEdge
  kb[TM_K_CELL_VERTEX] = nC * (1.0 + 8.0 + 10.0 * rs) + 4.0 * sumEoC + T * nC * 5.0 * rs +
                         nV * (12.0 + 12.0 + 24.0 + 8.0 + 8.0 + rs) + nE * rs * 3.0 +
                         (T > 0 ? 4.0 * sumEoC + nE * (rs + 8.0) : 0.0);
  kb[TM_K_EDGE_PV] = nE * (1.0 + 8.0 + fsz + 2.0 * rs) + eoe_pv + fsz * sumEoE + fsz * sumFitE + nV * (rs + 1.0) +
                     4.0 * sumFitV;
  kb[TM_K_EDGE_TEND] = nE * (1.0 + 8.0 + fsz + 9.0 * rs) + eoe_tend + fsz * sumEoE + nC * 2.0 * rs;
  // stages 1-3 of 4 also gather gd_next, read dvEdge, and write gd_edge, u,
  // fk and u dcEdge: averaged over the step's four launches, as the timings are
  if (c->xmask & kXTendPrep) kb[TM_K_EDGE_TEND] += 0.75 * (nC * rs + nE * (fsz + 5.0 * rs));
  kb[TM_K_REDUCE] = (double)kNumSums * c->part_stride * 16.0;
  kb[TM_K_FINALIZE] = 3.0 * rs * prog;
}

extern "C" int tm_ctx_create(tm_ctx **out, const tm_mesh_desc *m, const tm_fit_desc *f, const tm_opts *o) {
  if (!out || !m || !o) return TM_EINVAL;
  *out = nullptr;
  tm_ctx *c = new tm_ctx();
  c->o = *o;
  int rc = create_impl(c, m, f);
  if (rc) {
    std::fprintf(stderr, "tm_ctx_create: %s\n", c->err.c_str());
    tm_ctx_destroy(c);
    return rc;
  }
  *out = c;
  return TM_OK;
}

extern "C" int tm_ctx_set_layout(tm_ctx *c, int mode) {
  if (!c) return TM_EINVAL;
  if (mode < 0 || mode > 31) return c->fail(TM_EINVAL, "layout mode must be 0..31");
  int want = ((mode & 1) ? kXEoePv : 0) | ((mode & 2) ? kXEoeTend : 0) | ((mode & 4) ? kXTendPrep : 0) |
             ((mode & 8) ? kXTendLds : 0) | ((mode & 16) ? kXPvLds : 0);
  if (want & kXTendLds) want &= ~kXEoeTend;  // the LDS tiles supersede the cell rows
  if (want & kXPvLds) want &= ~kXEoePv;
  if ((want & (kXTendLds | kXPvLds)) && !c->md.tend_lds)
    return c->fail(TM_EINVAL, "a tile's edgesOnEdge halo exceeds the LDS tile");
  if ((want & (kXEoePv | kXEoeTend)) && !c->md.cell_rows)
    return c->fail(TM_EINVAL, "the mesh's edgesOnEdge rows do not follow the MPAS order");
  if (want && (c->xmask & kXAdv)) return c->fail(TM_EINVAL, "the layout modes do not run with adv_order >= 3");
  if ((want & (kXTendPrep | kXTendLds | kXPvLds)) && (c->xmask & kXExt))
    return c->fail(TM_EINVAL, "edge values from the tendency kernel and its LDS tile need CRK4 with pv_scheme 3 order <= 4");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  drop_graphs(c);  // captured steps hold the previous variant
  c->xmask = (c->xmask & ~(kXEoePv | kXEoeTend | kXTendPrep | kXTendLds | kXPvLds)) | want;
  return TM_OK;
}

extern "C" int tm_ctx_get_layout(tm_ctx *c, int *mode) {
  if (!c || !mode) return TM_EINVAL;
  *mode = ((c->xmask & kXEoePv) ? 1 : 0) | ((c->xmask & kXEoeTend) ? 2 : 0) | ((c->xmask & kXTendPrep) ? 4 : 0) |
          ((c->xmask & kXTendLds) ? 8 : 0) | ((c->xmask & kXPvLds) ? 16 : 0);
  return TM_OK;
}

extern "C" int tm_state_upload(tm_ctx *c, const double *gd, const double *u, const double *iap_u,
                               const double *ghs, const double *tracers) {
  if (!c) return TM_EINVAL;
  if (!gd || !ghs || (!u && !iap_u) || (c->nT > 0 && !tracers)) return c->fail(TM_EINVAL, "missing state array");
  if (!finite_all(gd, c->nC) || !finite_all(ghs, c->nC) || (iap_u && !finite_all(iap_u, c->nE)) ||
      (!iap_u && !finite_all(u, c->nE)) || (c->nT > 0 && !finite_all(tracers, (size_t)c->nT * c->nC)))
    return c->fail(TM_ENONFINITE, "non-finite value in the uploaded state");
  for (int i = 0; i < c->nC; ++i)
    if (!(gd[i] > 0.0)) return c->fail(TM_EINVAL, "gd must be positive (sqrt of gd_edge)");
  int rc;
  if ((rc = upload_state(c, c->gd_old, gd, c->pC, 1, c->nC))) return rc;
  if ((rc = upload_state(c, c->ghs, ghs, c->pC, 1, c->nC))) return rc;
  if (c->nT > 0 && (rc = upload_state(c, c->q_old, tracers, c->pC, c->nT, c->nC))) return rc;
  if (iap_u) {
    if ((rc = upload_state(c, c->iap_old, iap_u, c->pE, 1, c->nE))) return rc;
  } else {
    if ((rc = upload_state(c, c->scratch_e[0], u, c->pE, 1, c->nE))) return rc;
    const int nb = (c->nE + kBlock - 1) / kBlock;
    if (c->rsz == 8)
      hipLaunchKernelGGL(k_iap_from_u<double>, dim3(nb), dim3(kBlock), 0, c->stream, c->dm,
                         stage_bufs<double>(c, 0, false), (const double *)c->gd_old, (const double *)c->scratch_e[0],
                         (double *)c->iap_old);
    else
      hipLaunchKernelGGL(k_iap_from_u<float>, dim3(nb), dim3(kBlock), 0, c->stream, c->dm,
                         stage_bufs<float>(c, 0, false), (const float *)c->gd_old, (const float *)c->scratch_e[0],
                         (float *)c->iap_old);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  c->have_state = true;
  c->cl_step = 0;  // a held closure correction belongs to the previous state
  return TM_OK;
}

extern "C" int tm_state_download(tm_ctx *c, double *gd, double *u, double *iap_u, double *tracers) {
  if (!c) return TM_EINVAL;
  if (!c->have_state) return c->fail(TM_ESTATE, "no state uploaded");
  int rc;
  if (gd && (rc = download_state(c, gd, c->gd_old, c->pC, 1, c->nC))) return rc;
  if (iap_u && (rc = download_state(c, iap_u, c->iap_old, c->pE, 1, c->nE))) return rc;
  if (tracers && c->nT > 0 && (rc = download_state(c, tracers, c->q_old, c->pC, c->nT, c->nC))) return rc;
  if (u) {
    rc = c->rsz == 8 ? launch_stage<double>(c, stage_bufs<double>(c, 0, false), false)
                     : launch_stage<float>(c, stage_bufs<float>(c, 0, false), false);
    if (rc) return rc;
    if ((rc = download_state(c, u, c->u, c->pE, 1, c->nE))) return rc;
  }
  return TM_OK;
}

static int step_impl(tm_ctx *c, int nsteps, double *inv_out, double *beta_out) {
  if (!c) return TM_EINVAL;
  if (!c->have_state) return c->fail(TM_ESTATE, "tm_step before tm_state_upload");
  if (nsteps < 0) return c->fail(TM_EINVAL, "nsteps < 0");
  int done = 0;
  double ms_total = 0.0;
  while (done < nsteps) {
    const int n = std::min(nsteps - done, c->hist_cap);
    int rc = run_steps(c, n, true);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    ms_total += ms;
    std::vector<double> bh(n), ih(4 * (size_t)n);
    HIPCHK(c, hipMemcpy(bh.data(), c->beta_hist, n * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(ih.data(), c->inv_hist, 4 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
    if (beta_out) std::memcpy(beta_out + done, bh.data(), n * sizeof(double));
    if (inv_out) std::memcpy(inv_out + 4 * (size_t)done, ih.data(), 4 * (size_t)n * sizeof(double));
    if (!finite_all(ih.data(), ih.size()) || !finite_all(bh.data(), bh.size()))
      return c->fail(TM_ENONFINITE, "non-finite invariant or beta: the state blew up");
    done += n;
  }
  c->err.clear();
  (void)ms_total;
  return TM_OK;
}

extern "C" int tm_step(tm_ctx *c, int nsteps, double *invariants_out, double *beta_out) {
  return step_impl(c, nsteps, invariants_out, beta_out);
}

extern "C" int tm_step_async(tm_ctx *c, int nsteps) {
  if (!c) return TM_EINVAL;
  if (!c->have_state) return c->fail(TM_ESTATE, "tm_step_async before tm_state_upload");
  if (nsteps < 0 || nsteps > c->hist_cap) return c->fail(TM_EINVAL, "nsteps out of range for one batch");
  return run_steps(c, nsteps, false);
}

extern "C" int tm_sync(tm_ctx *c) {
  if (!c) return TM_EINVAL;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TM_OK;
}

extern "C" int tm_last_elapsed_ms(tm_ctx *c, double *ms) {
  if (!c || !ms) return TM_EINVAL;
  if (!c->timed) return c->fail(TM_ESTATE, "nothing timed yet");
  HIPCHK(c, hipEventSynchronize(c->ev1));
  float f = 0.f;
  HIPCHK(c, hipEventElapsedTime(&f, c->ev0, c->ev1));
  *ms = f;
  return TM_OK;
}

extern "C" int tm_step_bytes(tm_ctx *c, double bytes[3]) {
  if (!c || !bytes) return TM_EINVAL;
  for (int i = 0; i < 3; ++i) bytes[i] = c->step_bytes[i];
  return TM_OK;
}

extern "C" int tm_kernel_bytes(tm_ctx *c, double bytes[TM_K_COUNT]) {
  if (!c || !bytes) return TM_EINVAL;
  kernel_bytes(c, bytes);
  return TM_OK;
}

// Eager launches of one step with an event pair round every kernel.
template <typename R>
static int profile_one_step(tm_ctx *c, std::vector<hipEvent_t> &ev, double *ms, int *launches) {
  struct Rec {
    int kind, e0;
  };
  std::vector<Rec> recs;
  int next = 0;
  auto mark = [&](int kind) -> int {
    recs.push_back({kind, next});
    HIPCHK(c, hipEventRecord(ev[next++], c->stream));
    return TM_OK;
  };
  auto done = [&]() -> int {
    HIPCHK(c, hipEventRecord(ev[next++], c->stream));
    return TM_OK;
  };
  const DMesh &m = c->dm;
  int rc = TM_OK;
  for (int k = 1; k <= c->nstages; ++k) {
    StageBufs<R> b = stage_bufs<R>(c, k, k == 1);
    const int ids[4] = {TM_K_EDGE_PREP, TM_K_CELL_VERTEX, TM_K_EDGE_PV, TM_K_EDGE_TEND};
    for (int which = (c->xmask & kXTendPrep) && k >= 2 ? 1 : 0; which < 4; ++which) {
      if ((rc = mark(ids[which]))) return rc;
      b.rev = (which & 1) ? c->serpentine : 0;
      launch_part<R>(c, which, b);
      if ((rc = done())) return rc;
    }
  }
  if ((rc = mark(TM_K_REDUCE))) return rc;
  hipLaunchKernelGGL(k_reduce, dim3(kNumSums), dim3(kReduceBlock), 0, c->stream, reduce_args(c, 1, true));
  if ((rc = done())) return rc;
  if ((rc = mark(TM_K_FINALIZE))) return rc;
  hipLaunchKernelGGL(k_finalize<R>, dim3(m.nbE + m.nbC), dim3(kBlock), 0, c->stream, m, final_bufs<R>(c));
  if ((rc = done())) return rc;
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (const Rec &r : recs) {
    float t = 0.f;
    HIPCHK(c, hipEventElapsedTime(&t, ev[r.e0], ev[r.e0 + 1]));
    ms[r.kind] += t;
    launches[r.kind] += 1;
  }
  return TM_OK;
}

extern "C" int tm_profile_steps(tm_ctx *c, int nsteps, double ms[TM_K_COUNT], int launches[TM_K_COUNT]) {
  if (!c || !ms || !launches) return TM_EINVAL;
  if (!c->have_state) return c->fail(TM_ESTATE, "tm_profile_steps before tm_state_upload");
  for (int i = 0; i < TM_K_COUNT; ++i) {
    ms[i] = 0.0;
    launches[i] = 0;
  }
  std::vector<hipEvent_t> ev(40, nullptr);
  for (auto &e : ev) HIPCHK(c, hipEventCreate(&e));
  hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, c->stream, c->hist_idx);
  int rc = TM_OK;
  for (int s = 0; s < nsteps && rc == TM_OK; ++s)
    rc = c->rsz == 8 ? profile_one_step<double>(c, ev, ms, launches) : profile_one_step<float>(c, ev, ms, launches);
  for (auto &e : ev) (void)hipEventDestroy(e);
  return rc;
}

extern "C" int tm_invariants(tm_ctx *c, double out[4]) {
  if (!c || !out) return TM_EINVAL;
  if (!c->have_state) return c->fail(TM_ESTATE, "no state uploaded");
  // counter 1 so that the invariants-only pass writes slot 0
  int one = 1;
  HIPCHK(c, hipMemcpyAsync(c->hist_idx, &one, sizeof(int), hipMemcpyHostToDevice, c->stream));
  int rc = c->rsz == 8 ? enqueue_refresh<double>(c) : enqueue_refresh<float>(c);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(out, c->inv_hist, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TM_OK;
}

// The four invariants from the sums into inv[4].
static __global__ void k_inv_write(StepOut o) {
  if (threadIdx.x == 0 && blockIdx.x == 0) write_invariants(o, 0);
}

extern "C" int tm_invariants_noalloc(tm_ctx *c, double out[4]) {
  if (!c || !out) return TM_EINVAL;
  if (!c->have_state) return c->fail(TM_ESTATE, "no state uploaded");
  if (!c->noalloc_scratch) return c->fail(TM_ESTATE, "no device scratch for this context");
  if (c->part_stride > kNoAllocMaxBlocks)
    return c->fail(TM_EINVAL, "tm_invariants_noalloc: the mesh needs " + std::to_string(c->part_stride) +
                                  " workgroups per element kind, the static scratch holds " +
                                  std::to_string(kNoAllocMaxBlocks) +
                                  " (at most 4,194,304 cells, edges and vertices); use tm_invariants");
  unsigned char *base = static_cast<unsigned char *>(c->noalloc_scratch);
  double2 *part = reinterpret_cast<double2 *>(base);
  double *sums = reinterpret_cast<double *>(base + kNoAllocPartBytes);
  double *inv = sums + kNumSums;
  // every partial the reduction reads is written by this call's refresh
  // pass, so nothing needs clearing
  int rc = c->rsz == 8 ? enqueue_refresh_sums<double>(c, part, sums) : enqueue_refresh_sums<float>(c, part, sums);
  if (rc) return rc;
  StepOut o = step_out(c, true);
  o.sums = sums;
  o.inv_hist = inv;
  o.hist_cap = 1;
  hipLaunchKernelGGL(k_inv_write, dim3(1), dim3(64), 0, c->stream, o);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(out, inv, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TM_OK;
}

template <typename R>
static int stage_impl(tm_ctx *c, tm_stage_out *out) {
  StageBufs<R> b = stage_bufs<R>(c, 1, false);
  // diagnostics into scratch / dedicated buffers
  b.d_ke = (R *)c->scratch_c[0];
  b.d_vor = (R *)c->scratch_v[0];
  b.d_gdv = (R *)c->scratch_v[1];
  b.d_v = (R *)c->scratch_e[0];
  b.d_pvflx = (R *)c->scratch_e[1];
  b.d_utend = (R *)c->scratch_e[2];
  b.d_tiap = (R *)c->scratch_e[3];
  void *tq = nullptr;
  if (c->nT > 0) {
    if (!(tq = c->alloc_r((size_t)c->nT * c->nC))) return c->fail(TM_ENOMEM, "hipMalloc tend_q");
    b.d_tq = (R *)tq;
  }
  // the tracer tendency buffer lives for this call only, on every return path
  struct Release {
    tm_ctx *c;
    void *p;
    ~Release() { c->release(p); }
  } release_tq{c, tq};
  int rc = launch_stage<R>(c, b);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  struct {
    double *dst;
    void *src;
    const std::vector<int> *perm;
    int n, copies;
  } outs[] = {
      {out->tend_gd, c->tend_gd, &c->pC, c->nC, 1},   {out->tend_iap, c->scratch_e[3], &c->pE, c->nE, 1},
      {out->tend_q, tq, &c->pC, c->nC, c->nT},        {out->gd_edge, c->gd_e, &c->pE, c->nE, 1},
      {out->u, c->u, &c->pE, c->nE, 1},               {out->ke, c->scratch_c[0], &c->pC, c->nC, 1},
      {out->vor, c->scratch_v[0], &c->pV, c->nV, 1},  {out->gd_vertex, c->scratch_v[1], &c->pV, c->nV, 1},
      {out->pv_vertex, c->pv_v, &c->pV, c->nV, 1},    {out->v, c->scratch_e[0], &c->pE, c->nE, 1},
      {out->pv_edge, c->pv_e, &c->pE, c->nE, 1},      {out->pv_flx, c->scratch_e[1], &c->pE, c->nE, 1},
      {out->u_tend, c->scratch_e[2], &c->pE, c->nE, 1},
  };
  for (auto &x : outs)
    if (x.dst && x.src && x.copies > 0 && (rc = download_state(c, x.dst, x.src, *x.perm, x.copies, x.n))) return rc;
  return TM_OK;
}

extern "C" int tm_stage(tm_ctx *c, tm_stage_out *out) {
  if (!c || !out) return TM_EINVAL;
  if (!c->have_state) return c->fail(TM_ESTATE, "tm_stage before tm_state_upload");
  return c->rsz == 8 ? stage_impl<double>(c, out) : stage_impl<float>(c, out);
}

// ------------------------------------------------------------ history fields
namespace tmk {

template <typename R>
struct HistArgs {
  const R *u, *gd, *ghs;
  const double *rc, *trig;
  double inv_g;
  double *zonal, *merid, *h, *div;  // [nC] each, any may be null
};

// Per cell: mpas_reconstruct_1d (mpas_vector_reconstruction.F90:195-247) of u,
// h = (gd + ghs) / g and div_operator(u) (history_mod.F90:125,166-168).
template <typename R>
__global__ __launch_bounds__(kBlock) void k_hist(DMesh m, HistArgs<R> a) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= m.nC) return;
  double x = 0.0, y = 0.0, z = 0.0, s = 0.0;
  for (int q = 0; q < m.nEoC[c]; ++q) {
    const int idx = m.eocS[q * m.nC + c], e = unsign(idx);
    const double ue = (double)a.u[e];
    x = x + a.rc[(0 * kMaxEoC + q) * (size_t)m.nC + c] * ue;
    y = y + a.rc[(1 * kMaxEoC + q) * (size_t)m.nC + c] * ue;
    z = z + a.rc[(2 * kMaxEoC + q) * (size_t)m.nC + c] * ue;
    const double t = ue * m.dv[e];
    s += idx < 0 ? -t : t;
  }
  const double clat = a.trig[c], slat = a.trig[m.nC + c], clon = a.trig[2 * (size_t)m.nC + c],
               slon = a.trig[3 * (size_t)m.nC + c];
  if (a.zonal) a.zonal[c] = -x * slon + y * clon;
  if (a.merid) a.merid[c] = -(x * clon + y * slon) * slat + z * clat;
  if (a.h) a.h[c] = ((double)a.gd[c] + (double)a.ghs[c]) * a.inv_g;
  if (a.div) a.div[c] = (-s) / m.areaCell[c];
}

}  // namespace tmk

namespace {

template <typename R>
int history_impl(tm_ctx *c, double *u, double *zonal, double *merid, double *h, double *pv, double *vor,
                 double *div) {
  const int nC = c->nC, nE = c->nE, nV = c->nV;
  // the refreshed fields of state(old) (tmcore_swm_mod.F90:78-84): a
  // diagnostic stage writes u, pv_vertex and, here, the vorticity
  StageBufs<R> b = stage_bufs<R>(c, 0, false);
  if (!c->hist_vor && (!(c->hist_vor = c->alloc_r(nV)) || !(c->hist_out = c->alloc<double>((size_t)4 * nC))))
    return c->fail(TM_ENOMEM, "hipMalloc history fields");
  void *dvor = c->hist_vor, *dout = c->hist_out;
  b.d_vor = (R *)dvor;
  int rc = launch_stage<R>(c, b, false);
  if (!rc) {
    HistArgs<R> a{};
    a.u = (const R *)c->u;
    a.gd = (const R *)c->gd_old;
    a.ghs = (const R *)c->ghs;
    a.rc = c->hist_rc;
    a.trig = c->hist_trig;
    a.inv_g = 1.0 / 9.80616;  // g of const_mod.F90:12
    double *o = (double *)dout;
    a.zonal = o;
    a.merid = o + nC;
    a.h = o + 2 * (size_t)nC;
    a.div = o + 3 * (size_t)nC;
    hipLaunchKernelGGL(k_hist<R>, dim3((nC + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, c->dm, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = c->fail(TM_EDEVICE, std::string("history kernel: ") + hipGetErrorString(e));
  }
  auto cell = [&](double *dst, int k) -> int {
    if (!dst) return TM_OK;
    std::vector<double> hbuf(nC);
    HIPCHK(c, hipMemcpyAsync(hbuf.data(), (double *)dout + (size_t)k * nC, nC * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < nC; ++i) dst[c->pC[i]] = hbuf[i];
    return TM_OK;
  };
  if (!rc) rc = cell(zonal, 0);
  if (!rc) rc = cell(merid, 1);
  if (!rc) rc = cell(h, 2);
  if (!rc) rc = cell(div, 3);
  if (!rc && u) rc = download_state(c, u, c->u, c->pE, 1, nE);
  if (!rc && pv) rc = download_state(c, pv, c->pv_v, c->pV, 1, nV);
  if (!rc && vor) rc = download_state(c, vor, dvor, c->pV, 1, nV);
  return rc;
}

}  // namespace

extern "C" int tm_history_setup(tm_ctx *c, const double *coeffs, const double *latCell, const double *lonCell) {
  if (!c) return TM_EINVAL;
  if (!coeffs || !latCell || !lonCell) return c->fail(TM_EINVAL, "null history setup input");
  const int nC = c->nC, ME = c->maxEdges;
  std::vector<double> rc((size_t)3 * kMaxEoC * nC, 0.0), trig((size_t)4 * nC);
  for (int i = 0; i < nC; ++i) {
    const int oc = c->pC[i];
    for (int q = 0; q < kMaxEoC && q < ME; ++q)
      for (int d = 0; d < 3; ++d) rc[((size_t)d * kMaxEoC + q) * nC + i] = coeffs[((size_t)oc * ME + q) * 3 + d];
    trig[i] = std::cos(latCell[oc]);
    trig[(size_t)nC + i] = std::sin(latCell[oc]);
    trig[2 * (size_t)nC + i] = std::cos(lonCell[oc]);
    trig[3 * (size_t)nC + i] = std::sin(lonCell[oc]);
  }
  if (!c->hist_rc && (!(c->hist_rc = c->alloc<double>(rc.size())) || !(c->hist_trig = c->alloc<double>(trig.size()))))
    return c->fail(TM_ENOMEM, "hipMalloc history coefficients");
  HIPCHK(c, hipMemcpyAsync(c->hist_rc, rc.data(), rc.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->hist_trig, trig.data(), trig.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return TM_OK;
}

extern "C" int tm_history_fields(tm_ctx *c, double *u, double *zonal, double *meridional, double *h, double *pv,
                                 double *vor, double *div) {
  if (!c) return TM_EINVAL;
  if (!c->have_state) return c->fail(TM_ESTATE, "tm_history_fields before tm_state_upload");
  if ((zonal || meridional) && !c->hist_rc) return c->fail(TM_ESTATE, "tm_history_fields before tm_history_setup");
  if (!c->hist_trig) {  // h and div need no coefficients: zero ones keep the kernel's reads valid
    std::vector<double> zero((size_t)3 * kMaxEoC * c->nC, 0.0);
    if (!(c->hist_rc = c->alloc<double>(zero.size())) || !(c->hist_trig = c->alloc<double>((size_t)4 * c->nC)))
      return c->fail(TM_ENOMEM, "hipMalloc history coefficients");
    HIPCHK(c, hipMemcpyAsync(c->hist_rc, zero.data(), zero.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->hist_trig, zero.data(), (size_t)4 * c->nC * sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return c->rsz == 8 ? history_impl<double>(c, u, zonal, meridional, h, pv, vor, div)
                     : history_impl<float>(c, u, zonal, meridional, h, pv, vor, div);
}

// ------------------------------------------------------------ learned closure
namespace {

int cl_check(tm_ctx *c, tm_mlp *mlp) {
  if (!c) return TM_EINVAL;
  if (!mlp) return c->fail(TM_EINVAL, "null MLP");
  if (mlp_state(mlp).device != c->o.device) return c->fail(TM_EINVAL, "MLP and context are on different devices");
  return TM_OK;
}

int cl_reserve(tm_ctx *c) {
  if (c->cl_X) return TM_OK;
  if (!(c->cl_X = c->alloc<float>((size_t)c->nE * kMlpIn)) || !(c->cl_Y = c->alloc<float>((size_t)c->nE * kMlpOut)) ||
      !(c->cl_dY = c->alloc<float>((size_t)c->nE * kMlpOut)) || !(c->cl_r = c->alloc<float>(c->nC)) ||
      !(c->cl_sq = c->alloc_r(c->nE)) || !(c->cl_vor = c->alloc_r(c->nV)) || !(c->cl_div = c->alloc_r(c->nC)) ||
      !(c->cl_lpart = c->alloc<double>((size_t)c->dm.nbE + c->dm.nbC)))
    return c->fail(TM_ENOMEM, "hipMalloc closure buffers");
  return TM_OK;
}

int cl_loss_reserve(tm_ctx *c, int n) {
  if (n > c->cl_loss_cap) {
    double *p = c->alloc<double>(n);
    if (!p) return c->fail(TM_ENOMEM, "hipMalloc closure loss");
    HIPCHK(c, hipStreamSynchronize(c->stream));  // no queued work may still use the old buffer
    c->release(c->cl_loss);
    c->cl_loss = p;
    c->cl_loss_cap = n;
  }
  HIPCHK(c, hipMemsetAsync(c->cl_loss, 0, (size_t)n * sizeof(double), c->stream));
  return TM_OK;
}

// Inputs of the current state into X (new order): a diagnostic stage of the
// state followed by the input kernel, with mlp's normalisation.
template <typename R>
int cl_diag_stage(tm_ctx *c) {
  StageBufs<R> b = stage_bufs<R>(c, 0, false);
  b.d_vor = (R *)c->cl_vor;
  return launch_stage<R>(c, b, false);
}

int cl_features_now(tm_ctx *c, tm_mlp *mlp, float *X) {
  tm_mlp *saved = c->cl;
  c->cl = mlp;  // enqueue_cl_features reads the normalisation of c->cl
  int rc = c->rsz == 8 ? cl_diag_stage<double>(c) : cl_diag_stage<float>(c);
  if (!rc) rc = c->rsz == 8 ? enqueue_cl_features<double>(c, X, nullptr) : enqueue_cl_features<float>(c, X, nullptr);
  c->cl = saved;
  return rc;
}

ClLossArgs cl_loss_args(tm_ctx *c, tm_mlp *mlp, const tm_ctx::ClSample &smp, bool want_dy, double *loss) {
  ClLossArgs a{};
  a.Y = c->cl_Y;
  a.tu = smp.tu;
  a.tg = smp.tg;
  a.norm = mlp_state(mlp).norm;
  a.tgs = smp.tgs;
  a.r = c->cl_r;
  a.dY = want_dy ? c->cl_dY : nullptr;
  a.part = c->cl_lpart;
  a.loss = loss;
  return a;
}

int cl_enqueue_loss(tm_ctx *c, const ClLossArgs &a) {
  hipLaunchKernelGGL(k_cl_resid, dim3(c->dm.nbC), dim3(kBlock), 0, c->stream, c->dm, a);
  hipLaunchKernelGGL(k_cl_dy, dim3(c->dm.nbE), dim3(kBlock), 0, c->stream, c->dm, a);
  hipLaunchKernelGGL(k_cl_loss_sum, dim3(1), dim3(64), 0, c->stream, c->dm, (const double *)a.part, a.loss);
  HIPCHK(c, hipGetLastError());
  return TM_OK;
}

template <typename T>
int upload_perm_f32(tm_ctx *c, float *dst, const T *src, const std::vector<int> &perm) {
  std::vector<float> h(perm.size());
  for (size_t i = 0; i < perm.size(); ++i) h[i] = (float)src[perm[i]];
  return upload(c, dst, h);
}

}  // namespace

extern "C" int tm_closure_attach(tm_ctx *c, tm_mlp *mlp) {
  if (!c) return TM_EINVAL;
  if (mlp) {
    int rc = cl_check(c, mlp);
    if (rc || (rc = cl_reserve(c))) return rc;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  // the captured step changes shape: capture again at the next run
  drop_graphs(c);
  c->cl = mlp;
  c->cl_step = 0;
  return TM_OK;
}

extern "C" int tm_closure_set_gain(tm_ctx *c, double gain) {
  if (!c) return TM_EINVAL;
  if (!(gain >= 0.0) || !std::isfinite(gain)) return c->fail(TM_EINVAL, "closure gain must be finite and >= 0");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  drop_graphs(c);  // the applying graphs bake the gain into their time step
  c->cl_gain = gain;
  return TM_OK;
}

extern "C" int tm_closure_set_interval(tm_ctx *c, int nsteps, int lumped) {
  if (!c) return TM_EINVAL;
  if (nsteps < 1) return c->fail(TM_EINVAL, "closure interval must be >= 1");
  if (lumped != 0 && lumped != 1) return c->fail(TM_EINVAL, "lumped must be 0 or 1");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  drop_graphs(c);  // kind 4 bakes the interval into its time step
  c->cl_interval = nsteps;
  c->cl_lumped = lumped;
  c->cl_step = 0;
  return TM_OK;
}

extern "C" int tm_closure_features(tm_ctx *c, tm_mlp *mlp, float *x) {
  int rc = cl_check(c, mlp);
  if (rc) return rc;
  if (!x) return c->fail(TM_EINVAL, "null output");
  if (!c->have_state) return c->fail(TM_ESTATE, "no state uploaded");
  if ((rc = cl_reserve(c)) || (rc = cl_features_now(c, mlp, c->cl_X))) return rc;
  std::vector<float> h((size_t)c->nE * kMlpIn);
  HIPCHK(c, hipMemcpyAsync(h.data(), c->cl_X, h.size() * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < c->nE; ++i)
    std::memcpy(x + (size_t)c->pE[i] * kMlpIn, h.data() + (size_t)i * kMlpIn, kMlpIn * 4);
  return TM_OK;
}

extern "C" int tm_closure_dataset_add(tm_ctx *c, tm_mlp *mlp, const double *t_u, const double *t_gd, double tg_scale) {
  int rc = cl_check(c, mlp);
  if (rc) return rc;
  if (!t_u || !t_gd) return c->fail(TM_EINVAL, "null target");
  if (!(tg_scale > 0.0) || !std::isfinite(tg_scale)) return c->fail(TM_EINVAL, "tg_scale must be positive");
  if (!c->have_state) return c->fail(TM_ESTATE, "no state uploaded");
  if (!finite_all(t_u, c->nE) || !finite_all(t_gd, c->nC)) return c->fail(TM_ENONFINITE, "non-finite target");
  if ((rc = cl_reserve(c))) return rc;
  tm_ctx::ClSample smp{nullptr, nullptr, nullptr, (float)tg_scale};
  if (hipMalloc((void **)&smp.X, (size_t)c->nE * kMlpIn * 4) != hipSuccess ||
      hipMalloc((void **)&smp.tu, (size_t)c->nE * 4) != hipSuccess ||
      hipMalloc((void **)&smp.tg, (size_t)c->nC * 4) != hipSuccess) {
    (void)hipFree(smp.X);
    (void)hipFree(smp.tu);
    (void)hipFree(smp.tg);
    return c->fail(TM_ENOMEM, "hipMalloc closure sample");
  }
  if ((rc = cl_features_now(c, mlp, smp.X)) || (rc = upload_perm_f32(c, smp.tu, t_u, c->pE)) ||
      (rc = upload_perm_f32(c, smp.tg, t_gd, c->pC))) {
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(smp.X);
    (void)hipFree(smp.tu);
    (void)hipFree(smp.tg);
    return rc;
  }
  c->ds.push_back(smp);  // filled; owned by the context from here on
  return TM_OK;
}

extern "C" int tm_closure_dataset_size(tm_ctx *c, int *n) {
  if (!c || !n) return TM_EINVAL;
  *n = (int)c->ds.size();
  return TM_OK;
}

extern "C" int tm_closure_dataset_clear(tm_ctx *c) {
  if (!c) return TM_EINVAL;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (auto &smp : c->ds) {
    (void)hipFree(smp.X);
    (void)hipFree(smp.tu);
    (void)hipFree(smp.tg);
  }
  c->ds.clear();
  return TM_OK;
}

extern "C" int tm_closure_eval(tm_ctx *c, tm_mlp *mlp, int first, int count, double *loss) {
  int rc = cl_check(c, mlp);
  if (rc) return rc;
  if (!loss || first < 0 || count < 1 || (size_t)first + count > c->ds.size())
    return c->fail(TM_EINVAL, "sample range outside the dataset");
  MlpState &ms = mlp_state(mlp);
  HIPCHK(c, hipStreamSynchronize(ms.stream));
  if ((rc = cl_reserve(c)) || (rc = cl_loss_reserve(c, 2))) return rc;
  c->cl_step = 0;  // cl_Y is overwritten: an attached closure recomputes its correction next step
  for (int i = first; i < first + count; ++i) {
    if ((rc = mlp_forward(ms, c->ds[i].X, c->nE, c->cl_Y, false, c->stream))) return c->fail(rc, ms.err);
    if ((rc = cl_enqueue_loss(c, cl_loss_args(c, mlp, c->ds[i], false, c->cl_loss)))) return rc;
  }
  HIPCHK(c, hipMemcpyAsync(loss, c->cl_loss, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  loss[0] /= count;
  loss[1] /= count;
  return TM_OK;
}

extern "C" int tm_closure_train(tm_ctx *c, tm_mlp *mlp, int ntrain, int epochs, float lr, uint64_t seed,
                                double *loss_hist) {
  int rc = cl_check(c, mlp);
  if (rc) return rc;
  if (ntrain < 1 || (size_t)ntrain > c->ds.size() || epochs < 1) return c->fail(TM_EINVAL, "bad ntrain or epochs");
  if (!(lr > 0.f) || !std::isfinite(lr)) return c->fail(TM_EINVAL, "lr must be positive");
  MlpState &ms = mlp_state(mlp);
  HIPCHK(c, hipStreamSynchronize(ms.stream));
  if ((rc = cl_reserve(c)) || (rc = cl_loss_reserve(c, 2 * epochs))) return rc;
  c->cl_step = 0;  // cl_Y is overwritten: an attached closure recomputes its correction next step
  if ((rc = mlp_reserve(ms, c->nE)) || (rc = mlp_zero_grad(ms, c->stream))) return c->fail(rc, ms.err);
  std::vector<int> order(ntrain);
  std::iota(order.begin(), order.end(), 0);
  std::mt19937_64 rng(seed);
  for (int ep = 0; ep < epochs; ++ep) {
    std::shuffle(order.begin(), order.end(), rng);
    for (int i : order) {
      const tm_ctx::ClSample &smp = c->ds[i];
      if ((rc = mlp_forward(ms, smp.X, c->nE, c->cl_Y, true, c->stream))) return c->fail(rc, ms.err);
      if ((rc = cl_enqueue_loss(c, cl_loss_args(c, mlp, smp, true, c->cl_loss + 2 * ep)))) return rc;
      if ((rc = mlp_backward(ms, smp.X, c->cl_dY, c->nE, nullptr, c->stream)) ||
          (rc = mlp_adam(ms, lr, 0.9f, 0.999f, 1e-8f, c->stream)))
        return c->fail(rc, ms.err);
    }
  }
  std::vector<double> h(2 * (size_t)epochs);
  HIPCHK(c, hipMemcpyAsync(h.data(), c->cl_loss, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (auto &v : h) v /= ntrain;
  if (loss_hist) std::memcpy(loss_hist, h.data(), h.size() * sizeof(double));
  if (!finite_all(h.data(), h.size())) return c->fail(TM_ENONFINITE, "non-finite training loss");
  return TM_OK;
}

#include "ops_kernels.h"

// ------------------------------------------------------------- copy probe
// One 16-byte element per thread over a grid that covers the array (the
// fastest of the copy forms measured on MI355X: grid-stride loops with 2-8
// elements per thread, with and without non-temporal hints, were slower).
__global__ __launch_bounds__(256) void tm_copy_probe_kernel(const double2 *__restrict__ a, double2 *__restrict__ b,
                                                            size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

extern "C" int tm_hbm_copy_probe(int device, double bytes, int iters, double *gbps) {
  if (!gbps || !(bytes >= 1 << 20) || bytes > 64.0 * (1ull << 30) || iters < 1) return TM_EINVAL;
  const size_t n = (size_t)(bytes / sizeof(double2));
  void *a = nullptr, *b = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = TM_OK;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    rc = TM_EDEVICE;
  if (!rc && (hipMalloc(&a, n * sizeof(double2)) != hipSuccess || hipMalloc(&b, n * sizeof(double2)) != hipSuccess))
    rc = TM_ENOMEM;
  if (!rc) {
    (void)hipMemsetAsync(a, 0, n * sizeof(double2), s);
    const unsigned grid = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(tm_copy_probe_kernel, dim3(grid), dim3(256), 0, s, (const double2 *)a, (double2 *)b, n);
    (void)hipEventRecord(e0, s);
    for (int i = 0; i < iters; ++i)
      hipLaunchKernelGGL(tm_copy_probe_kernel, dim3(grid), dim3(256), 0, s, (const double2 *)a, (double2 *)b, n);
    (void)hipEventRecord(e1, s);
    float ms = 0.f;
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
      rc = TM_EDEVICE;
    else
      *gbps = 2.0 * (double)n * sizeof(double2) / ((double)ms * 1e-3 / iters) / 1e9;
  }
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  return rc;
}
