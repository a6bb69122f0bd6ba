// Fused TRiSK RK-stage kernels for gfx950, templated on the state type R
// (double: plain FP64 form; float: precision form, FP32 stencil arithmetic
// on FP32 copies of the edge weights and lengths, cell and vertex areas in
// FP64, and FP64 compensated global sums).
//
// One RK stage = four launches:
//   k_edge_prep     gd_edge, u, and the per-edge products the cell/vertex sums
//                   gather (scalar_c2e + inverse_iap, operators_mod.F90:93-104,
//                   168-176; flux of calc_gd_tend_on_cell tmcore_swm_mod.F90:
//                   135; tracer flux adv_scheme_sg11_mod.F90:42-45)
//   k_cell_vertex   cells: div_operator (:40-55), calc_kinetic_energy
//                   (:206-221), E = ke+gd+ghs (tmcore_swm_mod.F90:177), tracer
//                   div, RK stage update of gd/tracers, beta partial products;
//                   vertices: curl_operator (:76-91), scalar_c2v (:106-123),
//                   calc_pv_on_vertex (:262-270)
//   k_edge_pv       calc_tangent_wind (:223-237) + calc_pv_on_edge scheme 3
//                   (:359-412)
//   k_edge_tend     calc_tangent_vor_flux (:239-260), calc_u_tend_on_edge
//                   (tmcore_swm_mod.F90:155-201), RK stage update of iap_u
// and per step k_reduce (inner_product, beta, invariants) and k_finalize
// (update_state(beta*dt), time_scheme_mod.F90:152).
//
// Each expression keeps the reference's association so that the FP64 form,
// built with -ffp-contract=off, reproduces the reference arithmetic.
#pragma once

#include "tm_device.h"

namespace tmk {

__device__ __forceinline__ double tm_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float tm_sqrt(float x) { return sqrtf(x); }
template <typename R>
__device__ __forceinline__ R sign1(R x) {  // Fortran sign(1.0d0, x)
  return signbit(x) ? R(-1) : R(1);
}
__device__ __forceinline__ int unsign(int idx) { return idx ^ (idx >> 31); }

// Loads of per-element constants that only one kernel of a stage reads
// (mesh connectivity and weights streamed once per launch): non-temporal, so
// they do not evict the fields the stage's kernels hand to each other from
// the memory-side cache.  Measured at level 8: the vertex fit weights in
// k_edge_pv 1.552 -> 1.523 ms per FP64 step, the cell/vertex constants of
// k_cell_vertex 1.524 -> 1.508 ms.
template <typename T>
__device__ __forceinline__ T nt_ld(const T *p) {
  return __builtin_nontemporal_load(p);
}

// Edge geometry and weights read in the arithmetic type R: FP32 copies in
// FP32 contexts (half the bytes, and the very values a cast of the FP64 array
// at the point of use gives), the FP64 arrays otherwise.  NT: non-temporal
// load, for the stream's last reader in a stage.
template <typename R, bool NT = false>
__device__ __forceinline__ R geo_at(const double *d64, const float *d32, size_t i) {
  if constexpr (sizeof(R) == 4) {
    if constexpr (NT) return nt_ld(&d32[i]);
    return d32[i];
  } else {
    if constexpr (NT) return nt_ld(&d64[i]);
    return d64[i];
  }
}

template <typename R>
struct StageBufs {
  const R *gd, *iap, *q;                 // stage input state X_k
  const R *gd_base, *iap_base, *q_base;  // X_1 (state(one))
  R *gd_next, *iap_next, *q_next;        // X_{k+1}
  R *acc_gd, *acc_iap, *acc_q;           // t1 + 2 t2 + 2 t3 (+ t4)/6
  const R *tprev_gd, *tprev_iap;         // tendency of stage k-1
  R *tcur_gd, *tcur_iap;                 // tendency of stage k
  R *gd_e, *u, *udc, *E, *tend_gd, *pv_v, *pv_e;
  R *gd_e_n, *u_n;  // kXTendPrep: the next stage's gd_edge and u
  int prep_next;    // kXTendPrep: this stage's k_edge_tend forms the next stage's edge values
  // [nE][2]: u*gd_e*dvEdge (mass flux) and u^2*areaEdge, interleaved because
  // the cell sums gather both at the same edge (one load per slot)
  R *fk;
  const R *ghs;
  int nT;
  int k;          // 1..4: CRK4 stage; 0: diagnostics/invariants only
  double cdt;     // c_k * dt of update_state (c = 1/2, 1/2, 1)
  int want_inv;   // accumulate the invariants of X_k
  int pv_order;   // interp_pv_order (2..6) of pv_scheme 3
  int pv_scheme;  // 3: high-order upwind; 1: APVM; 2: conservative APVM
  double apvm_wdt;  // apvm_weight * dt (pv_scheme 1)
  double dt;
  R *pv_c, *tgd_v;  // pv_cell (schemes 1, 2), gd_tend_vertex (scheme 2), else null
  R *dpv, *pvf, *dpvf;  // scheme 2: APVM difference term and the two TRiSK fluxes
  double2 *cv_part;  // scheme 2: [3][nbV] block partials of the eps sums
  double *cv_eps;    // scheme 2: eps
  int adv_order;  // 2..6: add_upwind_correction_on_cell for orders >= 3
  double adv_coef2, adv_coef4;
  R *qe;          // [nT][nE] tracer edge values (adv_order >= 3), else null
  // stage descriptor of the kXExt kernels (time_scheme_mod.F90:65-184):
  // X_next = (base [+ cprev_dt T_prev]) + cdt T_cur unless ts_last; the
  // accumulator: ts_acc 0 untouched, 1 acc = T, 2 acc += ts_w T, 3 acc =
  // (acc + T) / 6; sums: T_prev.T into rid_prev, T_first.T into rid_first
  // (edge ids; cells +1; -1 = none), acc.acc into kR44 at the last stage
  int ts_last, ts_acc, rid_prev, rid_first;
  double ts_w, cprev_dt;
  const R *tfirst_gd, *tfirst_iap;
  double2 *part;  // [kNumSums][part_stride] block partial sums (hi, lo)
  int part_stride;
  int rev;        // sweep the tiles from the last one (see xcd_tile)
  // optional diagnostics (test path), NULL in timed runs
  R *d_ke, *d_vor, *d_gdv, *d_v, *d_pvflx, *d_utend, *d_tiap, *d_tq;
};

// ---------------------------------------------------------------- reductions
// Double-double accumulation (Knuth TwoSum) for the compensated global sums.
__device__ __forceinline__ double2 dd_add(double2 a, double2 b) {
  double s = a.x + b.x;
  double bb = s - a.x;
  double err = (a.x - (s - bb)) + (b.x - bb);
  double lo = err + (a.y + b.y);
  double hi = s + lo;
  return make_double2(hi, lo - (hi - s));
}

template <int N>
__device__ __forceinline__ void block_sums(const double (&v)[N], const int (&ids)[N], double2 *part, int stride,
                                           int tile) {
  __shared__ double2 sh[N][kBlock / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();  // a previous call with the same N may still be reading sh
#pragma unroll
  for (int i = 0; i < N; ++i) {
    double2 a = make_double2(v[i], 0.0);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      double2 o = make_double2(__shfl_xor(a.x, off, 64), __shfl_xor(a.y, off, 64));
      a = dd_add(a, o);
    }
    if (lane == 0) sh[i][w] = a;
  }
  __syncthreads();
  if (threadIdx.x < N) {
    const int i = threadIdx.x;
    double2 a = sh[i][0];
#pragma unroll
    for (int q = 1; q < kBlock / 64; ++q) a = dd_add(a, sh[i][q]);
    part[ids[i] * stride + tile] = a;
  }
}

// ----------------------------------------------------------- element code
// Values: one model's element is a scalar R.  Ensemble members of an element
// are stored side by side (member fastest) and a thread may carry W of them
// as one vector value (val_t<R, W>), so the element's mesh constants and
// stencil indices are loaded once for W members and every state gather is
// one W-wide load.  With W = 1 the value type is R itself and the code below
// is the plain scalar code.
template <typename R, int W>
struct val_s {
  typedef R type __attribute__((ext_vector_type(W)));
};
template <typename R>
struct val_s<R, 1> {
  typedef R type;
};
template <typename R, int W>
using val_t = typename val_s<R, W>::type;

// Slot pair p (slots 2p, 2p+1) of element i of a pair-stored table (pair_at).
template <typename T, bool NT = false>
__device__ __forceinline__ val_t<T, 2> ld_pair(const T *a, int p, int i, int n) {
  const val_t<T, 2> *q = reinterpret_cast<const val_t<T, 2> *>(a) + ((size_t)p * n + i);
  if constexpr (NT) return __builtin_nontemporal_load(q);
  return *q;
}
// Slots 2p, 2p+1 of element i of an edge weight table read in the arithmetic
// type R: the FP64 table is stored in slot pairs (one 16-byte load), its FP32
// copy slot-major (two 4-byte loads, which measured faster than one 8-byte
// pair load).  geo_slot: one slot.
template <typename R, bool NT = false>
__device__ __forceinline__ void geo_pair(const double *d64, const float *d32, int p, int i, int n, R &x, R &y) {
  if constexpr (sizeof(R) == 4) {
    x = geo_at<R, NT>(d64, d32, (size_t)(2 * p) * n + i);
    y = geo_at<R, NT>(d64, d32, (size_t)(2 * p + 1) * n + i);
  } else {
    const val_t<double, 2> v = ld_pair<double, NT>(d64, p, i, n);
    x = v.x, y = v.y;
  }
}
template <typename R>
__device__ __forceinline__ R geo_slot(const double *d64, const float *d32, int s, int i, int n) {
  return geo_at<R>(d64, d32, sizeof(R) == 4 ? (size_t)s * n + i : pair_at(s, i, n));
}
// The edgesOnEdge slots of edge e, read in pairs.
template <bool NT = false>
__device__ __forceinline__ void eoe_row(const DMesh &m, int e, int (&ep)[kMaxEoE]) {
#pragma unroll
  for (int p = 0; p < kMaxEoE / 2; ++p) {
    const val_t<int, 2> t = ld_pair<int, NT>(m.eoe, p, e, m.nE);
    ep[2 * p] = t.x, ep[2 * p + 1] = t.y;
  }
}

// Slot s (0..5) of a cell's edge row held in two int4 registers, s not a
// compile-time constant: selects on the bits of s, so the row never goes
// through private memory.
__device__ __forceinline__ int row_slot(const int4 &a, const int4 &b, int s) {
  const int lo = (s & 1) ? a.y : a.x, mid = (s & 1) ? a.w : a.z, hi = (s & 1) ? b.y : b.x;
  return s < 2 ? lo : (s < 4 ? mid : hi);
}

// The edgesOnEdge slots of edge e derived from the edge rows of its cells c1
// = cellsOnEdge(1) and c2 = cellsOnEdge(2) (the MPAS ordering checked at
// upload): the edges of c1 that follow e counter-clockwise, then those of c2;
// padding slots hold e itself (their weights are 0), as in the table.
__device__ __forceinline__ void eoe_row_cells(const DMesh &m, int e, int c1, int c2, int (&ep)[kMaxEoE]) {
  const int4 *r1 = reinterpret_cast<const int4 *>(m.eocR) + (size_t)c1 * 2;
  const int4 *r2 = reinterpret_cast<const int4 *>(m.eocR) + (size_t)c2 * 2;
  const int4 a0 = r1[0], a1 = r1[1], b0 = r2[0], b1 = r2[1];
  const int n1 = a1.z, n2 = b1.z;  // slot 6 of a row: nEdgesOnCell
  // e's slot in each row (slots past nEdgesOnCell hold 0 and are skipped)
  const int k1 = unsign(a0.x) == e ? 0 : unsign(a0.y) == e ? 1 : unsign(a0.z) == e ? 2 :
                 unsign(a0.w) == e ? 3 : unsign(a1.x) == e ? 4 : 5;
  const int k2 = unsign(b0.x) == e ? 0 : unsign(b0.y) == e ? 1 : unsign(b0.z) == e ? 2 :
                 unsign(b0.w) == e ? 3 : unsign(b1.x) == e ? 4 : 5;
#pragma unroll
  for (int j = 0; j < kMaxEoE; ++j) {
    int s1 = k1 + 1 + j;
    s1 -= s1 >= n1 ? n1 : 0;
    int s2 = k2 + 1 + (j - (n1 - 1));
    s2 -= s2 >= n2 ? n2 : 0;
    const int v = j < n1 - 1 ? row_slot(a0, a1, s1) : row_slot(b0, b1, s2);
    ep[j] = j < n1 + n2 - 2 ? unsign(v) : e;
  }
}

template <typename T, int W>
__device__ __forceinline__ val_t<T, W> tm_sqrt(val_t<T, W> x) {
  if constexpr (W == 1) {
    return tm_sqrt(x);
  } else {
    val_t<T, W> r;
#pragma unroll
    for (int w = 0; w < W; ++w) r[w] = tm_sqrt(T(x[w]));
    return r;
  }
}
// Component w of a value (the value itself when W = 1).
template <typename T, int W>
__device__ __forceinline__ T vget(const val_t<T, W> &v, int w) {
  if constexpr (W == 1) {
    (void)w;
    return v;
  } else {
    return v[w];
  }
}
template <typename T, int W>
__device__ __forceinline__ void vset(val_t<T, W> &v, int w, T x) {
  if constexpr (W == 1) {
    (void)w;
    v = x;
  } else {
    v[w] = x;
  }
}
template <typename R, int W>
__device__ __forceinline__ val_t<R, W> sign1v(val_t<R, W> x) {
  if constexpr (W == 1) {
    return sign1(x);
  } else {
    val_t<R, W> r;
#pragma unroll
    for (int w = 0; w < W; ++w) r[w] = sign1(R(x[w]));
    return r;
  }
}
template <typename R, int W>
__device__ __forceinline__ val_t<double, W> to_d(val_t<R, W> x) {
  if constexpr (W == 1)
    return (double)x;
  else
    return __builtin_convertvector(x, val_t<double, W>);
}

// Layouts: where element i of a state or intermediate array lives, and how
// a thread's W values of it are loaded and stored.  Solo is one model;
// Members<W> holds the M members of element i at [i*M, i*M + M) and a thread
// takes members [m0, m0 + W).
struct Solo {
  static constexpr int W = 1;
  __device__ __forceinline__ size_t operator()(size_t i) const { return i; }
  template <typename R>
  __device__ __forceinline__ R ld(const R *p, size_t i) const {
    return p[i];
  }
  template <typename R>
  __device__ __forceinline__ void st(R *p, size_t i, R x) const {
    p[i] = x;
  }
  // streamed state (the RK base, accumulator and stage tendencies, each read
  // or written once per launch by one kernel): non-temporal
  template <typename R>
  __device__ __forceinline__ R ld_s(const R *p, size_t i) const {
    return __builtin_nontemporal_load(p + i);
  }
  template <typename R>
  __device__ __forceinline__ void st_s(R *p, size_t i, R x) const {
    __builtin_nontemporal_store(x, p + i);
  }
  // two values of element i of an interleaved pair array ([n][2]): one load
  template <typename R>
  __device__ __forceinline__ void ld2(const R *p, size_t i, R &a, R &b) const {
    const val_t<R, 2> v = *reinterpret_cast<const val_t<R, 2> *>(p + 2 * i);
    a = v.x, b = v.y;
  }
  template <typename R>
  __device__ __forceinline__ void st2(R *p, size_t i, R a, R b) const {
    val_t<R, 2> v;
    v.x = a, v.y = b;
    *reinterpret_cast<val_t<R, 2> *>(p + 2 * i) = v;
  }
};
template <int NW>
struct Members {
  static constexpr int W = NW;
  int M, m0;
  __device__ __forceinline__ size_t operator()(size_t i) const { return i * (size_t)M + (size_t)m0; }
  template <typename R>
  __device__ __forceinline__ val_t<R, NW> ld(const R *p, size_t i) const {
    return *reinterpret_cast<const val_t<R, NW> *>(p + (*this)(i));
  }
  template <typename R>
  __device__ __forceinline__ void st(R *p, size_t i, val_t<R, NW> x) const {
    *reinterpret_cast<val_t<R, NW> *>(p + (*this)(i)) = x;
  }
  template <typename R>
  __device__ __forceinline__ val_t<R, NW> ld_s(const R *p, size_t i) const {
    return ld(p, i);
  }
  template <typename R>
  __device__ __forceinline__ void st_s(R *p, size_t i, val_t<R, NW> x) const {
    st(p, i, x);
  }
  // interleaved pair array: elements 2i and 2i+1 of a [2n][M] array
  template <typename R>
  __device__ __forceinline__ void ld2(const R *p, size_t i, val_t<R, NW> &a, val_t<R, NW> &b) const {
    a = ld(p, 2 * i), b = ld(p, 2 * i + 1);
  }
  template <typename R>
  __device__ __forceinline__ void st2(R *p, size_t i, val_t<R, NW> a, val_t<R, NW> b) const {
    st(p, 2 * i, a), st(p, 2 * i + 1, b);
  }
};
using Member = Members<1>;

// Compile-time variants of the stage code: X = 0 is the default configuration
// (adv_order 2, pv_scheme 3 with interp_pv_order <= 4, CRK4) and carries none
// of the code below, so its kernels keep their register budget; kXAdv adds the
// upwind cell correction (adv_order >= 3), kXAdv56 (with kXAdv) its
// fourth-order fits for adv_order 5/6, kXExt the PV schemes 1/2, the
// order-5/6 PV and the stage-descriptor time integration (CRK3, PC2; CRK4
// too in that variant), all still checked at run time inside.
constexpr int kXAdv = 1, kXExt = 2, kXAdv56 = 4;
// Cell-row variants (meshes whose edgesOnEdge rows follow the MPAS order
// checked at upload, dm.eocR set): k_edge_pv (kXEoePv) or k_edge_tend
// (kXEoeTend) derives edgesOnEdge from the edge rows of the edge's two cells
// instead of streaming the edgesOnEdge table.  Same slots in the same order:
// results are bit-identical.
constexpr int kXEoePv = 8, kXEoeTend = 16;
// kXTendPrep (CRK4 without the upwind correction): k_edge_tend of stages 1-3
// also forms the next stage's edge values from iap_next and the cells' next
// state, as k_edge_prep would (same expressions on the same values), so
// stages 2-4 have no edge-prep pass.  gd_edge and u ping-pong between two
// buffers (the tendency reads its neighbours' current values).
constexpr int kXTendPrep = 32;
// kXTendLds (CRK4 without the upwind correction): the tendency kernel stages
// u, gd_edge and pv_edge of its tile and the tile's halo in LDS and gathers
// the edgesOnEdge values from there (DMesh::eoeL, haloE); the arithmetic is
// unchanged.
constexpr int kXTendLds = 64;
// kXPvLds (same schemes): the PV kernel stages u of its tile and of the
// tile's edgesOnEdge halo (the tendency kernel's tiles and lists) in LDS for
// the tangent wind.
constexpr int kXPvLds = 128;

// sum_j w(j, side, e) * (f(fit(j, c)) - f(c)) over the NS padded slots of a
// fit stencil of element c (f at offset base; padded slots add an exact 0).
// c is itself gathered (an edge's cell or vertex), so its index row is read
// as whole int4s rather than one slot-major load per slot.
template <int NS, typename R, typename L>
__device__ __forceinline__ val_t<R, L::W> fit_sum(const DMesh &m, const R *f, size_t base, L at, const int *fit,
                                                  const double *w, const float *w32, int c, int side, int e) {
  typedef val_t<R, L::W> V;
  constexpr int RW = fit_row(NS);
  const V fc = at.ld(f, base + c);
  int idx[RW];
  const int4 *row = reinterpret_cast<const int4 *>(fit) + (size_t)c * (RW / 4);
#pragma unroll
  for (int j = 0; j < RW / 4; ++j) {
    const int4 a = row[j];
    idx[4 * j] = a.x, idx[4 * j + 1] = a.y, idx[4 * j + 2] = a.z, idx[4 * j + 3] = a.w;
  }
  V s = V(R(0));
#pragma unroll
  for (int j = 0; j < NS; ++j)
    s += geo_at<R>(w, w32, ((size_t)side * NS + j) * m.nE + e) * (at.ld(f, base + idx[j]) - fc);
  return s;
}

// add_upwind_correction_on_cell (operators_mod.F90:423-472) at edge e: fe is
// c2e(f) there, sgn = sign(1, u_e); returns fe unchanged at adv_order 2.
template <typename R, typename L, int X = 0>
__device__ __forceinline__ val_t<R, L::W> upwind_cell_e(const DMesh &m, const StageBufs<R> &b, L at, const R *f,
                                                        size_t base, int e, val_t<R, L::W> fe, val_t<R, L::W> sgn) {
  typedef val_t<R, L::W> V;
  if constexpr ((X & kXAdv) == 0) return fe;
  if (b.adv_order < 3) return fe;
  const int c1 = m.coe[e], c2 = m.coe[m.nE + e];
  const V d1 = fit_sum<kFit2C, R, L>(m, f, base, at, m.fit2C, m.d2c, m.d2c32, c1, 0, e);
  const V d2 = fit_sum<kFit2C, R, L>(m, f, base, at, m.fit2C, m.d2c, m.d2c32, c2, 1, e);
  const R dc = geo_at<R>(m.dc, m.dc32, e);
  if constexpr ((X & kXAdv56) == 0) return fe - ((dc * dc) * ((d1 + d2) - (sgn * (R)b.adv_coef2) * (d2 - d1))) / R(12);
  if (b.adv_order <= 4) return fe - ((dc * dc) * ((d1 + d2) - (sgn * (R)b.adv_coef2) * (d2 - d1))) / R(12);
  const V q1 = fit_sum<kFit4C, R, L>(m, f, base, at, m.fit4C, m.d4c, m.d4c32, c1, 0, e);
  const V q2 = fit_sum<kFit4C, R, L>(m, f, base, at, m.fit4C, m.d4c, m.d4c32, c2, 1, e);
  return (fe - ((dc * dc) * (d1 + d2)) / R(12)) +
         (((dc * dc) * (dc * dc)) * ((q1 + q2) - (sgn * (R)b.adv_coef4) * (q2 - q1))) / R(60);
}

// The order-3/4 upwind correction's per-edge data (both cells' fit2Cells
// rows, their deriv2OnCell weights, dcEdge), loaded once per edge and applied
// to every field the edge corrects: gd and each tracer.  apply() is
// upwind_cell_e at adv_order 3/4 term for term.
template <typename R>
struct UpwindFit2 {
  int c1, c2;
  int i1[kFit2C], i2[kFit2C];
  R w1[kFit2C], w2[kFit2C];
  R dc;
  __device__ __forceinline__ void load(const DMesh &m, int e) {
    constexpr int RW = fit_row(kFit2C);
    c1 = m.coe[e], c2 = m.coe[m.nE + e];
    const int4 *r1 = reinterpret_cast<const int4 *>(m.fit2C) + (size_t)c1 * (RW / 4);
    const int4 *r2 = reinterpret_cast<const int4 *>(m.fit2C) + (size_t)c2 * (RW / 4);
    int t1[RW], t2[RW];
#pragma unroll
    for (int j = 0; j < RW / 4; ++j) {
      const int4 a = r1[j], b = r2[j];
      t1[4 * j] = a.x, t1[4 * j + 1] = a.y, t1[4 * j + 2] = a.z, t1[4 * j + 3] = a.w;
      t2[4 * j] = b.x, t2[4 * j + 1] = b.y, t2[4 * j + 2] = b.z, t2[4 * j + 3] = b.w;
    }
#pragma unroll
    for (int j = 0; j < kFit2C; ++j) {
      i1[j] = t1[j], i2[j] = t2[j];
      w1[j] = geo_at<R>(m.d2c, m.d2c32, (size_t)j * m.nE + e);
      w2[j] = geo_at<R>(m.d2c, m.d2c32, (size_t)(kFit2C + j) * m.nE + e);
    }
    dc = geo_at<R>(m.dc, m.dc32, e);
  }
  template <typename L>
  __device__ __forceinline__ val_t<R, L::W> apply(const R *f, size_t base, L at, val_t<R, L::W> fe,
                                                  val_t<R, L::W> sgn, R coef2) const {
    typedef val_t<R, L::W> V;
    const V f1 = at.ld(f, base + c1), f2 = at.ld(f, base + c2);
    V d1 = V(R(0)), d2 = V(R(0));
#pragma unroll
    for (int j = 0; j < kFit2C; ++j) d1 += w1[j] * (at.ld(f, base + i1[j]) - f1);
#pragma unroll
    for (int j = 0; j < kFit2C; ++j) d2 += w2[j] * (at.ld(f, base + i2[j]) - f2);
    return fe - ((dc * dc) * ((d1 + d2) - (sgn * coef2) * (d2 - d1))) / R(12);
  }
};

// gd_edge, u and the per-edge products the cell/vertex sums gather; returns
// the edge part of the energy invariant when b.want_inv.
template <typename R, typename L, int X = 0>
__device__ __forceinline__ val_t<double, L::W> edge_prep_e(const DMesh &m, const StageBufs<R> &b, L at, int e) {
  typedef val_t<R, L::W> V;
  const int c1 = m.coe[e], c2 = m.coe[m.nE + e];
  const V iap = at.ld(b.iap, e);
  // adv_order 3/4 (kXAdv without kXAdv56): the fit data, loaded once,
  // serves gd and every tracer of this edge
  constexpr bool kHoist = (X & kXAdv) != 0 && (X & kXAdv56) == 0;
  V gde;
  [[maybe_unused]] UpwindFit2<R> uf;
  if constexpr (kHoist) {
    uf.load(m, e);
    gde = uf.apply(b.gd, 0, at, R(0.5) * (at.ld(b.gd, c1) + at.ld(b.gd, c2)), sign1v<R, L::W>(iap),
                   (R)b.adv_coef2);
  } else {
    gde = upwind_cell_e<R, L, X>(m, b, at, b.gd, 0, e, R(0.5) * (at.ld(b.gd, c1) + at.ld(b.gd, c2)),
                                 sign1v<R, L::W>(iap));
  }
  const V u = iap / tm_sqrt<R, L::W>(gde);
  const R dv = geo_at<R>(m.dv, m.dv32, e);
  at.st(b.gd_e, e, gde);
  at.st(b.u, e, u);
  at.st2(b.fk, e, (u * gde) * dv, (u * u) * geo_at<R>(m.aE, m.aE32, e));
  at.st(b.udc, e, u * geo_at<R>(m.dc, m.dc32, e));
  if constexpr ((X & kXAdv) != 0) {
    if (b.qe)  // tracer edge values with the upwind correction (adv_scheme_sg11_mod.F90:42-43)
      for (int t = 0; t < b.nT; ++t) {
        const size_t tb = (size_t)t * m.nC;
        const V qe = R(0.5) * (at.ld(b.q, tb + c1) + at.ld(b.q, tb + c2));
        V qc;
        if constexpr (kHoist)
          qc = uf.apply(b.q, tb, at, qe, sign1v<R, L::W>(u), (R)b.adv_coef2);
        else
          qc = upwind_cell_e<R, L, X>(m, b, at, b.q, tb, e, qe, sign1v<R, L::W>(u));
        at.st(b.qe, (size_t)t * m.nE + e, qc);
      }
  }
  const val_t<double, L::W> id = to_d<R, L::W>(iap);
  return b.want_inv ? (id * id) * m.aE[e] : val_t<double, L::W>(0.0);
}

// CRK4 bookkeeping of one prognostic value: next stage state, accumulator
// t1 + 2 t2 + 2 t3 (+ t4)/6, and the previous/current tendency for beta.
template <typename R, typename L>
__device__ __forceinline__ val_t<R, L::W> rk_accumulate(const StageBufs<R> &b, R *acc_arr, L at, size_t i,
                                                        val_t<R, L::W> t) {
  val_t<R, L::W> acc;
  if (b.k == 1)
    acc = t;
  else if (b.k < 4)
    acc = at.ld_s(acc_arr, i) + R(2) * t;
  else
    acc = (at.ld_s(acc_arr, i) + t) / R(6);
  at.st_s(acc_arr, i, acc);
  return acc;
}

template <typename D>
struct CellSums {
  D mass, enc, rp, r44, rf;
};

// One element's RK update of a stage in the kXExt kernels (stage descriptor
// above); rp, rf, r44 are its parts of the stage's inner products.
template <typename R, typename L>
__device__ __forceinline__ val_t<R, L::W> rk_update_ext(const StageBufs<R> &b, L at, size_t i, val_t<R, L::W> t,
                                                        const R *base, R *next, R *acc_arr, const R *tprev,
                                                        const R *tfirst, R *tcur, double w, val_t<double, L::W> &rp,
                                                        val_t<double, L::W> &rf, val_t<double, L::W> &r44) {
  typedef val_t<R, L::W> V;
  if (!b.ts_last) {
    V x = at.ld(base, i);
    if (b.cprev_dt != 0.0) x = x + (R)b.cprev_dt * at.ld(tprev, i);
    at.st(next, i, x + (R)b.cdt * t);
  }
  V acc = t;
  if (b.ts_acc == 1) {
    at.st(acc_arr, i, t);
  } else if (b.ts_acc == 2) {
    acc = at.ld(acc_arr, i) + (R)b.ts_w * t;
    at.st(acc_arr, i, acc);
  } else if (b.ts_acc == 3) {
    acc = (at.ld(acc_arr, i) + t) / R(6);
    at.st(acc_arr, i, acc);
  }
  if (b.rid_prev >= 0) rp = (to_d<R, L::W>(at.ld(tprev, i)) * to_d<R, L::W>(t)) * w;
  if (b.rid_first >= 0) rf = (to_d<R, L::W>(at.ld(tfirst, i)) * to_d<R, L::W>(t)) * w;
  if (b.ts_last) r44 = (to_d<R, L::W>(acc) * to_d<R, L::W>(acc)) * w;
  if (!b.ts_last) at.st(tcur, i, t);
  return acc;
}

template <typename R, typename L, int X = 0>
__device__ __forceinline__ CellSums<val_t<double, L::W>> cell_e(const DMesh &m, const StageBufs<R> &b, L at, int c) {
  typedef val_t<R, L::W> V;
  typedef val_t<double, L::W> D;
  CellSums<D> r{D(0.0), D(0.0), D(0.0), D(0.0), D(0.0)};
  // Unused slots of pentagons hold index 0 (a valid edge): every gather is
  // issued unconditionally so they overlap, and the unused terms are
  // replaced by an exact +0 afterwards.
  const int n = m.nEoC[c];
  int idx[kMaxEoC];
#pragma unroll
  for (int s = 0; s < kMaxEoC; ++s) idx[s] = nt_ld(&m.eocS[s * m.nC + c]);
  V fv[kMaxEoC], kv[kMaxEoC];
#pragma unroll
  for (int s = 0; s < kMaxEoC; ++s) {
    const int e = unsign(idx[s]);
    at.ld2(b.fk, e, fv[s], kv[s]);
  }
  V sdiv = V(R(0)), ske = V(R(0));
#pragma unroll
  for (int s = 0; s < kMaxEoC; ++s) {
    const V f = idx[s] < 0 ? -fv[s] : fv[s];
    sdiv += s < n ? f : V(R(0));
    ske += s < n ? kv[s] : V(R(0));
  }
  const double A = nt_ld(&m.areaCell[c]);
  const V tend = (-sdiv) / (R)A;
  const V ke = (ske * R(0.25)) / (R)A;
  const V gd = at.ld(b.gd, c);
  const V ghs = at.ld(b.ghs, c);
  at.st(b.E, c, (ke + gd) + ghs);
  at.st(b.tend_gd, c, tend);
  if (b.d_ke) at.st(b.d_ke, c, ke);
  if (b.want_inv) {
    const D g = to_d<R, L::W>(gd);
    r.mass = g * A;
    r.enc = ((g * g) + (2.0 * g) * to_d<R, L::W>(ghs)) * A;
  }
  if (b.k >= 1) {
    if constexpr ((X & kXExt) != 0) {
      (void)rk_update_ext<R, L>(b, at, c, tend, b.gd_base, b.gd_next, b.acc_gd, b.tprev_gd, b.tfirst_gd, b.tcur_gd,
                                A, r.rp, r.rf, r.r44);
    } else {
      if (b.k < 4) at.st(b.gd_next, c, at.ld_s(b.gd_base, c) + (R)b.cdt * tend);
      const V acc = rk_accumulate<R, L>(b, b.acc_gd, at, c, tend);
      if (b.k >= 2) r.rp = (to_d<R, L::W>(at.ld_s(b.tprev_gd, c)) * to_d<R, L::W>(tend)) * A;
      if (b.k == 4) r.r44 = (to_d<R, L::W>(acc) * to_d<R, L::W>(acc)) * A;
      if (b.k < 4) at.st_s(b.tcur_gd, c, tend);
    }
  }
  if (b.nT > 0) {
    int nb[kMaxEoC];
    V uv[kMaxEoC];
    R dvv[kMaxEoC];
#pragma unroll
    for (int s = 0; s < kMaxEoC; ++s) {
      const int e = unsign(idx[s]);
      nb[s] = nt_ld(&m.cocN[s * m.nC + c]);
      uv[s] = at.ld(b.u, e);
      dvv[s] = geo_at<R>(m.dv, m.dv32, e);
    }
    for (int t = 0; t < b.nT; ++t) {
      // adv_scheme_sg11_run: f_edge = c2e(q), flux = u*f_edge, div_operator
      const size_t tb = (size_t)t * m.nC;
      const V q0 = at.ld(b.q, tb + c);
      V qn[kMaxEoC];
#pragma unroll
      for (int s = 0; s < kMaxEoC; ++s) qn[s] = at.ld(b.q, tb + nb[s]);
      V sq = V(R(0));
#pragma unroll
      for (int s = 0; s < kMaxEoC; ++s) {
        // c2e is symmetric in its two cells; at adv_order >= 3 the corrected
        // edge value comes from the edge pass
        V qe = R(0.5) * (q0 + qn[s]);
        if constexpr ((X & kXAdv) != 0)
          if (b.qe) qe = at.ld(b.qe, (size_t)t * m.nE + unsign(idx[s]));
        const V fl = (uv[s] * qe) * dvv[s];
        const V f = idx[s] < 0 ? -fl : fl;
        sq += s < n ? f : V(R(0));
      }
      const V tq = (-sq) / (R)A;
      if (b.d_tq) at.st(b.d_tq, tb + c, tq);
      if (b.k >= 1) {
        if (b.k < 4) at.st(b.q_next, tb + c, at.ld_s(b.q_base, tb + c) + (R)b.cdt * tq);
        (void)rk_accumulate<R, L>(b, b.acc_q, at, tb + c, tq);
      }
    }
  }
  return r;
}

template <typename D>
struct VertSums {
  D ens, av;
};

// calc_pv_on_vertex (operators_mod.F90:262-270)
template <typename R, typename V>
__device__ __forceinline__ V pv_vertex_at(const DMesh &m, int v, V vor, V gdv) {
  return ((R)m.fV[v] + vor) / gdv;
}

template <typename R, typename L>
__device__ __forceinline__ VertSums<val_t<double, L::W>> vertex_e(const DMesh &m, const StageBufs<R> &b, L at,
                                                                  int v) {
  typedef val_t<R, L::W> V;
  typedef val_t<double, L::W> D;
  V s = V(R(0));
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int idx = nt_ld(&m.eovS[j * m.nV + v]);
    const V f = at.ld(b.udc, unsign(idx));
    s += idx < 0 ? -f : f;
  }
  const double At = nt_ld(&m.areaTri[v]);
  const V vor = (-s) / (R)At;
  V g = V(R(0));
#pragma unroll
  for (int j = 0; j < 3; ++j) g = g + at.ld(b.gd, nt_ld(&m.cov[j * m.nV + v])) * (R)nt_ld(&m.kite[j * m.nV + v]);
  const V gdv = g / (R)At;
  const V pv = pv_vertex_at<R, V>(m, v, vor, gdv);
  at.st(b.pv_v, v, pv);
  if (b.d_vor) at.st(b.d_vor, v, vor);
  if (b.d_gdv) at.st(b.d_gdv, v, gdv);
  VertSums<D> r{D(0.0), D(0.0)};
  if (b.want_inv) {
    const D p = to_d<R, L::W>(pv);
    r.ens = (to_d<R, L::W>(gdv) * (p * p)) * At;
    r.av = (to_d<R, L::W>(vor) + m.fV[v]) * At;
  }
  return r;
}

// PV on the edge (calc_pv_on_edge, operators_mod.F90:273-412) from b.pv_v
// and the tangent wind v: scheme 3 (high-order upwind, :359-412) in full;
// schemes 1/2 read pv_cell (b.pv_c) and u, and scheme 2 leaves v2e(pv) here
// and its correction to k_cvor_*.  Padded
// stencil slots point at the element itself with weight 0, so every load is
// issued unconditionally and a padded term adds an exact +0.
template <typename R, typename L, int X = 0>
__device__ __forceinline__ val_t<R, L::W> pv_edge_at(const DMesh &m, const StageBufs<R> &b, L at, int e,
                                                     val_t<R, L::W> v) {
  typedef val_t<R, L::W> V;
  const int v1 = nt_ld(&m.voe[e]), v2 = nt_ld(&m.voe[m.nE + e]);
  const V p1 = at.ld(b.pv_v, v1), p2 = at.ld(b.pv_v, v2);
  V pve = R(0.5) * (p1 + p2);
  if ((X & kXExt) != 0 && b.pv_scheme != 3) {
    // APVM (:296-316) and conservative APVM (:318-357): v2e(pv) minus a
    // multiple of dpv = v dpv/dt-along + u dpv/dn
    const int c1 = m.coe[e], c2 = m.coe[m.nE + e];
    const V dpv = (v * (p2 - p1)) / geo_at<R>(m.dv, m.dv32, e) +
                  (at.ld(b.u, e) * (at.ld(b.pv_c, c2) - at.ld(b.pv_c, c1))) / geo_at<R>(m.dc, m.dc32, e);
    if (b.pv_scheme == 2) {
      at.st(b.dpv, e, dpv);  // eps comes from global sums (k_cvor_*)
      return pve;
    }
    return b.apvm_wdt != 0.0 ? pve - (R)b.apvm_wdt * dpv : pve;
  }
  if (b.pv_order >= 3) {
    const R coef2 = b.pv_order == 3 ? R(1) : R(0);
    // each vertex's stencil is one 48-byte row: three 16-byte loads
    static_assert(kFit2 % 4 == 0, "fit2Vertices rows are whole int4s");
    int f1[kFit2], f2[kFit2];
    const int4 *r1 = reinterpret_cast<const int4 *>(m.fit2V) + (size_t)v1 * (kFit2 / 4);
    const int4 *r2 = reinterpret_cast<const int4 *>(m.fit2V) + (size_t)v2 * (kFit2 / 4);
#pragma unroll
    for (int j = 0; j < kFit2 / 4; ++j) {
      const int4 a1 = r1[j], a2 = r2[j];
      f1[4 * j] = a1.x, f1[4 * j + 1] = a1.y, f1[4 * j + 2] = a1.z, f1[4 * j + 3] = a1.w;
      f2[4 * j] = a2.x, f2[4 * j + 1] = a2.y, f2[4 * j + 2] = a2.z, f2[4 * j + 3] = a2.w;
    }
    V q1[kFit2], q2[kFit2];
#pragma unroll
    for (int j = 0; j < kFit2; ++j) {
      q1[j] = at.ld(b.pv_v, f1[j]);
      q2[j] = at.ld(b.pv_v, f2[j]);
    }
    // deriv2OnVertex, the largest stream of a stage (192 B per edge in FP64):
    // non-temporal, so that edgesOnEdge/weightsOnEdge, which k_edge_tend
    // reads again, stay cached
    R w1[kFit2], w2[kFit2];
#pragma unroll
    for (int j = 0; j < kFit2 / 2; ++j) {
      geo_pair<R, true>(m.d2v, m.d2v32, j, e, m.nE, w1[2 * j], w1[2 * j + 1]);
      geo_pair<R, true>(m.d2v, m.d2v32, kFit2 / 2 + j, e, m.nE, w2[2 * j], w2[2 * j + 1]);
    }
    V d1 = V(R(0)), d2 = V(R(0));
#pragma unroll
    for (int j = 0; j < kFit2; ++j) {
      d1 += w1[j] * (q1[j] - p1);
      d2 += w2[j] * (q2[j] - p2);
    }
    const R dv = geo_at<R, true>(m.dv, m.dv32, e);
    if ((X & kXExt) == 0 || b.pv_order <= 4) {
      pve = pve - ((dv * dv) * ((d1 + d2) - (sign1v<R, L::W>(v) * coef2) * (d2 - d1))) / R(12);
    } else {
      const V q1 = fit_sum<kFit4V, R, L>(m, b.pv_v, 0, at, m.fit4V, m.d4v, m.d4v32, v1, 0, e);
      const V q2 = fit_sum<kFit4V, R, L>(m, b.pv_v, 0, at, m.fit4V, m.d4v, m.d4v32, v2, 1, e);
      const R coef4 = b.pv_order == 6 ? R(0) : R(0.25);
      pve = (pve - ((dv * dv) * (d1 + d2)) / R(12)) +
            (((dv * dv) * (dv * dv)) * ((q1 + q2) - (sign1v<R, L::W>(v) * coef4) * (q2 - q1))) / R(60);
    }
  }
  return pve;
}

// Tangent wind (calc_tangent_wind) + scheme-3 PV on the edge.
template <typename R, typename L, int X = 0>
__device__ __forceinline__ void edge_pv_e(const DMesh &m, const StageBufs<R> &b, L at, int e,
                                          const R *su = nullptr) {
  typedef val_t<R, L::W> V;
  V v = V(R(0));
  {
    V uv[kMaxEoE];
    R w[kMaxEoE];
    if constexpr ((X & kXPvLds) != 0) {
      static_assert(L::W == 1, "the LDS tile holds one member");
#pragma unroll
      for (int p = 0; p < kMaxEoE / 2; ++p) {
        const uint32_t pr = nt_ld(&m.eoeL[(size_t)p * m.nE + e]);
        uv[2 * p] = su[pr & 0xffffu];
        uv[2 * p + 1] = su[pr >> 16];
      }
    } else {
      int ep[kMaxEoE];
      if constexpr ((X & kXEoePv) != 0)
        eoe_row_cells(m, e, m.coe[e], m.coe[m.nE + e], ep);
      else
        eoe_row(m, e, ep);
#pragma unroll
      for (int s = 0; s < kMaxEoE; ++s) uv[s] = at.ld(b.u, ep[s]);
    }
#pragma unroll
    for (int p = 0; p < kMaxEoE / 2; ++p) geo_pair<R>(m.woe, m.woe32, p, e, m.nE, w[2 * p], w[2 * p + 1]);
#pragma unroll
    for (int s = 0; s < kMaxEoE; ++s) v += uv[s] * w[s];
  }
  at.st(b.pv_e, e, pv_edge_at<R, L, X>(m, b, at, e, v));
  if (b.d_v) at.st(b.d_v, e, v);
}

template <typename D>
struct EdgeSums {
  D rp, r44, rf;
};

// calc_u_tend_on_edge (tmcore_swm_mod.F90:155-201, flux_4th_order_correct
// off) from E = ke + gd + ghs in b.E, gd_edge, u, the PV flux s and the edge
// thickness tendency tge: returns u_tend, sets the IAP tendency.
// gde and ue: the edge's own gd_edge and u.
template <typename R, typename L>
__device__ __forceinline__ val_t<R, L::W> u_tend_at(const DMesh &m, const StageBufs<R> &b, L at, int e,
                                                    val_t<R, L::W> s, val_t<R, L::W> tge, val_t<R, L::W> &it,
                                                    val_t<R, L::W> gde, val_t<R, L::W> ue) {
  typedef val_t<R, L::W> V;
  const int c1 = nt_ld(&m.coe[e]), c2 = nt_ld(&m.coe[m.nE + e]);
  const V dEdx = (at.ld(b.E, c2) - at.ld(b.E, c1)) / geo_at<R, true>(m.dc, m.dc32, e);
  const V ut = s - dEdx;
  const V sg = tm_sqrt<R, L::W>(gde);
  it = sg * ut + ((R(0.5) * ue) / sg) * tge;
  return ut;
}
template <typename R, typename L>
__device__ __forceinline__ val_t<R, L::W> u_tend_at(const DMesh &m, const StageBufs<R> &b, L at, int e,
                                                    val_t<R, L::W> s, val_t<R, L::W> tge, val_t<R, L::W> &it) {
  return u_tend_at<R, L>(m, b, at, e, s, tge, it, at.ld(b.gd_e, e), at.ld(b.u, e));
}

// The tile's staged rows (kXTendLds): u, gd_edge, pv_edge at local rows, the
// thread's own edge at row tid.
template <typename R>
struct TendTile {
  const R *u, *gd_e, *pv_e;
  int tid;
};

// TRiSK PV flux, edge gradient, IAP tendency and the RK update of iap_u.
template <typename R, typename L, int X = 0>
__device__ __forceinline__ EdgeSums<val_t<double, L::W>> edge_tend_e(const DMesh &m, const StageBufs<R> &b, L at,
                                                                     int e, TendTile<R> tl = {}) {
  typedef val_t<R, L::W> V;
  typedef val_t<double, L::W> D;
  constexpr bool kLds = (X & kXTendLds) != 0;
  static_assert(!kLds || L::W == 1, "the LDS tile holds one member");
  V pve, uv[kMaxEoE], gv[kMaxEoE], pv[kMaxEoE], gde_own, ue_own;
  R w[kMaxEoE];
  if constexpr (kLds) {
    pve = tl.pv_e[tl.tid];
    gde_own = tl.gd_e[tl.tid];
    ue_own = tl.u[tl.tid];
#pragma unroll
    for (int p = 0; p < kMaxEoE / 2; ++p) {
      const uint32_t pr = nt_ld(&m.eoeL[(size_t)p * m.nE + e]);
      const int l0 = (int)(pr & 0xffffu), l1 = (int)(pr >> 16);
      uv[2 * p] = tl.u[l0], gv[2 * p] = tl.gd_e[l0], pv[2 * p] = tl.pv_e[l0];
      uv[2 * p + 1] = tl.u[l1], gv[2 * p + 1] = tl.gd_e[l1], pv[2 * p + 1] = tl.pv_e[l1];
    }
  } else {
    pve = at.ld(b.pv_e, e);
    int ep[kMaxEoE];
    if constexpr ((X & kXEoeTend) != 0)
      eoe_row_cells(m, e, nt_ld(&m.coe[e]), nt_ld(&m.coe[m.nE + e]), ep);
    else
      eoe_row<true>(m, e, ep);
#pragma unroll
    for (int q = 0; q < kMaxEoE; ++q) {
      uv[q] = at.ld(b.u, ep[q]);
      gv[q] = at.ld(b.gd_e, ep[q]);
      pv[q] = at.ld(b.pv_e, ep[q]);
    }
    gde_own = at.ld(b.gd_e, e);
    ue_own = at.ld(b.u, e);
  }
#pragma unroll
  for (int p = 0; p < kMaxEoE / 2; ++p) geo_pair<R, true>(m.woe, m.woe32, p, e, m.nE, w[2 * p], w[2 * p + 1]);
  V s = V(R(0));
#pragma unroll
  for (int q = 0; q < kMaxEoE; ++q) s = s + (((w[q] * uv[q]) * gv[q]) * R(0.5)) * (pve + pv[q]);
  const int c1 = nt_ld(&m.coe[e]), c2 = nt_ld(&m.coe[m.nE + e]);
  const V tge = upwind_cell_e<R, L, X>(m, b, at, b.tend_gd, 0, e,
                                       R(0.5) * (at.ld(b.tend_gd, c1) + at.ld(b.tend_gd, c2)),
                                       sign1v<R, L::W>(at.ld(b.iap, e)));
  V it;
  const V ut = u_tend_at<R, L>(m, b, at, e, s, tge, it, gde_own, ue_own);
  if (b.d_pvflx) at.st(b.d_pvflx, e, s);
  if (b.d_utend) at.st(b.d_utend, e, ut);
  if (b.d_tiap) at.st(b.d_tiap, e, it);
  EdgeSums<D> r{D(0.0), D(0.0), D(0.0)};
  if (b.k >= 1) {
    const double aE = nt_ld(&m.aE[e]);
    if constexpr ((X & kXExt) != 0) {
      (void)rk_update_ext<R, L>(b, at, e, it, b.iap_base, b.iap_next, b.acc_iap, b.tprev_iap, b.tfirst_iap,
                                b.tcur_iap, aE, r.rp, r.rf, r.r44);
    } else {
      if (b.k < 4) {
        const V xn = at.ld_s(b.iap_base, e) + (R)b.cdt * it;
        at.st(b.iap_next, e, xn);
        if constexpr ((X & kXTendPrep) != 0) {
          if (b.prep_next) {
            // edge_prep_e of the next stage: its gd is gd_next (written by
            // this stage's k_cell_vertex), its iap_u is xn
            const V gde = R(0.5) * (at.ld(b.gd_next, c1) + at.ld(b.gd_next, c2));
            const V un = xn / tm_sqrt<R, L::W>(gde);
            at.st(b.gd_e_n, e, gde);
            at.st(b.u_n, e, un);
            at.st2(b.fk, e, (un * gde) * geo_at<R>(m.dv, m.dv32, e), (un * un) * geo_at<R>(m.aE, m.aE32, e));
            at.st(b.udc, e, un * geo_at<R>(m.dc, m.dc32, e));
          }
        }
      }
      const V acc = rk_accumulate<R, L>(b, b.acc_iap, at, e, it);
      if (b.k >= 2) r.rp = (to_d<R, L::W>(at.ld_s(b.tprev_iap, e)) * to_d<R, L::W>(it)) * aE;
      if (b.k == 4) r.r44 = (to_d<R, L::W>(acc) * to_d<R, L::W>(acc)) * aE;
      if (b.k < 4) at.st_s(b.tcur_iap, e, it);
    }
  }
  return r;
}

// ---------------------------------------------------------------- kernels
template <typename R, int X>
__global__ __launch_bounds__(kBlock) void k_edge_prep(DMesh m, StageBufs<R> b) {
  const int tile = xcd_tile(blockIdx.x, m.nbE, b.rev);
  const int e = tile * kBlock + threadIdx.x;
  double en = 0.0;
  if (e < m.nE) en = edge_prep_e<R, Solo, X>(m, b, Solo{}, e);
  if (b.want_inv) {
    const double v[1] = {en};
    const int ids[1] = {kEnergyE};
    block_sums<1>(v, ids, b.part, b.part_stride, tile);
  }
}

// Block partials of the stage's inner products in the kXExt kernels: the
// descriptor's sums, edge ids + off (0 edges, 1 cells).
template <typename R>
__device__ __forceinline__ void rk_sums_ext(const StageBufs<R> &b, double rp, double rf, double r44, int off,
                                            int tile) {
  double v[3];
  int ids[3], n = 0;
  if (b.rid_prev >= 0) v[n] = rp, ids[n++] = b.rid_prev + off;
  if (b.rid_first >= 0) v[n] = rf, ids[n++] = b.rid_first + off;
  if (b.ts_last) v[n] = r44, ids[n++] = kR44e + off;
  if (n == 1) {
    const double v1[1] = {v[0]};
    const int i1[1] = {ids[0]};
    block_sums<1>(v1, i1, b.part, b.part_stride, tile);
  } else if (n == 2) {
    const double v2[2] = {v[0], v[1]};
    const int i2[2] = {ids[0], ids[1]};
    block_sums<2>(v2, i2, b.part, b.part_stride, tile);
  } else if (n == 3) {
    const double v3[3] = {v[0], v[1], v[2]};
    const int i3[3] = {ids[0], ids[1], ids[2]};
    block_sums<3>(v3, i3, b.part, b.part_stride, tile);
  }
}

template <typename R, int X>
__device__ __forceinline__ void cell_part(const DMesh &m, const StageBufs<R> &b, int tile) {
  const int c = tile * kBlock + threadIdx.x;
  CellSums<double> r{0.0, 0.0, 0.0, 0.0, 0.0};
  if (c < m.nC) r = cell_e<R, Solo, X>(m, b, Solo{}, c);
  if (b.want_inv) {
    const double v[2] = {r.mass, r.enc};
    const int ids[2] = {kMass, kEnergyC};
    block_sums<2>(v, ids, b.part, b.part_stride, tile);
  }
  if constexpr ((X & kXExt) != 0) {
    if (b.k >= 1) rk_sums_ext<R>(b, r.rp, r.rf, r.r44, 1, tile);
  } else if (b.k >= 2) {
    const int rid = b.k == 2 ? kR12c : (b.k == 3 ? kR23c : kR34c);
    if (b.k == 4) {
      const double v[2] = {r.rp, r.r44};
      const int ids[2] = {rid, kR44c};
      block_sums<2>(v, ids, b.part, b.part_stride, tile);
    } else {
      const double v[1] = {r.rp};
      const int ids[1] = {rid};
      block_sums<1>(v, ids, b.part, b.part_stride, tile);
    }
  }
}

template <typename R>
__device__ __forceinline__ void vertex_part(const DMesh &m, const StageBufs<R> &b, int tile) {
  const int v = tile * kBlock + threadIdx.x;
  VertSums<double> r{0.0, 0.0};
  if (v < m.nV) r = vertex_e<R, Solo>(m, b, Solo{}, v);
  if (b.want_inv) {
    const double vv[2] = {r.ens, r.av};
    const int ids[2] = {kEnstrophy, kAbsVor};
    block_sums<2>(vv, ids, b.part, b.part_stride, tile);
  }
}

template <typename R, int X>
__global__ __launch_bounds__(kBlock) void k_cell_vertex(DMesh m, StageBufs<R> b) {
  if ((int)blockIdx.x < m.nbC)
    cell_part<R, X>(m, b, xcd_tile(blockIdx.x, m.nbC, b.rev));
  else
    vertex_part<R>(m, b, xcd_tile(blockIdx.x - m.nbC, m.nbV, b.rev));
}

template <typename R, int X>
__global__ __launch_bounds__(kBlock) void k_edge_pv(DMesh m, StageBufs<R> b) {
  const int tile = xcd_tile(blockIdx.x, m.nbE, b.rev);
  const int e = tile * kBlock + threadIdx.x;
  if constexpr ((X & kXPvLds) != 0) {
    // u of the tile's edges, then of its halo (the tendency kernel's tiles)
    __shared__ R su[kTileE + kTendHaloMax];
    if (e < m.nE) su[threadIdx.x] = b.u[e];
    const int h0 = m.haloOff[tile], h1 = m.haloOff[tile + 1];
    for (int h = h0 + (int)threadIdx.x; h < h1; h += kBlock) su[kTileE + h - h0] = b.u[m.haloE[h]];
    __syncthreads();
    if (e < m.nE) edge_pv_e<R, Solo, X>(m, b, Solo{}, e, su);
  } else {
    if (e < m.nE) edge_pv_e<R, Solo, X>(m, b, Solo{}, e);
  }
}

template <typename R, int X>
__global__ __launch_bounds__(kBlock) void k_edge_tend(DMesh m, StageBufs<R> b) {
  const int tile = xcd_tile(blockIdx.x, m.nbE, b.rev);
  const int e = tile * kBlock + threadIdx.x;
  EdgeSums<double> r{0.0, 0.0, 0.0};
  if constexpr ((X & kXTendLds) != 0) {
    static_assert(kBlock == kTileE, "a tendency tile is one workgroup");
    // the tile's rows, then its halo's, of the three gathered edge fields
    __shared__ R su[kTileE + kTendHaloMax], sg[kTileE + kTendHaloMax], sp[kTileE + kTendHaloMax];
    if (e < m.nE) {
      su[threadIdx.x] = b.u[e];
      sg[threadIdx.x] = b.gd_e[e];
      sp[threadIdx.x] = b.pv_e[e];
    }
    const int h0 = m.haloOff[tile], h1 = m.haloOff[tile + 1];
    for (int h = h0 + (int)threadIdx.x; h < h1; h += kBlock) {
      const int g = m.haloE[h];
      su[kTileE + h - h0] = b.u[g];
      sg[kTileE + h - h0] = b.gd_e[g];
      sp[kTileE + h - h0] = b.pv_e[g];
    }
    __syncthreads();
    if (e < m.nE) r = edge_tend_e<R, Solo, X>(m, b, Solo{}, e, TendTile<R>{su, sg, sp, (int)threadIdx.x});
  } else {
    if (e < m.nE) r = edge_tend_e<R, Solo, X>(m, b, Solo{}, e);
  }
  if constexpr ((X & kXExt) != 0) {
    if (b.k >= 1) rk_sums_ext<R>(b, r.rp, r.rf, r.r44, 0, tile);
  } else if (b.k >= 2) {
    const int rid = b.k == 2 ? kR12e : (b.k == 3 ? kR23e : kR34e);
    if (b.k == 4) {
      const double v[2] = {r.rp, r.r44};
      const int ids[2] = {rid, kR44e};
      block_sums<2>(v, ids, b.part, b.part_stride, tile);
    } else {
      const double v[1] = {r.rp};
      const int ids[1] = {rid};
      block_sums<1>(v, ids, b.part, b.part_stride, tile);
    }
  }
}

// ------------------------------------------------ pv_scheme 1 and 2 (FP R)
// pv_cell = scalar_v2c(pv_vertex) (operators_mod.F90:125-141; the reference's
// scatter over cellsOnVertex, as a gather over verticesOnCell) and, for scheme
// 2, gd_tend_vertex = scalar_c2v(tend_gd) (tmcore_swm_mod.F90:109).  Blocks
// [0, nbC) do cells, the rest vertices.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_pv_cell(DMesh m, StageBufs<R> b) {
  if ((int)blockIdx.x < m.nbC) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= m.nC) return;
    const R A = (R)m.areaCell[c];
    R s = R(0);
    for (int q = 0; q < m.nEoC[c]; ++q) s = s + (b.pv_v[m.vocS[q * m.nC + c]] * (R)m.kiteC[q * m.nC + c]) / A;
    b.pv_c[c] = s;
  } else {
    const int v = (blockIdx.x - m.nbC) * kBlock + threadIdx.x;
    if (v >= m.nV) return;
    R g = R(0);
#pragma unroll
    for (int j = 0; j < 3; ++j) g = g + b.tend_gd[m.cov[j * m.nV + v]] * (R)m.kite[j * m.nV + v];
    b.tgd_v[v] = g / (R)m.areaTri[v];
  }
}

// Conservative APVM (operators_mod.F90:318-357), after k_edge_pv left
// pv_e = v2e(pv) and dpv: the TRiSK fluxes of both (calc_tangent_vor_flux,
// :239-260) ...
template <typename R>
__global__ __launch_bounds__(kBlock) void k_cvor_flux(DMesh m, StageBufs<R> b) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= m.nE) return;
  const R p0 = b.pv_e[e], d0 = b.dpv[e];
  R s = R(0), sd = R(0);
#pragma unroll
  for (int q = 0; q < kMaxEoE; ++q) {
    const int ep = m.eoe[pair_at(q, e, m.nE)];
    const R f = ((geo_slot<R>(m.woe, m.woe32, q, e, m.nE) * b.u[ep]) * b.gd_e[ep]) * R(0.5);
    s = s + f * (p0 + b.pv_e[ep]);
    sd = sd + f * (d0 + b.dpv[ep]);
  }
  b.pvf[e] = s;
  b.dpvf[e] = sd;
}

// ... their vertex curls (calc_vor_tend_on_vertex, :414-421) and the block
// partials of sum(pv^2 gd_tend_v A), sum(pv vor_tend A), sum(pv dvor_tend A) ...
template <typename R>
__global__ __launch_bounds__(kBlock) void k_cvor_sums(DMesh m, StageBufs<R> b) {
  const int v = blockIdx.x * kBlock + threadIdx.x;
  double t[3] = {0.0, 0.0, 0.0};
  if (v < m.nV) {
    R s = R(0), sd = R(0);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int idx = m.eovS[j * m.nV + v], e = unsign(idx);
      const R dc = geo_at<R>(m.dc, m.dc32, e);
      const R f = b.pvf[e] * dc, fd = b.dpvf[e] * dc;
      s += idx < 0 ? -f : f;
      sd += idx < 0 ? -fd : fd;
    }
    const double At = m.areaTri[v];
    const double vt = (double)((-s) / (R)At), dvt = (double)((-sd) / (R)At);
    const double p = (double)b.pv_v[v];
    t[0] = ((p * p) * (double)b.tgd_v[v]) * At;
    t[1] = (p * vt) * At;
    t[2] = (p * dvt) * At;
  }
  const int ids[3] = {0, 1, 2};
  block_sums<3>(t, ids, b.cv_part, m.nbV, blockIdx.x);
}

// ... eps = -(s1 - 2 s2) / (2 dt s3) from the double-double partials (one block) ...
template <typename R>
__global__ __launch_bounds__(kBlock) void k_cvor_eps(DMesh m, StageBufs<R> b) {
  __shared__ double2 sh[3][kBlock / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double2 s[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double2 a = make_double2(0.0, 0.0);
    for (int k = threadIdx.x; k < m.nbV; k += kBlock) a = dd_add(a, b.cv_part[i * m.nbV + k]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
      a = dd_add(a, make_double2(__shfl_xor(a.x, off, 64), __shfl_xor(a.y, off, 64)));
    if (lane == 0) sh[i][w] = a;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      s[i] = sh[i][0];
#pragma unroll
      for (int q = 1; q < kBlock / 64; ++q) s[i] = dd_add(s[i], sh[i][q]);
    }
    const double s1 = s[0].x + s[0].y, s2 = s[1].x + s[1].y, s3 = s[2].x + s[2].y;
    b.cv_eps[0] = -(s1 - s2 * 2.0) / ((2.0 * b.dt) * s3);
  }
}

// ... and pv_edge = pv_edge - eps dt dpv.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_cvor_apply(DMesh m, StageBufs<R> b) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= m.nE) return;
  const R f = (R)(b.cv_eps[0] * b.dt);
  b.pv_e[e] = b.pv_e[e] - f * b.dpv[e];
}

// Global sums: block i of k_reduce folds the block partials of sum i
// (double-double) into sums[i].  beta and the invariants are formed from
// sums[] by whoever consumes them (k_finalize, k_inv_out), so no extra launch
// sits between the last stage and the state update.
constexpr int kReduceBlock = 512;

struct ReduceArgs {
  const double2 *part;
  int stride, nbE, nbC, nbV;
  int mode;      // 1: end of a CRK4 step (beta sums); 0: invariants only
  int have_inv;  // the stage-1 partials hold the invariants of the base state
  double *sums;  // [kNumSums]
};

static __global__ __launch_bounds__(kReduceBlock) void k_reduce(ReduceArgs a) {
  __shared__ double2 sh[kReduceBlock / 64];
  const int i = blockIdx.x;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const bool edge = i == kR12e || i == kR23e || i == kR34e || i == kR44e || i == kEnergyE;
  const bool vert = i == kEnstrophy || i == kAbsVor;
  const bool is_inv = i >= kMass;
  const int nb = edge ? a.nbE : (vert ? a.nbV : a.nbC);
  double2 acc = make_double2(0.0, 0.0);
  if ((is_inv && a.have_inv) || (!is_inv && a.mode == 1)) {
    // partials j = tid, tid + block, ... added in that order; loaded eight
    // at a time so the loads overlap instead of each waiting for the add
    constexpr int kB = 8;
    for (int j0 = threadIdx.x; j0 < nb; j0 += kB * kReduceBlock) {
      double2 v[kB];
#pragma unroll
      for (int k = 0; k < kB; ++k) {
        const int j = j0 + k * kReduceBlock;
        v[k] = j < nb ? a.part[i * a.stride + j] : make_double2(0.0, 0.0);
      }
#pragma unroll
      for (int k = 0; k < kB; ++k)
        if (j0 + k * kReduceBlock < nb) acc = dd_add(acc, v[k]);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    double2 o = make_double2(__shfl_xor(acc.x, off, 64), __shfl_xor(acc.y, off, 64));
    acc = dd_add(acc, o);
  }
  if (lane == 0) sh[w] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double2 t = sh[0];
    for (int q = 1; q < kReduceBlock / 64; ++q) t = dd_add(t, sh[q]);
    a.sums[i] = t.x + t.y;
  }
}

struct StepOut {
  const double *sums;
  double inv_total_area, inv_sum_area_tri;
  int conserve_energy;
  int time_scheme;  // 1 CRK3, 2 CRK4, 3 PC2
  int have_inv;
  int *hist_idx;      // device step counter of the current batch
  double *beta_hist;  // [cap]
  double *inv_hist;   // [cap][4]
  int hist_cap;
};

// inner_product_tend (operators_mod.F90:187-194) and beta of CRK4
// (time_scheme_mod.F90:140-150), CRK3 and PC2.
__device__ __forceinline__ double crk4_beta(const StepOut &o) {
  if (!o.conserve_energy) return 1.0;
  const double *r = o.sums;
  const double R44 = (r[kR44e] + r[kR44c]) * o.inv_total_area;
  const double R12 = (r[kR12e] + r[kR12c]) * o.inv_total_area;
  const double R23 = (r[kR23e] + r[kR23c]) * o.inv_total_area;
  const double R34 = (r[kR34e] + r[kR34c]) * o.inv_total_area;
  if (o.time_scheme == 1)  // CRK3 (time_scheme_mod.F90:94-100): R13 is held in the R34 slots
    return ((2.0 * R12 - R34) + 2.0 * R23) / (3.0 * R44);
  if (o.time_scheme == 3)  // PC2 (:174-177): ip1 in the R34 slots, ip2 in R44
    return (R34 != 0.0 && R44 != 0.0) ? R34 / R44 : 1.0;
  return (R12 + R23 + R34) / (3.0 * R44);
}

// calc_total_mass/energy/potential_enstropy/absolute_vorticity
// (tmcore_swm_mod.F90:215-255) into inv_hist[slot].
__device__ __forceinline__ void write_invariants(const StepOut &o, int slot) {
  if (slot < 0 || slot >= o.hist_cap) return;
  const double *r = o.sums;
  double *d = o.inv_hist + 4 * (size_t)slot;
  d[0] = r[kMass];
  d[1] = r[kEnergyE] + r[kEnergyC];
  d[2] = r[kEnstrophy];
  d[3] = r[kAbsVor] * o.inv_sum_area_tri;
}

// Invariants-only pass (the post-loop refresh): slot hist_idx - 1.
static __global__ void k_inv_out(StepOut o) {
  if (threadIdx.x == 0 && blockIdx.x == 0) write_invariants(o, *o.hist_idx - 1);
}

template <typename R>
struct FinalBufs {
  R *gd, *iap, *q;
  const R *acc_gd, *acc_iap, *acc_q;
  double dt;
  int nT;
  StepOut out;
};

// update_state(beta*dt, tend(new), state(old), state(new)) (time_scheme_mod.F90:152);
// tracers with beta = 1 (tmcore_adv_mod.F90:175 with conserve_energy off).
// Every block forms beta from the same sums; block 0 records the history.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_finalize(DMesh m, FinalBufs<R> f) {
  __shared__ double sbeta;
  // What does not need beta goes before the wait for it: each thread's loads,
  // and the tracer update (beta = 1).
  const bool edges = (int)blockIdx.x < m.nbE;
  const int i = edges ? xcd_tile(blockIdx.x, m.nbE) * kBlock + threadIdx.x
                      : xcd_tile(blockIdx.x - m.nbE, m.nbC) * kBlock + threadIdx.x;
  const bool live = i < (edges ? m.nE : m.nC);
  R s0 = R(0), a0 = R(0);
  if (live) {
    s0 = edges ? f.iap[i] : f.gd[i];
    a0 = edges ? f.acc_iap[i] : f.acc_gd[i];
  }
  if (!edges && live)
    for (int t = 0; t < f.nT; ++t) {
      const size_t o = (size_t)t * m.nC + i;
      f.q[o] = f.q[o] + (R)f.dt * f.acc_q[o];
    }
  if (threadIdx.x == 0) {
    const double beta = crk4_beta(f.out);
    sbeta = beta;
    if (blockIdx.x == 0) {
      const int hi = *f.out.hist_idx;
      if (hi < f.out.hist_cap) f.out.beta_hist[hi] = beta;
      if (f.out.have_inv) write_invariants(f.out, hi - 1);
      *f.out.hist_idx = hi + 1;
    }
  }
  __syncthreads();
  const double bdt = sbeta * f.dt;
  if (live) (edges ? f.iap : f.gd)[i] = s0 + (R)bdt * a0;
}

// iap_u = sqrt(c2e(gd) + upwind correction) * u (tmcore_swm_mod.F90:62-64).
// The reference applies the correction with iap_u before it is set; with a
// zeroed array that is sign(1, 0) = +1 on every edge, which b.adv_* and the
// +1 below reproduce.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_iap_from_u(DMesh m, StageBufs<R> b, const R *gd, const R *u, R *iap) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= m.nE) return;
  const R gde =
      upwind_cell_e<R, Solo, kXAdv | kXAdv56>(m, b, Solo{}, gd, 0, e, R(0.5) * (gd[m.coe[e]] + gd[m.coe[m.nE + e]]), R(1));
  iap[e] = tm_sqrt(gde) * u[e];
}

static __global__ void k_reset_counter(int *p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}

}  // namespace tmk
