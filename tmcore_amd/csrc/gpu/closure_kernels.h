// Learned-closure kernels on the TRiSK mesh (new numbering): the 40 MLP
// inputs of an edge, the state correction built from the MLP outputs, and
// the training loss with its gradient with respect to the outputs.
//
// Every edge is seen in a canonical frame: kappa(e) = +1 when the mesh normal
// cellsOnEdge(1) -> (2) points east (north if exactly meridional), else -1.
// Velocities and both outputs are taken along the canonical normal
// kappa(e) * n(e), and the input slots run over the west cell's other edges
// first (counter-clockwise from e), then the east cell's, so the inputs and
// the correction do not depend on the mesh's edge orientation.
// Inputs of edge e, j = field * 10 + slot, slot 0 = e, slot s = the s-th
// edge of that list (of up to 10, the last one is not used):
//   field 0  kappa(e) * sigma(e, e') * u(e')        normal velocity
//   field 1  gd_edge(e')                             thickness (g h)
//   field 2  (vor(v1(e')) + vor(v2(e'))) / 2         relative vorticity
//   field 3  (div(c1(e')) + div(c2(e'))) / 2         velocity divergence
// vor is the vertex curl of the stage (curl_operator), div the cell
// divergence of u, (1/A_c) sum_i nSign(i,c) u(e_i) dv(e_i).
// sigma(e, e') = nSign(e, c) * nSign(e', c) with c the cell e and e' share,
// nSign(e, c) = +1 when c = cellsOnEdge(1, e).
#pragma once

#include "mlp_dev.h"
#include "swm_kernels.h"

namespace tmk {

constexpr int kClSlots = 10;

__device__ __forceinline__ float cl_orient(const DMesh &m, int e, int j) {
  const int a = m.coe[e], b = m.coe[m.nE + e];
  const int c = m.coe[j], d = m.coe[m.nE + j];
  const bool on_a = (a == c) || (a == d);
  const int shared = on_a ? a : b;
  return (on_a ? 1.f : -1.f) * (shared == c ? 1.f : -1.f);
}

template <typename R>
struct ClFeatArgs {
  const R *u, *gd_e, *vor, *div;       // stage-1 (start of step) fields
  const float *norm;                   // [kNormLen] shift, scale, out scales
  float *X;                            // [nE][40]
  R *sq;                               // [nE] sqrt(gd_edge), or null
};

// Edge of canonical slot s >= 1 of edge e: edgesOnEdge lists cellsOnEdge(1)'s
// other edges (nEdgesOnCell - 1 of them) and then cellsOnEdge(2)'s; with
// kappa = -1 the second cell is the west one and its group comes first.
__device__ __forceinline__ int cl_slot_edge(const DMesh &m, int e, int kap, int n1, int n, int s) {
  const int q = s - 1;
  if (q >= n) return e;
  const int slot = kap > 0 ? q : (q + n1) % n;
  return m.eoe[pair_at(slot, e, m.nE)];
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_cl_features(DMesh m, ClFeatArgs<R> a) {
  const int e = xcd_tile(blockIdx.x, m.nbE) * kBlock + threadIdx.x;
  if (e >= m.nE) return;
  const int kap = m.kapE[e];
  const int n = m.nEoE[e], n1 = m.nEoC[m.coe[e]] - 1;
  int nb[kClSlots];
  nb[0] = e;
#pragma unroll
  for (int s = 1; s < kClSlots; ++s) nb[s] = cl_slot_edge(m, e, kap, n1, n, s);
  float f[kMlpIn];
#pragma unroll
  for (int s = 0; s < kClSlots; ++s) {
    const int j = nb[s];
    const int c1 = m.coe[j], c2 = m.coe[m.nE + j];
    const int v1 = m.voe[j], v2 = m.voe[m.nE + j];
    f[s] = ((float)kap * cl_orient(m, e, j)) * (float)a.u[j];
    f[kClSlots + s] = (float)((double)a.gd_e[j] - (double)a.norm[kNormShift + kClSlots + s]);
    f[2 * kClSlots + s] = (float)(R(0.5) * (a.vor[v1] + a.vor[v2]));
    f[3 * kClSlots + s] = (float)(R(0.5) * (a.div[c1] + a.div[c2]));
  }
  float4 *out = reinterpret_cast<float4 *>(a.X + (size_t)e * kMlpIn);
#pragma unroll
  for (int q = 0; q < kMlpIn / 4; ++q) {
    float v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int j = 4 * q + i;
      // gd_edge had its (large) shift removed in FP64 above
      const float sh = (j >= kClSlots && j < 2 * kClSlots) ? 0.f : a.norm[kNormShift + j];
      v[i] = (f[j] - sh) * a.norm[kNormScale + j];
    }
    out[q] = make_float4(v[0], v[1], v[2], v[3]);
  }
  if (a.sq) a.sq[e] = tm_sqrt(a.gd_e[e]);
}

// Velocity divergence of the cells, (1/A_c) sum_i nSign(i,c) u(e_i) dv(e_i).
template <typename R>
__global__ __launch_bounds__(kBlock) void k_cl_div(DMesh m, const R *__restrict__ u, R *__restrict__ div) {
  const int c = xcd_tile(blockIdx.x, m.nbC) * kBlock + threadIdx.x;
  if (c >= m.nC) return;
  const int n = m.nEoC[c];
  R s = R(0);
#pragma unroll
  for (int q = 0; q < kMaxEoC; ++q) {
    const int idx = m.eocS[q * m.nC + c];
    const int e = unsign(idx);
    const R f = u[e] * (R)m.dv[e];
    s += q < n ? (idx < 0 ? -f : f) : R(0);
  }
  div[c] = s / (R)m.areaCell[c];
}

template <typename R>
struct ClApplyArgs {
  const float *Y;     // [nE][2] MLP outputs of the start-of-step state
  const R *sq;        // [nE] sqrt(gd_edge) of the start-of-step state
  const float *norm;  // out scales at kNormOut
  double dt;
  R *iap, *gd;        // state after the CRK4 update, corrected in place
};

// Thickness tendency of the closure flux at cell c:
// P_c = -(1/A_c) sum_i nSign(i,c) * F(e_i) * dv(e_i), with the flux along the
// mesh normal F(e) = kappa(e) * out_scale[1] * y1(e).
template <typename R>
__device__ __forceinline__ R cl_flux_div(const DMesh &m, const float *Y, float s1, int c) {
  const int n = m.nEoC[c];
  R s = R(0);
#pragma unroll
  for (int q = 0; q < kMaxEoC; ++q) {
    const int idx = m.eocS[q * m.nC + c];
    const int e = unsign(idx);
    const R F = (R)((float)m.kapE[e] * Y[2 * (size_t)e + 1]) * (R)m.dv[e];
    s += q < n ? (idx < 0 ? -F : F) : R(0);
  }
  return ((-s) * (R)s1) / (R)m.areaCell[c];
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_cl_apply(DMesh m, ClApplyArgs<R> a) {
  const int b = blockIdx.x;
  if (b < m.nbE) {
    const int e = xcd_tile(b, m.nbE) * kBlock + threadIdx.x;
    if (e >= m.nE) return;
    const R du = (R)(((float)m.kapE[e] * a.norm[kNormOut]) * a.Y[2 * (size_t)e]);
    a.iap[e] = a.iap[e] + ((R)a.dt * a.sq[e]) * du;
  } else {
    const int c = xcd_tile(b - m.nbE, m.nbC) * kBlock + threadIdx.x;
    if (c >= m.nC) return;
    a.gd[c] = a.gd[c] + (R)a.dt * cl_flux_div<R>(m, a.Y, a.norm[kNormOut + 1], c);
  }
}

struct ClLossArgs {
  const float *Y;     // [nE][2]
  const float *tu;    // [nE] velocity target (m s^-2)
  const float *tg;    // [nC] thickness target (m^2 s^-3)
  const float *norm;
  float tgs;          // thickness loss scale
  float *r;           // [nC] normalised thickness residual
  float *dY;          // [nE][2] dL/dY, or null (loss only)
  double *part;       // [nbE + nbC] per-block loss partials (velocity, then thickness)
  double *loss;       // [2] += velocity and thickness loss terms (k_cl_loss_sum)
};

// The block's sum of v into its own partial slot (no atomics: the loss is
// summed in a fixed order and is the same from run to run).
__device__ __forceinline__ void cl_block_part(double v, double *slot) {
  __shared__ double sh[kBlock / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if (lane == 0) sh[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < kBlock / 64; ++q) s += sh[q];
    *slot = s;
  }
}

__global__ __launch_bounds__(kBlock) void k_cl_resid(DMesh m, ClLossArgs a) {
  const int c = xcd_tile(blockIdx.x, m.nbC) * kBlock + threadIdx.x;
  double l = 0.0;
  if (c < m.nC) {
    const float P = cl_flux_div<float>(m, a.Y, a.norm[kNormOut + 1], c);
    const float r = (P - a.tg[c]) / a.tgs;
    a.r[c] = r;
    l = (double)r * (double)r / (double)m.nC;
  }
  cl_block_part(l, a.part + m.nbE + blockIdx.x);
}

// Needs k_cl_resid of the same outputs first (reads r).
__global__ __launch_bounds__(kBlock) void k_cl_dy(DMesh m, ClLossArgs a) {
  const int e = xcd_tile(blockIdx.x, m.nbE) * kBlock + threadIdx.x;
  double l = 0.0;
  if (e < m.nE) {
    const float kap = (float)m.kapE[e];
    const float d = a.Y[2 * (size_t)e] - kap * (a.tu[e] / a.norm[kNormOut]);
    l = (double)d * (double)d / (double)m.nE;
    if (a.dY) {
      const int c1 = m.coe[e], c2 = m.coe[m.nE + e];
      // dP_c1/dy1 = -kappa s1 dv / A_c1 (nSign = +1), dP_c2/dy1 = +kappa s1 dv / A_c2
      const float g = kap * (2.f * a.norm[kNormOut + 1] * (float)m.dv[e] / ((float)m.nC * a.tgs));
      a.dY[2 * (size_t)e] = 2.f * d / (float)m.nE;
      a.dY[2 * (size_t)e + 1] = g * (a.r[c2] / (float)m.areaCell[c2] - a.r[c1] / (float)m.areaCell[c1]);
    }
  }
  cl_block_part(l, a.part + blockIdx.x);
}

// loss[0] += the velocity partials, loss[1] += the thickness partials, each
// in block order (one wave; lanes take fixed strides, then a fixed tree).
__global__ void k_cl_loss_sum(DMesh m, const double *part, double *loss) {
  const int lane = threadIdx.x;
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int n = t == 0 ? m.nbE : m.nbC;
    const double *p = part + (t == 0 ? 0 : m.nbE);
    double s = 0.0;
    for (int i = lane; i < n; i += 64) s += p[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) loss[t] += s;
  }
}

}  // namespace tmk
