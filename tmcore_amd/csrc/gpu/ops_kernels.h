// Single-operator entry points (test path): each runs one reference operator
// (operators_mod.F90) on the device in FP64 over the context's renumbered mesh,
// with a host round trip per call.  Included at the end of tmgpu.hip.
#pragma once

namespace tmk {

__global__ __launch_bounds__(kBlock) void k_op_div(DMesh m, const double *f, double *out) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= m.nC) return;
  double s = 0.0;
  for (int q = 0; q < m.nEoC[c]; ++q) {
    const int idx = m.eocS[q * m.nC + c], e = unsign(idx);
    const double t = f[e] * m.dv[e];
    s += idx < 0 ? -t : t;
  }
  out[c] = (-s) / m.areaCell[c];
}

__global__ __launch_bounds__(kBlock) void k_op_curl(DMesh m, const double *f, double *out) {
  const int v = blockIdx.x * kBlock + threadIdx.x;
  if (v >= m.nV) return;
  double s = 0.0;
  for (int j = 0; j < 3; ++j) {
    const int idx = m.eovS[j * m.nV + v], e = unsign(idx);
    const double t = f[e] * m.dc[e];
    s += idx < 0 ? -t : t;
  }
  out[v] = (-s) / m.areaTri[v];
}

__global__ __launch_bounds__(kBlock) void k_op_c2e(DMesh m, const double *f, double *out) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= m.nE) return;
  out[e] = 0.5 * (f[m.coe[e]] + f[m.coe[m.nE + e]]);
}

__global__ __launch_bounds__(kBlock) void k_op_c2v(DMesh m, const double *f, double *out) {
  const int v = blockIdx.x * kBlock + threadIdx.x;
  if (v >= m.nV) return;
  double s = 0.0;
  for (int j = 0; j < 3; ++j) s = s + f[m.cov[j * m.nV + v]] * m.kite[j * m.nV + v];
  out[v] = s / m.areaTri[v];
}

// scalar_v2c is a scatter in the reference (operators_mod.F90:133-139); here a
// gather over verticesOnCell with the kite of each (cell, vertex) pair.
__global__ __launch_bounds__(kBlock) void k_op_v2c(DMesh m, const double *f, double *out) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= m.nC) return;
  double s = 0.0;
  for (int q = 0; q < m.nEoC[c]; ++q)
    s = s + (f[m.vocS[q * m.nC + c]] * m.kiteC[q * m.nC + c]) / m.areaCell[c];
  out[c] = s;
}

__global__ __launch_bounds__(kBlock) void k_op_v2e(DMesh m, const double *f, double *out) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= m.nE) return;
  out[e] = 0.5 * (f[m.voe[e]] + f[m.voe[m.nE + e]]);
}

__global__ __launch_bounds__(kBlock) void k_op_ke(DMesh m, const double *u, double *out) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= m.nC) return;
  double s = 0.0;
  for (int q = 0; q < m.nEoC[c]; ++q) {
    const int e = unsign(m.eocS[q * m.nC + c]);
    s += (u[e] * u[e]) * m.aE[e];
  }
  out[c] = (s * 0.25) / m.areaCell[c];
}

__global__ __launch_bounds__(kBlock) void k_op_tangent_wind(DMesh m, const double *u, double *out) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= m.nE) return;
  double s = 0.0;
  for (int q = 0; q < m.nEoE[e]; ++q) s += u[m.eoe[pair_at(q, e, m.nE)]] * m.woe[pair_at(q, e, m.nE)];
  out[e] = s;
}

__global__ __launch_bounds__(kBlock) void k_op_tvf(DMesh m, const double *u, const double *gde, const double *pve,
                                                   double *out) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= m.nE) return;
  double s = 0.0;
  for (int q = 0; q < m.nEoE[e]; ++q) {
    const int ep = m.eoe[pair_at(q, e, m.nE)];
    s = s + (((m.woe[pair_at(q, e, m.nE)] * u[ep]) * gde[ep]) * 0.5) * (pve[e] + pve[ep]);
  }
  out[e] = s;
}

// add_upwind_correction_on_cell (operators_mod.F90:423-472) with the
// context's adv_order: fe_in is the c2e value, sign taken from u.
__global__ __launch_bounds__(kBlock) void k_op_upwind_cell(DMesh m, StageBufs<double> b, const double *f,
                                                           const double *u, const double *fe_in, double *out) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= m.nE) return;
  out[e] = upwind_cell_e<double, Solo, kXAdv | kXAdv56>(m, b, Solo{}, f, 0, e, fe_in[e], sign1v<double, 1>(u[e]));
}

// calc_pv_on_vertex: pv = (f + vor) / gd_vertex
__global__ __launch_bounds__(kBlock) void k_op_pv_vertex(DMesh m, const double *vor, const double *gdv,
                                                         double *out) {
  const int v = blockIdx.x * kBlock + threadIdx.x;
  if (v >= m.nV) return;
  out[v] = pv_vertex_at<double, double>(m, v, vor[v], gdv[v]);
}

// calc_pv_on_edge from pv_vertex (b.pv_v), the given v_edge and, for schemes
// 1/2, u (b.u) and pv_cell (b.pv_c)
__global__ __launch_bounds__(kBlock) void k_op_pv_edge(DMesh m, StageBufs<double> b, const double *v_edge,
                                                       double *out) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= m.nE) return;
  out[e] = pv_edge_at<double, Solo, kXExt>(m, b, Solo{}, e, v_edge[e]);
}

// E = ke + gd + ghs (the cell part of calc_u_tend_on_edge)
__global__ __launch_bounds__(kBlock) void k_op_energy_cell(DMesh m, const double *ke, const double *gd,
                                                           const double *ghs, double *E) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= m.nC) return;
  E[c] = (ke[c] + gd[c]) + ghs[c];
}

// calc_u_tend_on_edge: b.E, b.gd_e and b.u set; pv_flx and gd_tend_edge given
__global__ __launch_bounds__(kBlock) void k_op_u_tend(DMesh m, StageBufs<double> b, const double *pvflx,
                                                      const double *tge, double *ut, double *it) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= m.nE) return;
  double i;
  ut[e] = u_tend_at<double, Solo>(m, b, Solo{}, e, pvflx[e], tge[e], i);
  it[e] = i;
}

}  // namespace tmk

namespace {

enum Kind { kC = 0, kE = 1, kV = 2 };

struct OpIo {
  const double *host;
  Kind kind;
};

struct OpOut {
  double *host;  // null: device scratch
  Kind kind;
  int n;         // scratch size in doubles when > 0, else the count of kind
};

int kind_n(tm_ctx *c, Kind k) { return k == kC ? c->nC : (k == kE ? c->nE : c->nV); }
const std::vector<int> &kind_perm(tm_ctx *c, Kind k) { return k == kC ? c->pC : (k == kE ? c->pE : c->pV); }

// Copies the inputs in (reference order -> device order), runs launch(nb, ins,
// outs) with nb blocks over the first output, copies the outputs back.  An
// output with a null host pointer is device scratch that is not copied back.
template <typename Launch>
int run_op(tm_ctx *c, std::initializer_list<OpIo> ins, std::initializer_list<OpOut> outs, Launch launch) {
  if (!c || outs.size() == 0) return TM_EINVAL;
  std::vector<double *> dins, douts;
  std::vector<void *> tmp;
  auto cleanup = [&]() {
    for (void *p : tmp) (void)hipFree(p);
  };
  auto dev = [&](int n) -> double * {
    void *p = nullptr;
    if (hipMalloc(&p, (size_t)n * sizeof(double)) != hipSuccess) return nullptr;
    tmp.push_back(p);
    return (double *)p;
  };
  for (const OpIo &io : ins) {
    if (!io.host) {
      cleanup();
      return c->fail(TM_EINVAL, "null operator input");
    }
    const int n = kind_n(c, io.kind);
    const auto &perm = kind_perm(c, io.kind);
    std::vector<double> h(n);
    for (int i = 0; i < n; ++i) h[i] = io.host[perm[i]];
    double *p = dev(n);
    if (!p) {
      cleanup();
      return c->fail(TM_ENOMEM, "hipMalloc op input");
    }
    if (hipMemcpy(p, h.data(), n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
      cleanup();
      return c->fail(TM_EDEVICE, "hipMemcpy op input");
    }
    dins.push_back(p);
  }
  for (const OpOut &io : outs) {
    double *p = dev(io.n > 0 ? io.n : kind_n(c, io.kind));
    if (!p) {
      cleanup();
      return c->fail(TM_ENOMEM, "hipMalloc op output");
    }
    douts.push_back(p);
  }
  const int nb = (kind_n(c, outs.begin()->kind) + kBlock - 1) / kBlock;
  launch(nb, dins, douts);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  int k = 0;
  for (const OpOut &io : outs) {
    const int no = kind_n(c, io.kind);
    if (io.host && e == hipSuccess) {
      std::vector<double> h(no);
      e = hipMemcpy(h.data(), douts[k], no * sizeof(double), hipMemcpyDeviceToHost);
      const auto &perm = kind_perm(c, io.kind);
      if (e == hipSuccess)
        for (int i = 0; i < no; ++i) io.host[perm[i]] = h[i];
    }
    ++k;
  }
  cleanup();
  if (e != hipSuccess) return c->fail(TM_EDEVICE, std::string("operator launch: ") + hipGetErrorString(e));
  return TM_OK;
}

}  // namespace

#define TM_OP1(name, kern, kin, kout)                                                                   \
  extern "C" int tm_op_##name(tm_ctx *c, const double *in, double *out) {                              \
    if (!out) return c ? c->fail(TM_EINVAL, "null operator output") : TM_EINVAL;                      \
    return run_op(c, {{in, kin}}, {{out, kout, 0}},                                                       \
                  [&](int nb, std::vector<double *> &d, std::vector<double *> &o) {                   \
                    hipLaunchKernelGGL(tmk::kern, dim3(nb), dim3(tmk::kBlock), 0, c->stream, c->dm, d[0], \
                                       o[0]);                                                          \
                  });                                                                                  \
  }
TM_OP1(div, k_op_div, kE, kC)
TM_OP1(curl, k_op_curl, kE, kV)
TM_OP1(c2e, k_op_c2e, kC, kE)
TM_OP1(c2v, k_op_c2v, kC, kV)
TM_OP1(v2c, k_op_v2c, kV, kC)
TM_OP1(v2e, k_op_v2e, kV, kE)
TM_OP1(ke, k_op_ke, kE, kC)
TM_OP1(tangent_wind, k_op_tangent_wind, kE, kE)
#undef TM_OP1

extern "C" int tm_op_tangent_vor_flux(tm_ctx *c, const double *u, const double *gd_edge, const double *pv_edge,
                                      double *flux) {
  if (!flux) return c ? c->fail(TM_EINVAL, "null operator output") : TM_EINVAL;
  return run_op(c, {{u, kE}, {gd_edge, kE}, {pv_edge, kE}}, {{flux, kE, 0}},
                [&](int nb, std::vector<double *> &d, std::vector<double *> &o) {
                  hipLaunchKernelGGL(tmk::k_op_tvf, dim3(nb), dim3(tmk::kBlock), 0, c->stream, c->dm, d[0], d[1],
                                     d[2], o[0]);
                });
}

extern "C" int tm_op_upwind_cell(tm_ctx *c, const double *f_cell, const double *u_edge, double *f_edge) {
  if (!c) return TM_EINVAL;
  if (!f_edge) return c->fail(TM_EINVAL, "null operator output");
  tmk::StageBufs<double> b = {};
  set_adv(b, c->o);
  return run_op(c, {{f_cell, kC}, {u_edge, kE}, {f_edge, kE}}, {{f_edge, kE, 0}},
                [&](int nb, std::vector<double *> &d, std::vector<double *> &o) {
                  hipLaunchKernelGGL(tmk::k_op_upwind_cell, dim3(nb), dim3(tmk::kBlock), 0, c->stream, c->dm, b,
                                     d[0], d[1], d[2], o[0]);
                });
}

extern "C" int tm_op_pv_vertex(tm_ctx *c, const double *vor_vertex, const double *gd_vertex, double *pv_vertex) {
  if (!c) return TM_EINVAL;
  if (!pv_vertex) return c->fail(TM_EINVAL, "null operator output");
  return run_op(c, {{vor_vertex, kV}, {gd_vertex, kV}}, {{pv_vertex, kV, 0}},
                [&](int nb, std::vector<double *> &d, std::vector<double *> &o) {
                  hipLaunchKernelGGL(tmk::k_op_pv_vertex, dim3(nb), dim3(tmk::kBlock), 0, c->stream, c->dm, d[0],
                                     d[1], o[0]);
                });
}

// The context's pv_scheme; the conservative APVM (scheme 2) runs the same
// kernels as the step (k_cvor_*) on the given inputs.
extern "C" int tm_op_pv_edge(tm_ctx *c, const double *u_edge, const double *v_edge, const double *gd_edge,
                             const double *gd_tend_vertex, const double *pv_vertex, const double *pv_cell,
                             double *pv_edge) {
  if (!c) return TM_EINVAL;
  if (!pv_edge) return c->fail(TM_EINVAL, "null operator output");
  tmk::StageBufs<double> b = {};
  b.pv_order = c->o.pv_scheme == 3 ? c->o.interp_pv_order : 2;
  b.pv_scheme = c->o.pv_scheme;
  b.apvm_wdt = c->o.apvm_weight * c->o.dt;
  b.dt = c->o.dt;
  const tmk::DMesh &m = c->dm;
  return run_op(c, {{u_edge, kE}, {v_edge, kE}, {gd_edge, kE}, {gd_tend_vertex, kV}, {pv_vertex, kV},
                    {pv_cell, kC}},
                {{pv_edge, kE, 0}, {nullptr, kE, 0}, {nullptr, kE, 0}, {nullptr, kE, 0},
                 {nullptr, kV, 2 * 3 * c->dm.nbV}, {nullptr, kC, 1}},
                [&](int nb, std::vector<double *> &d, std::vector<double *> &o) {
                  b.u = d[0];
                  b.gd_e = d[2];
                  b.tgd_v = d[3];
                  b.pv_v = d[4];
                  b.pv_c = d[5];
                  b.pv_e = o[0];
                  b.dpv = o[1];
                  b.pvf = o[2];
                  b.dpvf = o[3];
                  b.cv_part = (double2 *)o[4];
                  b.cv_eps = o[5];
                  hipLaunchKernelGGL(tmk::k_op_pv_edge, dim3(nb), dim3(tmk::kBlock), 0, c->stream, m, b, d[1],
                                     o[0]);
                  if (b.pv_scheme == 2) {
                    hipLaunchKernelGGL(tmk::k_cvor_flux<double>, dim3(m.nbE), dim3(tmk::kBlock), 0, c->stream, m, b);
                    hipLaunchKernelGGL(tmk::k_cvor_sums<double>, dim3(m.nbV), dim3(tmk::kBlock), 0, c->stream, m, b);
                    hipLaunchKernelGGL(tmk::k_cvor_eps<double>, dim3(1), dim3(tmk::kBlock), 0, c->stream, m, b);
                    hipLaunchKernelGGL(tmk::k_cvor_apply<double>, dim3(m.nbE), dim3(tmk::kBlock), 0, c->stream, m, b);
                  }
                });
}

extern "C" int tm_op_u_tend(tm_ctx *c, const double *u_edge, const double *ke_cell, const double *gd_cell,
                            const double *gd_edge, const double *ghs_cell, const double *pv_flx_edge,
                            const double *gd_tend_edge, double *u_tend_edge, double *iap_u_tend_edge) {
  if (!c) return TM_EINVAL;
  if (!u_tend_edge || !iap_u_tend_edge) return c->fail(TM_EINVAL, "null operator output");
  tmk::StageBufs<double> b = {};
  return run_op(c, {{u_edge, kE}, {ke_cell, kC}, {gd_cell, kC}, {gd_edge, kE}, {ghs_cell, kC}, {pv_flx_edge, kE},
                    {gd_tend_edge, kE}},
                {{u_tend_edge, kE, 0}, {iap_u_tend_edge, kE, 0}, {nullptr, kC, 0}},
                [&](int nb, std::vector<double *> &d, std::vector<double *> &o) {
                  const int nbc = (c->nC + tmk::kBlock - 1) / tmk::kBlock;
                  hipLaunchKernelGGL(tmk::k_op_energy_cell, dim3(nbc), dim3(tmk::kBlock), 0, c->stream, c->dm,
                                     d[1], d[2], d[4], o[2]);
                  b.E = o[2];
                  b.gd_e = d[3];
                  b.u = d[0];
                  hipLaunchKernelGGL(tmk::k_op_u_tend, dim3(nb), dim3(tmk::kBlock), 0, c->stream, c->dm, b, d[5],
                                     d[6], o[0], o[1]);
                });
}
