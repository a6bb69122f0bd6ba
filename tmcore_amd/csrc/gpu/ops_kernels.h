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
  for (int q = 0; q < m.nEoE[e]; ++q) s += u[m.eoe[q * m.nE + e]] * m.woe[q * m.nE + e];
  out[e] = s;
}

__global__ __launch_bounds__(kBlock) void k_op_tvf(DMesh m, const double *u, const double *gde, const double *pve,
                                                   double *out) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= m.nE) return;
  double s = 0.0;
  for (int q = 0; q < m.nEoE[e]; ++q) {
    const int ep = m.eoe[q * m.nE + e];
    s = s + (((m.woe[q * m.nE + e] * u[ep]) * gde[ep]) * 0.5) * (pve[e] + pve[ep]);
  }
  out[e] = s;
}

// add_upwind_correction_on_cell (operators_mod.F90:423-472) with the
// context's adv_order: fe_in is the c2e value, sign taken from u.
__global__ __launch_bounds__(kBlock) void k_op_upwind_cell(DMesh m, StageBufs<double> b, const double *f,
                                                           const double *u, const double *fe_in, double *out) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= m.nE) return;
  out[e] = upwind_cell_e<double, Solo>(m, b, Solo{}, f, 0, e, fe_in[e], sign1v<double, 1>(u[e]));
}

}  // namespace tmk

namespace {

enum Kind { kC = 0, kE = 1, kV = 2 };

struct OpIo {
  const double *host;
  Kind kind;
};

int kind_n(tm_ctx *c, Kind k) { return k == kC ? c->nC : (k == kE ? c->nE : c->nV); }
const std::vector<int> &kind_perm(tm_ctx *c, Kind k) { return k == kC ? c->pC : (k == kE ? c->pE : c->pV); }

template <typename Launch>
int run_op(tm_ctx *c, std::initializer_list<OpIo> ins, Kind out_kind, double *out, Launch launch) {
  if (!c || !out) return TM_EINVAL;
  std::vector<double *> dins;
  std::vector<void *> tmp;
  auto cleanup = [&]() {
    for (void *p : tmp) (void)hipFree(p);
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
    void *p = nullptr;
    if (hipMalloc(&p, n * sizeof(double)) != hipSuccess) {
      cleanup();
      return c->fail(TM_ENOMEM, "hipMalloc op input");
    }
    tmp.push_back(p);
    if (hipMemcpy(p, h.data(), n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
      cleanup();
      return c->fail(TM_EDEVICE, "hipMemcpy op input");
    }
    dins.push_back((double *)p);
  }
  const int no = kind_n(c, out_kind);
  void *po = nullptr;
  if (hipMalloc(&po, no * sizeof(double)) != hipSuccess) {
    cleanup();
    return c->fail(TM_ENOMEM, "hipMalloc op output");
  }
  tmp.push_back(po);
  const int nb = (no + kBlock - 1) / kBlock;
  launch(nb, dins, (double *)po);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  std::vector<double> h(no);
  if (e == hipSuccess) e = hipMemcpy(h.data(), po, no * sizeof(double), hipMemcpyDeviceToHost);
  cleanup();
  if (e != hipSuccess) return c->fail(TM_EDEVICE, std::string("operator launch: ") + hipGetErrorString(e));
  const auto &perm = kind_perm(c, out_kind);
  for (int i = 0; i < no; ++i) out[perm[i]] = h[i];
  return TM_OK;
}

}  // namespace

#define TM_OP1(name, kern, kin, kout)                                                                   \
  extern "C" int tm_op_##name(tm_ctx *c, const double *in, double *out) {                              \
    return run_op(c, {{in, kin}}, kout, out, [&](int nb, std::vector<double *> &d, double *o) {          \
      hipLaunchKernelGGL(tmk::kern, dim3(nb), dim3(tmk::kBlock), 0, c->stream, c->dm, d[0], o);            \
    });                                                                                                \
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
  return run_op(c, {{u, kE}, {gd_edge, kE}, {pv_edge, kE}}, kE, flux,
                [&](int nb, std::vector<double *> &d, double *o) {
                  hipLaunchKernelGGL(tmk::k_op_tvf, dim3(nb), dim3(tmk::kBlock), 0, c->stream, c->dm, d[0], d[1],
                                     d[2], o);
                });
}

extern "C" int tm_op_upwind_cell(tm_ctx *c, const double *f_cell, const double *u_edge, double *f_edge) {
  if (!c) return TM_EINVAL;
  tmk::StageBufs<double> b = {};
  set_adv(b, c->o);
  return run_op(c, {{f_cell, kC}, {u_edge, kE}, {f_edge, kE}}, kE, f_edge,
                [&](int nb, std::vector<double *> &d, double *o) {
                  hipLaunchKernelGGL(tmk::k_op_upwind_cell, dim3(nb), dim3(tmk::kBlock), 0, c->stream, c->dm, b,
                                     d[0], d[1], d[2], o);
                });
}
