/*
 * trisk_oracle.c -- TEST INFRASTRUCTURE (the oracle), not product code.
 *
 * FP64 CPU restatement of the reference TMCORE shallow-water path.  Every
 * function cites the reference lines it follows.  Expressions keep the
 * reference's evaluation order (Fortran left-to-right association) and the
 * library is compiled with -ffp-contract=off, so each result is what the
 * reference's arithmetic gives up to the order of its global sums.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 */
#include "trisk_oracle.h"

/* Element loops run over OpenMP threads when the oracle is built with
 * -fopenmp (each element's arithmetic is unchanged; the global sums stay
 * serial, in element order). */
#define OM_FOR _Pragma("omp parallel for schedule(static)")

#ifdef _OPENMP
#include <omp.h>
#endif

/* Threads of the element loops (1 when built without OpenMP). */
int om_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OM_G 9.80616       /* const_mod.F90:12 */
#define OM_OMEGA 7.292e-5  /* const_mod.F90:13 */

struct om_ctx {
  int nC, nE, nV, ME, ME2, MF;
  om_params p;
  double radius;
  /* connectivity, 0-based */
  int *nEdgesOnCell, *nEdgesOnEdge, *edgesOnCell, *cellsOnCell, *verticesOnCell;
  int *cellsOnEdge, *verticesOnEdge, *edgesOnEdge, *cellsOnVertex, *edgesOnVertex;
  int *nSignEdge, *tSignEdge; /* mesh_mod.F90:182-203 */
  double *weightsOnEdge, *kiteAreasOnVertex;
  double *dvEdge, *dcEdge, *areaCell, *areaTriangle, *areaEdge, *fVertex;
  double totalArea;
  /* fit weights (poly_fit_mod.F90:35-57), 0-based stencil indices */
  int *nFit2Cells, *fit2Cells, *nFit4Cells, *fit4Cells;
  int *nFit2Vertices, *fit2Vertices, *nFit4Vertices, *fit4Vertices;
  double *deriv2OnCell, *deriv4OnCell, *deriv2OnVertex, *deriv4OnVertex;
  /* implicit-SAVE locals of the reference (operators_mod.F90:367-368,429-430) */
  double pv_coef4, adv_coef4;
  /* scratch for the stage */
  double *w_e[16], *w_c[8], *w_v[8];
};

static int *dup_idx(const int32_t *a, size_t n) {
  if (!a) return NULL;
  int *r = (int *)malloc(n * sizeof(int));
  for (size_t i = 0; i < n; ++i) r[i] = a[i] - 1;
  return r;
}
static int *dup_int(const int32_t *a, size_t n) {
  if (!a) return NULL;
  int *r = (int *)malloc(n * sizeof(int));
  for (size_t i = 0; i < n; ++i) r[i] = a[i];
  return r;
}
static double *dup_dbl(const double *a, size_t n, double scale) {
  if (!a) return NULL;
  double *r = (double *)malloc(n * sizeof(double));
  for (size_t i = 0; i < n; ++i) r[i] = a[i] * scale;
  return r;
}

om_ctx *om_create(const tm_mesh_desc *m, const tm_fit_desc *f, double radius, const om_params *p) {
  om_ctx *c = (om_ctx *)calloc(1, sizeof(om_ctx));
  c->nC = m->nCells;
  c->nE = m->nEdges;
  c->nV = m->nVertices;
  c->ME = m->maxEdges;
  c->ME2 = m->maxEdges2;
  c->radius = radius;
  c->p = *p;
  size_t nC = c->nC, nE = c->nE, nV = c->nV, ME = c->ME, ME2 = c->ME2;
  c->nEdgesOnCell = dup_int(m->nEdgesOnCell, nC);
  c->nEdgesOnEdge = dup_int(m->nEdgesOnEdge, nE);
  c->edgesOnCell = dup_idx(m->edgesOnCell, nC * ME);
  c->cellsOnCell = dup_idx(m->cellsOnCell, nC * ME);
  c->verticesOnCell = dup_idx(m->verticesOnCell, nC * ME);
  c->cellsOnEdge = dup_idx(m->cellsOnEdge, nE * 2);
  c->verticesOnEdge = dup_idx(m->verticesOnEdge, nE * 2);
  c->edgesOnEdge = dup_idx(m->edgesOnEdge, nE * ME2);
  c->cellsOnVertex = dup_idx(m->cellsOnVertex, nV * 3);
  c->edgesOnVertex = dup_idx(m->edgesOnVertex, nV * 3);
  c->weightsOnEdge = dup_dbl(m->weightsOnEdge, nE * ME2, 1.0);
  /* mesh_mod.F90:178,205-224: areaEdge = dv*dc (unit sphere) then * a^2 */
  c->areaEdge = (double *)malloc(nE * sizeof(double));
  for (size_t e = 0; e < nE; ++e) c->areaEdge[e] = (m->dvEdge[e] * m->dcEdge[e]) * (radius * radius);
  c->dvEdge = dup_dbl(m->dvEdge, nE, radius);
  c->dcEdge = dup_dbl(m->dcEdge, nE, radius);
  c->areaCell = dup_dbl(m->areaCell, nC, radius * radius);
  c->areaTriangle = dup_dbl(m->areaTriangle, nV, radius * radius);
  c->kiteAreasOnVertex = dup_dbl(m->kiteAreasOnVertex, nV * 3, radius * radius);
  /* mesh_mod.F90:180 */
  c->fVertex = (double *)malloc(nV * sizeof(double));
  for (size_t v = 0; v < nV; ++v) c->fVertex[v] = 2.0 * OM_OMEGA * sin(m->latVertex[v]);
  /* mesh_mod.F90:182-188 */
  c->nSignEdge = (int *)calloc(nC * ME, sizeof(int));
  for (size_t i = 0; i < nC; ++i)
    for (int k = 0; k < c->nEdgesOnCell[i]; ++k) {
      int e = c->edgesOnCell[i * ME + k];
      if ((int)i == c->cellsOnEdge[2 * e]) c->nSignEdge[i * ME + k] = 1;
      if ((int)i == c->cellsOnEdge[2 * e + 1]) c->nSignEdge[i * ME + k] = -1;
    }
  /* mesh_mod.F90:190-203 */
  c->tSignEdge = (int *)calloc(nV * 3, sizeof(int));
  for (size_t v = 0; v < nV; ++v)
    for (int k = 0; k < 3; ++k) {
      int e = c->edgesOnVertex[v * 3 + k];
      if ((int)v == c->verticesOnEdge[2 * e]) c->tSignEdge[v * 3 + k] = 1;
      if ((int)v == c->verticesOnEdge[2 * e + 1]) c->tSignEdge[v * 3 + k] = -1;
    }
  /* mesh_mod.F90:224 */
  double s = 0.0;
  for (size_t i = 0; i < nC; ++i) s += c->areaCell[i];
  c->totalArea = s;
  if (f) {
    c->MF = f->maxFit;
    size_t MF = c->MF;
    c->nFit2Cells = dup_int(f->nFit2Cells, nC);
    c->fit2Cells = dup_idx(f->fit2Cells, nC * MF);
    c->deriv2OnCell = dup_dbl(f->deriv2OnCell, nE * 2 * MF, 1.0);
    c->nFit4Cells = dup_int(f->nFit4Cells, nC);
    c->fit4Cells = dup_idx(f->fit4Cells, nC * MF);
    c->deriv4OnCell = dup_dbl(f->deriv4OnCell, nE * 2 * MF, 1.0);
    c->nFit2Vertices = dup_int(f->nFit2Vertices, nV);
    c->fit2Vertices = dup_idx(f->fit2Vertices, nV * MF);
    c->deriv2OnVertex = dup_dbl(f->deriv2OnVertex, nE * 2 * MF, 1.0);
    c->nFit4Vertices = dup_int(f->nFit4Vertices, nV);
    c->fit4Vertices = dup_idx(f->fit4Vertices, nV * MF);
    c->deriv4OnVertex = dup_dbl(f->deriv4OnVertex, nE * 2 * MF, 1.0);
  }
  c->pv_coef4 = 0.25;
  c->adv_coef4 = 0.25;
  for (int i = 0; i < 16; ++i) c->w_e[i] = (double *)calloc(nE, sizeof(double));
  for (int i = 0; i < 8; ++i) c->w_c[i] = (double *)calloc(nC, sizeof(double));
  for (int i = 0; i < 8; ++i) c->w_v[i] = (double *)calloc(nV, sizeof(double));
  return c;
}

int om_set_params(om_ctx *c, const om_params *p) {
  c->p = *p;
  return 0;
}

void om_destroy(om_ctx *c) {
  if (!c) return;
  void *ptrs[] = {c->nEdgesOnCell, c->nEdgesOnEdge, c->edgesOnCell, c->cellsOnCell, c->verticesOnCell,
                  c->cellsOnEdge, c->verticesOnEdge, c->edgesOnEdge, c->cellsOnVertex, c->edgesOnVertex,
                  c->nSignEdge, c->tSignEdge, c->weightsOnEdge, c->kiteAreasOnVertex, c->dvEdge,
                  c->dcEdge, c->areaCell, c->areaTriangle, c->areaEdge, c->fVertex, c->nFit2Cells,
                  c->fit2Cells, c->nFit4Cells, c->fit4Cells, c->nFit2Vertices, c->fit2Vertices,
                  c->nFit4Vertices, c->fit4Vertices, c->deriv2OnCell, c->deriv4OnCell,
                  c->deriv2OnVertex, c->deriv4OnVertex};
  for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); ++i) free(ptrs[i]);
  for (int i = 0; i < 16; ++i) free(c->w_e[i]);
  for (int i = 0; i < 8; ++i) free(c->w_c[i]);
  for (int i = 0; i < 8; ++i) free(c->w_v[i]);
  free(c);
}

/* ---------------------------------------------------------------- operators */

/* div_operator, operators_mod.F90:40-55 (returns minus the divergence) */
void om_div(const om_ctx *c, const double *f_edge, double *div) {
  OM_FOR
  for (int i = 0; i < c->nC; ++i) {
    double s = 0.0;
    for (int k = 0; k < c->nEdgesOnCell[i]; ++k) {
      int e = c->edgesOnCell[i * c->ME + k];
      s += ((double)c->nSignEdge[i * c->ME + k] * f_edge[e]) * c->dvEdge[e];
    }
    div[i] = -s;
  }
  OM_FOR
  for (int i = 0; i < c->nC; ++i) div[i] = div[i] / c->areaCell[i];
}

/* curl_operator, operators_mod.F90:76-91 */
void om_curl(const om_ctx *c, const double *f_edge, double *curl) {
  OM_FOR
  for (int v = 0; v < c->nV; ++v) {
    double s = 0.0;
    for (int k = 0; k < 3; ++k) {
      int e = c->edgesOnVertex[v * 3 + k];
      s += ((double)c->tSignEdge[v * 3 + k] * f_edge[e]) * c->dcEdge[e];
    }
    curl[v] = -s;
  }
  OM_FOR
  for (int v = 0; v < c->nV; ++v) curl[v] = curl[v] / c->areaTriangle[v];
}

/* scalar_c2e_interp_operator, operators_mod.F90:93-104 */
void om_c2e(const om_ctx *c, const double *f_cell, double *f_edge) {
  OM_FOR
  for (int e = 0; e < c->nE; ++e)
    f_edge[e] = 0.5 * (f_cell[c->cellsOnEdge[2 * e]] + f_cell[c->cellsOnEdge[2 * e + 1]]);
}

/* scalar_c2v_interp_operator, operators_mod.F90:106-123 */
void om_c2v(const om_ctx *c, const double *f_cell, double *f_vertex) {
  OM_FOR
  for (int v = 0; v < c->nV; ++v) {
    double s = 0.0;
    for (int k = 0; k < 3; ++k) s = s + f_cell[c->cellsOnVertex[v * 3 + k]] * c->kiteAreasOnVertex[v * 3 + k];
    f_vertex[v] = s;
  }
  OM_FOR
  for (int v = 0; v < c->nV; ++v) f_vertex[v] = f_vertex[v] / c->areaTriangle[v];
}

/* scalar_v2c_interp_operator, operators_mod.F90:125-141 (scatter order kept) */
void om_v2c(const om_ctx *c, const double *f_vertex, double *f_cell) {
  for (int i = 0; i < c->nC; ++i) f_cell[i] = 0.0;
  for (int v = 0; v < c->nV; ++v)
    for (int k = 0; k < 3; ++k) {
      int i = c->cellsOnVertex[v * 3 + k];
      f_cell[i] = f_cell[i] + (f_vertex[v] * c->kiteAreasOnVertex[v * 3 + k]) / c->areaCell[i];
    }
}

/* scalar_v2e_interp_operator, operators_mod.F90:143-156 */
void om_v2e(const om_ctx *c, const double *f_vertex, double *f_edge) {
  OM_FOR
  for (int e = 0; e < c->nE; ++e)
    f_edge[e] = 0.5 * (f_vertex[c->verticesOnEdge[2 * e]] + f_vertex[c->verticesOnEdge[2 * e + 1]]);
}

/* iap_swm_operator / inverse, operators_mod.F90:158-176 */
void om_iap(const om_ctx *c, const double *gd_edge, const double *u, double *iap) {
  OM_FOR
  for (int e = 0; e < c->nE; ++e) iap[e] = sqrt(gd_edge[e]) * u[e];
}
void om_inverse_iap(const om_ctx *c, const double *gd_edge, const double *iap, double *u) {
  OM_FOR
  for (int e = 0; e < c->nE; ++e) u[e] = iap[e] / sqrt(gd_edge[e]);
}

/* calc_kinetic_energy, operators_mod.F90:206-221 */
void om_ke(const om_ctx *c, const double *u, double *ke) {
  OM_FOR
  for (int i = 0; i < c->nC; ++i) {
    double s = 0.0;
    for (int k = 0; k < c->nEdgesOnCell[i]; ++k) {
      int e = c->edgesOnCell[i * c->ME + k];
      s += (u[e] * u[e]) * c->areaEdge[e];
    }
    ke[i] = s * 0.25;
  }
  OM_FOR
  for (int i = 0; i < c->nC; ++i) ke[i] = ke[i] / c->areaCell[i];
}

/* calc_tangent_wind, operators_mod.F90:223-237 */
void om_tangent_wind(const om_ctx *c, const double *u, double *v) {
  OM_FOR
  for (int e = 0; e < c->nE; ++e) {
    double s = 0.0;
    for (int k = 0; k < c->nEdgesOnEdge[e]; ++k)
      s += u[c->edgesOnEdge[e * c->ME2 + k]] * c->weightsOnEdge[e * c->ME2 + k];
    v[e] = s;
  }
}

/* calc_tangent_vor_flux, operators_mod.F90:239-260 */
void om_tangent_vor_flux(const om_ctx *c, const double *u, const double *gd_edge, const double *pv_edge,
                         double *flux) {
  OM_FOR
  for (int e = 0; e < c->nE; ++e) {
    double s = 0.0;
    for (int k = 0; k < c->nEdgesOnEdge[e]; ++k) {
      int ep = c->edgesOnEdge[e * c->ME2 + k];
      s = s + (((c->weightsOnEdge[e * c->ME2 + k] * u[ep]) * gd_edge[ep]) * 0.5) * (pv_edge[e] + pv_edge[ep]);
    }
    flux[e] = s;
  }
}

/* calc_pv_on_vertex, operators_mod.F90:262-270 */
void om_pv_vertex(const om_ctx *c, const double *vor, const double *gd_vertex, double *pv) {
  OM_FOR
  for (int v = 0; v < c->nV; ++v) pv[v] = (c->fVertex[v] + vor[v]) / gd_vertex[v];
}

static inline double fsign1(double x) { return signbit(x) ? -1.0 : 1.0; } /* sign(1.0d0, x) */

/* Weighted stencil difference sum(w(1:n-1) * (f(fit(1:n-1)) - f(centre))). */
static inline double stencil_d(const double *w, const int *fit, int nfit, const double *f, int centre) {
  double s = 0.0;
  for (int j = 1; j < nfit; ++j) s += w[j] * (f[fit[j]] - f[centre]);
  return s;
}

/* calc_pv_on_edge_high_order_upwind, operators_mod.F90:359-412 */
static void pv_edge_upwind(om_ctx *c, const double *v_edge, const double *pv_vertex, double *pv_edge) {
  const int o = c->p.interp_pv_order;
  const size_t MF = c->MF;
  OM_FOR
  for (int e = 0; e < c->nE; ++e)
    pv_edge[e] = 0.5 * (pv_vertex[c->verticesOnEdge[2 * e]] + pv_vertex[c->verticesOnEdge[2 * e + 1]]);
  double coef2 = 0.0;
  if (o == 3) coef2 = 1.0;
  if (o >= 4) coef2 = 0.0;
  if (o == 6) c->pv_coef4 = 0.0;
  if (o == 3 || o == 4) {
    OM_FOR
    for (int e = 0; e < c->nE; ++e) {
      int v1 = c->verticesOnEdge[2 * e], v2 = c->verticesOnEdge[2 * e + 1];
      double d1 = stencil_d(c->deriv2OnVertex + (e * 2 + 0) * MF, c->fit2Vertices + v1 * MF,
                            c->nFit2Vertices[v1], pv_vertex, v1);
      double d2 = stencil_d(c->deriv2OnVertex + (e * 2 + 1) * MF, c->fit2Vertices + v2 * MF,
                            c->nFit2Vertices[v2], pv_vertex, v2);
      double dv = c->dvEdge[e];
      pv_edge[e] = pv_edge[e] - (dv * dv) * ((d1 + d2) - (fsign1(v_edge[e]) * coef2) * (d2 - d1)) / 12.0;
    }
  } else if (o == 5 || o == 6) {
    OM_FOR
    for (int e = 0; e < c->nE; ++e) {
      int v1 = c->verticesOnEdge[2 * e], v2 = c->verticesOnEdge[2 * e + 1];
      double d1 = stencil_d(c->deriv2OnVertex + (e * 2 + 0) * MF, c->fit2Vertices + v1 * MF,
                            c->nFit2Vertices[v1], pv_vertex, v1);
      double d2 = stencil_d(c->deriv2OnVertex + (e * 2 + 1) * MF, c->fit2Vertices + v2 * MF,
                            c->nFit2Vertices[v2], pv_vertex, v2);
      double q1 = stencil_d(c->deriv4OnVertex + (e * 2 + 0) * MF, c->fit4Vertices + v1 * MF,
                            c->nFit4Vertices[v1], pv_vertex, v1);
      double q2 = stencil_d(c->deriv4OnVertex + (e * 2 + 1) * MF, c->fit4Vertices + v2 * MF,
                            c->nFit4Vertices[v2], pv_vertex, v2);
      double dv = c->dvEdge[e];
      pv_edge[e] = pv_edge[e] - (dv * dv) * (d1 + d2) / 12.0 +
                   ((dv * dv) * (dv * dv)) * ((q1 + q2) - (fsign1(v_edge[e]) * c->pv_coef4) * (q2 - q1)) / 60.0;
    }
  }
}

/* calc_pv_on_edge dispatch, operators_mod.F90:273-357 */
void om_pv_edge(om_ctx *c, const double *u, const double *v, const double *gd_edge, const double *gd_tend_vertex,
                const double *pv_vertex, const double *pv_cell, double *pv_edge) {
  const double dt = c->p.dt;
  if (c->p.pv_scheme == 1) { /* APVM 296-316 */
    om_v2e(c, pv_vertex, pv_edge);
    if (c->p.apvm_weight != 0.0)
      OM_FOR
      for (int e = 0; e < c->nE; ++e) {
        int v1 = c->verticesOnEdge[2 * e], v2 = c->verticesOnEdge[2 * e + 1];
        int c1 = c->cellsOnEdge[2 * e], c2 = c->cellsOnEdge[2 * e + 1];
        pv_edge[e] = pv_edge[e] - (c->p.apvm_weight * dt) *
                                      ((v[e] * (pv_vertex[v2] - pv_vertex[v1])) / c->dvEdge[e] +
                                       (u[e] * (pv_cell[c2] - pv_cell[c1])) / c->dcEdge[e]);
      }
  } else if (c->p.pv_scheme == 2) { /* conservative APVM 318-357 */
    double *dpv = (double *)malloc(c->nE * sizeof(double));
    double *pvf = (double *)malloc(c->nE * sizeof(double));
    double *dpvf = (double *)malloc(c->nE * sizeof(double));
    double *vt = (double *)malloc(c->nV * sizeof(double));
    double *dvt = (double *)malloc(c->nV * sizeof(double));
    om_v2e(c, pv_vertex, pv_edge);
    OM_FOR
    for (int e = 0; e < c->nE; ++e) {
      int v1 = c->verticesOnEdge[2 * e], v2 = c->verticesOnEdge[2 * e + 1];
      int c1 = c->cellsOnEdge[2 * e], c2 = c->cellsOnEdge[2 * e + 1];
      dpv[e] = (v[e] * (pv_vertex[v2] - pv_vertex[v1])) / c->dvEdge[e] +
               (u[e] * (pv_cell[c2] - pv_cell[c1])) / c->dcEdge[e];
    }
    om_tangent_vor_flux(c, u, gd_edge, pv_edge, pvf);
    om_tangent_vor_flux(c, u, gd_edge, dpv, dpvf);
    om_curl(c, pvf, vt);
    om_curl(c, dpvf, dvt);
    double s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (int k = 0; k < c->nV; ++k) s1 += ((pv_vertex[k] * pv_vertex[k]) * gd_tend_vertex[k]) * c->areaTriangle[k];
    for (int k = 0; k < c->nV; ++k) s2 += (pv_vertex[k] * vt[k]) * c->areaTriangle[k];
    for (int k = 0; k < c->nV; ++k) s3 += (pv_vertex[k] * dvt[k]) * c->areaTriangle[k];
    double eps = -(s1 - s2 * 2.0) / ((2.0 * dt) * s3);
    OM_FOR
    for (int e = 0; e < c->nE; ++e) pv_edge[e] = pv_edge[e] - (eps * dt) * dpv[e];
    free(dpv);
    free(pvf);
    free(dpvf);
    free(vt);
    free(dvt);
  } else {
    pv_edge_upwind(c, v, pv_vertex, pv_edge);
  }
}

/* add_upwind_correction_on_cell, operators_mod.F90:423-472 */
void om_upwind_cell(om_ctx *c, const double *f_cell, const double *u_edge, double *f_edge) {
  const int o = c->p.adv_order;
  const size_t MF = c->MF;
  double coef2 = 0.0;
  if (o == 3) coef2 = 1.0;
  if (o == 3 && c->p.adv_mono) coef2 = 0.25;
  if (o >= 4) coef2 = 0.0;
  if (o == 6) c->adv_coef4 = 0.0;
  if (o == 3 || o == 4) {
    OM_FOR
    for (int e = 0; e < c->nE; ++e) {
      int c1 = c->cellsOnEdge[2 * e], c2 = c->cellsOnEdge[2 * e + 1];
      double d1 = stencil_d(c->deriv2OnCell + (e * 2 + 0) * MF, c->fit2Cells + c1 * MF, c->nFit2Cells[c1], f_cell, c1);
      double d2 = stencil_d(c->deriv2OnCell + (e * 2 + 1) * MF, c->fit2Cells + c2 * MF, c->nFit2Cells[c2], f_cell, c2);
      double dc = c->dcEdge[e];
      f_edge[e] = f_edge[e] - (dc * dc) * ((d1 + d2) - (fsign1(u_edge[e]) * coef2) * (d2 - d1)) / 12.0;
    }
  } else if (o == 5 || o == 6) {
    OM_FOR
    for (int e = 0; e < c->nE; ++e) {
      int c1 = c->cellsOnEdge[2 * e], c2 = c->cellsOnEdge[2 * e + 1];
      double d1 = stencil_d(c->deriv2OnCell + (e * 2 + 0) * MF, c->fit2Cells + c1 * MF, c->nFit2Cells[c1], f_cell, c1);
      double d2 = stencil_d(c->deriv2OnCell + (e * 2 + 1) * MF, c->fit2Cells + c2 * MF, c->nFit2Cells[c2], f_cell, c2);
      double q1 = stencil_d(c->deriv4OnCell + (e * 2 + 0) * MF, c->fit4Cells + c1 * MF, c->nFit4Cells[c1], f_cell, c1);
      double q2 = stencil_d(c->deriv4OnCell + (e * 2 + 1) * MF, c->fit4Cells + c2 * MF, c->nFit4Cells[c2], f_cell, c2);
      double dc = c->dcEdge[e];
      f_edge[e] = f_edge[e] - (dc * dc) * (d1 + d2) / 12.0 +
                  ((dc * dc) * (dc * dc)) * ((q1 + q2) - (fsign1(u_edge[e]) * c->adv_coef4) * (q2 - q1)) / 60.0;
    }
  }
}

/* calc_gd_tend_on_cell, tmcore_swm_mod.F90:127-139 */
void om_gd_tend(const om_ctx *c, const double *u, const double *gd_edge, double *tend) {
  double *flux = (double *)malloc(c->nE * sizeof(double));
  OM_FOR
  for (int e = 0; e < c->nE; ++e) flux[e] = u[e] * gd_edge[e];
  om_div(c, flux, tend);
  free(flux);
}

/* calc_u_tend_on_edge, tmcore_swm_mod.F90:155-201 (flux_4th_order_correct = .false.) */
void om_u_tend(const om_ctx *c, const double *u, const double *ke, const double *gd, const double *gd_edge,
               const double *ghs, const double *pv_flx, const double *gd_tend_edge, double *u_tend,
               double *iap_tend) {
  double *E = (double *)malloc(c->nC * sizeof(double));
  OM_FOR
  for (int i = 0; i < c->nC; ++i) E[i] = (ke[i] + gd[i]) + ghs[i];
  OM_FOR
  for (int e = 0; e < c->nE; ++e) {
    double sg = sqrt(gd_edge[e]);
    double dEdx = (E[c->cellsOnEdge[2 * e + 1]] - E[c->cellsOnEdge[2 * e]]) / c->dcEdge[e];
    double ut = pv_flx[e] - dEdx;
    if (u_tend) u_tend[e] = ut;
    iap_tend[e] = sg * ut + ((0.5 * u[e]) / sg) * gd_tend_edge[e];
  }
  free(E);
}

/* adv_scheme_sg11_run, adv_scheme_sg11_mod.F90:33-48 */
void om_adv(om_ctx *c, const double *f_cell, const double *u, double *f_edge, double *tend) {
  double *flux = (double *)malloc(c->nE * sizeof(double));
  om_c2e(c, f_cell, f_edge);
  om_upwind_cell(c, f_cell, u, f_edge);
  OM_FOR
  for (int e = 0; e < c->nE; ++e) flux[e] = u[e] * f_edge[e];
  om_div(c, flux, tend);
  free(flux);
}

/* ------------------------------------------------------------- stage / step */

/* spatial_operators, tmcore_swm_mod.F90:92-125 (mpas_reconstruct at :101 is
 * omitted: no tendency reads its outputs). */
int om_spatial_operators(om_ctx *c, const double *gd, const double *iap_u, const double *ghs, int ntracers,
                         const double *q, double *tend_gd, double *tend_iap, double *tend_q, om_diag *dg) {
  double *gd_e = c->w_e[0], *u = c->w_e[1], *tgd_e = c->w_e[2], *v = c->w_e[3], *pv_e = c->w_e[4];
  double *pv_flx = c->w_e[5], *u_tend = c->w_e[6], *q_e = c->w_e[7];
  double *tgd_v = c->w_v[0], *gd_v = c->w_v[1], *vor = c->w_v[2], *pv_v = c->w_v[3];
  double *ke = c->w_c[0], *pv_c = c->w_c[1];
  om_c2e(c, gd, gd_e);                                   /* :98  */
  om_upwind_cell(c, gd, iap_u, gd_e);                    /* :99  */
  om_inverse_iap(c, gd_e, iap_u, u);                     /* :100 */
  om_gd_tend(c, u, gd_e, tend_gd);                       /* :104 */
  om_c2e(c, tend_gd, tgd_e);                             /* :107 */
  om_upwind_cell(c, tend_gd, iap_u, tgd_e);              /* :108 */
  om_c2v(c, tend_gd, tgd_v);                             /* :109 */
  om_c2v(c, gd, gd_v);                                   /* :110 */
  om_curl(c, u, vor);                                    /* :113 */
  om_tangent_wind(c, u, v);                              /* :114 */
  om_ke(c, u, ke);                                       /* :115 */
  om_pv_vertex(c, vor, gd_v, pv_v);                      /* :116 */
  om_v2c(c, pv_v, pv_c);                                 /* :117 */
  om_pv_edge(c, u, v, gd_e, tgd_v, pv_v, pv_c, pv_e);    /* :118 */
  om_tangent_vor_flux(c, u, gd_e, pv_e, pv_flx);         /* :121 */
  om_u_tend(c, u, ke, gd, gd_e, ghs, pv_flx, tgd_e, u_tend, tend_iap); /* :122 */
  for (int t = 0; t < ntracers; ++t)
    om_adv(c, q + (size_t)t * c->nC, u, q_e, tend_q + (size_t)t * c->nC);
  if (dg) {
#define CP(dst, src, n) \
  if (dg->dst) memcpy(dg->dst, src, (size_t)(n) * sizeof(double));
    CP(gd_edge, gd_e, c->nE) CP(u, u, c->nE) CP(tend_gd, tend_gd, c->nC) CP(tend_gd_edge, tgd_e, c->nE)
    CP(tend_gd_vertex, tgd_v, c->nV) CP(gd_vertex, gd_v, c->nV) CP(vor, vor, c->nV) CP(v, v, c->nE)
    CP(ke, ke, c->nC) CP(pv_vertex, pv_v, c->nV) CP(pv_cell, pv_c, c->nC) CP(pv_edge, pv_e, c->nE)
    CP(pv_flx, pv_flx, c->nE) CP(u_tend, u_tend, c->nE) CP(iap_tend, tend_iap, c->nE)
#undef CP
  }
  return 0;
}

/* tmcore_swm_mod.F90:62-64 */
void om_initial_iap(om_ctx *c, const double *gd, const double *u, double *iap_u) {
  double *gd_e = c->w_e[8];
  om_c2e(c, gd, gd_e);
  om_upwind_cell(c, gd, iap_u, gd_e); /* reads iap_u before it is set: no-op at adv_order 2 */
  om_iap(c, gd_e, u, iap_u);
}

/* Post-step refresh + diag_run: tmcore_swm_mod.F90:78-85, 215-255 */
void om_invariants(om_ctx *c, const double *gd, const double *iap_u, const double *ghs, double out[4]) {
  double *gd_e = c->w_e[8], *u = c->w_e[9];
  double *gd_v = c->w_v[4], *vor = c->w_v[5], *pv_v = c->w_v[6];
  om_c2e(c, gd, gd_e);
  om_upwind_cell(c, gd, iap_u, gd_e);
  om_inverse_iap(c, gd_e, iap_u, u);
  om_c2v(c, gd, gd_v);
  om_curl(c, u, vor);
  om_pv_vertex(c, vor, gd_v, pv_v);
  double mass = 0.0, e1 = 0.0, e2 = 0.0, ens = 0.0, av = 0.0, at = 0.0;
  for (int i = 0; i < c->nC; ++i) mass += gd[i] * c->areaCell[i];
  for (int e = 0; e < c->nE; ++e) e1 += (iap_u[e] * iap_u[e]) * c->areaEdge[e];
  for (int i = 0; i < c->nC; ++i) e2 += ((gd[i] * gd[i]) + (2.0 * gd[i]) * ghs[i]) * c->areaCell[i];
  for (int v = 0; v < c->nV; ++v) ens += (gd_v[v] * (pv_v[v] * pv_v[v])) * c->areaTriangle[v];
  for (int v = 0; v < c->nV; ++v) av += (vor[v] + c->fVertex[v]) * c->areaTriangle[v];
  for (int v = 0; v < c->nV; ++v) at += c->areaTriangle[v];
  out[0] = mass;
  out[1] = e1 + e2;
  out[2] = ens;
  out[3] = av / at;
}

/* inner_product_tend, operators_mod.F90:187-194 */
static double inner(const om_ctx *c, const double *a_gd, const double *a_iap, const double *b_gd,
                    const double *b_iap) {
  double s1 = 0.0, s2 = 0.0;
  for (int e = 0; e < c->nE; ++e) s1 += (a_iap[e] * b_iap[e]) * c->areaEdge[e];
  for (int i = 0; i < c->nC; ++i) s2 += (a_gd[i] * b_gd[i]) * c->areaCell[i];
  return (s1 + s2) / c->totalArea;
}

/* update_state, tmcore_swm_mod.F90:203-213 (and the tracer analogue of
 * tmcore_adv_mod.F90:175) */
static void axpy(int n, const double *old, double dt, const double *t, double *out) {
  for (int i = 0; i < n; ++i) out[i] = old[i] + dt * t[i];
}

int om_run(om_ctx *c, int nsteps, double *gd, double *iap, const double *ghs, int nT, double *q,
           double *inv_out, double *beta_out) {
  const int nC = c->nC, nE = c->nE;
  const double dt = c->p.dt;
  if (nT > 0 && c->p.time_scheme != 2) return TM_EINVAL;
  size_t nq = (size_t)nT * nC;
  /* slots one..four (state) and their tendencies (state_mod.F90:63, tend_mod.F90:44) */
  double *sg[5], *si[5], *sq[5], *tg[5], *ti[5], *tq[5];
  for (int k = 0; k < 5; ++k) {
    sg[k] = (double *)malloc(nC * sizeof(double));
    si[k] = (double *)malloc(nE * sizeof(double));
    sq[k] = (double *)malloc((nq ? nq : 1) * sizeof(double));
    tg[k] = (double *)malloc(nC * sizeof(double));
    ti[k] = (double *)malloc(nE * sizeof(double));
    tq[k] = (double *)malloc((nq ? nq : 1) * sizeof(double));
  }
  for (int step = 0; step < nsteps; ++step) {
    double beta = 1.0;
    if (c->p.time_scheme == 2) { /* time_integrate_CRK4, time_scheme_mod.F90:109-154 */
      memcpy(sg[1], gd, nC * sizeof(double));
      memcpy(si[1], iap, nE * sizeof(double));
      if (nq) memcpy(sq[1], q, nq * sizeof(double));
      const double cf[3] = {0.5, 0.5, 1.0};
      for (int k = 1; k <= 4; ++k) {
        om_spatial_operators(c, sg[k], si[k], ghs, nT, sq[k], tg[k], ti[k], tq[k], NULL);
        if (k < 4) {
          axpy(nC, sg[1], cf[k - 1] * dt, tg[k], sg[k + 1]);
          axpy(nE, si[1], cf[k - 1] * dt, ti[k], si[k + 1]);
          if (nq) axpy((int)nq, sq[1], cf[k - 1] * dt, tq[k], sq[k + 1]);
        }
      }
      for (int i = 0; i < nC; ++i) tg[0][i] = (((tg[1][i] + 2.0 * tg[2][i]) + 2.0 * tg[3][i]) + tg[4][i]) / 6.0;
      for (int e = 0; e < nE; ++e) ti[0][e] = (((ti[1][e] + 2.0 * ti[2][e]) + 2.0 * ti[3][e]) + ti[4][e]) / 6.0;
      for (size_t i = 0; i < nq; ++i) tq[0][i] = (((tq[1][i] + 2.0 * tq[2][i]) + 2.0 * tq[3][i]) + tq[4][i]) / 6.0;
      if (c->p.conserve_energy) {
        double R44 = inner(c, tg[0], ti[0], tg[0], ti[0]);
        double R12 = inner(c, tg[1], ti[1], tg[2], ti[2]);
        double R23 = inner(c, tg[2], ti[2], tg[3], ti[3]);
        double R34 = inner(c, tg[3], ti[3], tg[4], ti[4]);
        beta = (R12 + R23 + R34) / (3.0 * R44);
      }
      axpy(nC, gd, beta * dt, tg[0], gd);
      axpy(nE, iap, beta * dt, ti[0], iap);
      if (nq) axpy((int)nq, q, dt, tq[0], q);
    } else if (c->p.time_scheme == 1) { /* time_integrate_CRK3, time_scheme_mod.F90:65-107 */
      memcpy(sg[1], gd, nC * sizeof(double));
      memcpy(si[1], iap, nE * sizeof(double));
      om_spatial_operators(c, sg[1], si[1], ghs, 0, NULL, tg[1], ti[1], NULL, NULL);
      axpy(nC, sg[1], 0.5 * dt, tg[1], sg[2]);
      axpy(nE, si[1], 0.5 * dt, ti[1], si[2]);
      om_spatial_operators(c, sg[2], si[2], ghs, 0, NULL, tg[2], ti[2], NULL, NULL);
      axpy(nC, sg[1], -dt, tg[1], sg[3]);
      axpy(nE, si[1], -dt, ti[1], si[3]);
      axpy(nC, sg[3], 2.0 * dt, tg[2], sg[3]);
      axpy(nE, si[3], 2.0 * dt, ti[2], si[3]);
      om_spatial_operators(c, sg[3], si[3], ghs, 0, NULL, tg[3], ti[3], NULL, NULL);
      for (int i = 0; i < nC; ++i) tg[0][i] = ((tg[1][i] + 4.0 * tg[2][i]) + tg[3][i]) / 6.0;
      for (int e = 0; e < nE; ++e) ti[0][e] = ((ti[1][e] + 4.0 * ti[2][e]) + ti[3][e]) / 6.0;
      if (c->p.conserve_energy) {
        double R44 = inner(c, tg[0], ti[0], tg[0], ti[0]);
        double R12 = inner(c, tg[1], ti[1], tg[2], ti[2]);
        double R13 = inner(c, tg[1], ti[1], tg[3], ti[3]);
        double R23 = inner(c, tg[2], ti[2], tg[3], ti[3]);
        beta = ((2.0 * R12 - R13) + 2.0 * R23) / (3.0 * R44);
      }
      axpy(nC, gd, beta * dt, tg[0], gd);
      axpy(nE, iap, beta * dt, ti[0], iap);
    } else if (c->p.time_scheme == 3) { /* time_integrate_PC2, time_scheme_mod.F90:156-184 */
      om_spatial_operators(c, gd, iap, ghs, 0, NULL, tg[1], ti[1], NULL, NULL);
      axpy(nC, gd, 0.5 * dt, tg[1], sg[2]);
      axpy(nE, iap, 0.5 * dt, ti[1], si[2]);
      om_spatial_operators(c, sg[2], si[2], ghs, 0, NULL, tg[1], ti[1], NULL, NULL);
      axpy(nC, gd, 0.5 * dt, tg[1], sg[2]);
      axpy(nE, iap, 0.5 * dt, ti[1], si[2]);
      om_spatial_operators(c, sg[2], si[2], ghs, 0, NULL, tg[0], ti[0], NULL, NULL);
      if (c->p.conserve_energy) {
        double ip1 = inner(c, tg[1], ti[1], tg[0], ti[0]);
        double ip2 = inner(c, tg[0], ti[0], tg[0], ti[0]);
        beta = (ip1 != 0.0 && ip2 != 0.0) ? ip1 / ip2 : 1.0;
      }
      axpy(nC, gd, beta * dt, tg[0], gd);
      axpy(nE, iap, beta * dt, ti[0], iap);
    } else {
      return TM_EINVAL;
    }
    if (beta_out) beta_out[step] = beta;
    if (inv_out) om_invariants(c, gd, iap, ghs, inv_out + 4 * (size_t)step);
  }
  for (int k = 0; k < 5; ++k) {
    free(sg[k]);
    free(si[k]);
    free(sq[k]);
    free(tg[k]);
    free(ti[k]);
    free(tq[k]);
  }
  return 0;
}
