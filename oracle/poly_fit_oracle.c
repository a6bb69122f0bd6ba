/*
 * poly_fit_oracle.c -- TEST INFRASTRUCTURE (the oracle), not product code.
 *
 * Literal restatement of poly_fit_run (reference src/utils/poly_fit_mod.F90:
 * 119-609), math_inv_matrix / gaussian_elimination (src/utils/math_mod.F90:
 * 52-133) and calc_sphere_angle / calc_arc_length / norm_vector
 * (src/utils/sphere_geometry_mod.F90:186-252), including the explicit weight
 * matrix W = I and the products WTW, PTWTW, PTWTWP.  Fortran arrays are kept
 * 1-based through the F2/F3 macros.
 */
#include <math.h>
#include <string.h>

#include "trisk_oracle.h"

#define MAXFIT 30 /* maxFitCells, poly_fit_mod.F90:33 */
#define F2(a, i, j, ld) (a)[((i)-1) + ((j)-1) * (ld)]

static void norm_vector(const double *x, double *res) { /* sphere_geometry_mod.F90:186-200 */
  double n = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  if (n != 0.0)
    for (int i = 0; i < 3; ++i) res[i] = x[i] / n;
  else
    for (int i = 0; i < 3; ++i) res[i] = x[i];
}
static void cross_product(const double *x, const double *y, double *r) { /* math_mod.F90:16-26 */
  r[0] = x[1] * y[2] - x[2] * y[1];
  r[1] = x[2] * y[0] - x[0] * y[2];
  r[2] = x[0] * y[1] - x[1] * y[0];
}
static double dot_product(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

static double calc_sphere_angle(const double *a, const double *b, const double *c) { /* :204-220 */
  double t[3], nab[3], nac[3], ba[3], ca[3];
  cross_product(a, b, t);
  norm_vector(t, nab);
  cross_product(a, c, t);
  norm_vector(t, nac);
  double d = dot_product(nab, nac);
  double res = acos(fmax(fmin(d, 1.0), -1.0));
  for (int i = 0; i < 3; ++i) {
    ba[i] = b[i] - a[i];
    ca[i] = c[i] - a[i];
  }
  cross_product(ba, ca, t);
  if (dot_product(t, a) < 0.0) res = -res;
  return res;
}
static double calc_arc_length(const double *a, const double *b) { /* :245-252 */
  return acos(fmax(fmin(dot_product(a, b), 1.0), -1.0));
}

/* gaussian_elimination, math_mod.F90:90-133 */
static void gaussian_elimination(int n, double *A, int *ipiv) {
  double c[MAXFIT + 1];
  for (int i = 1; i <= n; ++i) ipiv[i] = i;
  for (int i = 1; i <= n; ++i) {
    double c1 = 0.0;
    for (int j = 1; j <= n; ++j) c1 = fmax(c1, fabs(F2(A, i, j, n)));
    c[i] = c1;
  }
  for (int j = 1; j <= n - 1; ++j) {
    double pi1 = 0.0;
    int k = j;
    for (int i = j; i <= n; ++i) {
      double pi = fabs(F2(A, ipiv[i], j, n)) / c[ipiv[i]];
      if (pi > pi1) {
        pi1 = pi;
        k = i;
      }
    }
    int itmp = ipiv[j];
    ipiv[j] = ipiv[k];
    ipiv[k] = itmp;
    for (int i = j + 1; i <= n; ++i) {
      double pj = F2(A, ipiv[i], j, n) / F2(A, ipiv[j], j, n);
      F2(A, ipiv[i], j, n) = pj;
      for (int kk = j + 1; kk <= n; ++kk) F2(A, ipiv[i], kk, n) = F2(A, ipiv[i], kk, n) - pj * F2(A, ipiv[j], kk, n);
    }
  }
}

/* math_inv_matrix, math_mod.F90:52-88 */
static void math_inv_matrix(int n, double *A, double *B) {
  double C[MAXFIT * MAXFIT];
  int ipiv[MAXFIT + 1];
  for (int i = 0; i < n * n; ++i) C[i] = 0.0;
  for (int i = 1; i <= n; ++i) F2(C, i, i, n) = 1.0;
  gaussian_elimination(n, A, ipiv);
  for (int i = 1; i <= n - 1; ++i)
    for (int j = i + 1; j <= n; ++j)
      for (int k = 1; k <= n; ++k)
        F2(C, ipiv[j], k, n) = F2(C, ipiv[j], k, n) - F2(A, ipiv[j], i, n) * F2(C, ipiv[i], k, n);
  for (int i = 1; i <= n; ++i) {
    F2(B, n, i, n) = F2(C, ipiv[n], i, n) / F2(A, ipiv[n], n, n);
    for (int j = n - 1; j >= 1; --j) {
      F2(B, j, i, n) = F2(C, ipiv[j], i, n);
      for (int k = j + 1; k <= n; ++k) F2(B, j, i, n) = F2(B, j, i, n) - F2(A, ipiv[j], k, n) * F2(B, k, i, n);
      F2(B, j, i, n) = F2(B, j, i, n) / F2(A, ipiv[j], j, n);
    }
  }
}

/* One least-squares fit of poly_fit_run: coordinates -> theta(1:m), B(1:n,1:m).
 * xyz[0..m] are the fit points divided by the radius (poly_fit_mod.F90:188-192). */
static void fit_one(int m, int n, double xyz[][3], double radius, double *theta, double *Bout) {
  double xp[MAXFIT + 1], yp[MAXFIT + 1];
  static double P[MAXFIT * MAXFIT], W[MAXFIT * MAXFIT], WTW[MAXFIT * MAXFIT], PT[MAXFIT * MAXFIT];
  static double PTWTW[MAXFIT * MAXFIT], PTWTWP[MAXFIT * MAXFIT], Binv[MAXFIT * MAXFIT];
  xp[0] = 0.0;
  yp[0] = 0.0;
  for (int i = 1; i <= m; ++i) {
    theta[i] = calc_sphere_angle(xyz[0], xyz[1], xyz[i]);
    double d = radius * calc_arc_length(xyz[0], xyz[i]);
    xp[i] = cos(theta[i]) * d;
    yp[i] = sin(theta[i]) * d;
  }
  memset(P, 0, sizeof(P));
  memset(W, 0, sizeof(W));
  for (int i = 1; i <= m; ++i) {
    double x = xp[i], y = yp[i];
    F2(P, i, 1, m) = x;
    F2(P, i, 2, m) = y;
    F2(P, i, 3, m) = x * x;
    F2(P, i, 4, m) = x * y;
    F2(P, i, 5, m) = y * y;
    if (n >= 9) {
      F2(P, i, 6, m) = x * x * x;
      F2(P, i, 7, m) = y * (x * x);
      F2(P, i, 8, m) = x * (y * y);
      F2(P, i, 9, m) = y * y * y;
    }
    if (n >= 14) {
      F2(P, i, 10, m) = (x * x) * (x * x);
      F2(P, i, 11, m) = y * (x * x * x);
      F2(P, i, 12, m) = (x * x) * (y * y);
      F2(P, i, 13, m) = x * (y * y * y);
      F2(P, i, 14, m) = (y * y) * (y * y);
    }
    F2(W, i, i, m) = 1.0;
  }
  /* WTW = matmul(transpose(W), W); PT = transpose(P); PTWTW = PT*WTW; PTWTWP = PTWTW*P */
  for (int i = 1; i <= m; ++i)
    for (int j = 1; j <= m; ++j) {
      double s = 0.0;
      for (int k = 1; k <= m; ++k) s += F2(W, k, i, m) * F2(W, k, j, m);
      F2(WTW, i, j, m) = s;
    }
  for (int i = 1; i <= n; ++i)
    for (int j = 1; j <= m; ++j) F2(PT, i, j, n) = F2(P, j, i, m);
  for (int i = 1; i <= n; ++i)
    for (int j = 1; j <= m; ++j) {
      double s = 0.0;
      for (int k = 1; k <= m; ++k) s += F2(PT, i, k, n) * F2(WTW, k, j, m);
      F2(PTWTW, i, j, n) = s;
    }
  for (int i = 1; i <= n; ++i)
    for (int j = 1; j <= n; ++j) {
      double s = 0.0;
      for (int k = 1; k <= m; ++k) s += F2(PTWTW, i, k, n) * F2(P, k, j, m);
      F2(PTWTWP, i, j, n) = s;
    }
  math_inv_matrix(n, PTWTWP, Binv);
  for (int i = 1; i <= n; ++i)
    for (int j = 1; j <= m; ++j) {
      double s = 0.0;
      for (int k = 1; k <= n; ++k) s += F2(Binv, i, k, n) * F2(PTWTW, k, j, n);
      F2(Bout, i, j, n) = s;
    }
}

static double deriv2(const double *B, int n, int j, double th) { /* :239-243 */
  double ct = cos(th), st = sin(th);
  double d2fdx2 = 2.0 * F2(B, 3, j, n) * (ct * ct);
  double d2fdxdy = F2(B, 4, j, n) * ct * st;
  double d2fdy2 = 2.0 * F2(B, 5, j, n) * (st * st);
  return d2fdx2 + 2.0 * d2fdxdy + d2fdy2;
}
static double deriv4(const double *B, int n, int j, double th) { /* :385-391 */
  double ct = cos(th), st = sin(th);
  double d4fdx4 = 24.0 * F2(B, 10, j, n) * ((ct * ct) * (ct * ct));
  double d4fdx3dy = 6.0 * F2(B, 11, j, n) * (ct * ct * ct) * st;
  double d4fdx2dy2 = 4.0 * F2(B, 12, j, n) * (ct * ct) * (st * st);
  double d4fdxdy3 = 6.0 * F2(B, 13, j, n) * ct * (st * st * st);
  double d4fdy4 = 24.0 * F2(B, 14, j, n) * ((st * st) * (st * st));
  return d4fdx4 + 4.0 * d4fdx3dy + 6.0 * d4fdx2dy2 + 4.0 * d4fdxdy3 + d4fdy4;
}

int om_poly_fit(const tm_mesh_desc *m, double radius, tm_fit_desc *f) {
  const int ME = m->maxEdges, MF = f->maxFit;
  int fit2[MAXFIT + 1], fit4[MAXFIT + 1];
  double xyz[MAXFIT + 1][3], theta[MAXFIT + 1];
  static double B[MAXFIT * MAXFIT];
#define CELL(x, i) ((x)[(i)-1] * radius)
  /* Fitting on cells, poly_fit_mod.F90:151-394 */
  for (int iCell = 1; iCell <= m->nCells; ++iCell) {
    int ne = m->nEdgesOnCell[iCell - 1];
    fit2[0] = fit4[0] = iCell;
    for (int i = 1; i <= ne; ++i) fit2[i] = fit4[i] = m->cellsOnCell[(iCell - 1) * ME + i - 1];
    int n2 = ne + 1, n4 = ne + 1;
    for (int i = 1; i <= ne; ++i) {
      int h = fit4[i];
      for (int j = 1; j <= m->nEdgesOnCell[h - 1]; ++j) {
        int ng = m->cellsOnCell[(h - 1) * ME + j - 1], rec = 0;
        for (int k = 0; k <= n4 - 1; ++k)
          if (ng == fit4[k]) {
            rec = 1;
            break;
          }
        if (!rec) {
          if (n4 > MAXFIT) return TM_ESTENCIL;
          fit4[n4++] = ng;
        }
      }
    }
    for (int order = 2; order <= 4; order += 2) {
      int nfit = order == 2 ? n2 : n4, *fit = order == 2 ? fit2 : fit4, n = order == 2 ? 5 : 14;
      int *nF = order == 2 ? f->nFit2Cells : f->nFit4Cells, *F = order == 2 ? f->fit2Cells : f->fit4Cells;
      double *D = order == 2 ? f->deriv2OnCell : f->deriv4OnCell;
      if (!D && !F && !nF) continue;
      if (nfit > MF) return TM_ESTENCIL;
      for (int i = 0; i < nfit; ++i) {
        xyz[i][0] = CELL(m->xCell, fit[i]) / radius;
        xyz[i][1] = CELL(m->yCell, fit[i]) / radius;
        xyz[i][2] = CELL(m->zCell, fit[i]) / radius;
      }
      fit_one(nfit - 1, n, xyz, radius, theta, B);
      if (nF) nF[iCell - 1] = nfit;
      if (F)
        for (int i = 0; i < MF; ++i) F[(iCell - 1) * MF + i] = i < nfit ? fit[i] : 0;
      if (D)
        for (int i = 1; i <= ne; ++i) {
          int iEdge = m->edgesOnCell[(iCell - 1) * ME + i - 1];
          int k = (iCell == m->cellsOnEdge[2 * (iEdge - 1)]) ? 1 : 2;
          double *w = D + ((size_t)(iEdge - 1) * 2 + (k - 1)) * MF;
          for (int j = 0; j < MF; ++j) w[j] = 0.0;
          for (int j = 1; j <= nfit - 1; ++j) w[j] = order == 2 ? deriv2(B, n, j, theta[i]) : deriv4(B, n, j, theta[i]);
        }
    }
  }
  /* Fitting on vertices, poly_fit_mod.F90:399-607 */
  for (int iVertex = 1; iVertex <= m->nVertices; ++iVertex) {
    int vov[4];
    for (int i = 1; i <= 3; ++i) { /* verticesOnVertex, mesh_mod.F90:190-203 */
      int e = m->edgesOnVertex[(iVertex - 1) * 3 + i - 1];
      int a = m->verticesOnEdge[2 * (e - 1)], b = m->verticesOnEdge[2 * (e - 1) + 1];
      if (iVertex == a) vov[i] = b;
      if (iVertex == b) vov[i] = a;
    }
    fit2[0] = iVertex;
    for (int i = 1; i <= 3; ++i) fit2[i] = vov[i];
    int n2 = 4;
    for (int i = 1; i <= 3; ++i) {
      int hc = m->cellsOnVertex[(iVertex - 1) * 3 + i - 1];
      for (int j = 1; j <= m->nEdgesOnCell[hc - 1]; ++j) {
        int ng = m->verticesOnCell[(hc - 1) * ME + j - 1], rec = 0;
        for (int k = 0; k <= n2 - 1; ++k)
          if (ng == fit2[k]) {
            rec = 1;
            break;
          }
        if (!rec) {
          if (n2 > MAXFIT) return TM_ESTENCIL;
          fit2[n2++] = ng;
        }
      }
    }
    int n4 = n2;
    memcpy(fit4, fit2, sizeof(fit2));
    for (int q = 1; q <= 3; ++q) {
      int hv = vov[q];
      for (int i = 1; i <= 3; ++i) {
        int hc = m->cellsOnVertex[(hv - 1) * 3 + i - 1];
        for (int j = 1; j <= m->nEdgesOnCell[hc - 1]; ++j) {
          int ng = m->verticesOnCell[(hc - 1) * ME + j - 1], rec = 0;
          for (int k = 0; k <= n4 - 1; ++k)
            if (ng == fit4[k]) {
              rec = 1;
              break;
            }
          if (!rec) {
            if (n4 > MAXFIT) return TM_ESTENCIL;
            fit4[n4++] = ng;
          }
        }
      }
    }
    for (int order = 2; order <= 4; order += 2) {
      int nfit = order == 2 ? n2 : n4, *fit = order == 2 ? fit2 : fit4, n = order == 2 ? 5 : 14;
      int *nF = order == 2 ? f->nFit2Vertices : f->nFit4Vertices, *F = order == 2 ? f->fit2Vertices : f->fit4Vertices;
      double *D = order == 2 ? f->deriv2OnVertex : f->deriv4OnVertex;
      if (!D && !F && !nF) continue;
      if (nfit > MF) return TM_ESTENCIL;
      for (int i = 0; i < nfit; ++i) {
        xyz[i][0] = CELL(m->xVertex, fit[i]) / radius;
        xyz[i][1] = CELL(m->yVertex, fit[i]) / radius;
        xyz[i][2] = CELL(m->zVertex, fit[i]) / radius;
      }
      fit_one(nfit - 1, n, xyz, radius, theta, B);
      if (nF) nF[iVertex - 1] = nfit;
      if (F)
        for (int i = 0; i < MF; ++i) F[(iVertex - 1) * MF + i] = i < nfit ? fit[i] : 0;
      if (D)
        for (int i = 1; i <= 3; ++i) {
          int iEdge = m->edgesOnVertex[(iVertex - 1) * 3 + i - 1];
          int k = (iVertex == m->verticesOnEdge[2 * (iEdge - 1)]) ? 1 : 2;
          double *w = D + ((size_t)(iEdge - 1) * 2 + (k - 1)) * MF;
          for (int j = 0; j < MF; ++j) w[j] = 0.0;
          for (int j = 1; j <= nfit - 1; ++j) w[j] = order == 2 ? deriv2(B, n, j, theta[i]) : deriv4(B, n, j, theta[i]);
        }
    }
  }
#undef CELL
  return 0;
}

/* Exported for the library-routine check of math_inv_matrix (numpy.linalg.inv).
 * A is overwritten (as the reference's intent(inout) dummy is). */
void om_inv_matrix(int n, double *A, double *B) { math_inv_matrix(n, A, B); }
