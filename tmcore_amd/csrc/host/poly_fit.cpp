// Init-time least-squares derivative weights for the upwind corrections.
//
// Restates poly_fit_run (reference src/utils/poly_fit_mod.F90:119-609) with
// its inverse math_inv_matrix / gaussian_elimination (src/utils/math_mod.F90:
// 52-133) and calc_sphere_angle / calc_arc_length / norm_vector
// (src/utils/sphere_geometry_mod.F90:186-252).  The operation order of those
// routines is kept, because the angle of the first stencil point,
// acos(nab.nab), is not exactly zero in floating point and the derivative
// weights inherit that rounding.  Loops over cells and vertices run in
// parallel; each fit is independent.

#include "tmcore_host.h"

#include <cmath>
#include <cstdint>
#include <vector>

namespace {

constexpr int kMaxFit = 31;  // reference maxFitCells = 30, slots 0..30

struct P3 {
  double v[3];
};

inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline void cross3(const double *x, const double *y, double *r) {
  r[0] = x[1] * y[2] - x[2] * y[1];
  r[1] = x[2] * y[0] - x[0] * y[2];
  r[2] = x[0] * y[1] - x[1] * y[0];
}
inline void normalise3(double *x) {
  double n = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  if (n != 0.0) {
    x[0] /= n;
    x[1] /= n;
    x[2] /= n;
  }
}
// Dihedral angle between planes AB and AC, signed by the turn at A
// (sphere_geometry_mod.F90:204-220).
double sphere_angle(const double *a, const double *b, const double *c) {
  double nab[3], nac[3];
  cross3(a, b, nab);
  normalise3(nab);
  cross3(a, c, nac);
  normalise3(nac);
  double d = dot3(nab, nac);
  d = d < 1.0 ? d : 1.0;
  d = d > -1.0 ? d : -1.0;
  double res = std::acos(d);
  double ba[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  double ca[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  double x[3];
  cross3(ba, ca, x);
  if (dot3(x, a) < 0.0) res = -res;
  return res;
}
double arc_length(const double *a, const double *b) {
  double d = dot3(a, b);
  d = d < 1.0 ? d : 1.0;
  d = d > -1.0 ? d : -1.0;
  return std::acos(d);
}

// Inverse by scaled partial pivoting, in place on a (n x n, column-major
// a[i + j*n]); restates math_inv_matrix + gaussian_elimination.
bool invert(int n, double *a, double *b) {
  int ipiv[kMaxFit];
  double c[kMaxFit];
  double C[kMaxFit * kMaxFit];
  for (int i = 0; i < n; ++i) ipiv[i] = i;
  for (int i = 0; i < n; ++i) {
    double c1 = 0.0;
    for (int j = 0; j < n; ++j) c1 = std::fmax(c1, std::fabs(a[i + j * n]));
    c[i] = c1;
  }
  for (int j = 0; j < n - 1; ++j) {
    double pi1 = 0.0;
    int k = -1;
    for (int i = j; i < n; ++i) {
      double pi = std::fabs(a[ipiv[i] + j * n]) / c[ipiv[i]];
      if (pi > pi1) {
        pi1 = pi;
        k = i;
      }
    }
    if (k < 0) return false;
    int t = ipiv[j];
    ipiv[j] = ipiv[k];
    ipiv[k] = t;
    for (int i = j + 1; i < n; ++i) {
      double pj = a[ipiv[i] + j * n] / a[ipiv[j] + j * n];
      a[ipiv[i] + j * n] = pj;
      for (int kk = j + 1; kk < n; ++kk) a[ipiv[i] + kk * n] -= pj * a[ipiv[j] + kk * n];
    }
  }
  for (int i = 0; i < n * n; ++i) C[i] = 0.0;
  for (int i = 0; i < n; ++i) C[i + i * n] = 1.0;
  for (int i = 0; i < n - 1; ++i)
    for (int j = i + 1; j < n; ++j)
      for (int k = 0; k < n; ++k) C[ipiv[j] + k * n] -= a[ipiv[j] + i * n] * C[ipiv[i] + k * n];
  for (int i = 0; i < n; ++i) {
    if (a[ipiv[n - 1] + (n - 1) * n] == 0.0) return false;
    b[(n - 1) + i * n] = C[ipiv[n - 1] + i * n] / a[ipiv[n - 1] + (n - 1) * n];
    for (int j = n - 2; j >= 0; --j) {
      double s = C[ipiv[j] + i * n];
      for (int k = j + 1; k < n; ++k) s -= a[ipiv[j] + k * n] * b[k + i * n];
      b[j + i * n] = s / a[ipiv[j] + j * n];
    }
  }
  return true;
}

// Least-squares operator of the fit: rows of Bm (n x m, column-major
// Bm[i + j*n]) map the m neighbour differences to polynomial coefficients.
// pts[0] is the centre, unit vectors; theta receives the stencil angles.
bool lsq_operator(int m, int n, const P3 *pts, double radius, double *theta, double *Bm) {
  double xp[kMaxFit], yp[kMaxFit];
  xp[0] = 0.0;
  yp[0] = 0.0;
  for (int i = 1; i <= m; ++i) {
    theta[i] = sphere_angle(pts[0].v, pts[1].v, pts[i].v);
    double d = radius * arc_length(pts[0].v, pts[i].v);
    xp[i] = std::cos(theta[i]) * d;
    yp[i] = std::sin(theta[i]) * d;
  }
  // P (m x n) column-major P[(i-1) + (q-1)*m]
  double P[kMaxFit * 14];
  for (int i = 1; i <= m; ++i) {
    double x = xp[i], y = yp[i];
    double row[14] = {x, y, x * x, x * y, y * y, x * x * x, y * (x * x), x * (y * y), y * y * y,
                      x * x * x * x, y * (x * x * x), (x * x) * (y * y), x * (y * y * y), y * y * y * y};
    for (int q = 0; q < n; ++q) P[(i - 1) + q * m] = row[q];
  }
  // x*x*x*x above must equal Fortran xp**4 = (x*x)*(x*x); fix the terms whose
  // association differs.
  for (int i = 1; i <= m && n == 14; ++i) {
    double x = xp[i], y = yp[i];
    P[(i - 1) + 5 * m] = (x * x) * x;
    P[(i - 1) + 8 * m] = (y * y) * y;
    P[(i - 1) + 9 * m] = (x * x) * (x * x);
    P[(i - 1) + 10 * m] = y * ((x * x) * x);
    P[(i - 1) + 12 * m] = x * ((y * y) * y);
    P[(i - 1) + 13 * m] = (y * y) * (y * y);
  }
  for (int i = 1; i <= m && n == 9; ++i) {
    double x = xp[i], y = yp[i];
    P[(i - 1) + 5 * m] = (x * x) * x;
    P[(i - 1) + 8 * m] = (y * y) * y;
  }
  double A[14 * 14], Binv[14 * 14];
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double s = 0.0;
      for (int k = 0; k < m; ++k) s += P[k + i * m] * P[k + j * m];
      A[i + j * n] = s;
    }
  if (!invert(n, A, Binv)) return false;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < m; ++j) {
      double s = 0.0;
      for (int k = 0; k < n; ++k) s += Binv[i + k * n] * P[j + k * m];
      Bm[i + j * n] = s;
    }
  return true;
}

inline double d2_weight(const double *Bm, int n, int j, double th) {
  double ct = std::cos(th), st = std::sin(th);
  double d2fdx2 = 2.0 * Bm[2 + j * n] * (ct * ct);
  double d2fdxdy = Bm[3 + j * n] * ct * st;
  double d2fdy2 = 2.0 * Bm[4 + j * n] * (st * st);
  return d2fdx2 + 2.0 * d2fdxdy + d2fdy2;
}
inline double d4_weight(const double *Bm, int n, int j, double th) {
  double ct = std::cos(th), st = std::sin(th);
  double c2 = ct * ct, s2 = st * st;
  double d4fdx4 = 24.0 * Bm[9 + j * n] * (c2 * c2);
  double d4fdx3dy = 6.0 * Bm[10 + j * n] * ((c2 * ct)) * st;
  double d4fdx2dy2 = 4.0 * Bm[11 + j * n] * c2 * s2;
  double d4fdxdy3 = 6.0 * Bm[12 + j * n] * ct * (s2 * st);
  double d4fdy4 = 24.0 * Bm[13 + j * n] * (s2 * s2);
  return d4fdx4 + 4.0 * d4fdx3dy + 6.0 * d4fdx2dy2 + 4.0 * d4fdxdy3 + d4fdy4;
}

}  // namespace

extern "C" int tm_poly_fit(const tm_mesh_desc *m, double radius, tm_fit_desc *f) {
  if (!m || !f || f->maxFit < 2 || f->maxFit > kMaxFit) return TM_EINVAL;
  const int32_t nC = m->nCells, nV = m->nVertices;
  const int ME = m->maxEdges, VD = m->vertexDegree, MF = f->maxFit;
  if (VD != 3) return TM_EINVAL;
  const bool want_c2 = f->deriv2OnCell || f->fit2Cells || f->nFit2Cells;
  const bool want_c4 = f->deriv4OnCell || f->fit4Cells || f->nFit4Cells;
  const bool want_v2 = f->deriv2OnVertex || f->fit2Vertices || f->nFit2Vertices;
  const bool want_v4 = f->deriv4OnVertex || f->fit4Vertices || f->nFit4Vertices;
  int status = TM_OK;

  if (want_c2 || want_c4) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int32_t c = 0; c < nC; ++c) {
      int32_t fit[kMaxFit];
      int nfit2, nfit4;
      fit[0] = c;
      const int ne = m->nEdgesOnCell[c];
      for (int i = 1; i <= ne; ++i) fit[i] = m->cellsOnCell[(size_t)c * ME + i - 1] - 1;
      nfit2 = ne + 1;
      nfit4 = ne + 1;
      for (int i = 1; i <= ne; ++i) {
        int32_t h = fit[i];
        for (int j = 0; j < m->nEdgesOnCell[h]; ++j) {
          int32_t ng = m->cellsOnCell[(size_t)h * ME + j] - 1;
          bool rec = false;
          for (int k = 0; k < nfit4; ++k)
            if (fit[k] == ng) {
              rec = true;
              break;
            }
          if (!rec) {
            if (nfit4 >= kMaxFit) {
#pragma omp atomic write
              status = TM_ESTENCIL;
              continue;
            }
            fit[nfit4++] = ng;
          }
        }
      }
      P3 pts[kMaxFit];
      for (int i = 0; i < nfit4; ++i) {
        int32_t q = fit[i];
        pts[i].v[0] = (m->xCell[q] * radius) / radius;
        pts[i].v[1] = (m->yCell[q] * radius) / radius;
        pts[i].v[2] = (m->zCell[q] * radius) / radius;
      }
      double theta[kMaxFit], Bm[14 * kMaxFit];
      if (want_c2) {
        if (nfit2 > MF) {
#pragma omp atomic write
          status = TM_ESTENCIL;
        } else if (!lsq_operator(nfit2 - 1, 5, pts, radius, theta, Bm)) {
#pragma omp atomic write
          status = TM_ESINGULAR;
        } else {
          if (f->nFit2Cells) f->nFit2Cells[c] = nfit2;
          if (f->fit2Cells)
            for (int i = 0; i < MF; ++i) f->fit2Cells[(size_t)c * MF + i] = i < nfit2 ? fit[i] + 1 : 0;
          if (f->deriv2OnCell)
            for (int i = 1; i <= ne; ++i) {
              int32_t E = m->edgesOnCell[(size_t)c * ME + i - 1] - 1;
              int k = (c == m->cellsOnEdge[2 * (size_t)E] - 1) ? 0 : 1;
              double *w = f->deriv2OnCell + ((size_t)E * 2 + k) * MF;
              w[0] = 0.0;
              for (int j = 1; j < nfit2; ++j) w[j] = d2_weight(Bm, 5, j - 1, theta[i]);
              for (int j = nfit2; j < MF; ++j) w[j] = 0.0;
            }
        }
      }
      if (want_c4) {
        if (nfit4 > MF) {
#pragma omp atomic write
          status = TM_ESTENCIL;
        } else if (!lsq_operator(nfit4 - 1, 14, pts, radius, theta, Bm)) {
#pragma omp atomic write
          status = TM_ESINGULAR;
        } else {
          if (f->nFit4Cells) f->nFit4Cells[c] = nfit4;
          if (f->fit4Cells)
            for (int i = 0; i < MF; ++i) f->fit4Cells[(size_t)c * MF + i] = i < nfit4 ? fit[i] + 1 : 0;
          if (f->deriv4OnCell)
            for (int i = 1; i <= ne; ++i) {
              int32_t E = m->edgesOnCell[(size_t)c * ME + i - 1] - 1;
              int k = (c == m->cellsOnEdge[2 * (size_t)E] - 1) ? 0 : 1;
              double *w = f->deriv4OnCell + ((size_t)E * 2 + k) * MF;
              w[0] = 0.0;
              for (int j = 1; j < nfit4; ++j) w[j] = d4_weight(Bm, 14, j - 1, theta[i]);
              for (int j = nfit4; j < MF; ++j) w[j] = 0.0;
            }
        }
      }
    }
  }

  if (want_v2 || want_v4) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int32_t v = 0; v < nV; ++v) {
      int32_t fit[kMaxFit], vov[3];
      // verticesOnVertex as mesh_init derives it (mesh_mod.F90:190-203)
      for (int i = 0; i < 3; ++i) {
        int32_t E = m->edgesOnVertex[(size_t)v * 3 + i] - 1;
        int32_t a = m->verticesOnEdge[2 * (size_t)E] - 1, b = m->verticesOnEdge[2 * (size_t)E + 1] - 1;
        vov[i] = (a == v) ? b : a;
      }
      fit[0] = v;
      for (int i = 0; i < 3; ++i) fit[i + 1] = vov[i];
      int nfit = 4;
      auto add_cells_of = [&](int32_t hv, int &cnt) {
        for (int i = 0; i < 3; ++i) {
          int32_t hc = m->cellsOnVertex[(size_t)hv * 3 + i] - 1;
          for (int j = 0; j < m->nEdgesOnCell[hc]; ++j) {
            int32_t ng = m->verticesOnCell[(size_t)hc * ME + j] - 1;
            bool rec = false;
            for (int k = 0; k < cnt; ++k)
              if (fit[k] == ng) {
                rec = true;
                break;
              }
            if (!rec) {
              if (cnt >= kMaxFit) return false;
              fit[cnt++] = ng;
            }
          }
        }
        return true;
      };
      bool ok = add_cells_of(v, nfit);
      int nfit2 = nfit, nfit4 = nfit;
      if (ok && want_v4)
        for (int q = 0; q < 3 && ok; ++q) ok = add_cells_of(vov[q], nfit4);
      if (!ok) {
#pragma omp atomic write
        status = TM_ESTENCIL;
        continue;
      }
      P3 pts[kMaxFit];
      const int ntot = want_v4 ? nfit4 : nfit2;
      for (int i = 0; i < ntot; ++i) {
        int32_t q = fit[i];
        pts[i].v[0] = (m->xVertex[q] * radius) / radius;
        pts[i].v[1] = (m->yVertex[q] * radius) / radius;
        pts[i].v[2] = (m->zVertex[q] * radius) / radius;
      }
      double theta[kMaxFit], Bm[14 * kMaxFit];
      if (want_v2) {
        if (nfit2 > MF) {
#pragma omp atomic write
          status = TM_ESTENCIL;
        } else if (!lsq_operator(nfit2 - 1, 5, pts, radius, theta, Bm)) {
#pragma omp atomic write
          status = TM_ESINGULAR;
        } else {
          if (f->nFit2Vertices) f->nFit2Vertices[v] = nfit2;
          if (f->fit2Vertices)
            for (int i = 0; i < MF; ++i) f->fit2Vertices[(size_t)v * MF + i] = i < nfit2 ? fit[i] + 1 : 0;
          if (f->deriv2OnVertex)
            for (int i = 1; i <= 3; ++i) {
              int32_t E = m->edgesOnVertex[(size_t)v * 3 + i - 1] - 1;
              int k = (v == m->verticesOnEdge[2 * (size_t)E] - 1) ? 0 : 1;
              double *w = f->deriv2OnVertex + ((size_t)E * 2 + k) * MF;
              w[0] = 0.0;
              for (int j = 1; j < nfit2; ++j) w[j] = d2_weight(Bm, 5, j - 1, theta[i]);
              for (int j = nfit2; j < MF; ++j) w[j] = 0.0;
            }
        }
      }
      if (want_v4) {
        if (nfit4 > MF) {
#pragma omp atomic write
          status = TM_ESTENCIL;
        } else if (!lsq_operator(nfit4 - 1, 14, pts, radius, theta, Bm)) {
#pragma omp atomic write
          status = TM_ESINGULAR;
        } else {
          if (f->nFit4Vertices) f->nFit4Vertices[v] = nfit4;
          if (f->fit4Vertices)
            for (int i = 0; i < MF; ++i) f->fit4Vertices[(size_t)v * MF + i] = i < nfit4 ? fit[i] + 1 : 0;
          if (f->deriv4OnVertex)
            for (int i = 1; i <= 3; ++i) {
              int32_t E = m->edgesOnVertex[(size_t)v * 3 + i - 1] - 1;
              int k = (v == m->verticesOnEdge[2 * (size_t)E] - 1) ? 0 : 1;
              double *w = f->deriv4OnVertex + ((size_t)E * 2 + k) * MF;
              w[0] = 0.0;
              for (int j = 1; j < nfit4; ++j) w[j] = d4_weight(Bm, 14, j - 1, theta[i]);
              for (int j = nfit4; j < MF; ++j) w[j] = 0.0;
            }
        }
      }
    }
  }
  return status;
}
