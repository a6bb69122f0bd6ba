/* Oracle (test infrastructure only): CPU restatement, in plain C, of the MPAS
 * vector reconstruction the reference uses for its history output:
 *   mpas_initialize_vectors   (src/core/vector_reconstruction/mpas_vector_operations.F90:493-557)
 *   mpas_init_reconstruct     (mpas_vector_reconstruction.F90:52-125)
 *   mpas_rbf_interp_func_3D_plane_vec_const_dir_comp_coeffs
 *                             (mpas_rbf_interpolation.F90:1076-1142)
 *   mpas_set_up_vector_dirichlet_rbf_matrix_and_rhs (:1524-1556), evaluate_rbf (:1366-1373)
 *   mpas_legs / elgs          (:1667-1697, :1777-1837; Tao Pang's scaled
 *                              partial-pivoting Gaussian elimination)
 *   mpas_reconstruct_1d       (mpas_vector_reconstruction.F90:195-247)
 * Geometry is radius-scaled before the coefficients are formed, as mesh_init
 * scales it (mesh_mod.F90:205-216) before tmcore_swm_init calls these
 * (tmcore_swm_mod.F90:44-45).  Mesh arrays follow tmcore_host.h (1-based,
 * [n][slot]).  Only tests/ may use this file. */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "trisk_oracle.h"

static void unit3(double *v) { /* mpas_unit_vec_in_r3 */
  double mag = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  v[0] = v[0] / mag;
  v[1] = v[1] / mag;
  v[2] = v[2] / mag;
}

static double rbf(double r2) { return 1.0 / sqrt(1.0 + r2); } /* inverse multiquadric */

/* elgs + mpas_legs on a column-major n x n matrix a (a[i + j*n] = A(i,j)) */
static void legs(double *a, int n, double *b, double *x) {
  int *indx = (int *)malloc(n * sizeof(int));
  double *c = (double *)malloc(n * sizeof(double));
#define A(i, j) a[(i) + (size_t)(j) * n]
  for (int i = 0; i < n; ++i) indx[i] = i;
  for (int i = 0; i < n; ++i) {
    double c1 = 0.0;
    for (int j = 0; j < n; ++j) c1 = fmax(c1, fabs(A(i, j)));
    c[i] = c1;
  }
  for (int j = 0; j < n - 1; ++j) {
    double pi1 = 0.0;
    int k = j;
    for (int i = j; i < n; ++i) {
      double pi = fabs(A(indx[i], j)) / c[indx[i]];
      if (pi > pi1) {
        pi1 = pi;
        k = i;
      }
    }
    int t = indx[j];
    indx[j] = indx[k];
    indx[k] = t;
    for (int i = j + 1; i < n; ++i) {
      double pj = A(indx[i], j) / A(indx[j], j);
      A(indx[i], j) = pj;
      for (int kk = j + 1; kk < n; ++kk) A(indx[i], kk) = A(indx[i], kk) - pj * A(indx[j], kk);
    }
  }
  for (int i = 0; i < n - 1; ++i)
    for (int j = i + 1; j < n; ++j) b[indx[j]] = b[indx[j]] - A(indx[j], i) * b[indx[i]];
  x[n - 1] = b[indx[n - 1]] / A(indx[n - 1], n - 1);
  for (int i = n - 2; i >= 0; --i) {
    x[i] = b[indx[i]];
    for (int j = i + 1; j < n; ++j) x[i] = x[i] - A(indx[i], j) * x[j];
    x[i] = x[i] / A(indx[i], i);
  }
#undef A
  free(indx);
  free(c);
}

/* coeffs[(c * maxEdges + i) * 3 + d] = coeffs_reconstruct(d, i, c) */
int om_reconstruct_coeffs(const tm_mesh_desc *m, double radius, double *coeffs) {
  const int nC = m->nCells, nE = m->nEdges, ME = m->maxEdges;
  double *en = (double *)malloc((size_t)3 * nE * sizeof(double));
  for (int e = 0; e < nE; ++e) { /* edgeNormalVectors: cell 1 -> cell 2 */
    int c1 = m->cellsOnEdge[2 * (size_t)e] - 1, c2 = m->cellsOnEdge[2 * (size_t)e + 1] - 1;
    double v[3] = {m->xCell[c2] * radius - m->xCell[c1] * radius, m->yCell[c2] * radius - m->yCell[c1] * radius,
                   m->zCell[c2] * radius - m->zCell[c1] * radius};
    unit3(v);
    memcpy(en + 3 * (size_t)e, v, sizeof(v));
  }
  memset(coeffs, 0, (size_t)nC * ME * 3 * sizeof(double));
  for (int c = 0; c < nC; ++c) {
    const int n = m->nEdgesOnCell[c], N = n + 2;
    double cc[3] = {m->xCell[c] * radius, m->yCell[c] * radius, m->zCell[c] * radius};
    /* cellTangentPlane: xHat toward the first edge's normal, yHat = rHat x xHat */
    double rh[3] = {cc[0], cc[1], cc[2]};
    unit3(rh);
    const double *n1 = en + 3 * (size_t)(m->edgesOnCell[(size_t)c * ME] - 1);
    double nd = n1[0] * rh[0] + n1[1] * rh[1] + n1[2] * rh[2];
    double xh[3] = {n1[0] - nd * rh[0], n1[1] - nd * rh[1], n1[2] - nd * rh[2]};
    unit3(xh);
    double yh[3] = {rh[1] * xh[2] - rh[2] * xh[1], rh[2] * xh[0] - rh[0] * xh[2], rh[0] * xh[1] - rh[1] * xh[0]};
    unit3(yh);
    double *loc = (double *)malloc((size_t)n * 3 * sizeof(double)), *nrm = (double *)malloc((size_t)n * 3 * sizeof(double));
    for (int i = 0; i < n; ++i) {
      int e = m->edgesOnCell[(size_t)c * ME + i] - 1;
      loc[3 * i + 0] = m->xEdge[e] * radius;
      loc[3 * i + 1] = m->yEdge[e] * radius;
      loc[3 * i + 2] = m->zEdge[e] * radius;
      memcpy(nrm + 3 * i, en + 3 * (size_t)e, 3 * sizeof(double));
    }
    double alpha = 0.0;
    for (int i = 0; i < n; ++i) {
      double d0 = cc[0] - loc[3 * i], d1 = cc[1] - loc[3 * i + 1], d2 = cc[2] - loc[3 * i + 2];
      alpha = alpha + sqrt((d0 * d0 + d1 * d1) + d2 * d2);
    }
    alpha = alpha / n;
    /* projections onto the plane (sum(a*b) in slot order) */
    double *ps = (double *)malloc((size_t)n * 2 * sizeof(double)), *pu = (double *)malloc((size_t)n * 2 * sizeof(double));
    const double *pb[2] = {xh, yh};
    for (int i = 0; i < n; ++i)
      for (int k = 0; k < 2; ++k) {
        ps[2 * i + k] = (loc[3 * i] * pb[k][0] + loc[3 * i + 1] * pb[k][1]) + loc[3 * i + 2] * pb[k][2];
        pu[2 * i + k] = (nrm[3 * i] * pb[k][0] + nrm[3 * i + 1] * pb[k][1]) + nrm[3 * i + 2] * pb[k][2];
      }
    double pd[2];
    for (int k = 0; k < 2; ++k) pd[k] = (cc[0] * pb[k][0] + cc[1] * pb[k][1]) + cc[2] * pb[k][2];
    double *mat = (double *)calloc((size_t)N * N, sizeof(double)), *rhs = (double *)calloc((size_t)N * 2, sizeof(double));
#define M(i, j) mat[(i) + (size_t)(j) * N]
    const double a2 = alpha * alpha;
    for (int j = 0; j < n; ++j)
      for (int i = j; i < n; ++i) {
        double d0 = ps[2 * i] - ps[2 * j], d1 = ps[2 * i + 1] - ps[2 * j + 1];
        double r2 = (d0 * d0 + d1 * d1) / a2;
        double dot = pu[2 * i] * pu[2 * j] + pu[2 * i + 1] * pu[2 * j + 1];
        M(i, j) = rbf(r2) * dot;
        M(j, i) = M(i, j);
      }
    for (int j = 0; j < n; ++j) {
      double d0 = pd[0] - ps[2 * j], d1 = pd[1] - ps[2 * j + 1];
      double v = rbf((d0 * d0 + d1 * d1) / a2);
      rhs[j] = v * pu[2 * j];
      rhs[N + j] = v * pu[2 * j + 1];
    }
    for (int i = 0; i < n; ++i) {
      M(i, n) = pu[2 * i];
      M(i, n + 1) = pu[2 * i + 1];
      M(n, i) = M(i, n);
      M(n + 1, i) = M(i, n + 1);
    }
#undef M
    rhs[n] = 1.0;
    rhs[N + n + 1] = 1.0;
    double *copy = (double *)malloc((size_t)N * N * sizeof(double)), *x = (double *)malloc((size_t)N * 2 * sizeof(double));
    memcpy(copy, mat, (size_t)N * N * sizeof(double));
    legs(mat, N, rhs, x);
    legs(copy, N, rhs + N, x + N);
    for (int i = 0; i < n; ++i)
      for (int d = 0; d < 3; ++d)
        coeffs[((size_t)c * ME + i) * 3 + d] = xh[d] * x[i] + yh[d] * x[N + i];
    free(loc);
    free(nrm);
    free(ps);
    free(pu);
    free(mat);
    free(rhs);
    free(copy);
    free(x);
  }
  free(en);
  return 0;
}

/* mpas_reconstruct_1d: out[5][nCells] = uX, uY, uZ, zonal, meridional */
int om_reconstruct(const tm_mesh_desc *m, const double *coeffs, const double *u, double *out) {
  const int nC = m->nCells, ME = m->maxEdges;
  for (int c = 0; c < nC; ++c) {
    double x = 0.0, y = 0.0, z = 0.0;
    for (int i = 0; i < m->nEdgesOnCell[c]; ++i) {
      const double ue = u[m->edgesOnCell[(size_t)c * ME + i] - 1];
      const double *k = coeffs + ((size_t)c * ME + i) * 3;
      x = x + k[0] * ue;
      y = y + k[1] * ue;
      z = z + k[2] * ue;
    }
    const double clat = cos(m->latCell[c]), slat = sin(m->latCell[c]);
    const double clon = cos(m->lonCell[c]), slon = sin(m->lonCell[c]);
    out[c] = x;
    out[nC + c] = y;
    out[2 * (size_t)nC + c] = z;
    out[3 * (size_t)nC + c] = -x * slon + y * clon;
    out[4 * (size_t)nC + c] = -(x * clon + y * slon) * slat + z * clat;
  }
  return 0;
}
