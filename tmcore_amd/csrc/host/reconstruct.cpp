// Coefficients of the MPAS vector reconstruction (cell-centre vectors from
// edge-normal components), computed at start-up:
//   mpas_initialize_vectors (reference src/core/vector_reconstruction/
//     mpas_vector_operations.F90:493-557): edge normals from cell 1 to cell 2,
//     and a tangent plane per cell whose first axis follows the normal of the
//     cell's first edge;
//   mpas_init_reconstruct (mpas_vector_reconstruction.F90:52-125) with
//   mpas_rbf_interp_func_3D_plane_vec_const_dir_comp_coeffs
//     (mpas_rbf_interpolation.F90:1076-1142): per cell, inverse-multiquadric
//     RBFs of length scale alpha (the mean centre-to-edge distance) on the
//     edge points projected into the plane, plus a constant vector; two
//     (n+2)-square systems solved by scaled partial-pivoting Gaussian
//     elimination (mpas_legs/elgs, :1667-1837).
// The geometry is radius-scaled first, as mesh_init does before these run.
// Cells are independent: the loop is spread over OpenMP threads, and each
// cell's arithmetic is the same whatever the thread count.
#include <cmath>
#include <cstring>
#include <vector>

#include "tmcore_host.h"

namespace {

void unit3(double *v) {
  const double mag = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  v[0] /= mag;
  v[1] /= mag;
  v[2] /= mag;
}

double dot3(const double *a, const double *b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

// Gaussian elimination with scaled partial pivoting (column-major a, size n);
// the factorisation is done in place and b is overwritten, as in elgs/legs.
void solve(std::vector<double> &a, int n, double *b, double *x) {
  std::vector<int> p(n);
  std::vector<double> s(n);
  auto A = [&](int i, int j) -> double & { return a[(size_t)i + (size_t)j * n]; };
  for (int i = 0; i < n; ++i) {
    p[i] = i;
    double m = 0.0;
    for (int j = 0; j < n; ++j) m = std::fmax(m, std::fabs(A(i, j)));
    s[i] = m;
  }
  for (int j = 0; j + 1 < n; ++j) {
    double best = 0.0;
    int k = j;
    for (int i = j; i < n; ++i) {
      const double r = std::fabs(A(p[i], j)) / s[p[i]];
      if (r > best) {
        best = r;
        k = i;
      }
    }
    std::swap(p[j], p[k]);
    for (int i = j + 1; i < n; ++i) {
      const double f = A(p[i], j) / A(p[j], j);
      A(p[i], j) = f;
      for (int q = j + 1; q < n; ++q) A(p[i], q) = A(p[i], q) - f * A(p[j], q);
    }
  }
  for (int i = 0; i + 1 < n; ++i)
    for (int j = i + 1; j < n; ++j) b[p[j]] = b[p[j]] - A(p[j], i) * b[p[i]];
  x[n - 1] = b[p[n - 1]] / A(p[n - 1], n - 1);
  for (int i = n - 2; i >= 0; --i) {
    double v = b[p[i]];
    for (int j = i + 1; j < n; ++j) v = v - A(p[i], j) * x[j];
    x[i] = v / A(p[i], i);
  }
}

}  // namespace

extern "C" int tm_reconstruct_coeffs(const tm_mesh_desc *m, double radius, double *coeffs) {
  if (!m || !coeffs || !(radius > 0.0) || m->nCells <= 0 || m->maxEdges < 3) return TM_EINVAL;
  const int nC = m->nCells, nE = m->nEdges, ME = m->maxEdges;
  for (int c = 0; c < nC; ++c)
    if (m->nEdgesOnCell[c] < 3 || m->nEdgesOnCell[c] > ME) return TM_ERANGE;
  std::vector<double> nrm((size_t)3 * nE);
  for (int e = 0; e < nE; ++e) {
    const int c1 = m->cellsOnEdge[2 * (size_t)e] - 1, c2 = m->cellsOnEdge[2 * (size_t)e + 1] - 1;
    if (c1 < 0 || c1 >= nC || c2 < 0 || c2 >= nC) return TM_ERANGE;
    double *v = &nrm[3 * (size_t)e];
    v[0] = m->xCell[c2] * radius - m->xCell[c1] * radius;
    v[1] = m->yCell[c2] * radius - m->yCell[c1] * radius;
    v[2] = m->zCell[c2] * radius - m->zCell[c1] * radius;
    unit3(v);
  }
  std::memset(coeffs, 0, (size_t)nC * ME * 3 * sizeof(double));
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(| : bad)
  for (int c = 0; c < nC; ++c) {
    const int n = m->nEdgesOnCell[c], N = n + 2;
    const double cc[3] = {m->xCell[c] * radius, m->yCell[c] * radius, m->zCell[c] * radius};
    double up[3] = {cc[0], cc[1], cc[2]};
    unit3(up);
    const int e0 = m->edgesOnCell[(size_t)c * ME] - 1;
    if (e0 < 0 || e0 >= nE) {
      bad = 1;
      continue;
    }
    const double *n0 = &nrm[3 * (size_t)e0];
    const double nd = dot3(n0, up);
    double ax[2][3];
    for (int d = 0; d < 3; ++d) ax[0][d] = n0[d] - nd * up[d];
    unit3(ax[0]);
    ax[1][0] = up[1] * ax[0][2] - up[2] * ax[0][1];
    ax[1][1] = up[2] * ax[0][0] - up[0] * ax[0][2];
    ax[1][2] = up[0] * ax[0][1] - up[1] * ax[0][0];
    unit3(ax[1]);
    // edge points and normals in the plane; alpha from the 3-D distances
    std::vector<double> ps(2 * (size_t)n), pu(2 * (size_t)n);
    double alpha = 0.0;
    for (int i = 0; i < n; ++i) {
      const int e = m->edgesOnCell[(size_t)c * ME + i] - 1;
      if (e < 0 || e >= nE) {
        bad = 1;
        break;
      }
      const double p[3] = {m->xEdge[e] * radius, m->yEdge[e] * radius, m->zEdge[e] * radius};
      const double dd[3] = {cc[0] - p[0], cc[1] - p[1], cc[2] - p[2]};
      alpha = alpha + std::sqrt((dd[0] * dd[0] + dd[1] * dd[1]) + dd[2] * dd[2]);
      for (int k = 0; k < 2; ++k) {
        ps[2 * i + k] = dot3(p, ax[k]);
        pu[2 * i + k] = dot3(&nrm[3 * (size_t)e], ax[k]);
      }
    }
    if (bad) continue;
    alpha = alpha / n;
    const double pd[2] = {dot3(cc, ax[0]), dot3(cc, ax[1])};
    const double a2 = alpha * alpha;
    std::vector<double> mat((size_t)N * N, 0.0), rhs(2 * (size_t)N, 0.0), x(2 * (size_t)N);
    auto M = [&](int i, int j) -> double & { return mat[(size_t)i + (size_t)j * N]; };
    for (int j = 0; j < n; ++j)
      for (int i = j; i < n; ++i) {
        const double d0 = ps[2 * i] - ps[2 * j], d1 = ps[2 * i + 1] - ps[2 * j + 1];
        const double phi = 1.0 / std::sqrt(1.0 + (d0 * d0 + d1 * d1) / a2);
        M(i, j) = phi * (pu[2 * i] * pu[2 * j] + pu[2 * i + 1] * pu[2 * j + 1]);
        M(j, i) = M(i, j);
      }
    for (int j = 0; j < n; ++j) {
      const double d0 = pd[0] - ps[2 * j], d1 = pd[1] - ps[2 * j + 1];
      const double phi = 1.0 / std::sqrt(1.0 + (d0 * d0 + d1 * d1) / a2);
      rhs[j] = phi * pu[2 * j];
      rhs[(size_t)N + j] = phi * pu[2 * j + 1];
    }
    for (int i = 0; i < n; ++i) {
      M(i, n) = M(n, i) = pu[2 * i];
      M(i, n + 1) = M(n + 1, i) = pu[2 * i + 1];
    }
    rhs[n] = 1.0;
    rhs[(size_t)N + n + 1] = 1.0;
    std::vector<double> mat2 = mat;
    solve(mat, N, rhs.data(), x.data());
    solve(mat2, N, rhs.data() + N, x.data() + N);
    for (int i = 0; i < n; ++i)
      for (int d = 0; d < 3; ++d) coeffs[((size_t)c * ME + i) * 3 + d] = ax[0][d] * x[i] + ax[1][d] * x[(size_t)N + i];
  }
  return bad ? TM_ERANGE : TM_OK;
}
