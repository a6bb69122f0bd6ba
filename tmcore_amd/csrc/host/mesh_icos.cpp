// Icosahedral Voronoi mesh generator with MPAS conventions and TRiSK weights.
//
// The reference never builds a mesh: mesh_init reads an MPAS grid file
// (reference src/core/mesh_mod.F90:79-167).  This file produces the same
// arrays for a recursively bisected icosahedron so the hot path has input at
// levels 0..11 without NetCDF.
//
// Conventions (the ones the reference's operators rely on):
//   * cellsOnEdge(1)->cellsOnEdge(2) is the positive normal n of an edge;
//   * verticesOnEdge(1)->verticesOnEdge(2) is the positive tangent t = k x n,
//     which makes curl_operator (operators_mod.F90:76-91) the counter-
//     clockwise circulation and the stream-function wind of
//     mountain_zonal_flow_test_mod.F90:50-52 eastward for psi = -a u0 sin(lat);
//   * edgesOnCell / cellsOnCell / verticesOnCell run counter-clockwise seen
//     from outside, cellsOnCell(i) lies across edgesOnCell(i), and
//     verticesOnCell(i) is the corner between edgesOnCell(i) and
//     edgesOnCell(i+1);
//   * cellsOnVertex / edgesOnVertex run counter-clockwise, edgesOnVertex(i)
//     joins cellsOnVertex(i) and cellsOnVertex(i+1), kiteAreasOnVertex(i)
//     belongs to cellsOnVertex(i);
//   * edgesOnEdge lists the other edges of cellsOnEdge(1) counter-clockwise
//     starting after the edge, then those of cellsOnEdge(2).
//
// TRiSK weights (Thuburn, Ringler, Skamarock & Klemp 2009, J. Comput. Phys.
// 228; Ringler et al. 2010, J. Comput. Phys. 229, eq. 33): for edge e and
// cell i on e, going counter-clockwise round i from e to edge j and summing
// the kite-area fractions R of the corners passed,
//     w(e,j) = n(e,i) * n(j,i) * (1/2 - R) * dvEdge(j) / dcEdge(e),
// with n(x,i) = +1 when i = cellsOnEdge(1,x), -1 otherwise.  The tangential
// wind is then v_e = sum_j w(e,j) u_j (operators_mod.F90:223-237), exact for
// a uniform flow on a regular hexagon, and dcEdge*dvEdge*w is antisymmetric,
// which makes the PV flux of operators_mod.F90:239-260 energy-neutral.

#include "tmcore_host.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <vector>

namespace {

struct V3 {
  double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline V3 unit(V3 a) { return (1.0 / std::sqrt(dot(a, a))) * a; }

// Great-circle angle between unit vectors, accurate for tiny angles.
inline double arc(V3 a, V3 b) {
  V3 c = cross(a, b);
  return std::atan2(std::sqrt(dot(c, c)), dot(a, b));
}

// Signed area of the spherical triangle abc on the unit sphere
// (Van Oosterom & Strackee 1983): tan(E/2) = a.(b x c) / (1 + a.b + b.c + c.a).
inline double tri_area(V3 a, V3 b, V3 c) {
  double num = dot(a, cross(b, c));
  double den = 1.0 + dot(a, b) + dot(b, c) + dot(c, a);
  return 2.0 * std::atan2(num, den);
}

struct Tri {
  std::vector<V3> p;                     // generator points (cells)
  std::vector<std::array<int32_t, 2>> e; // edge endpoints (a, b)
  std::vector<std::array<int32_t, 3>> tv;// triangle corners, counter-clockwise
  std::vector<std::array<int32_t, 3>> te;// te[k] joins tv[k] and tv[k+1]
};

Tri icosahedron() {
  Tri m;
  const double pi = std::acos(-1.0);
  const double lat = std::atan(0.5);
  m.p.push_back({0.0, 0.0, 1.0});
  for (int i = 0; i < 5; ++i) {
    double lon = 2.0 * pi * i / 5.0;
    m.p.push_back({std::cos(lat) * std::cos(lon), std::cos(lat) * std::sin(lon), std::sin(lat)});
  }
  for (int i = 0; i < 5; ++i) {
    double lon = 2.0 * pi * i / 5.0 + pi / 5.0;
    m.p.push_back({std::cos(lat) * std::cos(lon), std::cos(lat) * std::sin(lon), -std::sin(lat)});
  }
  m.p.push_back({0.0, 0.0, -1.0});
  auto U = [](int i) { return 1 + (i % 5); };
  auto L = [](int i) { return 6 + (i % 5); };
  for (int i = 0; i < 5; ++i) {
    m.tv.push_back({0, U(i), U(i + 1)});
    m.tv.push_back({U(i), L(i), U(i + 1)});
    m.tv.push_back({U(i + 1), L(i), L(i + 1)});
    m.tv.push_back({11, L(i + 1), L(i)});
  }
  for (auto &t : m.tv) {
    V3 a = m.p[t[0]], b = m.p[t[1]], c = m.p[t[2]];
    if (dot(cross(b - a, c - a), a + b + c) < 0.0) std::swap(t[1], t[2]);
  }
  m.te.resize(m.tv.size());
  for (size_t t = 0; t < m.tv.size(); ++t) {
    for (int k = 0; k < 3; ++k) {
      int32_t a = m.tv[t][k], b = m.tv[t][(k + 1) % 3];
      int32_t found = -1;
      for (size_t q = 0; q < m.e.size(); ++q)
        if ((m.e[q][0] == a && m.e[q][1] == b) || (m.e[q][0] == b && m.e[q][1] == a)) found = (int32_t)q;
      if (found < 0) {
        found = (int32_t)m.e.size();
        m.e.push_back({a, b});
      }
      m.te[t][k] = found;
    }
  }
  return m;
}

// One bisection: every edge gets its normalised midpoint, every triangle
// splits in four.  Indices are assigned by construction, so no hashing.
Tri bisect(const Tri &m) {
  const int32_t nP = (int32_t)m.p.size(), nE = (int32_t)m.e.size(), nT = (int32_t)m.tv.size();
  Tri r;
  r.p.resize(nP + nE);
  for (int32_t i = 0; i < nP; ++i) r.p[i] = m.p[i];
  for (int32_t i = 0; i < nE; ++i) r.p[nP + i] = unit(m.p[m.e[i][0]] + m.p[m.e[i][1]]);
  r.e.resize(2 * (size_t)nE + 3 * (size_t)nT);
  for (int32_t i = 0; i < nE; ++i) {
    r.e[2 * i] = {m.e[i][0], nP + i};
    r.e[2 * i + 1] = {nP + i, m.e[i][1]};
  }
  r.tv.resize(4 * (size_t)nT);
  r.te.resize(4 * (size_t)nT);
  for (int32_t t = 0; t < nT; ++t) {
    const int32_t v0 = m.tv[t][0], v1 = m.tv[t][1], v2 = m.tv[t][2];
    const int32_t E0 = m.te[t][0], E1 = m.te[t][1], E2 = m.te[t][2];
    const int32_t m0 = nP + E0, m1 = nP + E1, m2 = nP + E2;
    auto half = [&](int32_t E, int32_t v) { return m.e[E][0] == v ? 2 * E : 2 * E + 1; };
    const int32_t I0 = 2 * nE + 3 * t, I1 = I0 + 1, I2 = I0 + 2;
    r.e[I0] = {m0, m1};
    r.e[I1] = {m1, m2};
    r.e[I2] = {m2, m0};
    r.tv[4 * t + 0] = {v0, m0, m2};
    r.te[4 * t + 0] = {half(E0, v0), I2, half(E2, v0)};
    r.tv[4 * t + 1] = {m0, v1, m1};
    r.te[4 * t + 1] = {half(E0, v1), half(E1, v1), I0};
    r.tv[4 * t + 2] = {m2, m1, v2};
    r.te[4 * t + 2] = {I1, half(E1, v2), half(E2, v2)};
    r.tv[4 * t + 3] = {m0, m1, m2};
    r.te[4 * t + 3] = {I0, I1, I2};
  }
  return r;
}

inline void latlon(V3 a, double *lat, double *lon) {
  const double pi = std::acos(-1.0);
  double z = std::max(-1.0, std::min(1.0, a.z));
  *lat = std::asin(z);
  double l = std::atan2(a.y, a.x);
  if (l < 0.0) l += 2.0 * pi;
  *lon = l;
}

}  // namespace

extern "C" int tm_mesh_icos_dims(int level, int32_t *nCells, int32_t *nEdges, int32_t *nVertices) {
  if (level < 0 || level > 11 || !nCells || !nEdges || !nVertices) return TM_EINVAL;
  int64_t f = 1;
  for (int i = 0; i < level; ++i) f *= 4;
  *nCells = (int32_t)(10 * f + 2);
  *nEdges = (int32_t)(30 * f);
  *nVertices = (int32_t)(20 * f);
  return TM_OK;
}

extern "C" int tm_mesh_icos_generate(int level, tm_mesh_desc *m) {
  int32_t nC, nE, nV;
  if (!m || tm_mesh_icos_dims(level, &nC, &nE, &nV) != TM_OK) return TM_EINVAL;
  Tri t = icosahedron();
  for (int l = 0; l < level; ++l) t = bisect(t);
  if ((int32_t)t.p.size() != nC || (int32_t)t.e.size() != nE || (int32_t)t.tv.size() != nV) return TM_EINVAL;
  m->nCells = nC;
  m->nEdges = nE;
  m->nVertices = nV;
  m->maxEdges = 6;
  m->maxEdges2 = 12;
  m->vertexDegree = 3;
  const int ME = 6, ME2 = 12;

  // Voronoi vertices = circumcentres of the (counter-clockwise) triangles.
  std::vector<V3> vx(nV);
#pragma omp parallel for schedule(static)
  for (int32_t v = 0; v < nV; ++v) {
    V3 a = t.p[t.tv[v][0]], b = t.p[t.tv[v][1]], c = t.p[t.tv[v][2]];
    vx[v] = unit(cross(b - a, c - a));
  }
  // Edge points: normalised midpoint of the two generators.
  std::vector<V3> ex(nE);
#pragma omp parallel for schedule(static)
  for (int32_t e = 0; e < nE; ++e) ex[e] = unit(t.p[t.e[e][0]] + t.p[t.e[e][1]]);

  // verticesOnEdge: the triangle that runs a->b is on the left of n, i.e. on
  // the +t side, so it is vertex 2.
  std::vector<std::array<int32_t, 2>> voe(nE, {-1, -1});
  for (int32_t v = 0; v < nV; ++v)
    for (int k = 0; k < 3; ++k) {
      int32_t E = t.te[v][k];
      bool same = t.e[E][0] == t.tv[v][k];
      voe[E][same ? 1 : 0] = v;
    }

  // Kites: corner k of triangle v owns (p_k, mid(p_k,p_k+1), vertex, mid(p_k-1,p_k)).
  std::vector<std::array<double, 3>> kite(nV);
#pragma omp parallel for schedule(static)
  for (int32_t v = 0; v < nV; ++v)
    for (int k = 0; k < 3; ++k) {
      V3 pk = t.p[t.tv[v][k]];
      V3 mnext = ex[t.te[v][k]], mprev = ex[t.te[v][(k + 2) % 3]];
      kite[v][k] = tri_area(pk, mnext, vx[v]) + tri_area(pk, vx[v], mprev);
    }

  // Cell -> incident edges (CSR), then counter-clockwise order per cell.
  std::vector<int32_t> deg(nC, 0), off(nC + 1, 0);
  for (int32_t e = 0; e < nE; ++e) {
    deg[t.e[e][0]]++;
    deg[t.e[e][1]]++;
  }
  for (int32_t c = 0; c < nC; ++c) {
    if (deg[c] < 5 || deg[c] > ME) return TM_EINVAL;
    off[c + 1] = off[c] + deg[c];
  }
  std::vector<int32_t> adjE(off[nC]), fill(off.begin(), off.end() - 1);
  for (int32_t e = 0; e < nE; ++e) {
    adjE[fill[t.e[e][0]]++] = e;
    adjE[fill[t.e[e][1]]++] = e;
  }
  std::vector<int32_t> eoc((size_t)nC * ME, 0), coc((size_t)nC * ME, 0), vocl((size_t)nC * ME, 0);
#pragma omp parallel for schedule(static)
  for (int32_t c = 0; c < nC; ++c) {
    V3 p = t.p[c];
    V3 ref = std::fabs(p.z) < 0.9 ? V3{0.0, 0.0, 1.0} : V3{1.0, 0.0, 0.0};
    V3 b1 = unit(cross(ref, p)), b2 = cross(p, b1);
    std::array<std::pair<double, int32_t>, 6> s;
    int n = deg[c];
    for (int i = 0; i < n; ++i) {
      int32_t E = adjE[off[c] + i];
      int32_t nb = t.e[E][0] == c ? t.e[E][1] : t.e[E][0];
      V3 d = t.p[nb] - p;
      s[i] = {std::atan2(dot(d, b2), dot(d, b1)), E};
    }
    std::sort(s.begin(), s.begin() + n);
    for (int i = 0; i < n; ++i) {
      int32_t E = s[i].second;
      bool first = t.e[E][0] == c;
      eoc[(size_t)c * ME + i] = E;
      coc[(size_t)c * ME + i] = first ? t.e[E][1] : t.e[E][0];
      // corner between edge i and edge i+1 = triangle on the left of c->nb
      vocl[(size_t)c * ME + i] = first ? voe[E][1] : voe[E][0];
    }
  }

  // Cell areas as the sum of their kites (so kites partition cells exactly
  // as they partition triangles).
  std::vector<double> areaC(nC, 0.0);
  for (int32_t v = 0; v < nV; ++v)
    for (int k = 0; k < 3; ++k) areaC[t.tv[v][k]] += kite[v][k];

  // Fill the caller's arrays (1-based connectivity).
#pragma omp parallel for schedule(static)
  for (int32_t c = 0; c < nC; ++c) {
    V3 p = t.p[c];
    double la, lo;
    latlon(p, &la, &lo);
    if (m->latCell) m->latCell[c] = la;
    if (m->lonCell) m->lonCell[c] = lo;
    if (m->xCell) m->xCell[c] = p.x;
    if (m->yCell) m->yCell[c] = p.y;
    if (m->zCell) m->zCell[c] = p.z;
    if (m->areaCell) m->areaCell[c] = areaC[c];
    if (m->nEdgesOnCell) m->nEdgesOnCell[c] = deg[c];
    for (int i = 0; i < ME; ++i) {
      bool ok = i < deg[c];
      size_t s = (size_t)c * ME + i;
      if (m->edgesOnCell) m->edgesOnCell[s] = ok ? eoc[s] + 1 : 0;
      if (m->cellsOnCell) m->cellsOnCell[s] = ok ? coc[s] + 1 : 0;
      if (m->verticesOnCell) m->verticesOnCell[s] = ok ? vocl[s] + 1 : 0;
    }
  }
#pragma omp parallel for schedule(static)
  for (int32_t v = 0; v < nV; ++v) {
    double la, lo;
    latlon(vx[v], &la, &lo);
    if (m->latVertex) m->latVertex[v] = la;
    if (m->lonVertex) m->lonVertex[v] = lo;
    if (m->xVertex) m->xVertex[v] = vx[v].x;
    if (m->yVertex) m->yVertex[v] = vx[v].y;
    if (m->zVertex) m->zVertex[v] = vx[v].z;
    double at = 0.0;
    for (int k = 0; k < 3; ++k) {
      at += kite[v][k];
      if (m->cellsOnVertex) m->cellsOnVertex[(size_t)v * 3 + k] = t.tv[v][k] + 1;
      if (m->edgesOnVertex) m->edgesOnVertex[(size_t)v * 3 + k] = t.te[v][k] + 1;
      if (m->kiteAreasOnVertex) m->kiteAreasOnVertex[(size_t)v * 3 + k] = kite[v][k];
    }
    if (m->areaTriangle) m->areaTriangle[v] = at;
  }
#pragma omp parallel for schedule(static)
  for (int32_t e = 0; e < nE; ++e) {
    const int32_t c1 = t.e[e][0], c2 = t.e[e][1], v1 = voe[e][0], v2 = voe[e][1];
    double la, lo;
    latlon(ex[e], &la, &lo);
    if (m->latEdge) m->latEdge[e] = la;
    if (m->lonEdge) m->lonEdge[e] = lo;
    if (m->xEdge) m->xEdge[e] = ex[e].x;
    if (m->yEdge) m->yEdge[e] = ex[e].y;
    if (m->zEdge) m->zEdge[e] = ex[e].z;
    if (m->cellsOnEdge) {
      m->cellsOnEdge[2 * (size_t)e] = c1 + 1;
      m->cellsOnEdge[2 * (size_t)e + 1] = c2 + 1;
    }
    if (m->verticesOnEdge) {
      m->verticesOnEdge[2 * (size_t)e] = v1 + 1;
      m->verticesOnEdge[2 * (size_t)e + 1] = v2 + 1;
    }
    if (m->dcEdge) m->dcEdge[e] = arc(t.p[c1], t.p[c2]);
    if (m->dvEdge) m->dvEdge[e] = arc(vx[v1], vx[v2]);
    if (m->dv1Edge) m->dv1Edge[e] = arc(vx[v1], ex[e]);
    if (m->dv2Edge) m->dv2Edge[e] = arc(vx[v2], ex[e]);
    if (m->angleEdge) {
      V3 q = ex[e];
      V3 d = t.p[c2] - t.p[c1];
      V3 n = unit(d - dot(d, q) * q);
      V3 east = unit(V3{-q.y, q.x, 0.0});
      V3 north = cross(q, east);
      m->angleEdge[e] = std::atan2(dot(n, north), dot(n, east));
    }
  }

  // TRiSK edgesOnEdge / weightsOnEdge.
  if (m->edgesOnEdge || m->weightsOnEdge || m->nEdgesOnEdge) {
#pragma omp parallel for schedule(static)
    for (int32_t e = 0; e < nE; ++e) {
      int slot = 0;
      const double dce = arc(t.p[t.e[e][0]], t.p[t.e[e][1]]);
      for (int side = 0; side < 2; ++side) {
        const int32_t c = t.e[e][side];
        const double nei = side == 0 ? 1.0 : -1.0;
        const int n = deg[c];
        int pos = -1;
        for (int i = 0; i < n; ++i)
          if (eoc[(size_t)c * ME + i] == e) pos = i;
        double R = 0.0;
        for (int k = 1; k < n; ++k) {
          int32_t corner = vocl[(size_t)c * ME + (pos + k - 1) % n];
          int kk = t.tv[corner][0] == c ? 0 : (t.tv[corner][1] == c ? 1 : 2);
          R += kite[corner][kk] / areaC[c];
          int32_t j = eoc[(size_t)c * ME + (pos + k) % n];
          double nji = t.e[j][0] == c ? 1.0 : -1.0;
          double dvj = arc(vx[voe[j][0]], vx[voe[j][1]]);
          size_t s = (size_t)e * ME2 + slot;
          if (m->edgesOnEdge) m->edgesOnEdge[s] = j + 1;
          if (m->weightsOnEdge) m->weightsOnEdge[s] = nei * nji * (0.5 - R) * dvj / dce;
          ++slot;
        }
      }
      for (int s = slot; s < ME2; ++s) {
        if (m->edgesOnEdge) m->edgesOnEdge[(size_t)e * ME2 + s] = 0;
        if (m->weightsOnEdge) m->weightsOnEdge[(size_t)e * ME2 + s] = 0.0;
      }
      if (m->nEdgesOnEdge) m->nEdgesOnEdge[e] = slot;
    }
  }
  return TM_OK;
}
