// History output in the layout of the reference's history_mod
// (reference src/utils/history_mod.F90:27-174): one NetCDF file (classic
// 64-bit-offset format, CDF-2) with the mesh variables, and per output time
// one record of u, zonalWind, merdionalWind, h, pv, vor, div, tm and te.
// Names, dimension orders, long_name and units attributes follow the
// reference so that its plotting tools read the file.  Written here without a
// NetCDF library: a fixed header, the fixed-size variables, then records
// appended in place with numrecs rewritten after each one.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "tmcore_host.h"

namespace {

enum { kChar = 2, kInt = 4, kDouble = 6 };

struct Att {
  std::string name;
  int type;
  std::string text;
  double value;
};

struct Var {
  std::string name, long_name, units;
  std::vector<int> dims;  // dimension ids, slowest first; dim 0 = Time (record)
  int type;
  uint64_t vsize = 0, begin = 0;
  bool record() const { return !dims.empty() && dims[0] == 0; }
};

struct Out {
  std::vector<unsigned char> b;
  void u32(uint32_t v) {
    for (int k = 3; k >= 0; --k) b.push_back((unsigned char)(v >> (8 * k)));
  }
  void u64(uint64_t v) {
    u32((uint32_t)(v >> 32));
    u32((uint32_t)v);
  }
  void f64(double d) {
    uint64_t v;
    std::memcpy(&v, &d, 8);
    u64(v);
  }
  void pad() {
    while (b.size() % 4) b.push_back(0);
  }
  void name(const std::string &s) {
    u32((uint32_t)s.size());
    b.insert(b.end(), s.begin(), s.end());
    pad();
  }
  void att(const Att &a) {
    name(a.name);
    u32((uint32_t)a.type);
    if (a.type == kChar) {
      u32((uint32_t)a.text.size());
      b.insert(b.end(), a.text.begin(), a.text.end());
      pad();
    } else if (a.type == kInt) {
      u32(1);
      u32((uint32_t)(int32_t)a.value);
    } else {
      u32(1);
      f64(a.value);
    }
  }
  void atts(const std::vector<Att> &v) {
    if (v.empty()) {
      u32(0);
      u32(0);
      return;
    }
    u32(0x0C);
    u32((uint32_t)v.size());
    for (const Att &a : v) att(a);
  }
};

}  // namespace

struct tm_history {
  FILE *f = nullptr;
  std::vector<std::string> dim_names;
  std::vector<int64_t> dim_len;
  std::vector<Var> vars;
  uint64_t recsize = 0, rec0 = 0;
  uint32_t numrecs = 0;
  int32_t nC = 0, nE = 0, nV = 0;
};

namespace {

// big-endian copy of n values of type T (double or int32) into out
template <typename T>
void put_be(std::vector<unsigned char> &out, const T *v, size_t n) {
  const size_t w = sizeof(T);
  const size_t o = out.size();
  out.resize(o + n * w + ((4 - (n * w) % 4) % 4), 0);
  for (size_t i = 0; i < n; ++i) {
    uint64_t x = 0;
    std::memcpy(&x, v + i, w);
    for (size_t k = 0; k < w; ++k) out[o + i * w + k] = (unsigned char)(x >> (8 * (w - 1 - k)));
  }
}

}  // namespace

extern "C" int tm_history_create(tm_history **out, const char *path, const tm_mesh_desc *m, double radius, double dt,
                                 int time_scheme, const char *case_name) {
  if (!out || !path || !m || !(radius > 0.0)) return TM_EINVAL;
  *out = nullptr;
  tm_history *h = new tm_history();
  h->nC = m->nCells;
  h->nE = m->nEdges;
  h->nV = m->nVertices;
  h->dim_names = {"Time", "nCells", "nEdges", "nVertices", "TWO", "vertexDegree", "maxEdges", "maxEdges2"};
  h->dim_len = {0, m->nCells, m->nEdges, m->nVertices, 2, m->vertexDegree, m->maxEdges, m->maxEdges2};
  enum { T = 0, C = 1, E = 2, V = 3, TWO = 4, VD = 5, ME = 6, ME2 = 7 };
  auto add = [&](const char *n, const char *ln, const char *u, std::vector<int> d, int t) {
    h->vars.push_back(Var{n, ln, u, d, t});
  };
  add("Time", "Time", "seconds since start", {T}, kDouble);
  add("lonCell", "Longitude on the cell", "radian", {C}, kDouble);
  add("latCell", "Latitude on the cell", "radian", {C}, kDouble);
  add("xCell", "Cartesian X on the cell", "m", {C}, kDouble);
  add("yCell", "Cartesian Y on the cell", "m", {C}, kDouble);
  add("zCell", "Cartesian Z on the cell", "m", {C}, kDouble);
  add("indexToCellID", "Global cell ID", "1", {C}, kInt);
  add("lonEdge", "Longitude on the edge", "radian", {E}, kDouble);
  add("latEdge", "Latitude on the edge", "radian", {E}, kDouble);
  add("xEdge", "Cartesian X on the edge", "m", {E}, kDouble);
  add("yEdge", "Cartesian Y on the edge", "m", {E}, kDouble);
  add("zEdge", "Cartesian Z on the edge", "m", {E}, kDouble);
  add("indexToEdgeID", "Global edge ID", "1", {E}, kInt);
  add("lonVertex", "Longitude on the vertex", "radian", {V}, kDouble);
  add("latVertex", "Latitude on the vertex", "radian", {V}, kDouble);
  add("xVertex", "Cartesian X on the vertex", "m", {V}, kDouble);
  add("yVertex", "Cartesian Y on the vertex", "m", {V}, kDouble);
  add("zVertex", "Cartesian Z on the vertex", "m", {V}, kDouble);
  add("indexToVertexID", "Global vertex ID", "1", {V}, kInt);
  add("areaCell", "Primary cell area", "m2", {C}, kDouble);
  add("areaTriangle", "Dual cell area", "m2", {V}, kDouble);
  add("areaEdge", "Defined edge area", "m2", {E}, kDouble);
  add("nEdgesOnCell", "Edge number on the cell", "1", {C}, kInt);
  add("nEdgesOnEdge", "Edge number to reconstruct tangent velocity", "1", {E}, kInt);
  add("cellsOnCell", "Cell indices that surround cell", "1", {C, ME}, kInt);
  add("cellsOnEdge", "Cell indices that saddle cell", "1", {E, TWO}, kInt);
  add("cellsOnVertex", "Cell indices that surround vertex", "1", {V, VD}, kInt);
  add("edgesOnCell", "Edge indices on the cell", "1", {C, ME}, kInt);
  add("edgesOnEdge", "Edge indices to reconstruct tangent velocity", "1", {E, ME2}, kInt);
  add("edgesOnVertex", "Edge indices on the vertex", "1", {V, VD}, kInt);
  add("verticesOnCell", "Vertex indices on the cell", "1", {C, ME}, kInt);
  add("verticesOnEdge", "Vertex indices on the edge", "1", {E, TWO}, kInt);
  add("u", "Normal wind on the edge", "m s-1", {T, E}, kDouble);
  add("zonalWind", "Zonal Wind", "m s-1", {T, C}, kDouble);
  add("merdionalWind", "Merdional Wind", "m s-1", {T, C}, kDouble);
  add("h", "Geopotential height on the cell", "m", {T, C}, kDouble);
  add("pv", "Potential vorticity on the vertex", "s-1", {T, V}, kDouble);
  add("vor", "relative vorticity on the vertex", "s-1", {T, V}, kDouble);
  add("div", "Divergence", "s-1", {T, C}, kDouble);
  add("tm", "total mass", "m2 s-2", {T}, kDouble);
  add("te", "total energy", "m4 s-4", {T}, kDouble);
  for (Var &v : h->vars) {
    uint64_t n = v.type == kDouble ? 8 : 4;
    for (size_t k = v.record() ? 1 : 0; k < v.dims.size(); ++k) n *= (uint64_t)h->dim_len[v.dims[k]];
    v.vsize = (n + 3) / 4 * 4;
  }
  const std::vector<Att> gatts = {{"source", kChar, "tmcore_amd", 0.0},
                                  {"dt", kDouble, "", dt},
                                  {"time_scheme", kInt, "", (double)time_scheme},
                                  {"author", kChar, "N/A", 0.0},
                                  {"on_a_sphere", kChar, "YES", 0.0},
                                  {"sphere_radius", kDouble, "", radius},
                                  {"title", kChar, case_name ? case_name : "", 0.0}};
  // the header is written twice: first to learn its size, then with offsets
  auto header = [&](Out &o) {
    o.b.clear();
    for (unsigned char ch : {'C', 'D', 'F', '\x02'}) o.b.push_back(ch);
    o.u32(h->numrecs);
    o.u32(0x0A);
    o.u32((uint32_t)h->dim_names.size());
    for (size_t i = 0; i < h->dim_names.size(); ++i) {
      o.name(h->dim_names[i]);
      o.u32((uint32_t)h->dim_len[i]);
    }
    o.atts(gatts);
    o.u32(0x0B);
    o.u32((uint32_t)h->vars.size());
    for (const Var &v : h->vars) {
      o.name(v.name);
      o.u32((uint32_t)v.dims.size());
      for (int d : v.dims) o.u32((uint32_t)d);
      o.atts({{"long_name", kChar, v.long_name, 0.0}, {"units", kChar, v.units, 0.0}});
      o.u32((uint32_t)v.type);
      o.u32((uint32_t)v.vsize);
      o.u64(v.begin);
    }
  };
  Out o;
  header(o);
  uint64_t off = o.b.size();
  for (Var &v : h->vars)
    if (!v.record()) v.begin = off, off += v.vsize;
  h->rec0 = off;
  for (Var &v : h->vars)
    if (v.record()) v.begin = off, off += v.vsize, h->recsize += v.vsize;
  header(o);
  h->f = std::fopen(path, "wb+");
  if (!h->f) {
    delete h;
    return TM_EIO;
  }
  // fixed-size variables: the mesh, radius-scaled as mesh_init leaves it
  const int32_t nC = m->nCells, nE = m->nEdges, nV = m->nVertices;
  auto scaled = [&](const double *a, int32_t n, double s) {
    std::vector<double> v(a, a + n);
    for (double &x : v) x *= s;
    return v;
  };
  auto ids = [](int32_t n) {
    std::vector<int32_t> v(n);
    for (int32_t i = 0; i < n; ++i) v[i] = i + 1;
    return v;
  };
  std::vector<double> areaEdge(nE);
  for (int32_t e = 0; e < nE; ++e) areaEdge[e] = (m->dvEdge[e] * radius) * (m->dcEdge[e] * radius);
  const double r2 = radius * radius;
  std::vector<unsigned char> &d = o.b;
  for (const Var &v : h->vars) {
    if (v.record()) continue;
    const std::string &n = v.name;
    if (n == "lonCell") put_be(d, m->lonCell, nC);
    else if (n == "latCell") put_be(d, m->latCell, nC);
    else if (n == "xCell") { auto x = scaled(m->xCell, nC, radius); put_be(d, x.data(), nC); }
    else if (n == "yCell") { auto x = scaled(m->yCell, nC, radius); put_be(d, x.data(), nC); }
    else if (n == "zCell") { auto x = scaled(m->zCell, nC, radius); put_be(d, x.data(), nC); }
    else if (n == "indexToCellID") { auto x = ids(nC); put_be(d, x.data(), nC); }
    else if (n == "lonEdge") put_be(d, m->lonEdge, nE);
    else if (n == "latEdge") put_be(d, m->latEdge, nE);
    else if (n == "xEdge") { auto x = scaled(m->xEdge, nE, radius); put_be(d, x.data(), nE); }
    else if (n == "yEdge") { auto x = scaled(m->yEdge, nE, radius); put_be(d, x.data(), nE); }
    else if (n == "zEdge") { auto x = scaled(m->zEdge, nE, radius); put_be(d, x.data(), nE); }
    else if (n == "indexToEdgeID") { auto x = ids(nE); put_be(d, x.data(), nE); }
    else if (n == "lonVertex") put_be(d, m->lonVertex, nV);
    else if (n == "latVertex") put_be(d, m->latVertex, nV);
    else if (n == "xVertex") { auto x = scaled(m->xVertex, nV, radius); put_be(d, x.data(), nV); }
    else if (n == "yVertex") { auto x = scaled(m->yVertex, nV, radius); put_be(d, x.data(), nV); }
    else if (n == "zVertex") { auto x = scaled(m->zVertex, nV, radius); put_be(d, x.data(), nV); }
    else if (n == "indexToVertexID") { auto x = ids(nV); put_be(d, x.data(), nV); }
    else if (n == "areaCell") { auto x = scaled(m->areaCell, nC, r2); put_be(d, x.data(), nC); }
    else if (n == "areaTriangle") { auto x = scaled(m->areaTriangle, nV, r2); put_be(d, x.data(), nV); }
    else if (n == "areaEdge") put_be(d, areaEdge.data(), nE);
    else if (n == "nEdgesOnCell") put_be(d, m->nEdgesOnCell, nC);
    else if (n == "nEdgesOnEdge") put_be(d, m->nEdgesOnEdge, nE);
    else if (n == "cellsOnCell") put_be(d, m->cellsOnCell, (size_t)nC * m->maxEdges);
    else if (n == "cellsOnEdge") put_be(d, m->cellsOnEdge, (size_t)nE * 2);
    else if (n == "cellsOnVertex") put_be(d, m->cellsOnVertex, (size_t)nV * m->vertexDegree);
    else if (n == "edgesOnCell") put_be(d, m->edgesOnCell, (size_t)nC * m->maxEdges);
    else if (n == "edgesOnEdge") put_be(d, m->edgesOnEdge, (size_t)nE * m->maxEdges2);
    else if (n == "edgesOnVertex") put_be(d, m->edgesOnVertex, (size_t)nV * m->vertexDegree);
    else if (n == "verticesOnCell") put_be(d, m->verticesOnCell, (size_t)nC * m->maxEdges);
    else if (n == "verticesOnEdge") put_be(d, m->verticesOnEdge, (size_t)nE * 2);
  }
  if (d.size() != h->rec0 || std::fwrite(d.data(), 1, d.size(), h->f) != d.size() || std::fflush(h->f) != 0) {
    std::fclose(h->f);
    delete h;
    return TM_EIO;
  }
  *out = h;
  return TM_OK;
}

extern "C" int tm_history_write(tm_history *h, double time_seconds, const double *u, const double *zonal,
                                const double *meridional, const double *h_cell, const double *pv, const double *vor,
                                const double *div, double total_mass, double total_energy) {
  if (!h || !u || !zonal || !meridional || !h_cell || !pv || !vor || !div) return TM_EINVAL;
  std::vector<unsigned char> r;
  r.reserve(h->recsize);
  for (const Var &v : h->vars) {
    if (!v.record()) continue;
    const std::string &n = v.name;
    if (n == "Time") put_be(r, &time_seconds, 1);
    else if (n == "u") put_be(r, u, h->nE);
    else if (n == "zonalWind") put_be(r, zonal, h->nC);
    else if (n == "merdionalWind") put_be(r, meridional, h->nC);
    else if (n == "h") put_be(r, h_cell, h->nC);
    else if (n == "pv") put_be(r, pv, h->nV);
    else if (n == "vor") put_be(r, vor, h->nV);
    else if (n == "div") put_be(r, div, h->nC);
    else if (n == "tm") put_be(r, &total_mass, 1);
    else if (n == "te") put_be(r, &total_energy, 1);
  }
  if (r.size() != h->recsize) return TM_EIO;
  const uint64_t at = h->rec0 + (uint64_t)h->numrecs * h->recsize;
  if (fseeko(h->f, (off_t)at, SEEK_SET) != 0 || std::fwrite(r.data(), 1, r.size(), h->f) != r.size()) return TM_EIO;
  h->numrecs += 1;
  unsigned char nr[4] = {(unsigned char)(h->numrecs >> 24), (unsigned char)(h->numrecs >> 16),
                         (unsigned char)(h->numrecs >> 8), (unsigned char)h->numrecs};
  if (fseeko(h->f, 4, SEEK_SET) != 0 || std::fwrite(nr, 1, 4, h->f) != 4 || std::fflush(h->f) != 0) return TM_EIO;
  return TM_OK;
}

extern "C" int tm_history_close(tm_history *h) {
  if (!h) return TM_EINVAL;
  const int rc = std::fclose(h->f) == 0 ? TM_OK : TM_EIO;
  delete h;
  return rc;
}
