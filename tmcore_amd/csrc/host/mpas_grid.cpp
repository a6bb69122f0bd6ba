// MPAS grid files (NetCDF classic: CDF-1, CDF-2 64-bit offset, CDF-5 64-bit
// data) read into a tm_mesh_desc: the reads of mesh_init
// (reference src/core/mesh_mod.F90:79-164) without a NetCDF library.
//
// The classic format is a big-endian header (dimensions, global attributes,
// variables with their shape, type and file offset) followed by the data of
// each fixed-size variable stored contiguously in C order.  A grid variable
// such as cellsOnEdge(nEdges, TWO) is therefore already laid out as the
// Fortran array cellsOnEdge(2, nEdges) that tm_mesh_desc carries.  Record
// (unlimited-dimension) variables are not needed and not read.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "tmcore_host.h"

namespace {

enum NcType { kByte = 1, kChar, kShort, kInt, kFloat, kDouble, kUByte, kUShort, kUInt, kInt64, kUInt64 };

int type_size(int t) {
  switch (t) {
    case kByte: case kChar: case kUByte: return 1;
    case kShort: case kUShort: return 2;
    case kInt: case kFloat: case kUInt: return 4;
    case kDouble: case kInt64: case kUInt64: return 8;
    default: return 0;
  }
}

struct NcVar {
  std::string name;
  std::vector<int64_t> dims;  // dimension ids
  int type = 0;
  uint64_t begin = 0;
};

struct NcFile {
  FILE *f = nullptr;
  int version = 0;  // 1, 2 or 5
  std::vector<std::string> dim_names;
  std::vector<int64_t> dim_len;  // 0 = the record dimension
  std::vector<NcVar> vars;
  double sphere_radius = 1.0;
  std::string err;

  ~NcFile() {
    if (f) std::fclose(f);
  }

  bool rd(void *p, size_t n) { return std::fread(p, 1, n, f) == n; }
  bool u32(uint32_t &v) {
    unsigned char b[4];
    if (!rd(b, 4)) return false;
    v = (uint32_t)b[0] << 24 | (uint32_t)b[1] << 16 | (uint32_t)b[2] << 8 | b[3];
    return true;
  }
  bool u64(uint64_t &v) {
    uint32_t hi, lo;
    if (!u32(hi) || !u32(lo)) return false;
    v = (uint64_t)hi << 32 | lo;
    return true;
  }
  // NON_NEG / nelems / dim_length: 4 bytes (CDF-1, CDF-2), 8 bytes (CDF-5)
  bool count(uint64_t &v) {
    if (version == 5) return u64(v);
    uint32_t x;
    if (!u32(x)) return false;
    v = x;
    return true;
  }
  bool name(std::string &s) {
    uint64_t n;
    if (!count(n) || n > 4096) return false;
    s.resize(n);
    if (n && !rd(&s[0], n)) return false;
    return skip_pad(n);
  }
  bool skip_pad(uint64_t n) {
    const uint64_t pad = (4 - n % 4) % 4;
    char buf[4];
    return pad == 0 || rd(buf, pad);
  }
  bool fail(const std::string &m) {
    err = m;
    return false;
  }

  // gatt_list / vatt_list; keeps the global sphere_radius
  bool attributes(bool global) {
    uint32_t tag;
    uint64_t n;
    if (!u32(tag) || !count(n)) return fail("truncated attribute list");
    if (tag == 0 && n == 0) return true;  // ABSENT
    if (tag != 0x0C) return fail("bad attribute list tag");
    for (uint64_t i = 0; i < n; ++i) {
      std::string an;
      uint32_t t;
      uint64_t ne;
      if (!name(an) || !u32(t) || !count(ne)) return fail("truncated attribute");
      const int ts = type_size((int)t);
      if (!ts || ne > (1ull << 31)) return fail("bad attribute type or size");
      std::vector<unsigned char> v(ne * ts);
      if (!v.empty() && !rd(v.data(), v.size())) return fail("truncated attribute values");
      if (!skip_pad(v.size())) return fail("truncated attribute padding");
      if (global && an == "sphere_radius" && ne >= 1 && (t == kDouble || t == kFloat)) {
        if (t == kDouble) {
          uint64_t x = 0;
          for (int k = 0; k < 8; ++k) x = x << 8 | v[k];
          std::memcpy(&sphere_radius, &x, 8);
        } else {
          uint32_t x = 0;
          for (int k = 0; k < 4; ++k) x = x << 8 | v[k];
          float fx;
          std::memcpy(&fx, &x, 4);
          sphere_radius = fx;
        }
      }
    }
    return true;
  }

  bool open(const char *path) {
    f = std::fopen(path, "rb");
    if (!f) return fail(std::string("cannot open ") + path);
    unsigned char magic[4];
    if (!rd(magic, 4) || magic[0] != 'C' || magic[1] != 'D' || magic[2] != 'F')
      return fail("not a NetCDF classic file (NetCDF-4/HDF5 files are not read)");
    version = magic[3];
    if (version != 1 && version != 2 && version != 5) return fail("unknown NetCDF classic version");
    uint64_t numrecs;
    if (!count(numrecs)) return fail("truncated header");
    // dim_list
    uint32_t tag;
    uint64_t n;
    if (!u32(tag) || !count(n)) return fail("truncated dimension list");
    if (!(tag == 0 && n == 0)) {
      if (tag != 0x0A) return fail("bad dimension list tag");
      for (uint64_t i = 0; i < n; ++i) {
        std::string dn;
        uint64_t len;
        if (!name(dn) || !count(len)) return fail("truncated dimension");
        dim_names.push_back(dn);
        dim_len.push_back((int64_t)len);
      }
    }
    if (!attributes(true)) return false;
    // var_list
    if (!u32(tag) || !count(n)) return fail("truncated variable list");
    if (tag == 0 && n == 0) return true;
    if (tag != 0x0B) return fail("bad variable list tag");
    for (uint64_t i = 0; i < n; ++i) {
      NcVar v;
      uint64_t nd;
      if (!name(v.name) || !count(nd) || nd > 64) return fail("truncated variable");
      for (uint64_t k = 0; k < nd; ++k) {
        uint64_t id;
        if (!count(id) || id >= dim_len.size()) return fail("bad dimension id in " + v.name);
        v.dims.push_back((int64_t)id);
      }
      if (!attributes(false)) return false;
      uint32_t t;
      uint64_t vsize;
      if (!u32(t) || !count(vsize)) return fail("truncated variable " + v.name);
      v.type = (int)t;
      if (version == 1) {
        uint32_t b;
        if (!u32(b)) return fail("truncated variable " + v.name);
        v.begin = b;
      } else if (!u64(v.begin)) {
        return fail("truncated variable " + v.name);
      }
      vars.push_back(v);
    }
    return true;
  }

  int64_t dim(const char *nm) const {
    for (size_t i = 0; i < dim_names.size(); ++i)
      if (dim_names[i] == nm) return dim_len[i];
    return -1;
  }

  const NcVar *var(const char *nm) const {
    for (const NcVar &v : vars)
      if (v.name == nm) return &v;
    return nullptr;
  }

  // Reads the fixed-size variable `nm` of shape (n0[, n1]) into out (double
  // or int32 by want); float data widens to double, int to int32 only.
  bool read(const char *nm, int64_t n0, int64_t n1, void *out, bool want_double) {
    if (!out) return true;
    const NcVar *v = var(nm);
    if (!v) return fail(std::string("variable ") + nm + " missing");
    const size_t nd = v->dims.size();
    if (nd != (n1 > 0 ? 2u : 1u)) return fail(std::string("variable ") + nm + " has the wrong rank");
    if (dim_len[v->dims[0]] != n0 || (n1 > 0 && dim_len[v->dims[1]] != n1))
      return fail(std::string("variable ") + nm + " has the wrong shape");
    const bool ok_type = want_double ? (v->type == kDouble || v->type == kFloat) : v->type == kInt;
    if (!ok_type) return fail(std::string("variable ") + nm + " has an unsupported type");
    const size_t count = (size_t)n0 * (size_t)(n1 > 0 ? n1 : 1), ts = (size_t)type_size(v->type);
    std::vector<unsigned char> buf(count * ts);
    if (fseeko(f, (off_t)v->begin, SEEK_SET) != 0 || !rd(buf.data(), buf.size()))
      return fail(std::string("cannot read the data of ") + nm);
    for (size_t i = 0; i < count; ++i) {
      const unsigned char *p = &buf[i * ts];
      if (ts == 8) {
        uint64_t x = 0;
        for (int k = 0; k < 8; ++k) x = x << 8 | p[k];
        std::memcpy((double *)out + i, &x, 8);
      } else {
        uint32_t x = (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3];
        if (want_double) {
          float fx;
          std::memcpy(&fx, &x, 4);
          ((double *)out)[i] = fx;
        } else {
          std::memcpy((int32_t *)out + i, &x, 4);
        }
      }
    }
    return true;
  }
};

thread_local std::string g_err;

}  // namespace

extern "C" const char *tm_mesh_mpas_error(void) { return g_err.c_str(); }

extern "C" int tm_mesh_mpas_dims(const char *path, int32_t dims[6], double *sphere_radius) {
  if (!path || !dims) return TM_EINVAL;
  NcFile nc;
  if (!nc.open(path)) {
    g_err = nc.err;
    return TM_EIO;
  }
  const char *names[6] = {"nCells", "nEdges", "nVertices", "maxEdges", "maxEdges2", "vertexDegree"};
  for (int i = 0; i < 6; ++i) {
    const int64_t d = nc.dim(names[i]);
    if (d <= 0 || d > INT32_MAX) {
      g_err = std::string("dimension ") + names[i] + " missing";
      return TM_EIO;
    }
    dims[i] = (int32_t)d;
  }
  if (sphere_radius) *sphere_radius = nc.sphere_radius;
  return TM_OK;
}

extern "C" int tm_mesh_mpas_read(const char *path, tm_mesh_desc *m) {
  if (!path || !m) return TM_EINVAL;
  int32_t d[6];
  int rc = tm_mesh_mpas_dims(path, d, nullptr);
  if (rc) return rc;
  if (m->nCells != d[0] || m->nEdges != d[1] || m->nVertices != d[2] || m->maxEdges != d[3] ||
      m->maxEdges2 != d[4] || m->vertexDegree != d[5]) {
    g_err = "tm_mesh_desc dimensions differ from the file's";
    return TM_EINVAL;
  }
  NcFile nc;
  if (!nc.open(path)) {
    g_err = nc.err;
    return TM_EIO;
  }
  const int64_t nC = d[0], nE = d[1], nV = d[2], ME = d[3], ME2 = d[4], VD = d[5];
  struct Item {
    const char *name;
    int64_t n0, n1;
    void *out;
    bool dbl;
  } items[] = {
      {"latCell", nC, 0, m->latCell, true},         {"lonCell", nC, 0, m->lonCell, true},
      {"xCell", nC, 0, m->xCell, true},             {"yCell", nC, 0, m->yCell, true},
      {"zCell", nC, 0, m->zCell, true},             {"latEdge", nE, 0, m->latEdge, true},
      {"lonEdge", nE, 0, m->lonEdge, true},         {"xEdge", nE, 0, m->xEdge, true},
      {"yEdge", nE, 0, m->yEdge, true},             {"zEdge", nE, 0, m->zEdge, true},
      {"latVertex", nV, 0, m->latVertex, true},     {"lonVertex", nV, 0, m->lonVertex, true},
      {"xVertex", nV, 0, m->xVertex, true},         {"yVertex", nV, 0, m->yVertex, true},
      {"zVertex", nV, 0, m->zVertex, true},         {"nEdgesOnCell", nC, 0, m->nEdgesOnCell, false},
      {"nEdgesOnEdge", nE, 0, m->nEdgesOnEdge, false},
      {"cellsOnCell", nC, ME, m->cellsOnCell, false},
      {"cellsOnEdge", nE, 2, m->cellsOnEdge, false},
      {"edgesOnCell", nC, ME, m->edgesOnCell, false},
      {"edgesOnEdge", nE, ME2, m->edgesOnEdge, false},
      {"verticesOnCell", nC, ME, m->verticesOnCell, false},
      {"verticesOnEdge", nE, 2, m->verticesOnEdge, false},
      {"edgesOnVertex", nV, VD, m->edgesOnVertex, false},
      {"cellsOnVertex", nV, VD, m->cellsOnVertex, false},
      {"weightsOnEdge", nE, ME2, m->weightsOnEdge, true},
      {"dvEdge", nE, 0, m->dvEdge, true},           {"dv1Edge", nE, 0, m->dv1Edge, true},
      {"dv2Edge", nE, 0, m->dv2Edge, true},         {"dcEdge", nE, 0, m->dcEdge, true},
      {"angleEdge", nE, 0, m->angleEdge, true},     {"areaCell", nC, 0, m->areaCell, true},
      {"areaTriangle", nV, 0, m->areaTriangle, true},
      {"kiteAreasOnVertex", nV, VD, m->kiteAreasOnVertex, true},
  };
  for (const Item &it : items)
    if (!nc.read(it.name, it.n0, it.n1, it.out, it.dbl)) {
      g_err = nc.err;
      return TM_EIO;
    }
  return TM_OK;
}
