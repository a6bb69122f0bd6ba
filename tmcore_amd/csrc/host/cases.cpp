// Initial states for the shallow-water runs.
//
// tm_case_seeded is the seeded recipe the benchmarks use (no reference
// counterpart; the stream-function wind and the balanced height follow
// reference src/test_cases/swm/mountain_zonal_flow_test_mod.F90:43-59).
// tm_case_steady_geostrophic restates Williamson case 2
// (src/test_cases/swm/steady_geostrophic_flow_test_mod.F90:20-54, alpha = 0).
//
// Random numbers: SplitMix64 (Steele, Lea & Flood 2014), the top 53 bits of
// each output scaled by 2^-53 give a double in [0,1).  Draw order for one
// seed: 8 bumps x (z, lon, sign), then the mountain centre (z, lon).

#include "tmcore_host.h"

#include <cmath>
#include <cstdint>
#include <vector>

namespace {

constexpr double kG = 9.80616;       // const_mod.F90:12
constexpr double kOmega = 7.292e-5;  // const_mod.F90:13

struct SplitMix64 {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};

// Great-circle angle between (lat1,lon1) and a unit vector.
inline double angle_to(double lat, double lon, const double *c) {
  double x = std::cos(lat) * std::cos(lon), y = std::cos(lat) * std::sin(lon), z = std::sin(lat);
  double cx = y * c[2] - z * c[1], cy = z * c[0] - x * c[2], cz = x * c[1] - y * c[0];
  return std::atan2(std::sqrt(cx * cx + cy * cy + cz * cz), x * c[0] + y * c[1] + z * c[2]);
}

void random_point(SplitMix64 &r, double *c) {
  const double pi = std::acos(-1.0);
  double z = 2.0 * r.uniform() - 1.0;
  double lon = 2.0 * pi * r.uniform();
  double s = std::sqrt(std::fmax(0.0, 1.0 - z * z));
  c[0] = s * std::cos(lon);
  c[1] = s * std::sin(lon);
  c[2] = z;
}

}  // namespace

extern "C" int tm_case_seeded(const tm_mesh_desc *m, double radius, uint64_t seed, int ntracers,
                              const uint64_t *tracer_seeds, double *gd, double *u, double *ghs,
                              double *tracers) {
  if (!m || !gd || !u || !ghs || ntracers < 0 || (ntracers > 0 && (!tracers || !tracer_seeds)))
    return TM_EINVAL;
  const double u0 = 20.0, h_mean = 5000.0;
  const double bump_amp = 50.0, bump_len = 500.0e3;
  const double mount_h = 2000.0, mount_half = 1500.0e3;
  const double coef = radius * kOmega * u0 + 0.5 * u0 * u0;
  // area mean of sin^2(lat) is 1/3, so this h0 gives a 5000 m mean surface
  const double h0 = h_mean + coef / (3.0 * kG);

  SplitMix64 r{seed};
  double bc[8][3], ba[8];
  for (int b = 0; b < 8; ++b) {
    random_point(r, bc[b]);
    ba[b] = r.uniform() < 0.5 ? -bump_amp : bump_amp;
  }
  double mc[3];
  random_point(r, mc);

  const int32_t nC = m->nCells, nE = m->nEdges, nV = m->nVertices;
  std::vector<double> psi(nV);
#pragma omp parallel for schedule(static)
  for (int32_t v = 0; v < nV; ++v) psi[v] = -radius * u0 * std::sin(m->latVertex[v]);
#pragma omp parallel for schedule(static)
  for (int32_t e = 0; e < nE; ++e) {
    int32_t v1 = m->verticesOnEdge[2 * (size_t)e] - 1, v2 = m->verticesOnEdge[2 * (size_t)e + 1] - 1;
    u[e] = -(psi[v2] - psi[v1]) / (m->dvEdge[e] * radius);
  }
#pragma omp parallel for schedule(static)
  for (int32_t c = 0; c < nC; ++c) {
    double la = m->latCell[c], lo = m->lonCell[c];
    double dm = radius * angle_to(la, lo, mc) / mount_half;
    ghs[c] = kG * mount_h * std::exp(-std::log(2.0) * dm * dm);
    double hs = h0;
    for (int b = 0; b < 8; ++b) {
      double d = radius * angle_to(la, lo, bc[b]) / bump_len;
      hs += ba[b] * std::exp(-d * d);
    }
    double s = std::sin(la);
    gd[c] = kG * hs - coef * s * s - ghs[c];
  }
  if (ntracers > 0) {
    std::vector<double> a(nC), b(nC);
    for (int t = 0; t < ntracers; ++t) {
      SplitMix64 rt{tracer_seeds[t]};
      for (int32_t c = 0; c < nC; ++c) a[c] = rt.uniform();
      for (int pass = 0; pass < 4; ++pass) {
#pragma omp parallel for schedule(static)
        for (int32_t c = 0; c < nC; ++c) {
          int n = m->nEdgesOnCell[c];
          double s = 0.0;
          for (int i = 0; i < n; ++i) s += a[m->cellsOnCell[(size_t)c * m->maxEdges + i] - 1];
          b[c] = 0.5 * a[c] + 0.5 * (s / n);
        }
        a.swap(b);
      }
      for (int32_t c = 0; c < nC; ++c) tracers[(size_t)t * nC + c] = a[c];
    }
  }
  return TM_OK;
}

extern "C" int tm_case_steady_geostrophic(const tm_mesh_desc *m, double radius, double *gd, double *u,
                                          double *ghs) {
  if (!m || !gd || !u || !ghs) return TM_EINVAL;
  const double pi = std::acos(-1.0);
  const double u0 = 2.0 * pi * radius / (12.0 * 86400.0);
  const double gd0 = 2.94e4;
  for (int32_t e = 0; e < m->nEdges; ++e) {
    int32_t v1 = m->verticesOnEdge[2 * (size_t)e] - 1, v2 = m->verticesOnEdge[2 * (size_t)e + 1] - 1;
    double p1 = -radius * u0 * std::sin(m->latVertex[v1]);
    double p2 = -radius * u0 * std::sin(m->latVertex[v2]);
    u[e] = -(p2 - p1) / (m->dvEdge[e] * radius);
  }
  for (int32_t c = 0; c < m->nCells; ++c) {
    double s = std::sin(m->latCell[c]);
    gd[c] = gd0 - (radius * kOmega * u0 + 0.5 * u0 * u0) * (s * s);
    ghs[c] = 0.0;
  }
  return TM_OK;
}
