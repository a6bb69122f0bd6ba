// Device-side data of the TRiSK stage: mesh constants in slot-major padded
// SoA (renumbered along a Hilbert curve) and the per-precision state buffers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tmk {

constexpr int kBlock = 256;     // threads per workgroup (4 wave64)
constexpr int kXcd = 8;         // MI355X: 8 XCDs, workgroups dealt round-robin
constexpr int kMaxEoC = 6;      // maxEdges
constexpr int kMaxEoE = 10;     // edgesOnEdge slots used by hex/pent Voronoi meshes
constexpr int kFit2 = 12;       // vertex second-order fit stencil (excluding centre)
constexpr int kFit2C = 6;       // cell second-order fit stencil (excluding centre)
constexpr int kFit4C = 18;      // cell fourth-order fit stencil (excluding centre)
constexpr int kFit4V = 21;      // vertex fourth-order fit stencil (excluding centre)
// Streamed slot tables of the edges (edgesOnEdge, and the FP64
// weightsOnEdge and deriv2OnVertex) are stored in slot pairs: element
// (slot, i) of an n-element table sits at pair_at(slot, i, n), so one 8- or
// 16-byte load per thread reads two slots and a wavefront's load of a pair is
// still one contiguous block.  (The FP32 copies of the weights stay
// slot-major: 8-byte float pairs measured slower than two 4-byte loads.)
__host__ __device__ inline size_t pair_at(int slot, int i, int n) {
  return ((size_t)(slot >> 1) * n + i) * 2 + (slot & 1);
}
static_assert(kMaxEoE % 2 == 0 && kFit2 % 2 == 0, "slot pairs");

// Row length of a fit-index table (vertex- or cell-major): the stencil's
// slots rounded up to whole 16-byte int4 loads; extra slots are never summed.
constexpr int fit_row(int ns) { return (ns + 3) & ~3; }
constexpr int kNumSums = 13;    // global sums per step (see SumId)

enum SumId {
  kR12e = 0, kR12c, kR23e, kR23c, kR34e, kR34c, kR44e, kR44c,
  kMass, kEnergyE, kEnergyC, kEnstrophy, kAbsVor
};

// Mesh constants, new numbering, 0-based.  Slot-major: array[slot * n + i].
// Signed connectivity packs nSignEdge / tSignEdge (mesh_mod.F90:182-203) into
// the index: e for +1, ~e for -1.
constexpr int kTileE = 256;         // edges per tendency-kernel tile (= the kernels' workgroup size)
constexpr int kTendHaloMax = 512;   // halo rows one tile may stage

struct DMesh {
  int nC, nE, nV;
  int nbC, nbE, nbV;  // workgroups per element kind, multiples of kXcd
  const int8_t *nEoC;
  const int *eocS;        // [6][nC] signed edgesOnCell
  const double *areaCell; // [nC]
  const int *coe;         // [2][nE] cellsOnEdge
  const int *voe;         // [2][nE] verticesOnEdge
  const double *dv, *dc, *aE;  // [nE] dvEdge, dcEdge, areaEdge = dv*dc (radius-scaled)
  const int8_t *nEoE;
  const int8_t *kapE;     // [nE] +1 if the normal cellsOnEdge(1)->(2) points east, else -1
  const int *eoe;         // [5][nE][2] edgesOnEdge (slot pairs, pair_at)
  const double *woe;      // [5][nE][2] weightsOnEdge (slot pairs)
  const double *d2v;      // [2][6][nE][2] deriv2OnVertex slots 1..12 per side (slot pairs)
  const float *d2v32;     // [24][nE] the same in FP32, slot-major (precision-32 contexts), else null
  // cell fits of add_upwind_correction_on_cell (adv_order 3..6), else null:
  // unused slots point at the cell itself with weight 0
  const int *fit2C;       // [nC][fit_row(6)] fit2Cells slots 1..6
  const double *d2c;      // [2][6][nE] deriv2OnCell slots 1..6
  const int *fit4C;       // [nC][fit_row(18)] fit4Cells slots 1..18 (adv_order 5, 6)
  const double *d4c;      // [2][18][nE] deriv4OnCell slots 1..18
  const int *eovS;        // [3][nV] signed edgesOnVertex
  const int *cov;         // [3][nV] cellsOnVertex
  const double *kite;     // [3][nV] kiteAreasOnVertex
  const double *areaTri;  // [nV]
  const double *fV;       // [nV] 2*Omega*sin(latVertex)
  const int8_t *nFit2V;   // [nV] nFit2Vertices
  const int *fit2V;       // [nV][12] fit2Vertices slots 1..12: vertex-major, since it is
                          // reached through verticesOnEdge (a gather), not streamed
  const int *fit4V;       // [nV][fit_row(21)] fit4Vertices slots 1..21 (interp_pv_order 5, 6), else null
  const double *d4v;      // [2][21][nE] deriv4OnVertex slots 1..21
  const int *eocR;        // [nC][8] cell-major edge rows: signed edgesOnCell slots 0-5, nEdgesOnCell, 0
                          // (null when the mesh breaks the MPAS orderings they stand for)
  // LDS tiles of the tendency kernel (kXTendLds; null when a tile's halo
  // exceeds kTendHaloMax): tile t is edges [kTileE t, kTileE (t+1)) of the
  // Hilbert order, staged with its halo (the other edges its edgesOnEdge
  // name, ascending) at local rows kTileE.. ; eoeL holds every edgesOnEdge
  // slot as a local row, two 16-bit slots per word
  const uint32_t *eoeL;   // [5][nE]
  const int *haloE;       // halo edges of tile t at [haloOff[t], haloOff[t + 1])
  const int *haloOff;     // [nbE + 1]
  const int *vocS;        // [6][nC] verticesOnCell (for the v2c gather)
  const int *cocN;        // [6][nC] cellsOnCell, unused slots = the cell itself
  const double *kiteC;    // [6][nC] kite area of (cell, verticesOnCell(slot))
  // FP32 copies of the edge geometry and weights that FP32 arithmetic reads
  // (precision-32 contexts, else null): each element is the round-to-nearest
  // float of the FP64 one, the value a cast at the point of use gives
  const float *woe32;     // [10][nE] slot-major
  const float *dv32, *dc32, *aE32;  // [nE]
  const float *d2c32, *d4c32, *d4v32;  // FP32 copies of d2c, d4c, d4v where those exist
};

// Workgroup id -> logical tile so that each XCD (round-robin by workgroup id)
// gets one contiguous stretch of the Hilbert order, keeping a tile's stencil
// neighbours in the same L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  const int per = nb / kXcd;
  return (b % kXcd) * per + b / kXcd;
}
// The same tiles swept from the other end (rev = 1): consecutive kernels
// alternate direction, so each starts on the data the previous one touched
// last, which is still in the memory-side (Infinity) cache.
__device__ __forceinline__ int xcd_tile(int b, int nb, int rev) {
  const int t = xcd_tile(b, nb);
  return rev ? nb - 1 - t : t;
}

}  // namespace tmk
