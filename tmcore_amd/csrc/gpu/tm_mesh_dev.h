// Host-side description of one mesh uploaded to the device: its constants in
// the renumbered slot-major layout (DMesh) and the permutations between the
// caller's (reference) numbering and the Hilbert numbering.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "tm_device.h"
#include "tmcore_gpu.h"

namespace tmk {

struct MeshDev {
  DMesh dm{};
  std::vector<int> pC, pE, pV;  // new -> old
  std::vector<int> iC, iE, iV;  // old -> new
  double totalArea = 0.0, sumAreaTri = 0.0;
  double sumEoE = 0.0, sumFitE = 0.0, sumFitV = 0.0, sumEoC = 0.0;  // stencil nnz
  // every edgesOnEdge row is derivable from its cells' edgesOnCell rows (the
  // MPAS order the cell-row kernels rely on); dm.eocR is set
  bool cell_rows = false;
  // every tendency-kernel tile's halo fits kTendHaloMax; dm.eoeL, haloE and
  // haloOff are set.  haloTotal: halo rows over all tiles
  bool tend_lds = false;
  double haloTotal = 0.0;
};

// Validates the reference-layout mesh (and the fit when o.interp_pv_order
// >= 3), renumbers along a Hilbert curve, derives nSignEdge/tSignEdge/
// areaEdge/fVertex (mesh_mod.F90:169-224) and uploads the constants on
// `stream`.  Device allocations are appended to `allocs`.  Returns TM_OK or
// a TM_E* code with `err` set.  Must be called with the device current.
int build_mesh_dev(MeshDev &md, const tm_mesh_desc *m, const tm_fit_desc *f, const tm_opts &o,
                   hipStream_t stream, std::vector<void *> &allocs, std::string &err);

// Host-side validation only (no device work): the checks build_mesh_dev runs
// before touching the device.
int validate_mesh(const tm_mesh_desc *m, const tm_fit_desc *f, const tm_opts &o, std::string &err);

}  // namespace tmk
