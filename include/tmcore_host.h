/*
 * tmcore_host.h -- host-side C-ABI of libtmhost.so: mesh generation, the
 * init-time least-squares fit weights and the seeded initial state.
 *
 * Layout conventions (the ones the reference's Fortran arrays have, so a
 * Fortran caller passes its arrays through iso_c_binding unchanged):
 *   - connectivity values are 1-based int32;
 *   - 2-D arrays are column-major with the slot index fastest, i.e.
 *     edgesOnCell(maxEdges,nCells) is a C array [nCells][maxEdges];
 *   - geometry is on the unit sphere, exactly as an MPAS grid file holds it;
 *     the consumer scales by the radius as mesh_init does
 *     (reference src/core/mesh_mod.F90:205-224).
 * Every entry point returns 0 on success and a negative TM_E* code on error.
 */
#ifndef TMCORE_HOST_H
#define TMCORE_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TM_OK 0
#define TM_EINVAL (-1)   /* bad argument (null pointer, out-of-range size)   */
#define TM_ERANGE (-2)   /* a mesh index is outside 1..n                     */
#define TM_ESTENCIL (-3) /* a fit stencil does not fit in the given stride   */
#define TM_ESINGULAR (-4)/* a least-squares normal matrix is singular        */
#define TM_EDEVICE (-5)  /* HIP runtime error                                */
#define TM_ENOMEM (-6)   /* allocation failed                                */
#define TM_ENONFINITE (-7) /* NaN/Inf found in a state or invariant          */
#define TM_ESTATE (-8)   /* call out of order (e.g. step before upload)      */
#define TM_EIO (-9)      /* file unreadable, not NetCDF classic, or a grid
                            variable missing / of the wrong shape or type    */

/*
 * MPAS mesh arrays read by the hot path.  Replaces the NetCDF reads of
 * mesh_init (reference src/core/mesh_mod.F90:79-167); the derived arrays of
 * mesh_mod.F90:169-203 (nSignEdge, tSignEdge, verticesOnVertex, areaEdge,
 * fCell, fVertex) are computed by each consumer.  All pointers are owned by
 * the caller; a NULL member is skipped by tm_mesh_icos_generate.
 */
typedef struct tm_mesh_desc {
  int32_t nCells, nEdges, nVertices;
  int32_t maxEdges;      /* 6  */
  int32_t maxEdges2;     /* 12 */
  int32_t vertexDegree;  /* 3  */
  double *latCell, *lonCell, *xCell, *yCell, *zCell;             /* [nCells]   */
  double *latEdge, *lonEdge, *xEdge, *yEdge, *zEdge;             /* [nEdges]   */
  double *latVertex, *lonVertex, *xVertex, *yVertex, *zVertex;   /* [nVertices]*/
  int32_t *nEdgesOnCell;                                         /* [nCells]   */
  int32_t *nEdgesOnEdge;                                         /* [nEdges]   */
  int32_t *cellsOnCell, *edgesOnCell, *verticesOnCell;  /* [nCells][maxEdges] */
  int32_t *cellsOnEdge, *verticesOnEdge;                /* [nEdges][2]        */
  int32_t *edgesOnEdge;                                 /* [nEdges][maxEdges2]*/
  double  *weightsOnEdge;                               /* [nEdges][maxEdges2]*/
  int32_t *cellsOnVertex, *edgesOnVertex;               /* [nVertices][3]     */
  double  *kiteAreasOnVertex;                           /* [nVertices][3]     */
  double *dvEdge, *dv1Edge, *dv2Edge, *dcEdge, *angleEdge;       /* [nEdges]   */
  double *areaCell;                                              /* [nCells]   */
  double *areaTriangle;                                          /* [nVertices]*/
} tm_mesh_desc;

/*
 * Least-squares fit stencils and derivative weights of poly_fit_run
 * (reference src/utils/poly_fit_mod.F90:119-609), in the reference layout
 * fitXCells(0:maxFit-1, nCells) -> C [nCells][maxFit] (slot 0 is the centre),
 * derivXOnCell(0:maxFit-1, 2, nEdges) -> C [nEdges][2][maxFit].  maxFit is the
 * caller's stride (the reference uses 31); NULL families are skipped.
 */
typedef struct tm_fit_desc {
  int32_t maxFit;
  int32_t *nFit2Cells, *fit2Cells;        double *deriv2OnCell;
  int32_t *nFit4Cells, *fit4Cells;        double *deriv4OnCell;
  int32_t *nFit2Vertices, *fit2Vertices;  double *deriv2OnVertex;
  int32_t *nFit4Vertices, *fit4Vertices;  double *deriv4OnVertex;
} tm_fit_desc;

/* Element counts of the level-L icosahedral Voronoi mesh:
 * nCells = 10*4^L+2, nEdges = 30*4^L, nVertices = 20*4^L (L in 0..11). */
int tm_mesh_icos_dims(int level, int32_t *nCells, int32_t *nEdges,
                      int32_t *nVertices);

/* Generate the level-L mesh by recursive bisection of the icosahedron
 * (generators = bisection points, Voronoi vertices = triangle circumcentres),
 * with MPAS orderings and TRiSK weightsOnEdge.  No reference counterpart: the
 * reference only reads such meshes (mesh_mod.F90:79-167). */
int tm_mesh_icos_generate(int level, tm_mesh_desc *m);

/*
 * MPAS grid files (NetCDF classic CDF-1, CDF-2 and CDF-5; NetCDF-4/HDF5 is not
 * read), replacing the io_get_dim / io_input calls of mesh_init
 * (mesh_mod.F90:79-164) without a NetCDF library.
 * tm_mesh_mpas_dims: dims = {nCells, nEdges, nVertices, maxEdges, maxEdges2,
 * vertexDegree} and the global attribute sphere_radius (1.0 when absent:
 * geometry on the unit sphere).
 * tm_mesh_mpas_read: the arrays mesh_init reads into the caller-allocated
 * members of m (its dimensions must equal the file's; NULL members skipped).
 * Integer variables must be NC_INT, real ones NC_DOUBLE or NC_FLOAT (widened).
 * tm_mesh_mpas_error: text of the last TM_EIO of the calling thread.
 */
int tm_mesh_mpas_dims(const char *path, int32_t dims[6], double *sphere_radius);
int tm_mesh_mpas_read(const char *path, tm_mesh_desc *m);
const char *tm_mesh_mpas_error(void);

/*
 * Coefficients of the MPAS vector reconstruction used by the history output
 * (mpas_initialize_vectors, mpas_vector_operations.F90:493-557, and
 * mpas_init_reconstruct, mpas_vector_reconstruction.F90:52-125, with the
 * RBF solve of mpas_rbf_interpolation.F90:1076-1142), on the radius-scaled
 * geometry: coeffs[nCells][maxEdges][3] = coeffs_reconstruct(1:3, slot, cell)
 * (unused slots 0).  A cell vector is sum_slot coeffs * u(edgesOnCell).
 */
int tm_reconstruct_coeffs(const tm_mesh_desc *m, double radius, double *coeffs);

/*
 * History files in the layout of history_mod (history_mod.F90:27-174): a
 * NetCDF classic (CDF-2) file with the radius-scaled mesh variables and, per
 * tm_history_write, one record of Time (s), u [nEdges], zonalWind,
 * merdionalWind (the reference's spelling), h = (gd + ghs)/g, div [nCells],
 * pv, vor [nVertices], tm and te.  Arrays in reference order.  The file is
 * complete after every write (numrecs is rewritten).
 */
typedef struct tm_history tm_history;
int tm_history_create(tm_history **h, const char *path, const tm_mesh_desc *m, double radius, double dt,
                      int time_scheme, const char *case_name);
int tm_history_write(tm_history *h, double time_seconds, const double *u, const double *zonal,
                     const double *meridional, const double *h_cell, const double *pv, const double *vor,
                     const double *div, double total_mass, double total_energy);
int tm_history_close(tm_history *h);

/* poly_fit_run (poly_fit_mod.F90:119-609) on radius-scaled geometry.
 * radius: the scale applied to the unit-sphere mesh (const_mod.F90:11). */
int tm_poly_fit(const tm_mesh_desc *m, double radius, tm_fit_desc *f);

/*
 * Seeded initial state (no reference counterpart; the template is
 * mountain_zonal_flow_test_mod.F90:25-61): solid-body zonal flow u0 = 20 m/s
 * from a vertex stream function, balanced free surface with area mean
 * 5000 m, plus 8 Gaussian bumps (+-50 m, 500 km e-folding) and one Gaussian
 * mountain (2000 m, 1500 km half-width), all positions/signs from SplitMix64
 * seeded with `seed`.  Tracer k is U[0,1] from seed tracer_seeds[k], smoothed
 * by 4 passes of 0.5*self + 0.5*neighbour mean.  Outputs (reference order):
 * gd[nCells] (geopotential depth, m^2 s^-2), u[nEdges] (m/s), ghs[nCells],
 * tracers[ntracers][nCells] (may be NULL when ntracers == 0).
 */
int tm_case_seeded(const tm_mesh_desc *m, double radius, uint64_t seed,
                   int ntracers, const uint64_t *tracer_seeds, double *gd,
                   double *u, double *ghs, double *tracers);

/* Williamson et al. (1992) case 2, steady geostrophic flow
 * (reference src/test_cases/swm/steady_geostrophic_flow_test_mod.F90:20-54
 * restated with alpha = 0): gd, u, ghs (= 0). */
int tm_case_steady_geostrophic(const tm_mesh_desc *m, double radius,
                               double *gd, double *u, double *ghs);

#ifdef __cplusplus
}
#endif
#endif
