/*
 * tmcore_cpu.h -- C-ABI of libtmcpu.so, the FP64 CPU path in Fortran
 * (tmcore_amd/fortran/tm_swm_cpu_mod.F90, bind(c)): the CPU baseline the
 * device path is timed against, and the path the Fortran host runs with
 * use_gpu = .false.
 *
 * Layouts are those of tmcore_host.h: the reference-order mesh and fits
 * (1-based connectivity, unit-sphere geometry), state arrays in reference
 * order, tracers [ntracers][nCells].
 */
#ifndef TMCORE_CPU_H
#define TMCORE_CPU_H

#include "tmcore_host.h"

#ifdef __cplusplus
extern "C" {
#endif

/*
 * nsteps CRK4 steps with the energy-conserving beta on the host, the
 * reference's time loop (src/core/tmcore_swm_mod.F90:75-88 with
 * time_integrate_CRK4, src/core/time_scheme_mod.F90:109-154), its stage
 * (tmcore_swm_mod.F90:92-125) with the operators of
 * src/core/operators_mod.F90 for pv_scheme 3 at interp_pv_order 2-4 and
 * adv_order 2, the tracer advection (adv_scheme_sg11_mod.F90:33-48) and the
 * invariants of diag_run after every step (tmcore_swm_mod.F90:215-255).
 * gd[nCells], iap_u[nEdges] and tracers (unused when ntracers is 0) are
 * advanced in place; inv
 * [nsteps][4] and beta[nsteps] receive every step's invariants and beta;
 * *seconds the wall time of the step loop.  nthreads > 0 sets the OpenMP
 * threads (omp_set_num_threads: it stays the process's default for later
 * OpenMP regions of this runtime); 0 keeps the current one.  Returns 0, or -1 for an order outside 2-4, missing vertex fits
 * at orders 3-4, or negative counts.  Not re-entrant (one mesh at a time).
 */
int tm_cpu_run(const tm_mesh_desc *mesh, const tm_fit_desc *fit, double radius, double dt, int interp_pv_order,
               int ntracers, int nsteps, int nthreads, double *gd, double *iap_u, const double *ghs,
               double *tracers, double *inv, double *beta, double *seconds);

#ifdef __cplusplus
}
#endif

#endif /* TMCORE_CPU_H */
