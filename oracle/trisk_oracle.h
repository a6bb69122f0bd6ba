/*
 * trisk_oracle.h -- TEST INFRASTRUCTURE, not product code.
 *
 * Plain-C FP64 restatement of the reference TMCORE TRiSK shallow-water path,
 * used only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * as the checker of the HIP path.  Nothing under tmcore_amd/ links or loads it.
 *
 * Arrays are in the reference's layout: 1-based int32 connectivity,
 * Fortran (slot, n) column-major = C [n][slot], unit-sphere geometry that
 * om_create scales by the radius (reference src/core/mesh_mod.F90:205-224).
 */
#ifndef TRISK_ORACLE_H
#define TRISK_ORACLE_H

#include <stdint.h>

#include "../include/tmcore_host.h"

typedef struct om_params {
  int time_scheme;       /* 1 CRK3, 2 CRK4, 3 PC2      (params_mod.F90:35)  */
  int pv_scheme;         /* 1 APVM, 2 cons. APVM, 3 upwind (operators_mod.F90:283-292) */
  int interp_pv_order;   /* 2..6                        (params_mod.F90:44)  */
  int adv_order;         /* 2..6                        (params_mod.F90:42)  */
  int adv_mono;          /* logical                     (params_mod.F90:43)  */
  int conserve_energy;   /* logical                     (params_mod.F90:38)  */
  double apvm_weight;    /*                             (params_mod.F90:40)  */
  double dt;             /* seconds                     (params_mod.F90:13)  */
} om_params;

typedef struct om_ctx om_ctx;

/* mesh_init derived quantities (mesh_mod.F90:169-224) + the fit weights. */
om_ctx *om_create(const tm_mesh_desc *m, const tm_fit_desc *f, double radius, const om_params *p);
void om_destroy(om_ctx *c);
int om_set_params(om_ctx *c, const om_params *p);

/* poly_fit_run restated literally (poly_fit_mod.F90:119-609) with
 * math_inv_matrix (math_mod.F90:52-133); same layout as tm_poly_fit. */
int om_poly_fit(const tm_mesh_desc *m, double radius, tm_fit_desc *f);

/* MPAS vector reconstruction (reconstruct_oracle.c): the RBF coefficients
 * coeffs[nCells][maxEdges][3] of mpas_init_reconstruct on radius-scaled
 * geometry, and mpas_reconstruct_1d of an edge field u into
 * out[5][nCells] = uX, uY, uZ, zonal, meridional. */
int om_reconstruct_coeffs(const tm_mesh_desc *m, double radius, double *coeffs);
int om_reconstruct(const tm_mesh_desc *m, const double *coeffs, const double *u, double *out);

/* Single operators, same argument order as operators_mod.F90. */
void om_div(const om_ctx *c, const double *f_edge, double *div);                 /* 40-55   */
void om_curl(const om_ctx *c, const double *f_edge, double *curl);               /* 76-91   */
void om_c2e(const om_ctx *c, const double *f_cell, double *f_edge);              /* 93-104  */
void om_c2v(const om_ctx *c, const double *f_cell, double *f_vertex);            /* 106-123 */
void om_v2c(const om_ctx *c, const double *f_vertex, double *f_cell);            /* 125-141 */
void om_v2e(const om_ctx *c, const double *f_vertex, double *f_edge);            /* 143-156 */
void om_iap(const om_ctx *c, const double *gd_edge, const double *u, double *iap);          /* 158-166 */
void om_inverse_iap(const om_ctx *c, const double *gd_edge, const double *iap, double *u);  /* 168-176 */
void om_ke(const om_ctx *c, const double *u_edge, double *ke_cell);              /* 206-221 */
void om_tangent_wind(const om_ctx *c, const double *u_edge, double *v_edge);     /* 223-237 */
void om_tangent_vor_flux(const om_ctx *c, const double *u, const double *gd_edge,
                         const double *pv_edge, double *flux);                    /* 239-260 */
void om_pv_vertex(const om_ctx *c, const double *vor, const double *gd_vertex, double *pv); /* 262-270 */
void om_pv_edge(om_ctx *c, const double *u, const double *v, const double *gd_edge,
                const double *gd_tend_vertex, const double *pv_vertex, const double *pv_cell,
                double *pv_edge);                                                 /* 273-412 */
void om_upwind_cell(om_ctx *c, const double *f_cell, const double *u_edge, double *f_edge); /* 423-472 */
/* calc_gd_tend_on_cell (tmcore_swm_mod.F90:127-139) */
void om_gd_tend(const om_ctx *c, const double *u, const double *gd_edge, double *tend);
/* calc_u_tend_on_edge (tmcore_swm_mod.F90:155-201), flux_4th_order_correct off */
void om_u_tend(const om_ctx *c, const double *u, const double *ke, const double *gd,
               const double *gd_edge, const double *ghs, const double *pv_flx,
               const double *gd_tend_edge, double *u_tend, double *iap_tend);
/* adv_scheme_sg11_run (adv_scheme_sg11_mod.F90:33-48) */
void om_adv(om_ctx *c, const double *f_cell, const double *u, double *f_edge, double *tend);

/* One call of spatial_operators (tmcore_swm_mod.F90:92-125) on (gd, iap_u)
 * plus, per tracer, adv_scheme_sg11_run with the stage wind.  Outputs the
 * two prognostic tendencies and every diagnostic named in the reference. */
typedef struct om_diag {
  double *gd_edge, *u, *tend_gd, *tend_gd_edge, *tend_gd_vertex, *gd_vertex, *vor, *v, *ke,
      *pv_vertex, *pv_cell, *pv_edge, *pv_flx, *u_tend, *iap_tend; /* any may be NULL */
} om_diag;
int om_spatial_operators(om_ctx *c, const double *gd, const double *iap_u, const double *ghs,
                         int ntracers, const double *q, double *tend_gd, double *tend_iap,
                         double *tend_q, om_diag *diag);

/* Initial diagnostic refresh of tmcore_swm_run (tmcore_swm_mod.F90:62-64):
 * iap_u = sqrt(gd_edge) * u. */
void om_initial_iap(om_ctx *c, const double *gd, const double *u, double *iap_u);

/* The four invariants after the post-step refresh (tmcore_swm_mod.F90:78-85,
 * 215-255): [mass, energy, potential enstrophy, absolute vorticity]. */
void om_invariants(om_ctx *c, const double *gd, const double *iap_u, const double *ghs,
                   double out[4]);

/* nsteps of time_integrate (time_scheme_mod.F90:47-184) + time_advance,
 * in place on (gd, iap_u, q[ntracers][nCells]).  Tracers use the stage
 * winds with beta = 1.  invariants_out[4*nsteps] and beta_out[nsteps] may
 * be NULL. */
int om_run(om_ctx *c, int nsteps, double *gd, double *iap_u, const double *ghs, int ntracers,
           double *q, double *invariants_out, double *beta_out);

/* math_inv_matrix (math_mod.F90:52-88), column-major n x n; A is overwritten. */
void om_inv_matrix(int n, double *A, double *B);

#endif
