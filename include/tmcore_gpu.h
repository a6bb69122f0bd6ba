/*
 * tmcore_gpu.h -- C-ABI of libtmgpu.so, the MI355X (gfx950) HIP path of the
 * TRiSK shallow-water RK stage.
 *
 * The reference has no FFI: its seams are Fortran procedures.  Each entry
 * point below names the reference interface it replaces.  Host arrays are in
 * reference order (cell/edge/vertex numbering of the mesh descriptor) and are
 * borrowed only for the duration of a call; the context owns all device
 * memory.  Every function returns 0 (TM_OK) or a negative TM_E* code from
 * tmcore_host.h; tm_last_error() gives the text.  A context is used from one
 * host thread at a time.
 */
#ifndef TMCORE_GPU_H
#define TMCORE_GPU_H

#include <stdint.h>

#include "tmcore_host.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tm_ctx tm_ctx;

/* Scheme switches: the reference's &tmcore namelist group
 * (src/core/params_mod.F90:35-58) plus the device options. */
typedef struct tm_opts {
  int time_scheme;      /* 1 = CRK3, 2 = CRK4, 3 = PC2 (time_scheme_mod.F90:65-184);
                           tracers need CRK4                                      */
  int pv_scheme;        /* 1 = APVM, 2 = conservative APVM (global eps),
                           3 = high-order upwind PV (operators_mod.F90:283-412)  */
  int interp_pv_order;  /* 2..6 for pv_scheme 3 (5 and 6 need the order-4 fits)  */
  int adv_order;        /* 2..6: add_upwind_correction_on_cell (operators_mod.F90:
                           423-472), a no-op at 2; 3..6 need the cell fits      */
  int adv_mono;         /* logical, only read for adv_order 3                    */
  int conserve_energy;  /* logical: CRK4 beta rescale                            */
  double apvm_weight;   /* read by pv_scheme 1                                   */
  double dt;            /* seconds                                               */
  double radius;        /* metres; the reference's is 6371220 (const_mod.F90:11) */
  int ntracers;         /* passive tracers advected by the stage winds           */
  int precision;        /* 64: FP64 state and stencils; 32: FP32 state/stencils
                           with FP64 geometry, weights and invariant sums        */
  int device;           /* HIP device ordinal                                    */
  int renumber;         /* 1: Hilbert-curve renumbering of cells/edges/vertices  */
  int geometry_scaled;  /* 1: x/y/z, dvEdge, dcEdge, areas and kites are already
                           multiplied by radius / radius^2, as mesh_init leaves
                           them (mesh_mod.F90:205-224); 0: unit sphere        */
} tm_opts;

/* Fill *o with the reference namelist defaults (run/namelist.tmcore_swm:12-22):
 * CRK4, pv_scheme 3, interp_pv_order 3, adv_order 2, conserve_energy, FP64. */
void tm_opts_default(tm_opts *o);

/* Replaces mesh_init + poly_fit_init + state_init/tend_init
 * (reference tmcore_swm_mod.F90:37-41): uploads the mesh (reference layout of
 * tmcore_host.h), derives nSignEdge/tSignEdge/areaEdge/fVertex as
 * mesh_mod.F90:169-224 does, renumbers along a Hilbert curve and allocates the
 * device state.  f may be NULL unless pv_scheme 3 with interp_pv_order >= 3
 * (vertex fits) or adv_order >= 3 (cell fits) is selected. */
int tm_ctx_create(tm_ctx **out, const tm_mesh_desc *m, const tm_fit_desc *f, const tm_opts *o);
int tm_ctx_destroy(tm_ctx *ctx);

/* Kernel layout of the stage (no counterpart in the reference; results are
 * bit-identical in every mode).  Bit 1 (the PV kernel) and bit 2 (the
 * tendency kernel) derive each edge's edgesOnEdge slots from the edge rows
 * of its two cells instead of streaming the edgesOnEdge table.  That needs
 * the MPAS order of edgesOnEdge (the edges of cellsOnEdge(1) after the edge
 * counter-clockwise, then those of cellsOnEdge(2)), which tm_ctx_create
 * checks for every edge.  Bit 4: the tendency kernel of RK stages 1-3 also
 * forms the next stage's edge values (gd_edge, u and their products), so
 * stages 2-4 run no edge-prep pass.  Bit 8: the tendency kernel stages u,
 * gd_edge and pv_edge of its 256-edge tile and of the tile's edgesOnEdge
 * halo in LDS and gathers from there (supersedes bit 2; needs every tile's
 * halo within 512 edges, checked by tm_ctx_create).  Bit 16: the PV kernel
 * does the same with u for the tangent wind (supersedes bit 1).  Bits 4, 8
 * and 16 need time_scheme 2 (CRK4) with pv_scheme 3 and interp_pv_order
 * <= 4.  Every bit needs adv_order 2.
 * tm_ctx_create picks the fastest mode measured for the context's precision
 * that the mesh and options allow.  Setting a mode they do not allow
 * returns TM_EINVAL.  For tests and timing comparisons. */
int tm_ctx_set_layout(tm_ctx *ctx, int mode);
int tm_ctx_get_layout(tm_ctx *ctx, int *mode);
const char *tm_last_error(const tm_ctx *ctx);

/* Prognostic upload (reference state(old), tmcore_swm_mod.F90:62-64):
 * gd[nCells] and ghs[nCells]; either iap_u[nEdges], or u[nEdges] from which
 * iap_u = sqrt(gd_edge)*u is formed on the device as the reference's run
 * prologue does; tracers[ntracers][nCells] may be NULL when ntracers == 0.
 * Upload and download are synchronous; both stage through a pinned host
 * buffer the context allocates at its first transfer (sized to the largest
 * state array) and keeps until tm_ctx_destroy. */
int tm_state_upload(tm_ctx *ctx, const double *gd, const double *u, const double *iap_u,
                    const double *ghs, const double *tracers);
/* Download of state(old): any pointer may be NULL.  u is the inverse IAP
 * wind (tmcore_swm_mod.F90:80). */
int tm_state_download(tm_ctx *ctx, double *gd, double *u, double *iap_u, double *tracers);

/* Replaces the time loop body of tmcore_swm_run (tmcore_swm_mod.F90:75-87):
 * nsteps of time_integrate_CRK4 + time_advance + the diagnostic refresh.
 * invariants_out[4*nsteps] (mass, energy, potential enstrophy, absolute
 * vorticity after each step, tmcore_swm_mod.F90:215-255) and beta_out[nsteps]
 * may be NULL.  Synchronous at return. */
int tm_step(tm_ctx *ctx, int nsteps, double *invariants_out, double *beta_out);

/* Same as tm_step but asynchronous and without outputs: enqueues nsteps on
 * the context's stream (replayed from a captured hipGraph).  tm_sync waits. */
int tm_step_async(tm_ctx *ctx, int nsteps);
int tm_sync(tm_ctx *ctx);

/* The four invariants of the current state (diag_run, diag_mod.F90:76-91). */
int tm_invariants(tm_ctx *ctx, double out[4]);

/* tm_invariants without device allocation or context scratch: no counterpart
 * in the reference.  For callers that cannot allocate (code under graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The block
 * partials, sums and result go through a static device array of 3.25 MiB per
 * GPU (kept until the process ends), so the context's mesh may have at most
 * 4,194,304 cells, edges and vertices (16384 workgroups of 256 per element
 * kind); larger meshes return TM_EINVAL with a message.  The refresh pass
 * still writes the context's own edge, cell and vertex work arrays, as
 * tm_invariants does.  Results equal tm_invariants'.  Every call on one GPU
 * shares that GPU's array: calls that overlap in time on one GPU race. */
int tm_invariants_noalloc(tm_ctx *ctx, double out[4]);

/* Test path: one spatial_operators call (tmcore_swm_mod.F90:92-125) on the
 * current state plus the tracer tendencies, outputs in reference order; any
 * output pointer may be NULL.  Diagnostics mirror state/tend fields of the
 * reference: gd_edge, u (edge), ke (cell), vor/gd_vertex/pv_vertex (vertex),
 * v/pv_edge/pv_flx/u_tend (edge), tend_gd (cell), tend_iap (edge). */
typedef struct tm_stage_out {
  double *tend_gd, *tend_iap, *tend_q;
  double *gd_edge, *u, *ke, *vor, *gd_vertex, *pv_vertex, *v, *pv_edge, *pv_flx, *u_tend;
} tm_stage_out;
int tm_stage(tm_ctx *ctx, tm_stage_out *out);

/* Test path: single operators with the argument order of operators_mod.F90,
 * device round trip per call (not used in timed runs). */
int tm_op_div(tm_ctx *ctx, const double *f_edge, double *div);                 /* :40-55   */
int tm_op_curl(tm_ctx *ctx, const double *f_edge, double *curl);               /* :76-91   */
int tm_op_c2e(tm_ctx *ctx, const double *f_cell, double *f_edge);              /* :93-104  */
int tm_op_c2v(tm_ctx *ctx, const double *f_cell, double *f_vertex);            /* :106-123 */
int tm_op_v2c(tm_ctx *ctx, const double *f_vertex, double *f_cell);            /* :125-141 */
int tm_op_v2e(tm_ctx *ctx, const double *f_vertex, double *f_edge);            /* :143-156 */
int tm_op_ke(tm_ctx *ctx, const double *u_edge, double *ke_cell);              /* :206-221 */
int tm_op_tangent_wind(tm_ctx *ctx, const double *u_edge, double *v_edge);     /* :223-237 */
int tm_op_tangent_vor_flux(tm_ctx *ctx, const double *u_edge, const double *gd_edge,
                           const double *pv_edge, double *flux);               /* :239-260 */
/* f_edge += the upwind correction of f_cell at the context's adv_order
 * (unchanged at adv_order 2), upwinded by sign(u_edge).                        :423-472 */
int tm_op_upwind_cell(tm_ctx *ctx, const double *f_cell, const double *u_edge, double *f_edge);
int tm_op_pv_vertex(tm_ctx *ctx, const double *vor_vertex, const double *gd_vertex,
                    double *pv_vertex);                                        /* :262-270 */
/* The context's pv_scheme (1 APVM, 2 conservative APVM, 3 high-order upwind
 * with its interp_pv_order); each scheme reads the inputs the reference's does. :273-412 */
int tm_op_pv_edge(tm_ctx *ctx, const double *u_edge, const double *v_edge, const double *gd_edge,
                  const double *gd_tend_vertex, const double *pv_vertex, const double *pv_cell,
                  double *pv_edge);
/* calc_u_tend_on_edge (tmcore_swm_mod.F90:155-201, flux_4th_order_correct off) */
int tm_op_u_tend(tm_ctx *ctx, const double *u_edge, const double *ke_cell, const double *gd_cell,
                 const double *gd_edge, const double *ghs_cell, const double *pv_flx_edge,
                 const double *gd_tend_edge, double *u_tend_edge, double *iap_u_tend_edge);

/* History fields of the current state, in reference order (any pointer may be
 * NULL): the refresh of tmcore_swm_mod.F90:78-84 (u on edges, pv and vor on
 * vertices), mpas_reconstruct_1d (mpas_vector_reconstruction.F90:195-247) of
 * u into zonal and meridional cell winds, h = (gd + ghs)/g and
 * div_operator(u) (history_mod.F90:125,166-168).  tm_history_setup uploads
 * the reconstruction coefficients of tm_reconstruct_coeffs
 * ([nCells][maxEdges][3], reference order) with latCell/lonCell; it is
 * needed before zonal or meridional are asked for. */
int tm_history_setup(tm_ctx *ctx, const double *coeffs, const double *latCell, const double *lonCell);
int tm_history_fields(tm_ctx *ctx, double *u, double *zonal, double *meridional, double *h, double *pv,
                      double *vor, double *div);

/* Device timing of the last tm_step/tm_step_async batch, in milliseconds,
 * measured with HIP events on the context's stream. */
int tm_last_elapsed_ms(tm_ctx *ctx, double *ms);

/* Element counts and algorithmic bytes of one CRK4 step for this context:
 * bytes[0] = per stage, bytes[1] = finalize, bytes[2] = per step. */
int tm_step_bytes(tm_ctx *ctx, double bytes[3]);

/* Instrumentation (no reference counterpart): runs nsteps CRK4 steps with
 * HIP events around every kernel launch on the context's stream and returns,
 * per kernel kind (TM_K_*), the summed device time in ms and the number of
 * launches; kernel_bytes gives the compulsory bytes one launch of each kind
 * moves (each array it reads or writes counted once). */
#define TM_K_EDGE_PREP 0
#define TM_K_CELL_VERTEX 1
#define TM_K_EDGE_PV 2
#define TM_K_EDGE_TEND 3
#define TM_K_REDUCE 4
#define TM_K_FINALIZE 5
#define TM_K_COUNT 6
int tm_profile_steps(tm_ctx *ctx, int nsteps, double ms[TM_K_COUNT], int launches[TM_K_COUNT]);
int tm_kernel_bytes(tm_ctx *ctx, double bytes[TM_K_COUNT]);

/* Instrumentation (no reference counterpart): a plain device-to-device copy
 * (kernel tm_copy_probe_kernel) of `bytes` bytes, timed with HIP events over
 * `iters` launches; *gbps = 2*bytes / average launch time (read + write).
 * The practical HBM ceiling the stage kernels are compared with, and the
 * known-size kernel that scales rocprof's FETCH_SIZE/WRITE_SIZE counters. */
int tm_hbm_copy_probe(int device, double bytes, int iters, double *gbps);

/* ------------------------------------------------------------------------
 * Batched ensemble (no reference counterpart): many members on meshes of
 * uneven size, each level group advanced with its own dt, all of them in one
 * persistent kernel launch (replayed from a hipGraph) per tm_ens_run.
 * Members of a level share its mesh constants.  A level's members are cut
 * into slices of consecutive members (tm_ens_set_slices), each slice's state
 * stored member-fastest; the slices of a level are spread over the GPU's
 * XCDs, so that each XCD's L2 holds the member rows of its own slice.
 * Every member runs the same CRK4 + beta scheme as tm_step (precision from
 * tm_opts, FP64 double-double global sums and invariants per member).
 * ------------------------------------------------------------------------ */
typedef struct tm_ens tm_ens;

typedef struct tm_ens_level {
  const tm_mesh_desc *mesh;  /* reference-layout mesh of this level group   */
  const tm_fit_desc *fit;    /* poly_fit weights (interp_pv_order >= 3)     */
  int members;               /* 1..256                                      */
  double dt;                 /* seconds                                     */
} tm_ens_level;

int tm_ens_create(tm_ens **out, int nlevels, const tm_ens_level *levels, const tm_opts *o);
int tm_ens_destroy(tm_ens *e);
const char *tm_ens_last_error(const tm_ens *e);
/* State of one member in reference order (u -> iap_u as tm_state_upload). */
int tm_ens_upload(tm_ens *e, int level, int member, const double *gd, const double *u, const double *ghs,
                  const double *tracers);
int tm_ens_download(tm_ens *e, int level, int member, double *gd, double *iap_u, double *tracers);
/* Advances level l by nsteps[l] CRK4 steps; levels with fewer steps are
 * spread evenly over the run.  *ms = device time of the launch.  The four
 * invariants of every member after the run come back through
 * tm_ens_invariants. */
int tm_ens_run(tm_ens *e, const int *nsteps, double *ms);
/* Member slices of level l: 1..8, dividing its member count; set before the
 * first tm_ens_upload (it re-lays out the level's storage).  tm_ens_create
 * picks the most slices, up to eight, that keep 16 members in a slice. */
int tm_ens_set_slices(tm_ens *e, int level, int slices);
int tm_ens_get_slices(tm_ens *e, int level, int *slices);
int tm_ens_invariants(tm_ens *e, int level, int member, double out[4]);
/* Tuning: elements per work chunk, workgroups per CU of the persistent grid
 * and members per thread (1, 2 or 4; it must divide every slice's member
 * count and span at most 16 bytes).  tm_ens_autotune times candidates over
 * `ticks` ticks on a saved copy of the state (restored afterwards), keeps the
 * fastest and returns it in best = {chunk, blocks_per_cu, members_per_thread}. */
int tm_ens_set_tuning(tm_ens *e, int chunk, int blocks_per_cu, int members_per_thread);
int tm_ens_autotune(tm_ens *e, int ticks, int best[3], double *best_ms_per_tick);

/* ------------------------------------------------------------------------
 * Learned closure (no reference counterpart): a per-edge MLP
 * 40 -> 128 -> 128 -> 128 -> 2 with GELU, BF16 weights on MFMA and FP32
 * accumulation, trained and run by the library's own kernels.
 * Frame: kappa(e) = +1 when the mesh normal cellsOnEdge(1) -> (2) points
 * east (north if exactly meridional), else -1; velocities and both outputs
 * are taken along kappa(e) * n(e).
 * Inputs of edge e, j = field*10 + slot: 4 fields on the edge (slot 0) and on
 * 9 edgesOnEdge neighbours (slots 1..9: the west cell's other edges
 * counter-clockwise from e, then the east cell's; the 10th is not used):
 *   0: kappa(e) * nSign(e,c) * nSign(e',c) * u(e'), c the shared cell
 *   1: gd_edge(e')
 *   2: relative vorticity at the edge, (vor(v1(e')) + vor(v2(e')))/2
 *   3: velocity divergence at the edge, (div(c1(e')) + div(c2(e')))/2,
 *      div_c = (1/A_c) sum_i nSign(i,c) u(e_i) dv(e_i)
 * so the inputs do not depend on the mesh's edge orientation; each is
 * normalised as (value - in_shift[j]) * in_scale[j].
 * Outputs: y0, a normal-velocity tendency kappa * out_scale[0] * y0 (m s^-2),
 * and y1, an edge flux F = kappa * out_scale[1] * y1 along the mesh normal
 * whose divergence corrects the thickness,
 *   P_c = -(1/A_c) sum_i nSign(i,c) F(e_i) dv(e_i),
 * so total mass is unchanged.  Attached to a context, every tm_step computes
 * the inputs from the state at the start of the step and, after the CRK4
 * update, adds
 *   iap_u += dt * sqrt(gd_edge_start) * kappa * out_scale[0] * y0,
 *   gd    += dt * P.
 * Without it (or with mlp == NULL) the plain method runs unchanged.
 * ------------------------------------------------------------------------ */
#define TM_MLP_NIN 40
#define TM_MLP_NHID 128
#define TM_MLP_NOUT 2
#define TM_MLP_NPARAMS 38530
typedef struct tm_mlp tm_mlp;

/* Parameters: W1[40][128] b1[128] W2[128][128] b2 W3[128][128] b3 W4[128][2]
 * b4[2], FP32 masters (BF16 MFMA operands are packed from them). */
int tm_mlp_create(tm_mlp **out, int device, uint64_t seed);
int tm_mlp_destroy(tm_mlp *p);
const char *tm_mlp_last_error(const tm_mlp *p);
int tm_mlp_get_params(tm_mlp *p, float *params);
int tm_mlp_set_params(tm_mlp *p, const float *params);
int tm_mlp_set_scales(tm_mlp *p, const float *in_shift /*[40]*/, const float *in_scale /*[40]*/,
                      const float *out_scale /*[2]*/);
int tm_mlp_get_scales(tm_mlp *p, float *in_shift, float *in_scale, float *out_scale);
/* y[n][2] = MLP(x[n][40]), host arrays (x already scaled). */
int tm_mlp_forward(tm_mlp *p, const float *x, int n, float *y);
/* Gradients of L = sum(dy * y) w.r.t. the parameters (grad[TM_MLP_NPARAMS])
 * and, if dx != NULL, w.r.t. x (dx[n][40]). */
int tm_mlp_backward(tm_mlp *p, const float *x, const float *dy, int n, float *grad, float *dx);
/* One Adam step (beta 0.9/0.999, eps 1e-8) with the given gradient. */
int tm_mlp_adam_step(tm_mlp *p, const float *grad, float lr);

/* Normalised inputs of the context's current state, reference edge order:
 * x[nEdges][40]. */
int tm_closure_features(tm_ctx *ctx, tm_mlp *mlp, float *x);
/* Applies the closure in every following tm_step (mlp == NULL detaches).
 * The MLP must outlive the attachment; its parameters and scales may change
 * between steps (training) without re-attaching. */
int tm_closure_attach(tm_ctx *ctx, tm_mlp *mlp);
/* The MLP runs on every nsteps-th step (counted from the attach or the last
 * tm_state_upload), on the state at the start of that step: the form of a
 * closure trained on the mean tendency error over nsteps steps.  Its
 * correction (with the sqrt(gd_edge) of that state) is added
 *   lumped = 0: after every step of the interval, unchanged;
 *   lumped = 1: once, after the last step of the interval, times nsteps
 *               (the interval's whole correction, as the training target
 *               b - Plain^n(a) defines it).
 * Default 1 step (MLP and correction every step). */
int tm_closure_set_interval(tm_ctx *ctx, int nsteps, int lumped);
/* Factor (>= 0, default 1) on every correction the attached closure applies:
 * the deployment's damping of a closure whose full correction, fitted to
 * short-horizon targets, feeds enstrophy into long runs. */
int tm_closure_set_gain(tm_ctx *ctx, double gain);
/* Adds one training sample held on the device: the inputs of the context's
 * current state (normalised with mlp's scales at this call) and its one-step
 * targets in reference order, along the mesh normal,
 *   t_u[nEdges]  = (iap_target - iap_plain) / (dt * sqrt(gd_edge)),
 *   t_gd[nCells] = (gd_target - gd_plain) / dt,
 * where *_plain is one plain step from the current state and gd_edge is of
 * the current state.  The loss of a sample is
 *   mean_e (y0 - kappa t_u/out_scale[0])^2 + mean_c ((P_c - t_gd_c)/tg_scale)^2. */
int tm_closure_dataset_add(tm_ctx *ctx, tm_mlp *mlp, const double *t_u, const double *t_gd, double tg_scale);
int tm_closure_dataset_size(tm_ctx *ctx, int *n);
int tm_closure_dataset_clear(tm_ctx *ctx);
/* Mean loss of every sample in [first, first + count) under the current
 * parameters, summed per term: loss[0] (velocity) and loss[1] (thickness).
 * tm_closure_eval and tm_closure_train reuse the buffer that holds an
 * attached closure's correction between steps: after either, the next
 * tm_step computes the correction afresh from its state. */
int tm_closure_eval(tm_ctx *ctx, tm_mlp *mlp, int first, int count, double *loss);
/* epochs passes over samples [0, ntrain), one Adam step (lr, beta 0.9/0.999)
 * per sample in an order shuffled per epoch from seed; loss_hist[2*epochs]
 * receives the two loss terms averaged over each epoch (before each step). */
int tm_closure_train(tm_ctx *ctx, tm_mlp *mlp, int ntrain, int epochs, float lr, uint64_t seed,
                     double *loss_hist);

#ifdef __cplusplus
}
#endif
#endif
