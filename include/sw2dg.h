/*
 * sw2dg.h -- C-ABI of the MI355X SW2D single-porosity solver (libsw2dg.so).
 *
 * The reference has no FFI: its seams are C++ virtuals and public methods.
 * Each entry point below names the reference interface it stands in for.
 * Host arrays are caller-owned and copied in/out; no pointer is retained.
 * Every call returns SW2DG_OK (0) or an error code; sw2dg_last_error()
 * gives the message.  One ctx = one HIP stream on one device; calls on one
 * ctx must come from one host thread at a time.  Arithmetic is IEEE FP64.
 *
 * Index convention: every array crossing this boundary is in the
 * reference's ordering (cells, boundary edges and internal edges numbered
 * as in geometry.sw2d).  The device reorders internally.
 */
#ifndef SW2DG_H
#define SW2DG_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define SW2DG_OK 0
#define SW2DG_ERR_ARG 1         /* invalid argument / shape                 */
#define SW2DG_ERR_NAN 2         /* NaN mass flux (swSpHyp.cpp:374-392)      */
#define SW2DG_ERR_NEG_DEPTH 3   /* negative depth with detection active     */
#define SW2DG_ERR_FLOW_CONFIG 4 /* impossible HLLC regime (swSpHyp.cpp:729) */
#define SW2DG_ERR_BC 5          /* bad boundary type / value                */
#define SW2DG_ERR_HIP 6         /* HIP runtime failure                      */
#define SW2DG_ERR_NO_DEVICE 7   /* no GPU visible                           */

/* option values (input.solv keys num_scheme, topo_discretization, model_friction) */
#define SW2DG_GODUNOV 0
#define SW2DG_MUSCL_EVR 1
#define SW2DG_TOPO_DISCONTINUOUS 0
#define SW2DG_TOPO_BLENDING 1
#define SW2DG_FRIC_NONE 0
#define SW2DG_FRIC_MANNING 1
#define SW2DG_FRIC_STRICKLER 2
#define SW2DG_FRIC_CHEZY 3
/* model_runoff and model_infiltration values (scs and h_proportional are
 * industrial-only in the reference and are refused there too) */
#define SW2DG_RUNOFF_NONE 0
#define SW2DG_RUNOFF_FULL 1          /* full_rainfall            */
#define SW2DG_RUNOFF_INFIL_COEFF 2   /* infiltration_coefficient */
#define SW2DG_RUNOFF_RUNOFF_COEFF 3  /* runoff_coefficient       */
#define SW2DG_INFIL_NONE 0
#define SW2DG_INFIL_HORTON 1
/* boundary series types ("fr", "h", "z", "q") */
#define SW2DG_BC_FR 0
#define SW2DG_BC_H 1
#define SW2DG_BC_Z 2
#define SW2DG_BC_Q 3

typedef struct sw2dg_ctx sw2dg_ctx;

/* Mesh tables in reference order: swParent::setMeshTables
 * (src/sw2dSolver/models/swParent.cpp:391-501). */
typedef struct sw2dg_mesh {
    int64_t n_cell, n_bnd, n_int;
    const double *cell_area, *cell_x, *cell_y, *cell_zb; /* [n_cell] */
    const int64_t *bnd_cell;                             /* [n_bnd]  */
    const double *bnd_len, *bnd_nx, *bnd_ny;             /* [n_bnd]  */
    const int32_t *bnd_series; /* [n_bnd] column of the BC series table used by the edge */
    const int64_t *int_L, *int_R;                        /* [n_int]  */
    const double *int_len, *int_nx, *int_ny, *int_xm, *int_ym; /* [n_int] */
} sw2dg_mesh;

/* Solver options: the swSp attributes imported by
 * swSp::importAttributesFromSolverParam (swSp.cpp:673-717). */
typedef struct sw2dg_params {
    int32_t scheme, topo, friction, head_loss, neg_depth_detection;
    double h_min, phi_min, cfl_max, dt_max, dt_map, dt_time_series, t0, t_max, v_max, fr_max;
} sw2dg_params;

/* Replaces swParent::create("sp") + swSp::init (swSp.cpp:720-969) for the
 * mesh/parameter part: builds the device tables.  device < 0 = current. */
int sw2dg_ctx_create(const sw2dg_mesh *mesh, const sw2dg_params *params, int32_t device,
                     sw2dg_ctx **out);
/* Replaces swSp::setBoussinesqParameter / setPorosityParameter /
 * setFrictionParameters / setSingularHeadLossParameters (swSp.cpp:998-1368).
 * NULL arrays mean "not used" (friction / head loss off). Angles in degrees. */
int sw2dg_set_cell_params(sw2dg_ctx *ctx, const double *phi, const double *beta,
                          const double *fric_c1, const double *fric_c2,
                          const double *fric_alpha_deg, const double *hl_hsing,
                          const double *hl_c1, const double *hl_c2, const double *hl_alpha_deg);
/* Replaces swParent::initBoundaryConditions (swParent.cpp:192-328) for the
 * time-series case: n_series columns sharing n_times rows; values are
 * column-major ([series][time]); type[] uses SW2DG_BC_*. */
int sw2dg_set_bc_series(sw2dg_ctx *ctx, int32_t n_series, int32_t n_times, const int32_t *type,
                        const double *t, const double *v);
/* Replaces swParent::setWindStationCodes and setWindVelocityTimeSeriesTables
 * (swParent.cpp:571-619) and switches on swSp::sourceWind for model_wind
 * smithbanke (swSpSourceWind.cpp:11-85, called from swSp.cpp:582-587): each
 * step interpolates the station velocities at t + dt/2 and adds the Smith &
 * Banke stress after friction and head loss.  code[n_cell] (reference order)
 * is each cell's 1-based station; t[n_times] non-decreasing record times;
 * vx, vy[n_station * n_times] the Cartesian wind velocity per station
 * ([station][record]; the polar-to-Cartesian step of swSourceWind.cpp:9-38
 * stays with the caller).  n_station = 0 turns the source off.  The 4D-Var
 * layer (sw2dg_assim_*) has no wind source and refuses a ctx that has one
 * (likewise for runoff and infiltration below). */
int sw2dg_set_wind(sw2dg_ctx *ctx, int32_t n_station, int32_t n_times, const int32_t *code,
                   const double *t, const double *vx, const double *vy);
/* Replaces swSp::setRunoffParameters (swSp.cpp:1276-1320) and
 * swParent::setPrecipitationStationCodes / setPrecipitationTimeSeriesTable
 * (swParent.cpp:516-569) and switches on the runoff step of swSp::execute
 * (swSp.cpp:561-573): sourceRunoffEverywhere when water_only = 0
 * (model_porosity_rainfall everywhere), sourceRunoffWateronly when 1
 * (swSpSourceRunoffInfiltration.cpp:43-263).  model is SW2DG_RUNOFF_*;
 * code[n_cell] (reference order) each cell's 1-based rain station; t[n_times]
 * non-decreasing; rate[n_station * n_times] the rain rate (m/s, negative =
 * evaporation) per station ([station][record]), interpolated at t + dt/2;
 * coeff[n_cell] the infiltration or runoff coefficient in [0, 1] (unused,
 * may be NULL, for SW2DG_RUNOFF_FULL).  SW2DG_RUNOFF_NONE turns it off. */
int sw2dg_set_runoff(sw2dg_ctx *ctx, int32_t model, int32_t water_only, int32_t n_station,
                     int32_t n_times, const int32_t *code, const double *t, const double *rate,
                     const double *coeff);
/* Replaces swSp::setInfiltrationParameters (swSp.cpp:1075-1115) and switches
 * on swSp::sourceInfiltration (swSpSourceRunoffInfiltration.cpp:12-40, after
 * the runoff step): model SW2DG_INFIL_HORTON with rate[n_cell] >= 0 (m/s,
 * reference order); SW2DG_INFIL_NONE turns it off. */
int sw2dg_set_infiltration(sw2dg_ctx *ctx, int32_t model, const double *rate);
/* Switches on swSp::parabComputeExplicit (swSpParab.cpp:10-96) for
 * model_momentum_diffusion velocity_independent (model 1; 0 off): after the
 * hyperbolic balance and clamp and before friction (swSp.cpp:535-545), qx
 * and qy diffuse with the edge coefficient of the harmonic mean of
 * 1e-6 h phi of its cells (the reference's fixed coefficient, swSp.h:244)
 * and the factor length / |centre distance along the normal|
 * (swSp.cpp:1236-1246), in n = int(dt / min(t_max - t_0, dt, area phi /
 * (2 D))) explicit substeps of dt / n.  A step with n > 1 (cells of about a
 * micron or less) ends the device time loop's launch; the library runs its
 * substeps and resumes, so the results do not depend on it.  n above 2^20
 * stops the run with SW2DG_ERR_ARG.  The reference divides q by h of both
 * cells, so a dry cell makes q NaN next to it and the next step stops with
 * SW2DG_ERR_NAN, where the reference stops (swSpHyp.cpp:391). */
int sw2dg_set_momentum_diffusion(sw2dg_ctx *ctx, int32_t model);
/* Replaces swSp::setInitialConditions (swSp.cpp:1119-1224): sets the state
 * and the clock (t, previous dt = 0, step count 0, BC cursors rewound). */
int sw2dg_set_state(sw2dg_ctx *ctx, const double *h, const double *qx, const double *qy,
                    double t);
int sw2dg_get_state(sw2dg_ctx *ctx, double *h, double *qx, double *qy);
/* Replaces the body of swSp::execute's time loop (swSp.cpp:522-628): runs
 * until t >= t_stop, t >= t_max, max_steps steps were taken, or (when
 * pause_on_store != 0) a map / time-series storage instant is reached
 * (swParent::checkStore, swParent.cpp:46-88).  *n_done = steps taken. */
int sw2dg_run(sw2dg_ctx *ctx, double t_stop, int64_t max_steps, int32_t pause_on_store,
              int64_t *n_done);
/* Clock: current time, last step dt, steps since set_state, and the store
 * flags raised by the last step (bit 0 map, bit 1 time series). */
int sw2dg_get_clock(sw2dg_ctx *ctx, double *t, double *dt, int64_t *steps, int32_t *store_flags);
/* Boundary mass fluxes F1 of the last step, reference order [n_bnd]
 * (feeds swSp::storeResultBoundaryFluxes, swSp.cpp:1401-1427). */
int sw2dg_get_bnd_f1(sw2dg_ctx *ctx, double *f1);
/* Average device duration (ms) of each step kernel over n_steps timed with
 * HIP events on the ctx stream.  names may be NULL.  k_max entries max. */
int sw2dg_profile_kernels(sw2dg_ctx *ctx, int64_t n_steps, int32_t k_max, double *avg_ms,
                          const char **names, int32_t *n_kernels);
/* ---- 4D-Var layer: the reference's differentiable ("_E") forward and its
 * adjoint (assim/swSpE_checkpoint.cpp).  Godunov reconstruction. ---- */

/* Observations.  Gauges: cost sum_g w_g (h[cell_g] - gauge_h[k*n_gauges+g])^2
 * (compute_cost_gauges_ptr, assim/swSpE_gauges.cpp:166-179).  Frames: cost
 * sum_i w_h (h - frame_h)^2 + w_q ((qx - frame_qx)^2 + (qy - frame_qy)^2)
 * (compute_cost_frame_ptr, assim/swSpE_ad.cpp:273-308); frame arrays are
 * [n_obs][n_cell] in reference order, w_h / w_q may be NULL (1).  Both
 * divided by n_obs when normalize != 0 (cost_normalize). */
typedef struct sw2dg_obs {
    int32_t n_obs;
    const double *t;            /* [n_obs] observation times, ascending */
    int32_t n_gauges;
    const int64_t *gauge_cell;  /* [n_gauges] cells (reference order)     */
    const double *gauge_w;      /* [n_gauges] weights or NULL (1)         */
    const double *gauge_h;      /* [n_obs * n_gauges] observed depths     */
    const double *frame_h, *frame_qx, *frame_qy, *w_h, *w_q;
    int32_t normalize;
} sw2dg_obs;

/* Allocates the adjoint buffers: up to max_steps steps, a state checkpoint
 * every `stride` steps (0 = ceil(sqrt(max_steps))), the window
 * [t_window_start, t_max] (segments starting before it are not reversed,
 * -1e300 = whole run) split into n_tiles tiles whose adjoint is reset at the
 * tile top (get_spatial_gradient_checkpointed, swSpE_checkpoint.cpp:785-953). */
int sw2dg_assim_setup(sw2dg_ctx *ctx, const sw2dg_obs *obs, int32_t stride, int32_t n_tiles,
                      double t_window_start, int64_t max_steps);
/* Dry run of the _E forward from the state given to sw2dg_set_state
 * (build_checkpoint_records, swSpE_checkpoint.cpp:97-455): fixes dt, the
 * per-series boundary values and observation triggers of every step and
 * keeps the checkpoints.  Controls in reference order; NULL = the ctx's
 * porosity / friction maps.  The run stops at t_max or after max_steps
 * steps (the window then ends at t_seq[n_steps] < t_max). */
int sw2dg_assim_record(sw2dg_ctx *ctx, const double *phi, const double *K1, const double *K2,
                       int64_t *n_steps);
/* The recorded schedule: t_seq [n+1], dt_seq [n], bc_seq [n*n_series], obs_trig [n] (-1 = none). */
int sw2dg_assim_get_record(sw2dg_ctx *ctx, double *t_seq, double *dt_seq, double *bc_seq,
                           int32_t *obs_trig);
/* Replace the recorded boundary values (hydrograph controls) [n*n_series]. */
int sw2dg_assim_set_bc_seq(sw2dg_ctx *ctx, const double *bc_seq);
/* Cost on the frozen schedule (replay_trajectory_cost, swSpE_checkpoint.cpp:1311-1433). */
int sw2dg_assim_cost(sw2dg_ctx *ctx, const double *phi, const double *K1, const double *K2,
                     double *J);
/* Dry run at the given controls, then the adjoint (replaces the Enzyme reverse
 * pass of get_spatial_gradient_checkpointed, swSpE_checkpoint.cpp:722-1281):
 * J, dJ/dphi, dJ/dK1, dJ/dK2 per cell (reference order), dJ/d(boundary value)
 * per step and series [n*n_series] (sized for max_steps), the number of
 * recorded steps.  Any output may be NULL. */
int sw2dg_assim_gradient(sw2dg_ctx *ctx, const double *phi, const double *K1, const double *K2,
                         double *J, double *dJ_dphi, double *dJ_dK1, double *dJ_dK2,
                         double *dJ_dbc, int64_t *n_steps);
/* Depths at the gauge cells sampled by the last dry run at each observation
 * step, [n_obs * n_gauges] (0 for observations the run did not reach): the
 * series a twin experiment observes.  No reference counterpart (the
 * reference reads observations from files, swSpE_ad.cpp:47-154). */
int sw2dg_assim_gauge_series(sw2dg_ctx *ctx, double *out);
/* Change of variable of the porosity control (phi_correct / dphi_dm /
 * update_phi_norm_bounds, swSpE.h:59-107, swSpE_core.cpp:229-266): with mode
 * != 0 the `phi` argument of record / cost / gradient is the control m and
 * phi = correct(m): 1 clamp to [0, 1.05], 2 logistic 1 / (1 + exp(-lambda m)),
 * 3 normalise (m - lo) / (hi - lo) with lo / hi = min / max of m over
 * free_cells (all cells if n_free = 0), taken at each dry run and frozen for
 * later costs; dJ_dphi is then dJ/dm.  mask_wall != 0 zeroes dJ/dm where
 * phi <= phi_min (mask_wall_gradient, swSpE_checkpoint.cpp:1206-1215).
 * Mode 0 (default): phi is used as given. */
int sw2dg_assim_set_phi_transform(sw2dg_ctx *ctx, int32_t mode, double lambda,
                                  const int64_t *free_cells, int64_t n_free, int32_t mask_wall);
/* The same change of variable on host arrays (no device needed): phi and
 * dphi/dm of n controls for the given bounds; either output may be NULL. */
int sw2dg_phi_transform_eval(int32_t mode, double lambda, double lo, double hi, int64_t n,
                             const double *m, double *phi, double *dphi);
/* Adjoint explosion guard of the reverse sweep (swSpE_checkpoint.cpp:
 * 1035-1117): after each segment, ||adjoint of h|| of its start state is
 * compared with the first non-zero one; above `ratio` times it, the gradient
 * returns to the last watermark (taken after milestone segments -- every
 * 50th, the first and last, those with observations -- or, tiled, before
 * each tile) and the sweep stops.  Later sw2dg_assim_cost calls then start
 * at the first kept segment, as the reference's replay does (:1355-1400),
 * and the gradient's J is that truncated cost.  Default 1e8 (the
 * reference's); ratio <= 0 turns the guard off. */
int sw2dg_assim_set_explosion_guard(sw2dg_ctx *ctx, double ratio);
/* Truncation of the last gradient: first kept segment and its first step
 * (both 0 when the gradient covers the whole trajectory; the window's first
 * segment when an assimilation window or a frozen segment starts later, as
 * the reference's adjoint_truncation_seg_ + 1). */
int sw2dg_assim_truncation(sw2dg_ctx *ctx, int64_t *first_kept_step, int64_t *first_kept_segment);
/* Frozen truncation window (the reference's forced_first_seg_,
 * swSpE_checkpoint.cpp:871-881, set by its L-BFGS-B driver,
 * swSpE_minimize.cpp:881-903): segment >= 0 makes every later gradient and
 * cost start at min(segment, K - 1) -- the reverse sweep stops there and the
 * cost keeps the observations from there on -- with the explosion guard's
 * dynamic truncation off, so that a line search compares costs over one
 * window; segment < 0 restores the dynamic truncation (default). */
int sw2dg_assim_set_first_segment(sw2dg_ctx *ctx, int64_t segment);
/* Thin-film cut of the reverse sweep: the adjoint of a cell's state after a
 * step is dropped where that state's depth is below h_cut (m), so no
 * sensitivity is carried back through films that are all but dry.  0 (the
 * default) keeps the exact adjoint of the _E scheme, i.e. the reference's
 * Enzyme gradient (swSpE_checkpoint.cpp:722-1281).  No reference
 * counterpart: over long windows with wet/dry fronts the exact adjoint
 * grows without bound through such films while the misfit itself varies
 * smoothly; the cut gradient follows finite differences of the misfit. */
int sw2dg_assim_set_adjoint_cut(sw2dg_ctx *ctx, double h_cut);
/* One sw2dg_assim_gradient at the given controls with an event pair around
 * every kernel launch on the ctx stream: average device duration (ms) and
 * launch count per kernel (names are static strings), up to k_max entries.
 * Counterpart of the reference's per-phase timers (swSp.cpp:652-660). */
int sw2dg_assim_profile(sw2dg_ctx *ctx, const double *phi, const double *K1, const double *K2,
                        int32_t k_max, double *avg_ms, int64_t *launches, const char **names,
                        int32_t *n_kernels);
/* Largest L2 norm of the adjoint of h at a segment start over the reverse
 * sweep of the last sw2dg_assim_gradient (0 before any): the dh_peak
 * diagnostic of get_spatial_gradient_checkpointed (swSpE_checkpoint.cpp:
 * 1044-1049, 1242), a column of the descent log. */
int sw2dg_assim_adjoint_peak(sw2dg_ctx *ctx, double *dh_peak);
/* Keep the full state (h, qx, qy) that each later dry run reaches at every
 * observation time (on != 0 holds n_obs x 3 x n_cell doubles on the device):
 * the maps the reference's _E forward dumps at each observation trigger
 * (storeReferenceMap, assim/swSpE_core.cpp:300-341, called from
 * build_checkpoint_records, swSpE_checkpoint.cpp:366-375) for gen-ref-e and
 * the forward mode.  The next sw2dg_assim_record fills them. */
int sw2dg_assim_keep_obs_states(sw2dg_ctx *ctx, int32_t on);
/* The kept states of the last dry run, [n_obs * n_cell] each (reference
 * order; any may be NULL), reached[n_obs] (may be NULL) 1 for the
 * observations the run reached (others are 0-filled). */
int sw2dg_assim_obs_states(sw2dg_ctx *ctx, double *h, double *qx, double *qy, int32_t *reached);
/* State at the end of the recorded run (reference order). */
int sw2dg_assim_final_state(sw2dg_ctx *ctx, double *h, double *qx, double *qy);
/* Resident trajectory (Godunov), mode 1 (default) or 0: the dry run of
 * sw2dg_assim_gradient keeps, for every segment it can give a device block
 * (72 bytes per cell and step; blocks allocated on first use while they take
 * at most 60 % of the memory free at the first one), each step's state,
 * balanced state and boundary fluxes, and the reverse sweep reads them in
 * place of replaying that segment from its checkpoint (the replay
 * recomputes the same values bit for bit).  The reference replays every
 * segment (ckpt_segment_E, swSpE_checkpoint.cpp:584-715); segments,
 * checkpoints, tiles and the explosion guard are unchanged.
 * sw2dg_assim_resident: the number of segments the last gradient read from
 * resident blocks (0 = every segment was replayed). */
int sw2dg_assim_set_resident(sw2dg_ctx *ctx, int32_t mode);
int sw2dg_assim_resident(sw2dg_ctx *ctx, int32_t *active);

/* ---- Batched ensemble: many differentiable-forward runs over uneven meshes
 * in one persistent launch.  The reference runs one simulation per process
 * (the _E dry run of build_checkpoint_records, swSpE_checkpoint.cpp:97-455,
 * one sw2d_enzyme <case> forward call each); run r of an ensemble gives the
 * same result as sw2dg_assim_record on its own ctx.  Godunov, no head loss. ---- */
typedef struct sw2dg_ens sw2dg_ens;
/* meshes[n_runs] in reference order; params shared by every run (t0, t_max,
 * dt_max, cfl_max, h_min, phi_min, friction). */
int sw2dg_ens_create(int32_t n_runs, const sw2dg_mesh *meshes, const sw2dg_params *params,
                     int32_t device, sw2dg_ens **out);
/* Controls, boundary series, initial state and gauges of one run (reference
 * order; may be called again to replace them). */
int sw2dg_ens_set_run(sw2dg_ens *ens, int32_t run, const double *phi, const double *K1,
                      const double *K2, const double *fric_alpha_deg, int32_t n_series,
                      int32_t n_times, const int32_t *type, const double *t, const double *v,
                      const double *h, const double *qx, const double *qy, int32_t n_gauges,
                      const int64_t *gauge_cell);
/* Runs every run from its initial state to t_max (or max_steps steps) with
 * workgroups of `block` threads (256, 512, 1024), landing on the observation
 * times obs_t; gauge_out [n_runs][n_obs][max n_gauges] (0 where not reached),
 * steps_out [n_runs], device_ms = duration of the launch.  Any output may be NULL. */
int sw2dg_ens_run(sw2dg_ens *ens, int32_t n_obs, const double *obs_t, int32_t block,
                  int64_t max_steps, double *gauge_out, int64_t *steps_out, double *device_ms);
/* Final state of one run after sw2dg_ens_run (reference order). */
int sw2dg_ens_get_state(sw2dg_ens *ens, int32_t run, double *h, double *qx, double *qy);
/* Upper bound on the workgroups one run may get when the runs fit the GPU
 * at once (cooperative launch, runs split into clusters; default 64).  1:
 * one workgroup per run always.  Results do not depend on it. */
int sw2dg_ens_set_cluster_limit(sw2dg_ens *ens, int32_t max_wg_per_run);
/* Shape of the last sw2dg_ens_run launch: workgroups launched and the
 * largest cluster.  Either output may be NULL. */
int sw2dg_ens_last_launch(sw2dg_ens *ens, int32_t *workgroups, int32_t *max_wg_per_run);
int64_t sw2dg_ens_device_bytes(sw2dg_ens *ens);
const char *sw2dg_ens_last_error(sw2dg_ens *ens);
void sw2dg_ens_destroy(sw2dg_ens *ens);

/* ---- Learned warm start: a dense MLP (GELU between layers) trained and run
 * with BF16 MFMA kernels, FP32 accumulation, FP32 master weights.  No
 * reference counterpart (the reference starts its minimiser from a
 * porosity_ebauche_map, input.solv; swSpE_minimize.cpp:267-1073). ---- */
typedef struct sw2dg_mlp sw2dg_mlp;
/* dims[n_layers + 1]: input width, hidden widths, output width. */
int sw2dg_mlp_create(int32_t n_layers, const int32_t *dims, int32_t max_batch, int32_t device,
                     sw2dg_mlp **out);
/* Parameters flattened layer by layer: W_l [dims[l] x dims[l+1]] row-major, then b_l. */
int64_t sw2dg_mlp_n_params(sw2dg_mlp *mlp);
int sw2dg_mlp_set_params(sw2dg_mlp *mlp, const float *params);  /* also resets Adam */
int sw2dg_mlp_get_params(sw2dg_mlp *mlp, float *params);
/* y [batch x dims[L]] = MLP(x [batch x dims[0]]) (inputs rounded to BF16). */
int sw2dg_mlp_forward(sw2dg_mlp *mlp, int32_t batch, const float *x, float *y);
/* sw2dg_mlp_forward without any allocation, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops).
 * No counterpart in the reference.  x [batch x dims[0]] and y [batch x
 * dims[L]] are DEVICE pointers on the mlp's device; the BF16 input and the
 * hidden activations live in a fixed 64 MiB device array of the library (one
 * per GPU, held until the process ends), so batch is bounded by that array,
 * not by max_batch: a batch that does not fit returns SW2DG_ERR_ARG with the
 * limit in sw2dg_mlp_last_error.  stream: a hipStream_t to enqueue on and
 * return without waiting, or NULL for the mlp's own stream, waited for.  The
 * output equals sw2dg_mlp_forward's bit for bit.  Every call on one GPU
 * shares the array: calls that overlap on a GPU (other threads, other
 * streams, other mlp objects) race. */
int sw2dg_mlp_forward_noalloc(sw2dg_mlp *mlp, int32_t batch, const float *x, float *y, void *stream);
/* Mean squared error against t and its gradient (flattened like the
 * parameters; may be NULL). */
int sw2dg_mlp_grad(sw2dg_mlp *mlp, int32_t batch, const float *x, const float *t, float *loss,
                   float *grads);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) on a batch; loss may be NULL. */
int sw2dg_mlp_train_step(sw2dg_mlp *mlp, int32_t batch, const float *x, const float *t, float lr,
                         float *loss);
/* A training set kept on the device: n rows of x [n x dims[0]] (stored in
 * BF16) and t [n x dims[L]]; replaces the previous one. */
int sw2dg_mlp_set_data(sw2dg_mlp *mlp, int64_t n, const float *x, const float *t);
/* n_steps Adam steps on the resident training set, step k on its rows
 * rows[k * batch .. (k + 1) * batch), with one wait at the end: the same
 * results as n_steps calls of sw2dg_mlp_train_step on those rows, bit for
 * bit.  step_loss [n_steps] (may be NULL): each step's loss. */
int sw2dg_mlp_train_epoch(sw2dg_mlp *mlp, int32_t batch, int64_t n_steps, const int32_t *rows, float lr,
                          float *step_loss);
const char *sw2dg_mlp_last_error(sw2dg_mlp *mlp);
void sw2dg_mlp_destroy(sw2dg_mlp *mlp);

/* The device's exp (fn 0), log (1) or cbrt (2) of n values: the functions
 * every kernel uses (exp and cbrt equal to glibc's, which the reference runs
 * on; log correctly rounded).  No reference counterpart: it lets a host build
 * of the same code check the device bit for bit.  device < 0 = current. */
int sw2dg_math_eval(int32_t fn, int64_t n, const double *x, double *y, int32_t device);

/* Bytes of device memory held by the ctx. */
int64_t sw2dg_device_bytes(sw2dg_ctx *ctx);
const char *sw2dg_last_error(sw2dg_ctx *ctx);
void sw2dg_ctx_destroy(sw2dg_ctx *ctx);
/* Library build information (no device call). */
const char *sw2dg_version(void);

#ifdef __cplusplus
}
#endif
#endif
