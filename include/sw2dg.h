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
