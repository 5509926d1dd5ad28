/*
 * sp_oracle.h -- CPU oracle for the SW2D single-porosity (SP) forward step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * CPU solver's algorithm, used as the checker for the HIP implementation.
 * Only tests, the smoke check and the CPU-baseline leg of the benchmark may
 * load it.  It is never part of the shipped GPU path.
 *
 * Semantics follow the reference's native solver ("N" semantics):
 *   src/sw2dSolver/models/sp/swSp.cpp          (time loop, edge factors,
 *                                               correctVelocity, lambdas)
 *   src/sw2dSolver/models/sp/swSpHyp.cpp       (fluxes, HLLC, CFL, balance)
 *   src/sw2dSolver/models/sp/swSpHypGodunov.cpp, swSpHypMusclEvr.cpp
 *   src/sw2dSolver/models/sp/swSpSourceFriction.cpp,
 *   src/sw2dSolver/models/common/swSourceFriction.cpp
 *   src/sw2dSolver/models/common/swBoundaryConditionTimeSeries.cpp
 *   src/sw2dSolver/models/sp/swSpSourceWind.cpp, swParent.cpp (wind series)
 *   src/sw2dSolver/models/sp/swSpSourceRunoffInfiltration.cpp, swSpParab.cpp
 * All loops run in the reference's edge order with its scatter updates, so
 * the summation order is the reference's by construction.
 */
#ifndef SP_ORACLE_H
#define SP_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* numeric scheme / options (values shared with include/sw2dg.h) */
enum { ORC_GODUNOV = 0, ORC_MUSCL_EVR = 1 };
enum { ORC_TOPO_DISCONTINUOUS = 0, ORC_TOPO_BLENDING = 1 };
enum { ORC_FRIC_NONE = 0, ORC_FRIC_MANNING = 1, ORC_FRIC_STRICKLER = 2, ORC_FRIC_CHEZY = 3 };
enum { ORC_BC_FR = 0, ORC_BC_H = 1, ORC_BC_Z = 2, ORC_BC_Q = 3 };

/* error codes */
enum {
    ORC_OK = 0,
    ORC_ERR_ARG = 1,
    ORC_ERR_NAN = 2,          /* NaN mass flux on an internal edge   */
    ORC_ERR_NEG_DEPTH = 3,    /* negative depth with detection on    */
    ORC_ERR_FLOW_CONFIG = 4,  /* impossible HLLC flow configuration  */
    ORC_ERR_BC = 5            /* invalid boundary value / type       */
};

typedef struct orc_case {
    int64_t n_cell, n_bnd, n_int;
    /* cells [n_cell] */
    const double *cell_area, *cell_x, *cell_y, *cell_zb;
    const double *phi, *beta;
    const double *fric_c1, *fric_c2, *fric_alpha_deg;      /* may be NULL if no friction  */
    const double *hl_hsing, *hl_c1, *hl_c2, *hl_alpha_deg;  /* may be NULL if no head loss */
    /* boundary edges [n_bnd] */
    const int64_t *bnd_cell;
    const double *bnd_len, *bnd_nx, *bnd_ny;
    const int32_t *bnd_series;   /* index into the BC series table */
    /* internal edges [n_int] */
    const int64_t *int_L, *int_R;
    const double *int_len, *int_nx, *int_ny, *int_xm, *int_ym;
    /* boundary time series: n_series columns sharing n_times rows */
    int32_t n_series, n_times;
    const int32_t *series_type;  /* ORC_BC_* */
    const double *series_t;      /* [n_times] */
    const double *series_v;      /* [n_series * n_times], column-major per series */
    /* options */
    int32_t scheme, topo, friction, head_loss, neg_depth_detection;
    double h_min, phi_min, cfl_max, dt_max, dt_map, t0, t_max, v_max, fr_max;
    /* wind, model_wind smithbanke (n_wind_station = 0: off) */
    int32_t n_wind_station, n_wind_times;
    const int32_t *wind_code;        /* [n_cell] 1-based station of each cell */
    const double *wind_t;            /* [n_wind_times] */
    const double *wind_x, *wind_y;   /* [n_wind_station * n_wind_times], per station */
    /* runoff (model_runoff: 0 none, 1 full_rainfall, 2 infiltration_coefficient,
     * 3 runoff_coefficient; rain_water_only = model_porosity_rainfall water_only)
     * and Horton infiltration (infil_model 1) */
    int32_t runoff_model, rain_water_only, n_rain_station, n_rain_times;
    const int32_t *rain_code;        /* [n_cell] 1-based station of each cell */
    const double *rain_t;            /* [n_rain_times] */
    const double *rain_v;            /* [n_rain_station * n_rain_times] rate (m/s) */
    const double *runoff_coeff;      /* [n_cell], models 2 and 3 */
    int32_t infil_model;
    const double *infil_rate;        /* [n_cell] */
    int32_t momentum_diffusion;      /* model_momentum_diffusion other than none */
} orc_case;

typedef struct orc_solver orc_solver;

orc_solver *orc_create(const orc_case *c, const double *h0, const double *qx0, const double *qy0);
void orc_destroy(orc_solver *s);
/* One reference time step (swSp::execute loop body).  Returns ORC_* status. */
int orc_step(orc_solver *s);
/* Run while t < t_end (and t < t_max), at most max_steps steps.  *n_done gets the count. */
int orc_run(orc_solver *s, double t_end, int64_t max_steps, int64_t *n_done);
/* threads of the OpenMP build (liboracle_omp.so); 1 in the serial ones */
int orc_threads(void);
void orc_get_state(const orc_solver *s, double *h, double *qx, double *qy);
void orc_set_state(orc_solver *s, const double *h, const double *qx, const double *qy);
double orc_time(const orc_solver *s);
double orc_last_dt(const orc_solver *s);
/* boundary mass flux F1 per boundary edge from the last step */
void orc_get_bnd_f1(const orc_solver *s, double *f1);

/* ---- per-function known-answer entry points (scalar kernels) ---- */
void orc_hllc(int blending, double h_min, double phi_min,
              double bL, double bR, double phiL, double phiR, double zbL, double zbR,
              double zL, double qL, double rL, double zR, double qR, double rR,
              double lmL, double lmR, double lpL, double lpR,
              double out[5], int *status);
void orc_bnd_flux(double h_min, double bR, double phiR, double zbR, double zR, double qR, double rR,
                  double lpR, int bc_type, double bc_val, double out[4], int *status);
double orc_friction_factor(int model, double coeff, double h, double u_norm, double dt);
double orc_interp_series(const double *t, const double *v, int64_t n, double at);

#ifdef __cplusplus
}
#endif
#endif
