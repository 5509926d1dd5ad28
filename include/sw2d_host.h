/*
 * sw2d_host.h -- C-ABI of the host-side case tooling (libsw2d_host.so; the
 * same symbols are also exported by libsw2dg.so).  No GPU is touched here.
 *
 * These functions read and write the reference's case folders so the GPU
 * path consumes exactly the files the reference's sw2dSolver reads:
 *   geometry.sw2d v2  -- swMesh::readGeometryFileV2 (src/common/geometry/swMesh.cpp:213-471)
 *   parameter maps    -- swIO::readVectorParameterFile (src/common/swIO.cpp:1005-1191)
 *   BC time series    -- swIO::readBoundaryTimeSeriesFile (src/common/swIO.cpp:687-817)
 *   input.solv        -- swSolverParam::readInputFile (src/sw2dSolver/common/swSolverParam.cpp:664-880)
 *   ResultHydro maps  -- swIO::createTxtFileMapDecimal (src/common/swIO.cpp:182-223)
 * Errors: functions return 0 on success, nonzero otherwise; sw2dh_error()
 * gives the last message of the calling thread.
 */
#ifndef SW2D_HOST_H
#define SW2D_HOST_H
#include <stdint.h>

#include "sw2dg.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sw2dh_geom sw2dh_geom;
typedef struct sw2dh_case sw2dh_case;

const char *sw2dh_error(void);

/* ---- meshes ---- */
/* nx x ny rectangle of quads split along one diagonal (2*nx*ny triangles) */
int sw2dh_geom_rect(int32_t nx, int32_t ny, double Lx, double Ly, sw2dh_geom **out);
/* seeded jittered (px+1)x(py+1) lattice, Delaunay by Lawson flips */
int sw2dh_geom_delaunay(int32_t px, int32_t py, double Lx, double Ly, double jitter,
                        uint64_t seed, sw2dh_geom **out, int64_t *n_flips);
/* any polygon mesh from its nodes and cell->node lists: cell c has nodes
 * cell_nodes[cell_ptr[c] .. cell_ptr[c+1]) in order around it (either way);
 * edges, normals, areas and centres as the reference's geometry classes
 * derive them; every boundary edge gets code 1 (sw2dh_geom_set_bnd_codes) */
int sw2dh_geom_cells(int64_t n_node, const double *x, const double *y, const double *z, int64_t n_cell,
                     const int64_t *cell_ptr, const int64_t *cell_nodes, sw2dh_geom **out);
int sw2dh_geom_read(const char *path, sw2dh_geom **out);
int sw2dh_geom_write(const sw2dh_geom *g, const char *path);
int sw2dh_geom_set_node_z(sw2dh_geom *g, const double *z);
int sw2dh_geom_set_bnd_codes(sw2dh_geom *g, const int32_t *codes);
int sw2dh_geom_sizes(const sw2dh_geom *g, int64_t *n_cell, int64_t *n_bnd, int64_t *n_int,
                     int64_t *n_node);
/* dtype: 0 = float64, 1 = int64 (uint64 indices), 2 = int32.  Names:
 * cell_area cell_x cell_y cell_zb node_x node_y node_z bnd_cell bnd_len bnd_nx
 * bnd_ny bnd_xm bnd_ym bnd_code int_L int_R int_len int_nx int_ny int_xm int_ym */
const void *sw2dh_geom_array(const sw2dh_geom *g, const char *name, int64_t *n, int32_t *dtype);
void sw2dh_geom_free(sw2dh_geom *g);

/* Write a complete SP case folder: input/geometry.sw2d, input/input.solv (text
 * given verbatim), initial_conditions_map.txt (h, q, r), Boussinesq_map.txt,
 * SP_cell_porosity_map.txt, the friction map of friction_model ("none",
 * "manning", "strickler", "chezy"; c1 = c2 = fric, alpha = 0) and
 * hydro_boundary_time_series.txt.  phi / beta / fric may be NULL (uniform 1 /
 * 1 / none).  BC values are column-major [series][time]. */
int sw2dh_write_case(const char *folder, const sw2dh_geom *g, const char *solv_text,
                     const double *h0, const double *qx0, const double *qy0, const double *phi,
                     const double *beta, const char *friction_model, const double *fric,
                     int32_t n_series, const int32_t *codes, const char *types_ws,
                     int32_t n_times, const double *t, const double *v);

/* ---- case folders (what swSp::init loads, swSp.cpp:720-969) ---- */
int sw2dh_case_load(const char *folder, sw2dh_case **out);
int sw2dh_case_sizes(const sw2dh_case *c, int64_t *n_cell, int64_t *n_bnd, int64_t *n_int,
                     int32_t *n_series, int32_t *n_times);
int sw2dh_case_params(const sw2dh_case *c, sw2dg_params *out);
/* Geometry names as above plus: h0 qx0 qy0 phi beta fc1 fc2 falpha hsing hk1
 * hk2 halpha (empty when unused) bnd_series series_type series_t series_v
 * probe_cell probe_x probe_y (input/probe_locations.txt, read as
 * swSolverParam::readProbes, swSolverParam.cpp:958-1022; each probe's cell
 * located as swSp::init does, swSp.cpp:754-771) */
const void *sw2dh_case_array(const sw2dh_case *c, const char *name, int64_t *n, int32_t *dtype);
/* value of an input.solv key as text (native, assimilation or gpu_* keys), NULL if absent */
const char *sw2dh_case_key(const sw2dh_case *c, const char *key);
int sw2dh_case_n_decimal(const sw2dh_case *c);
/* ResultHydro map for a state (reference order) into path */
int sw2dh_write_result(const sw2dh_case *c, const char *path, const double *h, const double *qx,
                       const double *qy, int32_t n_decimal, const char *mode);
/* ResultHydro_p....txt file name for time t (swIO::computeFileNameFromTime) */
int sw2dh_result_name(double t, char *buf, int32_t buflen);
void sw2dh_case_free(sw2dh_case *c);
/* name of probe i of the case, NULL out of range */
const char *sw2dh_case_probe_name(const sw2dh_case *c, int32_t i);
/* Cell containing each point (x[i], y[i]), -1 outside the mesh:
 * swMesh::getCellFromCoordinates (src/common/geometry/swMesh.cpp:695-759),
 * the first cell in file order whose vertex angle sum exceeds 6 in
 * magnitude; any polygon (triangles, quads, mixed meshes).  Used for probes
 * and for locate_gauges (assim/swSpE_gauges.cpp:41-77). */
int sw2dh_geom_locate(const sw2dh_geom *g, int64_t n, const double *x, const double *y,
                      int64_t *cell);
int sw2dh_case_locate(const sw2dh_case *c, int64_t n, const double *x, const double *y,
                      int64_t *cell);

/* ---- box-constrained minimiser ----
 * Replaces LBFGSpp::LBFGSBSolver<double>::minimize as the reference's
 * run_minimization drives it (assim/swSpE_minimize.cpp:832-912, parameters
 * :857-862, call :873-898): L-BFGS-B (generalized Cauchy point, subspace
 * minimisation, More-Thuente line search).  The objective returns the value
 * in *f and the gradient in grad[n] at x; its return code is 0 to go on, > 0
 * to stop after this evaluation (the point is kept when it is the best), < 0
 * for an error.  A non-finite *f marks x as unusable (the line search steps
 * back).  lb / ub may be NULL (unbounded).  On return x holds the last
 * accepted iterate (the reference keeps the last iterate when its library
 * throws, :895-902); the status is 0 unless the arguments are invalid (1) or
 * the objective fails or is not finite at the start (2). */
typedef int (*sw2dh_objective)(const double *x, double *f, double *grad, void *user);
typedef struct sw2dh_lbfgsb_param {
    int32_t m;              /* correction pairs kept (lbfgsb_m)                      */
    double epsilon;         /* projected-gradient max-norm tolerance (lbfgsb_epsilon) */
    double epsilon_rel;     /* ... relative to ||x||_2                                */
    int32_t past;           /* objective-change test over `past` iterations (0 off)   */
    double delta;           /* ... relative tolerance                                 */
    int32_t max_iterations; /* 0 = no limit (max_it)                                  */
    int32_t max_linesearch; /* evaluations per line search (lbfgsb_max_linesearch)    */
    double ftol, wolfe;     /* sufficient decrease and curvature constants            */
    double max_step;        /* largest step along a direction                         */
    int32_t first_step;     /* first direction (no correction pairs yet): 0 = unit
                               length, LBFGSpp's rule; 1 = the full step to the
                               Cauchy point, the 1997 Fortran code's rule for boxed
                               problems                                            */
    int32_t max_submin;     /* passes of the box-constrained subspace minimisation  */
    int32_t restart;        /* after a failed line search: 0 = stop (LBFGSpp throws,
                               the reference keeps the last iterate); 1 = drop the
                               correction pairs and search again from the last
                               iterate along the Cauchy direction, the Fortran
                               L-BFGS-B 3.0 rule (stop if that search fails too) */
} sw2dh_lbfgsb_param;
#define SW2DH_LBFGSB_CONVERGED_GRAD 0
#define SW2DH_LBFGSB_CONVERGED_F 1
#define SW2DH_LBFGSB_MAX_ITER 2
#define SW2DH_LBFGSB_LINESEARCH_FAILED 3
#define SW2DH_LBFGSB_STOPPED 4
/* Library defaults: m 6, epsilon 1e-5, epsilon_rel 1e-5, past 1, delta
 * 1e-10, no iteration limit, 20 line-search evaluations, ftol 1e-4, wolfe
 * 0.9, max_step 1e20, first_step 0, max_submin 10, restart 0. */
void sw2dh_lbfgsb_default(sw2dh_lbfgsb_param *p);
int sw2dh_lbfgsb(int64_t n, double *x, const double *lb, const double *ub, sw2dh_objective fg,
                 void *user, const sw2dh_lbfgsb_param *param, double *fx, int32_t *n_iter,
                 int32_t *n_eval, double *proj_grad_norm, int32_t *stop_reason);
const char *sw2dh_lbfgsb_error(void);

#ifdef __cplusplus
}
#endif
#endif
