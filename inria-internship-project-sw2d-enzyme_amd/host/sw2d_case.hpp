// sw2d_case.hpp -- C++ host side of the MI355X SW2D solver: the reference's
// case-folder formats (geometry.sw2d v2, Unif/Distrib maps, boundary time
// series, input.solv, ResultHydro maps) and synthetic mesh generators.
// No HIP here: this part builds with g++ as well as hipcc.
#pragma once
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/sw2dg.h"

namespace sw2d {

// geometry.sw2d version "2" (writer src/common/geometry/swMesh.cpp:1221-1543,
// reader :194-471).  Values are used as stored (area, centre, normals...).
struct Geometry {
    // cells
    std::vector<uint64_t> cell_sms;
    std::vector<double> cell_area, cell_perim, cell_x, cell_y, cell_zb;
    std::vector<std::vector<uint64_t>> cell_nodes, cell_bnd, cell_int, cell_neigh;
    // nodes
    std::vector<uint64_t> node_sms;
    std::vector<double> node_x, node_y, node_z;
    std::vector<std::vector<uint64_t>> node_cells;
    std::vector<uint64_t> node_nedge;
    // boundary edges
    std::vector<uint64_t> bnd_n1, bnd_n2, bnd_cell;
    std::vector<double> bnd_xm, bnd_ym, bnd_zm, bnd_dnorm, bnd_len, bnd_nx, bnd_ny;
    std::vector<int32_t> bnd_code;
    // internal edges
    std::vector<uint64_t> int_n1, int_n2, int_L, int_R;
    std::vector<double> int_xm, int_ym, int_zm, int_dnorm, int_len, int_nx, int_ny;
    // boundaries: edge lists per boundary and their codes
    std::vector<std::vector<uint64_t>> boundaries;
    std::vector<int32_t> boundary_codes;

    size_t n_cell() const { return cell_area.size(); }
    size_t n_bnd() const { return bnd_len.size(); }
    size_t n_int() const { return int_len.size(); }
    size_t n_node() const { return node_x.size(); }
};

bool read_geometry(const std::string &path, Geometry &g, std::string &err);
bool write_geometry(const std::string &path, const Geometry &g, std::string &err);

// Build every derived table of a triangle/polygon mesh from node coordinates
// and cell->node lists (counter-clockwise or not; normals are oriented as the
// reference does: internal L->R, boundary towards the cell).  bc_code(x, y)
// gives the code of a boundary edge from its midpoint.
using BcCodeFn = int32_t (*)(double xm, double ym, void *user);
void build_geometry(const std::vector<double> &x, const std::vector<double> &y,
                    const std::vector<double> &z, const std::vector<std::vector<uint64_t>> &cells,
                    BcCodeFn bc_code, void *user, Geometry &g);

// Re-assign boundary codes per edge and rebuild the boundary lists.
void set_bnd_codes(Geometry &g, const std::vector<int32_t> &codes);
// Replace node elevations and recompute cell zb (vertex mean) and edge zm.
void set_node_z(Geometry &g, const std::vector<double> &z);

// Unif/Distrib parameter map (src/common/swIO.cpp:1005-1191)
bool read_map(const std::string &path, size_t n_params, size_t n_records,
              std::vector<std::string> &names, std::vector<std::vector<double>> &cols,
              std::string &err);
bool write_map_uniform(const std::string &path, const std::vector<std::string> &names,
                       const std::vector<double> &vals, std::string &err);
bool write_map_distrib(const std::string &path, const std::vector<std::string> &names,
                       const std::vector<double> &defaults,
                       const std::vector<std::vector<double>> &cols, std::string &err);

// boundary time series (src/common/swIO.cpp:687-817)
struct BcSeries {
    std::vector<int32_t> codes;
    std::vector<std::string> types;
    std::vector<double> t;
    std::vector<std::vector<double>> v;  // [series][time]
};
bool read_bc_series(const std::string &path, BcSeries &bc, std::string &err);
bool write_bc_series(const std::string &path, const BcSeries &bc, std::string &err);

// wind time series (swIO::readTimeSeries2dVectorFile, src/common/swIO.cpp:906-1010):
// two comment lines, the station count, two comment lines, then one row per
// record: time, then (norm, angle from North in degrees) per station.
// scalar time series (swIO::readTimeSeriesScalarFile, src/common/swIO.cpp:820-904):
// two comment lines, the series count, two comment lines, then one row per
// record: time, then one value per series.
bool read_scalar_series(const std::string &path, std::vector<double> &t,
                        std::vector<std::vector<double>> &v, std::string &err);
bool read_wind_series(const std::string &path, std::vector<double> &t,
                      std::vector<std::vector<double>> &norm,
                      std::vector<std::vector<double>> &angle_deg, std::string &err);

// input.solv: key/value, '#' or '=' comment lines (swSolverParam.cpp:664-880).
// Keys of the 4D-Var layer (assim/assimConfig.cpp:27-39) and "gpu_*" keys are
// kept aside, with '#' tail comments removed as the assimilation layer does.
struct SolverConfig {
    std::map<std::string, std::string> native, assim, gpu;
    // typed values with the reference defaults (swSolverParam.h:137-194)
    std::string model_hyp = "swes", model_friction = "none", model_singular_head_loss = "none";
    std::string num_scheme = "godunov", topo = "discontinuous", store_flag = "default";
    std::string negative_depth_detection = "none", div_corr = "none";
    std::string model_runoff = "none", model_infiltration = "none", model_wind = "none";
    std::string model_momentum_diffusion = "none", model_porosity_rainfall = "water_only";
    double cfl_max = 1, dt_map = 100, dt_max = 10, dt_time_series = 100, fr_max = 100;
    double h_min = 1e-5, phi_min = 1e-16, t0 = 0, t_max = 100, v_max = 100;
    int n_decimal = 10;
};
bool read_solver_config(const std::string &path, SolverConfig &cfg, std::string &err);

// input/probe_locations.txt (swSolverParam::readProbes, swSolverParam.cpp:
// 958-1022; file name swFile.h:45): a header line, then "x y name" lines up to
// the first empty line.  An absent file means no probes.
struct Probes {
    std::vector<std::string> names;
    std::vector<double> x, y;
};
bool read_probes(const std::string &path, Probes &p, std::string &err);
// Index of the first cell (file order) whose vertex polygon winds around
// (x, y), -1 if none: swMesh::getCellFromCoordinates (swMesh.cpp:695-759),
// the angle sum over the cell's vertices with |sum| > 6 meaning inside.
int64_t locate_cell(const Geometry &g, double x, double y);

// A loaded case folder (what swSp::init assembles, swSp.cpp:720-969).
struct Case {
    std::string folder;
    Geometry g;
    SolverConfig cfg;
    BcSeries bc;
    std::vector<int32_t> bnd_series;  // boundary edge -> series column
    std::vector<double> h, qx, qy;    // initial state
    std::vector<double> phi, beta;
    std::vector<double> fc1, fc2, falpha;        // friction (degrees)
    std::vector<double> hsing, hk1, hk2, halpha; // singular head loss (degrees)
    // wind (model_wind smithbanke): station code per cell (1-based), record
    // times, Cartesian components per station ([station][record] row-major)
    std::vector<int32_t> wind_code;
    std::vector<double> wind_t, wind_x, wind_y;
    int32_t wind_stations = 0;
    // runoff (model_runoff; SW2DG_RUNOFF_*) and infiltration (model_infiltration;
    // SW2DG_INFIL_*): rain station code per cell (1-based), record times, rate
    // per station ([station][record], m/s), runoff / infiltration coefficient
    // and Horton rate per cell
    int32_t runoff_model = 0, rain_water_only = 1, infil_model = 0, rain_stations = 0;
    int32_t momentum_diffusion = 0;  // model_momentum_diffusion other than none
    std::vector<int32_t> rain_code;
    std::vector<double> rain_t, rain_v, runoff_coeff, infil_rate;
    // probes (swSp::init, swSp.cpp:754-783): names, coordinates and the cell
    // of each (a probe outside the mesh stops the load, as the reference)
    Probes probes;
    std::vector<int64_t> probe_cell;
    sw2dg_params params() const;
    sw2dg_mesh mesh() const;                     // views into g (int64 copies kept here)
    std::vector<int64_t> bcell64, L64, R64;
    std::vector<int32_t> series_types() const;   // SW2DG_BC_* per series
    std::vector<double> series_values() const;   // column-major
};
bool load_case(const std::string &folder, Case &c, std::string &err);

// output/ProbesTimeSeries.txt (swFile.h:100): header of swSp::init
// (swSp.cpp:754-783) and one line per storage instant of
// swSp::storeResultProbes (swSp.cpp:1611-1639): time, then zb, h, q, r of
// each probe's cell, %.10e, tab separated (a tab after every value).
std::string probe_header(const Case &c);
std::string probe_line(const Case &c, double t, const std::vector<double> &h,
                       const std::vector<double> &qx, const std::vector<double> &qy);

// ResultHydro_pDDDDdHHhMMmSSsMMM.txt naming (swIO.cpp:51-86) and writer
// (swIO::createTxtFileMapDecimal, swIO.cpp:182-223; columns of
// swSp::storeResultMaps, swSp.cpp:1515-1580).
std::string result_file_name(double t);
bool write_result_map(const std::string &path, const Case &c, const std::vector<double> &h,
                      const std::vector<double> &qx, const std::vector<double> &qy,
                      int n_decimal, const std::string &mode, std::string &err);

// ---- synthetic cases ----
// Structured rectangle nx x ny quads, each split along the same diagonal.
void make_rect_tri_mesh(int nx, int ny, double Lx, double Ly, const std::vector<double> *node_z,
                        Geometry &g);
// Seeded jittered lattice (jitter fraction of the spacing) triangulated and
// made Delaunay by edge flips.  z(x, y) given per node by zfun.
using ZFn = double (*)(double x, double y, void *user);
void make_jittered_delaunay(int px, int py, double Lx, double Ly, double jitter, uint64_t seed,
                            ZFn zfun, void *zuser, Geometry &g, int64_t *n_flips);

}  // namespace sw2d
