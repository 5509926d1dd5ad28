// host_capi.cpp -- C-ABI over the case tooling (include/sw2d_host.h).
#include <cstring>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <string>

#include "../../include/sw2d_host.h"
#include "sw2d_case.hpp"

struct sw2dh_geom {
    sw2d::Geometry g;
    std::vector<int64_t> bcell, L, R, tri;  // tri: 3 node indices per cell (triangle meshes)
    void refresh() {
        bcell.assign(g.bnd_cell.begin(), g.bnd_cell.end());
        L.assign(g.int_L.begin(), g.int_L.end());
        R.assign(g.int_R.begin(), g.int_R.end());
        tri.clear();
        for (const auto &cn : g.cell_nodes) {
            if (cn.size() != 3) {
                tri.clear();
                break;
            }
            tri.insert(tri.end(), cn.begin(), cn.end());
        }
    }
};

struct sw2dh_case {
    sw2d::Case c;
    std::vector<int32_t> stype;
    std::vector<int32_t> models;  // runoff, rain on water only, infiltration, momentum diffusion
    std::vector<double> sval;
};

namespace {
thread_local std::string g_err;
int fail(const std::string &m) {
    g_err = m;
    return 1;
}

template <class V>
const void *view(const V &v, int64_t *n, int32_t *dtype, int32_t dt) {
    if (n) *n = (int64_t)v.size();
    if (dtype) *dtype = dt;
    return v.data();
}

const void *geom_array(const sw2d::Geometry &g, const std::vector<int64_t> &bcell,
                       const std::vector<int64_t> &L, const std::vector<int64_t> &R,
                       const std::string &k, int64_t *n, int32_t *dt) {
    if (k == "cell_area") return view(g.cell_area, n, dt, 0);
    if (k == "cell_x") return view(g.cell_x, n, dt, 0);
    if (k == "cell_y") return view(g.cell_y, n, dt, 0);
    if (k == "cell_zb") return view(g.cell_zb, n, dt, 0);
    if (k == "node_x") return view(g.node_x, n, dt, 0);
    if (k == "node_y") return view(g.node_y, n, dt, 0);
    if (k == "node_z") return view(g.node_z, n, dt, 0);
    if (k == "bnd_cell") return view(bcell, n, dt, 1);
    if (k == "bnd_len") return view(g.bnd_len, n, dt, 0);
    if (k == "bnd_nx") return view(g.bnd_nx, n, dt, 0);
    if (k == "bnd_ny") return view(g.bnd_ny, n, dt, 0);
    if (k == "bnd_xm") return view(g.bnd_xm, n, dt, 0);
    if (k == "bnd_ym") return view(g.bnd_ym, n, dt, 0);
    if (k == "bnd_code") return view(g.bnd_code, n, dt, 2);
    if (k == "int_L") return view(L, n, dt, 1);
    if (k == "int_R") return view(R, n, dt, 1);
    if (k == "int_len") return view(g.int_len, n, dt, 0);
    if (k == "int_nx") return view(g.int_nx, n, dt, 0);
    if (k == "int_ny") return view(g.int_ny, n, dt, 0);
    if (k == "int_xm") return view(g.int_xm, n, dt, 0);
    if (k == "int_ym") return view(g.int_ym, n, dt, 0);
    return nullptr;
}
}  // namespace

extern "C" {

const char *sw2dh_error(void) { return g_err.c_str(); }

int sw2dh_geom_rect(int32_t nx, int32_t ny, double Lx, double Ly, sw2dh_geom **out) {
    if (!out || nx <= 0 || ny <= 0 || !(Lx > 0) || !(Ly > 0)) return fail("bad arguments");
    auto *h = new sw2dh_geom();
    sw2d::make_rect_tri_mesh(nx, ny, Lx, Ly, nullptr, h->g);
    h->refresh();
    *out = h;
    return 0;
}

int sw2dh_geom_delaunay(int32_t px, int32_t py, double Lx, double Ly, double jitter,
                        uint64_t seed, sw2dh_geom **out, int64_t *n_flips) {
    if (!out || px <= 0 || py <= 0 || !(Lx > 0) || !(Ly > 0) || jitter < 0 || jitter > 0.3)
        return fail("bad arguments (jitter must be in [0, 0.3])");
    auto *h = new sw2dh_geom();
    sw2d::make_jittered_delaunay(px, py, Lx, Ly, jitter, seed, nullptr, nullptr, h->g, n_flips);
    h->refresh();
    *out = h;
    return 0;
}

int sw2dh_geom_cells(int64_t n_node, const double *x, const double *y, const double *z, int64_t n_cell,
                     const int64_t *cell_ptr, const int64_t *cell_nodes, sw2dh_geom **out) {
    if (!out || n_node < 3 || n_cell < 1 || !x || !y || !z || !cell_ptr || !cell_nodes)
        return fail("bad arguments");
    if (cell_ptr[0] != 0) return fail("cell_ptr[0] must be 0");
    std::vector<std::vector<uint64_t>> cells((size_t)n_cell);
    for (int64_t c = 0; c < n_cell; ++c) {
        const int64_t b = cell_ptr[c], e = cell_ptr[c + 1];
        if (e - b < 3) return fail("cell " + std::to_string(c) + " has fewer than 3 nodes");
        for (int64_t k = b; k < e; ++k) {
            if (cell_nodes[k] < 0 || cell_nodes[k] >= n_node)
                return fail("cell " + std::to_string(c) + " names a node outside [0, n_node)");
            cells[(size_t)c].push_back((uint64_t)cell_nodes[k]);
        }
    }
    auto *h = new sw2dh_geom();
    sw2d::build_geometry(std::vector<double>(x, x + n_node), std::vector<double>(y, y + n_node),
                         std::vector<double>(z, z + n_node), cells, nullptr, nullptr, h->g);
    h->refresh();
    *out = h;
    return 0;
}

int sw2dh_geom_read(const char *path, sw2dh_geom **out) {
    if (!out || !path) return fail("bad arguments");
    auto *h = new sw2dh_geom();
    std::string err;
    if (!sw2d::read_geometry(path, h->g, err)) {
        delete h;
        return fail(err);
    }
    h->refresh();
    *out = h;
    return 0;
}

int sw2dh_geom_write(const sw2dh_geom *g, const char *path) {
    if (!g || !path) return fail("bad arguments");
    std::string err;
    return sw2d::write_geometry(path, g->g, err) ? 0 : fail(err);
}

int sw2dh_geom_set_node_z(sw2dh_geom *g, const double *z) {
    if (!g || !z) return fail("bad arguments");
    sw2d::set_node_z(g->g, std::vector<double>(z, z + g->g.n_node()));
    return 0;
}

int sw2dh_geom_set_bnd_codes(sw2dh_geom *g, const int32_t *codes) {
    if (!g || (!codes && g->g.n_bnd())) return fail("bad arguments");
    sw2d::set_bnd_codes(g->g, std::vector<int32_t>(codes, codes + g->g.n_bnd()));
    return 0;
}

int sw2dh_geom_sizes(const sw2dh_geom *g, int64_t *nc, int64_t *nb, int64_t *ni, int64_t *nn) {
    if (!g) return fail("bad arguments");
    if (nc) *nc = (int64_t)g->g.n_cell();
    if (nb) *nb = (int64_t)g->g.n_bnd();
    if (ni) *ni = (int64_t)g->g.n_int();
    if (nn) *nn = (int64_t)g->g.n_node();
    return 0;
}

const void *sw2dh_geom_array(const sw2dh_geom *g, const char *name, int64_t *n, int32_t *dtype) {
    if (!g || !name) return nullptr;
    if (std::string(name) == "cell_tri") return view(g->tri, n, dtype, 1);
    return geom_array(g->g, g->bcell, g->L, g->R, name, n, dtype);
}

void sw2dh_geom_free(sw2dh_geom *g) { delete g; }

int sw2dh_write_case(const char *folder, const sw2dh_geom *g, const char *solv_text,
                     const double *h0, const double *qx0, const double *qy0, const double *phi,
                     const double *beta, const char *friction_model, const double *fric,
                     int32_t n_series, const int32_t *codes, const char *types_ws,
                     int32_t n_times, const double *t, const double *v) {
    if (!folder || !g || !solv_text || !h0 || !qx0 || !qy0 || n_series < 0 || n_times < 1 ||
        (n_series > 0 && (!codes || !types_ws || !t || !v)))
        return fail("bad arguments");
    namespace fs = std::filesystem;
    const std::string in = (fs::path(folder) / "input").string();
    std::error_code ec;
    fs::create_directories(in, ec);
    if (ec) return fail("cannot create " + in);
    const std::string p = in + "/";
    std::string err;
    if (!sw2d::write_geometry(p + "geometry.sw2d", g->g, err)) return fail(err);
    {
        std::ofstream f(p + "input.solv");
        f << solv_text;
        if (!f) return fail("cannot write input.solv");
    }
    const size_t nc = g->g.n_cell();
    std::vector<double> H(h0, h0 + nc), QX(qx0, qx0 + nc), QY(qy0, qy0 + nc);
    if (!sw2d::write_map_distrib(p + "initial_conditions_map.txt", {"h", "q", "r"}, {0, 0, 0},
                                 {H, QX, QY}, err))
        return fail(err);
    if (beta) {
        if (!sw2d::write_map_distrib(p + "Boussinesq_map.txt", {"beta"}, {1.0},
                                     {std::vector<double>(beta, beta + nc)}, err))
            return fail(err);
    } else if (!sw2d::write_map_uniform(p + "Boussinesq_map.txt", {"beta"}, {1.0}, err)) {
        return fail(err);
    }
    if (phi) {
        if (!sw2d::write_map_distrib(p + "SP_cell_porosity_map.txt", {"phi"}, {1.0},
                                     {std::vector<double>(phi, phi + nc)}, err))
            return fail(err);
    } else if (!sw2d::write_map_uniform(p + "SP_cell_porosity_map.txt", {"phi"}, {1.0}, err)) {
        return fail(err);
    }
    const std::string fm = friction_model ? friction_model : "none";
    if (fm != "none") {
        if (!fric) return fail("friction values missing");
        const std::string file = fm == "manning"     ? "friction_manning_map.txt"
                                 : fm == "strickler" ? "friction_strickler_map.txt"
                                 : fm == "chezy"     ? "friction_chezy_map.txt"
                                                     : "";
        if (file.empty()) return fail("unknown friction model " + fm);
        std::vector<double> f1(fric, fric + nc), zero(nc, 0.0);
        if (!sw2d::write_map_distrib(p + file, {"c1", "c2", "alpha"}, {0, 0, 0}, {f1, f1, zero},
                                     err))
            return fail(err);
    }
    sw2d::BcSeries bc;
    std::istringstream ts(types_ws ? types_ws : "");
    for (int k = 0; k < n_series; ++k) {
        bc.codes.push_back(codes[k]);
        std::string ty;
        ts >> ty;
        if (ty.empty()) return fail("missing BC type");
        bc.types.push_back(ty);
        bc.v.emplace_back(v + (size_t)k * n_times, v + (size_t)(k + 1) * n_times);
    }
    bc.t.assign(t, t + n_times);
    if (!sw2d::write_bc_series(p + "hydro_boundary_time_series.txt", bc, err)) return fail(err);
    return 0;
}

int sw2dh_case_load(const char *folder, sw2dh_case **out) {
    if (!folder || !out) return fail("bad arguments");
    auto *h = new sw2dh_case();
    std::string err;
    if (!sw2d::load_case(folder, h->c, err)) {
        delete h;
        return fail(err);
    }
    h->stype = h->c.series_types();
    h->sval = h->c.series_values();
    h->models = {h->c.runoff_model, h->c.rain_water_only, h->c.infil_model, h->c.momentum_diffusion};
    *out = h;
    return 0;
}

int sw2dh_case_sizes(const sw2dh_case *c, int64_t *nc, int64_t *nb, int64_t *ni, int32_t *ns,
                     int32_t *nt) {
    if (!c) return fail("bad arguments");
    if (nc) *nc = (int64_t)c->c.g.n_cell();
    if (nb) *nb = (int64_t)c->c.g.n_bnd();
    if (ni) *ni = (int64_t)c->c.g.n_int();
    if (ns) *ns = (int32_t)c->c.bc.codes.size();
    if (nt) *nt = (int32_t)c->c.bc.t.size();
    return 0;
}

int sw2dh_case_params(const sw2dh_case *c, sw2dg_params *out) {
    if (!c || !out) return fail("bad arguments");
    *out = c->c.params();
    return 0;
}

const void *sw2dh_case_array(const sw2dh_case *h, const char *name, int64_t *n, int32_t *dt) {
    if (!h || !name) return nullptr;
    const sw2d::Case &c = h->c;
    const std::string k = name;
    if (k == "h0") return view(c.h, n, dt, 0);
    if (k == "qx0") return view(c.qx, n, dt, 0);
    if (k == "qy0") return view(c.qy, n, dt, 0);
    if (k == "phi") return view(c.phi, n, dt, 0);
    if (k == "beta") return view(c.beta, n, dt, 0);
    if (k == "fc1") return view(c.fc1, n, dt, 0);
    if (k == "fc2") return view(c.fc2, n, dt, 0);
    if (k == "falpha") return view(c.falpha, n, dt, 0);
    if (k == "hsing") return view(c.hsing, n, dt, 0);
    if (k == "hk1") return view(c.hk1, n, dt, 0);
    if (k == "hk2") return view(c.hk2, n, dt, 0);
    if (k == "halpha") return view(c.halpha, n, dt, 0);
    if (k == "bnd_series") return view(c.bnd_series, n, dt, 2);
    if (k == "series_type") return view(h->stype, n, dt, 2);
    if (k == "series_t") return view(c.bc.t, n, dt, 0);
    if (k == "series_v") return view(h->sval, n, dt, 0);
    if (k == "wind_code") return view(c.wind_code, n, dt, 2);
    if (k == "wind_t") return view(c.wind_t, n, dt, 0);
    if (k == "wind_x") return view(c.wind_x, n, dt, 0);
    if (k == "wind_y") return view(c.wind_y, n, dt, 0);
    if (k == "rain_code") return view(c.rain_code, n, dt, 2);
    if (k == "rain_t") return view(c.rain_t, n, dt, 0);
    if (k == "rain_v") return view(c.rain_v, n, dt, 0);
    if (k == "runoff_coeff") return view(c.runoff_coeff, n, dt, 0);
    if (k == "infil_rate") return view(c.infil_rate, n, dt, 0);
    if (k == "source_models") return view(h->models, n, dt, 2);
    if (k == "probe_cell") return view(c.probe_cell, n, dt, 1);
    if (k == "probe_x") return view(c.probes.x, n, dt, 0);
    if (k == "probe_y") return view(c.probes.y, n, dt, 0);
    return geom_array(c.g, c.bcell64, c.L64, c.R64, k, n, dt);
}

const char *sw2dh_case_key(const sw2dh_case *h, const char *key) {
    if (!h || !key) return nullptr;
    for (const auto *m : {&h->c.cfg.native, &h->c.cfg.assim, &h->c.cfg.gpu}) {
        auto it = m->find(key);
        if (it != m->end()) return it->second.c_str();
    }
    return nullptr;
}

int sw2dh_case_n_decimal(const sw2dh_case *h) { return h ? h->c.cfg.n_decimal : 10; }

int sw2dh_write_result(const sw2dh_case *h, const char *path, const double *hh, const double *qx,
                       const double *qy, int32_t n_decimal, const char *mode) {
    if (!h || !path || !hh || !qx || !qy) return fail("bad arguments");
    const size_t nc = h->c.g.n_cell();
    std::string err;
    if (!sw2d::write_result_map(path, h->c, std::vector<double>(hh, hh + nc),
                                std::vector<double>(qx, qx + nc), std::vector<double>(qy, qy + nc),
                                n_decimal, mode ? mode : "default", err))
        return fail(err);
    return 0;
}

int sw2dh_result_name(double t, char *buf, int32_t buflen) {
    const std::string s = sw2d::result_file_name(t);
    if (!buf || buflen <= (int32_t)s.size()) return fail("buffer too small");
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return 0;
}

void sw2dh_case_free(sw2dh_case *c) { delete c; }

const char *sw2dh_case_probe_name(const sw2dh_case *h, int32_t i) {
    if (!h || i < 0 || (size_t)i >= h->c.probes.names.size()) return nullptr;
    return h->c.probes.names[i].c_str();
}

int sw2dh_geom_locate(const sw2dh_geom *g, int64_t n, const double *x, const double *y,
                      int64_t *cell) {
    if (!g || n < 0 || (n > 0 && (!x || !y || !cell))) return fail("bad arguments");
    for (int64_t i = 0; i < n; ++i) cell[i] = sw2d::locate_cell(g->g, x[i], y[i]);
    return 0;
}

int sw2dh_case_locate(const sw2dh_case *h, int64_t n, const double *x, const double *y,
                      int64_t *cell) {
    if (!h || n < 0 || (n > 0 && (!x || !y || !cell))) return fail("bad arguments");
    for (int64_t i = 0; i < n; ++i) cell[i] = sw2d::locate_cell(h->c.g, x[i], y[i]);
    return 0;
}

}  // extern "C"
