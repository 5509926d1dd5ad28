// sw2dSolver_gpu <folder> -- drop-in for the reference's `sw2dSolver <folder>`
// (src/sw2dSolver/main.cpp:18-188) for model_hyperb sp: reads the same case
// folder, runs swSp::execute's time loop (swSp.cpp:482-670) on the GPU through
// the C-ABI, and writes the same output/ResultHydro_p*.txt maps and
// output/ResultBoundaryFluxes.txt and output/ProbesTimeSeries.txt at the same
// storage instants.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/sw2dg.h"
#include "sw2d_case.hpp"

namespace {

std::string e10(double v) {
    char b[64];
    std::snprintf(b, sizeof b, "%.10e", v);
    return b;
}

int die(const std::string &m) {
    std::cerr << "sw2dSolver_gpu: " << m << std::endl;
    return 1;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 2) {
        std::cerr << "sw2dSolver_gpu should receive only 1 parameter: the simulation folder\n";
        return 1;
    }
    const std::string folder = argv[1];
    if (std::string(folder) == "-h" || std::string(folder) == "--help") {
        std::cout << "sw2dSolver_gpu <folder>: single-porosity solver on an MI355X GPU\n";
        return 0;
    }
    namespace fs = std::filesystem;
    if (!fs::exists(folder)) return die("Folder " + folder + " does not exist");
    sw2d::Case cs;
    std::string err;
    if (!sw2d::load_case(folder, cs, err)) return die(err);
    const sw2dg_params prm = cs.params();
    const sw2dg_mesh mesh = cs.mesh();
    sw2dg_ctx *ctx = nullptr;
    int st = sw2dg_ctx_create(&mesh, &prm, -1, &ctx);
    if (st) return die("sw2dg_ctx_create failed (" + std::to_string(st) + ")");
    auto check = [&](int s, const char *what) {
        if (s) {
            std::cerr << "sw2dSolver_gpu: " << what << ": " << sw2dg_last_error(ctx) << std::endl;
            sw2dg_ctx_destroy(ctx);
            std::exit(1);
        }
    };
    const bool fric = prm.friction != SW2DG_FRIC_NONE, hl = prm.head_loss != 0;
    check(sw2dg_set_cell_params(ctx, cs.phi.data(), cs.beta.data(), fric ? cs.fc1.data() : nullptr,
                                fric ? cs.fc2.data() : nullptr, fric ? cs.falpha.data() : nullptr,
                                hl ? cs.hsing.data() : nullptr, hl ? cs.hk1.data() : nullptr,
                                hl ? cs.hk2.data() : nullptr, hl ? cs.halpha.data() : nullptr),
          "set_cell_params");
    const auto types = cs.series_types();
    const auto vals = cs.series_values();
    check(sw2dg_set_bc_series(ctx, (int32_t)types.size(), (int32_t)cs.bc.t.size(), types.data(),
                              cs.bc.t.data(), vals.data()),
          "set_bc_series");
    if (cs.runoff_model != SW2DG_RUNOFF_NONE)
        check(sw2dg_set_runoff(ctx, cs.runoff_model, cs.rain_water_only, cs.rain_stations,
                               (int32_t)cs.rain_t.size(), cs.rain_code.data(), cs.rain_t.data(),
                               cs.rain_v.data(), cs.runoff_coeff.empty() ? nullptr : cs.runoff_coeff.data()),
              "set_runoff");
    if (cs.infil_model != SW2DG_INFIL_NONE)
        check(sw2dg_set_infiltration(ctx, cs.infil_model, cs.infil_rate.data()), "set_infiltration");
    if (cs.momentum_diffusion)
        check(sw2dg_set_momentum_diffusion(ctx, 1), "set_momentum_diffusion");
    if (cs.wind_stations > 0)
        check(sw2dg_set_wind(ctx, cs.wind_stations, (int32_t)cs.wind_t.size(), cs.wind_code.data(),
                             cs.wind_t.data(), cs.wind_x.data(), cs.wind_y.data()),
              "set_wind");
    check(sw2dg_set_state(ctx, cs.h.data(), cs.qx.data(), cs.qy.data(), prm.t0), "set_state");

    const std::string out = (fs::path(folder) / "output").string() + "/";
    fs::create_directories(out);
    fs::create_directories(fs::path(folder) / "log");
    const size_t nc = cs.g.n_cell(), nb = cs.g.n_bnd();
    std::vector<double> h(nc), qx(nc), qy(nc), f1(nb, 0.0);
    std::ofstream prb;
    const bool probes = !cs.probe_cell.empty();
    if (probes) {
        prb.open(out + "ProbesTimeSeries.txt");
        prb << sw2d::probe_header(cs) << std::flush;
    }
    auto store_probes = [&](double t) {
        if (!probes) return;
        check(sw2dg_get_state(ctx, h.data(), qx.data(), qy.data()), "get_state");
        prb << sw2d::probe_line(cs, t, h, qx, qy) << std::flush;
    };
    std::ofstream bfl(out + "ResultBoundaryFluxes.txt");
    bfl << "time";
    for (int32_t code : cs.g.boundary_codes) bfl << "\t" << code;
    bfl << std::endl;
    const std::string mode = cs.cfg.store_flag;
    auto store_maps = [&](double t) {
        if (mode == "none") return;
        check(sw2dg_get_state(ctx, h.data(), qx.data(), qy.data()), "get_state");
        std::string e;
        if (!sw2d::write_result_map(out + sw2d::result_file_name(t), cs, h, qx, qy,
                                    cs.cfg.n_decimal, mode, e))
            check(1, e.c_str());
    };
    auto store_fluxes = [&](double t) {
        bfl << e10(t);
        for (const auto &b : cs.g.boundaries) {
            double sum = 0;
            for (uint64_t id : b) sum += f1[id] * cs.g.bnd_len[id];
            bfl << "\t" << e10(sum);
        }
        bfl << std::endl;
    };
    const auto t0w = std::chrono::steady_clock::now();
    double t = prm.t0;
    store_maps(t);
    store_fluxes(t);
    store_probes(t);
    int64_t total = 0;
    while (t < prm.t_max) {
        int64_t n = 0;
        check(sw2dg_run(ctx, prm.t_max, INT64_MAX / 4, 1, &n), "run");
        total += n;
        double dt;
        int64_t steps;
        int32_t flags;
        check(sw2dg_get_clock(ctx, &t, &dt, &steps, &flags), "get_clock");
        if (n == 0) break;
        if (flags & 1) store_maps(t);
        if (flags & 2) {
            check(sw2dg_get_bnd_f1(ctx, f1.data()), "get_bnd_f1");
            store_fluxes(t);
            store_probes(t);
        }
    }
    check(sw2dg_get_bnd_f1(ctx, f1.data()), "get_bnd_f1");
    store_maps(t);
    store_fluxes(t);
    store_probes(t);
    const double wall =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0w).count();
    std::ofstream log((fs::path(folder) / "log" / "console_Solver.log").string());
    log << "sw2dSolver_gpu " << sw2dg_version() << "\n";
    log << "cells " << nc << " steps " << total << " final time " << e10(t) << "\n";
    log << "Total clock time = " << wall << "\n";
    std::cout << "steps " << total << " t " << e10(t) << " wall " << wall << " s\n";
    sw2dg_ctx_destroy(ctx);
    return 0;
}
