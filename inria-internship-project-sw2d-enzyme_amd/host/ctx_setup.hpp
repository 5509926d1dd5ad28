// ctx_setup.hpp -- a loaded case folder onto a solver context: what
// swSp::init hands the model after reading the folder (swSp.cpp:720-969),
// through the C-ABI.  Shared by the drop-in executables.
#pragma once
#include <string>

#include "../../include/sw2dg.h"
#include "sw2d_case.hpp"

namespace sw2d {

// Creates the context and sets its cell parameters, boundary series, the
// native source terms the case switches on (when with_sources) and the
// initial state.  On failure returns a status, *out is destroyed and err
// says what failed.
inline int make_ctx(const Case &cs, bool with_sources, sw2dg_ctx **out, std::string &err) {
    *out = nullptr;
    const sw2dg_params prm = cs.params();
    const sw2dg_mesh mesh = cs.mesh();
    sw2dg_ctx *ctx = nullptr;
    int st = sw2dg_ctx_create(&mesh, &prm, -1, &ctx);
    if (st) {
        err = "sw2dg_ctx_create failed (" + std::to_string(st) + ")";
        return st;
    }
    auto chk = [&](int s, const char *what) {
        if (s && err.empty()) err = std::string(what) + ": " + sw2dg_last_error(ctx);
        return s;
    };
    const bool fric = prm.friction != SW2DG_FRIC_NONE, hl = prm.head_loss != 0;
    st = chk(sw2dg_set_cell_params(ctx, cs.phi.data(), cs.beta.data(), fric ? cs.fc1.data() : nullptr,
                                   fric ? cs.fc2.data() : nullptr, fric ? cs.falpha.data() : nullptr,
                                   hl ? cs.hsing.data() : nullptr, hl ? cs.hk1.data() : nullptr,
                                   hl ? cs.hk2.data() : nullptr, hl ? cs.halpha.data() : nullptr),
             "set_cell_params");
    const auto types = cs.series_types();
    const auto vals = cs.series_values();
    if (!st)
        st = chk(sw2dg_set_bc_series(ctx, (int32_t)types.size(), (int32_t)cs.bc.t.size(), types.data(),
                                     cs.bc.t.data(), vals.data()),
                 "set_bc_series");
    if (!st && with_sources) {
        if (cs.runoff_model != SW2DG_RUNOFF_NONE)
            st = chk(sw2dg_set_runoff(ctx, cs.runoff_model, cs.rain_water_only, cs.rain_stations,
                                      (int32_t)cs.rain_t.size(), cs.rain_code.data(), cs.rain_t.data(),
                                      cs.rain_v.data(),
                                      cs.runoff_coeff.empty() ? nullptr : cs.runoff_coeff.data()),
                     "set_runoff");
        if (!st && cs.infil_model != SW2DG_INFIL_NONE)
            st = chk(sw2dg_set_infiltration(ctx, cs.infil_model, cs.infil_rate.data()),
                     "set_infiltration");
        if (!st && cs.momentum_diffusion)
            st = chk(sw2dg_set_momentum_diffusion(ctx, 1), "set_momentum_diffusion");
        if (!st && cs.wind_stations > 0)
            st = chk(sw2dg_set_wind(ctx, cs.wind_stations, (int32_t)cs.wind_t.size(), cs.wind_code.data(),
                                    cs.wind_t.data(), cs.wind_x.data(), cs.wind_y.data()),
                     "set_wind");
    }
    if (!st) st = chk(sw2dg_set_state(ctx, cs.h.data(), cs.qx.data(), cs.qy.data(), prm.t0), "set_state");
    if (st) {
        sw2dg_ctx_destroy(ctx);
        return st;
    }
    *out = ctx;
    return SW2DG_OK;
}

}  // namespace sw2d
