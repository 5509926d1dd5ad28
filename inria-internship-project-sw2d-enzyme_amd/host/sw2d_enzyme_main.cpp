// sw2d_enzyme_gpu <folder> [mode] [phi_init] [alpha_reg] -- drop-in for the
// reference's assimilation driver `sw2d_enzyme` (assim/main_enzyme.cpp:33-269)
// on the GPU: the same case folder, input.solv with its assimilation keys
// (assim/assimConfig.cpp:27-149), observations in reference/, and the same
// output files.  Modes:
//   forward (default)  the _E forward; maps output/ResultHydro_p*.txt every
//                      dt_map (run_forward_E, swSpE_checkpoint.cpp:524-573)
//   gen-ref-e          reference/ResultHydro_p*.txt every dt_map_ad in
//                      [t_state_assimilation, t_max] at porosity_ref_map
//                      (generate_references_E, :470-515)
//   taylor             one checkpointed gradient, the Taylor test
//                      (run_taylor_test_checkpointed, :1454-1663) and
//                      gradient_porosity.csv (main_enzyme.cpp:252-264)
//   minimize           run_minimization (swSpE_minimize.cpp:584-1073):
//                      descent/descent_log.csv, descent/phi_iter/,
//                      optimized_porosity.txt, alpha_sweep.csv
//   gen-ref            refused: the reference's classic generation writes no
//                      map (its native loop never stores one) and then fails
// Every forward, cost and gradient runs in the HIP kernels through the C-ABI.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/sw2d_host.h"
#include "../../include/sw2dg.h"
#include "ctx_setup.hpp"
#include "sw2d_case.hpp"

namespace fs = std::filesystem;

namespace {

constexpr double kSmall = 1e-16;  // small_real (src/common/param.h)

std::ofstream g_log;

void info(const std::string &m) {
    std::cout << m << std::endl;
    if (g_log) g_log << "[Info] " << m << std::endl;
}

int fail(const std::string &m) {
    std::cerr << "sw2d_enzyme_gpu: " << m << std::endl;
    if (g_log) g_log << "[Err.] " << m << std::endl;
    return EXIT_FAILURE;
}

// Assimilation keys with the reference's defaults (assim/assimConfig.h:18-52)
struct AdConfig {
    std::string minimize_algo = "bfgs";
    double adam_step_size = 0.05;
    bool adam_safe = false;
    double alpha_l1 = 0.0, alpha_phi_ref = 0.0, alpha_spatial = 0.0, lambda_tanh = 1.0;
    double lbfgsb_epsilon = 1e-5;
    int lbfgsb_m = 6, lbfgsb_max_linesearch = 20;
    std::string phi_transform = "none";
    double phi_lower_bound = 0.0, phi_upper_bound = 1.0;
    double t_state_assimilation = 0.0, dt_map_ad = -1.0;
    bool windowed_assimilation = false;
    double assimilation_window_duration = -1.0;
    long assimilation_window_steps = -1, n_assimilation_windows = 1;
    bool zonation = false, mask_dry_cells = false, mask_wall_gradient = false;
    bool cost_normalize = false;
    std::string obs_type = "cells", h_weight = "1.0", q_weight = "1.0";
    std::string porosity_ref_map, porosity_free_zone_map, porosity_ebauche_map;
    double alpha_reg = 0.0;
    size_t checkpoint_stride = 30, max_it = 3;
};

bool parse_bool(std::string v) {
    std::transform(v.begin(), v.end(), v.begin(), [](unsigned char ch) { return std::tolower(ch); });
    return v == "1" || v == "true" || v == "yes" || v == "on";
}

bool read_ad(const sw2d::SolverConfig &cfg, AdConfig &a, std::string &err) {
    try {
        for (const auto &kv : cfg.assim) {
            const std::string &k = kv.first, &v = kv.second;
            if (k == "minimize_algo") a.minimize_algo = v;
            else if (k == "adam_step_size") a.adam_step_size = std::stod(v);
            else if (k == "adam_safe") a.adam_safe = parse_bool(v);
            else if (k == "alpha_l1") a.alpha_l1 = std::stod(v);
            else if (k == "alpha_phi_ref") a.alpha_phi_ref = std::stod(v);
            else if (k == "alpha_reg") a.alpha_reg = std::stod(v);
            else if (k == "alpha_spatial") a.alpha_spatial = std::stod(v);
            else if (k == "lambda_tanh") a.lambda_tanh = std::stod(v);
            else if (k == "lbfgsb_epsilon") a.lbfgsb_epsilon = std::stod(v);
            else if (k == "lbfgsb_m") a.lbfgsb_m = std::stoi(v);
            else if (k == "lbfgsb_max_linesearch") a.lbfgsb_max_linesearch = std::stoi(v);
            else if (k == "phi_transform") a.phi_transform = v;
            else if (k == "phi_lower_bound") a.phi_lower_bound = std::stod(v);
            else if (k == "phi_upper_bound") a.phi_upper_bound = std::stod(v);
            else if (k == "t_state_assimilation") a.t_state_assimilation = std::stod(v);
            else if (k == "dt_map_ad") a.dt_map_ad = std::stod(v);
            else if (k == "windowed_assimilation") a.windowed_assimilation = parse_bool(v);
            else if (k == "assimilation_window_duration") a.assimilation_window_duration = std::stod(v);
            else if (k == "assimilation_window_steps") a.assimilation_window_steps = std::stol(v);
            else if (k == "n_assimilation_windows") a.n_assimilation_windows = std::stol(v);
            else if (k == "zonation") a.zonation = parse_bool(v);
            else if (k == "mask_dry_cells") a.mask_dry_cells = parse_bool(v);
            else if (k == "mask_wall_gradient") a.mask_wall_gradient = parse_bool(v);
            else if (k == "cost_normalize") a.cost_normalize = parse_bool(v);
            else if (k == "obs_type") a.obs_type = v;
            else if (k == "h_weight") a.h_weight = v;
            else if (k == "q_weight") a.q_weight = v;
            else if (k == "porosity_ref_map") a.porosity_ref_map = v;
            else if (k == "porosity_free_zone_map") a.porosity_free_zone_map = v;
            else if (k == "porosity_ebauche_map") a.porosity_ebauche_map = v;
            else if (k == "checkpoint_stride") a.checkpoint_stride = (size_t)std::stoul(v);
            else if (k == "max_it") a.max_it = (size_t)std::stoul(v);
        }
    } catch (const std::exception &e) {
        err = std::string("assimilation key with a bad value: ") + e.what();
        return false;
    }
    return true;
}

// A porosity (or weight) map: "Unif"/"Distrib" format, or CSV lines
// "cell_id,value" with an optional header (swSpE::load_phi_csv,
// swSpE_minimize.cpp:130-200).
bool load_phi_map(const std::string &path, size_t nc, std::vector<double> &out, std::string &err) {
    std::ifstream f(path);
    if (!f) {
        err = "load_phi_csv: cannot open " + path;
        return false;
    }
    std::string first;
    while (std::getline(f, first)) {
        const size_t p = first.find_first_not_of(" \t\r\n");
        if (p != std::string::npos) {
            first = first.substr(p);
            break;
        }
    }
    if (first.rfind("Unif", 0) == 0 || first.rfind("unif", 0) == 0) {
        std::vector<std::string> names;
        std::vector<std::vector<double>> cols;
        if (!sw2d::read_map(path, 1, nc, names, cols, err)) return false;
        out = cols[0];
        return true;
    }
    f.clear();
    f.seekg(0);
    std::string line;
    const std::streampos start = f.tellg();
    if (std::getline(f, line)) {
        const bool numeric = !line.empty() && (std::isdigit((unsigned char)line[0]) ||
                                               line[0] == '-' || line[0] == '+');
        if (numeric) f.seekg(start);
    }
    out.assign(nc, 0.0);
    size_t count = 0;
    while (std::getline(f, line)) {
        if (line.empty()) continue;
        std::replace(line.begin(), line.end(), ',', ' ');
        std::stringstream ss(line);
        long id;
        double v;
        if (!(ss >> id >> v)) continue;
        if (id < 0 || (size_t)id >= nc) {
            err = "load_phi_csv: cell_id out of range (" + std::to_string(id) + ") in " + path;
            return false;
        }
        out[id] = v;
        ++count;
    }
    if (count != nc) {
        err = "load_phi_csv: " + std::to_string(count) + " value(s) read, " + std::to_string(nc) +
              " cell(s) expected (" + path + ")";
        return false;
    }
    return true;
}

// SP_cell_porosity_map.txt layout (write_sp_porosity_map, swSpE_minimize.cpp:254-265)
void write_sp_map(const std::string &path, const std::vector<double> &phi) {
    std::ofstream o(path);
    o << "Unif\t0\n==== Default Param\nPhi\n1\n==== Distrib\n" << std::setprecision(10);
    for (double v : phi) o << v << "\n";
}

// time of a ResultHydro_p.... file name (swIO::timeFromFileName, swIO.cpp:1194-1226)
bool time_of(const std::string &name, double &t) {
    if (name.rfind("ResultHydro_p", 0) != 0 || name.size() < 13 + 17) return false;
    const std::string s = name.substr(13, 17);
    try {
        const size_t d = std::stoul(s.substr(0, 4)), h = std::stoul(s.substr(5, 2)),
                     m = std::stoul(s.substr(8, 2)), sec = std::stoul(s.substr(11, 2)),
                     ms = std::stoul(s.substr(14, 3));
        t = ms / 1000. + (double)(sec + 60 * (m + 60 * (h + 24 * d)));
    } catch (...) {
        return false;
    }
    return true;
}

// float, or a map path (swSpE::resolve_cost_weight, swSpE_core.cpp:268-298)
std::vector<double> cost_weight(const std::string &spec, const std::string &folder, size_t nc) {
    try {
        size_t pos = 0;
        const double v = std::stod(spec, &pos);
        while (pos < spec.size() && std::isspace((unsigned char)spec[pos])) ++pos;
        if (pos == spec.size()) return std::vector<double>(nc, v);
    } catch (...) {
    }
    std::vector<double> w;
    std::string err;
    if (!load_phi_map((fs::path(folder) / spec).string(), nc, w, err)) {
        info("cost weight: " + err + " -> uniform weight 1");
        return std::vector<double>(nc, 1.0);
    }
    return w;
}

struct Observations {
    std::vector<double> t;
    std::vector<double> fh, fqx, fqy;  // frames [n_obs][nc]
    std::vector<int64_t> gcell;
    std::vector<double> gh;            // gauges [n_obs][n_gauges]
    std::vector<double> wh, wq;
};

// reference/ResultHydro_p*.txt in [t_state_assimilation, t_max]
// (swSpE::load_all_observations / load_single_observation_file, swSpE_ad.cpp:47-154)
bool load_frames(const std::string &folder, size_t nc, double t_begin, double t_end,
                 Observations &o, std::string &err) {
    const fs::path dir = fs::path(folder) / "reference";
    if (!fs::is_directory(dir)) {
        err = "observation folder not found: " + dir.string();
        return false;
    }
    std::vector<std::pair<double, std::string>> cand;
    for (const auto &e : fs::directory_iterator(dir)) {
        if (!e.is_regular_file()) continue;
        const std::string n = e.path().filename().string();
        if (n.rfind("ResultHydro_p", 0) != 0 || e.path().extension() != ".txt") continue;
        double t;
        if (!time_of(n, t)) continue;
        if (t < t_begin - 1e-6 || t > t_end + 1e-6) continue;
        cand.emplace_back(t, e.path().string());
    }
    std::sort(cand.begin(), cand.end());
    for (const auto &c : cand) {
        std::ifstream f(c.second);
        std::string line;
        std::getline(f, line);
        std::vector<double> h(nc, 0.0), qx(nc, 0.0), qy(nc, 0.0);
        size_t k = 0;
        while (k < nc && std::getline(f, line)) {
            if (line.empty()) continue;
            std::stringstream ss(line);
            double v;
            for (int col = 0; col <= 5; ++col) {
                if (!(ss >> v)) break;
                if (col == 3) h[k] = v;
                if (col == 4) qx[k] = v;
                if (col == 5) qy[k] = v;
            }
            ++k;
        }
        if (!o.t.empty() && o.t.back() == c.first) {  // one frame per time: the later file wins
            o.t.pop_back();
            o.fh.resize(o.fh.size() - nc), o.fqx.resize(o.fqx.size() - nc), o.fqy.resize(o.fqy.size() - nc);
        }
        o.t.push_back(c.first);
        o.fh.insert(o.fh.end(), h.begin(), h.end());
        o.fqx.insert(o.fqx.end(), qx.begin(), qx.end());
        o.fqy.insert(o.fqy.end(), qy.begin(), qy.end());
    }
    if (o.t.empty()) {
        err = "no reference/ResultHydro_p*.txt observation in the window";
        return false;
    }
    return true;
}

// gauges: probes located in the mesh (locate_gauges, swSpE_gauges.cpp:41-77:
// probes outside are skipped) and reference/gauges.txt in the ProbesTimeSeries
// layout, h of probe g in column 2 + 4 g (load_gauge_observations, :94-158)
bool load_gauges(const sw2d::Case &cs, double t_begin, double t_end, Observations &o,
                 std::string &err) {
    std::vector<size_t> src;
    for (size_t i = 0; i < cs.probes.x.size(); ++i) {
        const int64_t c = sw2d::locate_cell(cs.g, cs.probes.x[i], cs.probes.y[i]);
        if (c < 0) {
            info("obs_type=gauges: probe " + std::to_string(i) + " outside the mesh -> skipped");
            continue;
        }
        o.gcell.push_back(c);
        src.push_back(i);
    }
    if (o.gcell.empty()) {
        err = "obs_type=gauges: no probe inside the mesh (input/probe_locations.txt)";
        return false;
    }
    std::ifstream f((fs::path(cs.folder) / "reference" / "gauges.txt").string());
    if (!f) {
        err = "obs_type=gauges: cannot open reference/gauges.txt";
        return false;
    }
    std::string line;
    bool header = false;
    while (std::getline(f, line)) {
        if (line.empty()) continue;
        if (!header) {
            header = line.rfind("time", 0) == 0;
            continue;
        }
        std::stringstream ss(line);
        std::vector<double> col;
        double v;
        while (ss >> v) col.push_back(v);
        if (col.empty() || col[0] < t_begin - 1e-6 || col[0] > t_end + 1e-6) continue;
        std::vector<double> fr;
        for (size_t g = 0; g < src.size(); ++g) {
            const size_t hc = 2 + 4 * src[g];
            if (hc >= col.size()) break;
            fr.push_back(col[hc]);
        }
        if (fr.size() != src.size()) continue;
        if (!o.t.empty() && o.t.back() == col[0]) {  // a later line for the same time wins
            std::copy(fr.begin(), fr.end(), o.gh.end() - fr.size());
            continue;
        }
        o.t.push_back(col[0]);
        o.gh.insert(o.gh.end(), fr.begin(), fr.end());
    }
    if (o.t.empty()) {
        err = "obs_type=gauges: no observation in the window from reference/gauges.txt";
        return false;
    }
    return true;
}

int phi_mode(const std::string &s) {
    std::string l = s;
    std::transform(l.begin(), l.end(), l.begin(), [](unsigned char ch) { return std::tolower(ch); });
    return l == "clamp" ? 1 : l == "tanh" ? 2 : l == "normalize" ? 3 : 0;
}

struct Run {
    sw2d::Case cs;
    AdConfig ad;
    sw2dg_ctx *ctx = nullptr;
    size_t nc = 0;
    int pmode = 0;
    double t_window = -1e300;
    // step budget of the recorded runs (device buffers are sized for it):
    // input.solv key gpu_max_steps, default 100000; a run that needs more stops
    // with an error instead of a truncated window
    int64_t max_steps = 100000;
    ~Run() {
        if (ctx) sw2dg_ctx_destroy(ctx);
    }
    int check(int st, const char *what) {
        if (st) fail(std::string(what) + ": " + sw2dg_last_error(ctx));
        return st;
    }
    // host phi_correct with the normalisation bounds of `m` over `free`
    void phi_of(const std::vector<double> &m, const std::vector<int64_t> &free_idx,
                std::vector<double> &phi, std::vector<double> *dphi = nullptr) const {
        double lo = 0.0, hi = 1.0;
        if (pmode == 3 && !free_idx.empty()) {
            lo = m[free_idx[0]], hi = lo;
            for (int64_t i : free_idx) lo = std::min(lo, m[i]), hi = std::max(hi, m[i]);
        }
        phi.resize(m.size());
        std::vector<double> d(m.size());
        sw2dg_phi_transform_eval(pmode, ad.lambda_tanh, lo, hi, (int64_t)m.size(), m.data(), phi.data(),
                                 d.data());
        if (dphi) *dphi = d;
    }
};

// assimilation window start (get_spatial_gradient_checkpointed, swSpE_checkpoint.cpp:785-840)
double window_start(const Run &r, const std::vector<double> &t_seq) {
    const AdConfig &a = r.ad;
    if (!a.windowed_assimilation) return -1e300;
    if (a.assimilation_window_steps > 0) {
        // segment starts of the recorded run, counted back from the end
        const size_t S = std::max<size_t>(a.checkpoint_stride, 1);
        const long n = (long)t_seq.size() - 1;
        const long K = (n + (long)S - 1) / (long)S;
        long acc = 0;
        for (long k = K - 1; k >= 0; --k) {
            acc += std::min<long>((long)S, n - k * (long)S);
            if (acc >= a.assimilation_window_steps) return t_seq[(size_t)(k * (long)S)];
        }
        return -1e300;
    }
    if (a.assimilation_window_duration > 0.0) return r.cs.params().t_max - a.assimilation_window_duration;
    if (a.t_state_assimilation > r.cs.params().t0) return a.t_state_assimilation;
    return -1e300;
}

int setup_assim(Run &r, const Observations &o, double t_window) {
    sw2dg_obs ob{};
    ob.n_obs = (int32_t)o.t.size();
    ob.t = o.t.data();
    ob.n_gauges = (int32_t)o.gcell.size();
    ob.gauge_cell = o.gcell.empty() ? nullptr : o.gcell.data();
    ob.gauge_h = o.gh.empty() ? nullptr : o.gh.data();
    if (!o.fh.empty()) {
        ob.frame_h = o.fh.data(), ob.frame_qx = o.fqx.data(), ob.frame_qy = o.fqy.data();
        ob.w_h = o.wh.empty() ? nullptr : o.wh.data();
        ob.w_q = o.wq.empty() ? nullptr : o.wq.data();
    }
    ob.normalize = r.ad.cost_normalize ? 1 : 0;
    if (r.check(sw2dg_assim_setup(r.ctx, &ob, (int32_t)r.ad.checkpoint_stride,
                                  (int32_t)std::max<long>(r.ad.n_assimilation_windows, 1), t_window,
                                  r.max_steps),
                "assim_setup"))
        return 1;
    r.t_window = t_window;
    return 0;
}

// a dry run that used the whole step budget without reaching t_max
bool reached_max(Run &r, int64_t n) {
    if (n < r.max_steps) return false;
    fail("the run needs more than " + std::to_string(r.max_steps) +
         " time steps to reach t_max: raise gpu_max_steps in input.solv");
    return true;
}

// the observation-time grid of the dumps (generate_references_E / run_forward_E)
std::vector<double> grid(double dt, double t_begin, double t_end) {
    std::vector<double> t;
    long k0 = (long)std::ceil((t_begin - kSmall) / dt);
    if (k0 < 0) k0 = 0;
    for (long k = k0;; ++k) {
        const double tt = (double)k * dt;
        if (tt > t_end + kSmall) break;
        if (tt >= t_begin - kSmall) t.push_back(tt);
    }
    return t;
}

// the _E forward dumping the state at each grid time into folder/sub
int dump_forward(Run &r, const std::vector<double> &phi, double dt, double t_begin,
                 const std::string &sub, bool wipe) {
    const sw2dg_params p = r.cs.params();
    Observations o;
    o.t = grid(dt, t_begin, p.t_max);
    if (o.t.empty()) return fail("no output time in [" + std::to_string(t_begin) + ", " +
                                 std::to_string(p.t_max) + "]");
    if (setup_assim(r, o, -1e300)) return 1;
    if (r.check(sw2dg_assim_keep_obs_states(r.ctx, 1), "keep_obs_states")) return 1;
    int64_t n = 0;
    if (r.check(sw2dg_assim_record(r.ctx, phi.data(), nullptr, nullptr, &n), "assim_record")) return 1;
    if (reached_max(r, n)) return 1;
    const size_t no = o.t.size(), nc = r.nc;
    std::vector<double> h(no * nc), qx(no * nc), qy(no * nc);
    std::vector<int32_t> hit(no);
    if (r.check(sw2dg_assim_obs_states(r.ctx, h.data(), qx.data(), qy.data(), hit.data()), "obs_states"))
        return 1;
    const fs::path dir = fs::path(r.cs.folder) / sub;
    std::error_code ec;
    if (wipe) fs::remove_all(dir, ec);
    fs::create_directories(dir, ec);
    size_t written = 0;
    for (size_t k = 0; k < no; ++k) {
        if (!hit[k]) continue;
        std::string err;
        const std::vector<double> H(h.begin() + k * nc, h.begin() + (k + 1) * nc),
            QX(qx.begin() + k * nc, qx.begin() + (k + 1) * nc), QY(qy.begin() + k * nc, qy.begin() + (k + 1) * nc);
        if (!sw2d::write_result_map((dir / sw2d::result_file_name(o.t[k])).string(), r.cs, H, QX, QY,
                                    r.cs.cfg.n_decimal, "default", err))
            return fail(err);
        ++written;
    }
    info(std::to_string(n) + " steps of the _E forward, " + std::to_string(written) + " map(s) in " +
         dir.string());
    if (written == 0) return fail("no map written: check the window and the output interval");
    return 0;
}

// ---------------------------------------------------------------------------
// the porosity problem (PorosityProblem, swSpE_minimize.cpp:267-581)
// ---------------------------------------------------------------------------
struct Problem {
    Run &r;
    double alpha;
    std::vector<double> phi_b, phi_ref;
    std::vector<int64_t> free_idx;
    bool zon = false;
    std::vector<size_t> zone_of_free;
    std::vector<double> zone_value;
    std::vector<double> m_buf, gm;
    std::string descent_dir;
    std::ofstream csv;
    double J_init = 0.0, J_prev = 0.0;
    bool J_init_set = false, J_prev_set = false;
    int grad_calls = 0;
    int64_t n_steps = 0;
    bool error = false;

    Problem(Run &run, double a, const std::vector<double> &pb, std::vector<int64_t> fidx, bool z)
        : r(run), alpha(a), phi_b(pb), free_idx(std::move(fidx)), zon(z), m_buf(pb), gm(pb.size()) {
        if (zon) {
            std::map<long, size_t> key2zone;
            for (size_t j = 0; j < free_idx.size(); ++j) {
                const double v = phi_b[free_idx[j]];
                const long key = (long)std::llround(v * 1e6);
                auto it = key2zone.find(key);
                size_t zz;
                if (it == key2zone.end()) {
                    zz = zone_value.size();
                    key2zone[key] = zz;
                    zone_value.push_back(v);
                } else {
                    zz = it->second;
                }
                zone_of_free.push_back(zz);
            }
        }
    }
    size_t size() const { return zon ? zone_value.size() : free_idx.size(); }
    void map_theta(const double *theta) {
        m_buf = phi_b;
        for (size_t j = 0; j < free_idx.size(); ++j) m_buf[free_idx[j]] = theta[zon ? zone_of_free[j] : j];
    }
    std::vector<double> phi_now(std::vector<double> *dphi = nullptr) const {
        std::vector<double> p;
        r.phi_of(m_buf, free_idx, p, dphi);
        return p;
    }
    struct Pen {
        double ridge = 0, l1 = 0, spatial = 0, ref = 0;
    };
    Pen penalties(const std::vector<double> &phi) const {
        Pen p;
        const AdConfig &a = r.ad;
        for (int64_t i : free_idx) {
            const double d = phi[i] - phi_b[i];
            if (alpha > 0.0) p.ridge += d * d;
            if (a.alpha_l1 > 0.0) p.l1 += std::fabs(d);
            if (a.alpha_phi_ref > 0.0 && !phi_ref.empty()) {
                const double e = phi[i] - phi_ref[i];
                p.ref += e * e;
            }
        }
        p.ridge *= 0.5 * alpha;
        p.l1 *= a.alpha_l1;
        p.ref *= 0.5 * a.alpha_phi_ref;
        if (a.alpha_spatial > 0.0) {
            const auto &L = r.cs.g.int_L, &R = r.cs.g.int_R;
            double s = 0.0;
            for (size_t e = 0; e < L.size(); ++e) {
                const double d = phi[L[e]] - phi[R[e]];
                s += d * d;
            }
            p.spatial = 0.5 * a.alpha_spatial * s;
        }
        return p;
    }
    double sse_ref(const std::vector<double> &phi) const {
        if (phi_ref.empty()) return -1.0;
        double s = 0.0;
        for (size_t i = 0; i < phi.size(); ++i) {
            const double v = free_is(i) ? phi[i] : phi_b[i];
            const double d = v - phi_ref[i];
            s += d * d;
        }
        return s;
    }
    std::vector<char> free_mask;
    bool free_is(size_t i) const { return free_mask.empty() || free_mask[i]; }

    // f = J_data (frozen replay) + all penalties
    double func(const double *theta) {
        map_theta(theta);
        double J = 0.0;
        if (r.check(sw2dg_assim_cost(r.ctx, m_buf.data(), nullptr, nullptr, &J), "assim_cost")) {
            error = true;
            return NAN;
        }
        const Pen p = penalties(phi_now());
        return J + p.ridge + p.l1 + p.spatial + p.ref;
    }
    // J (without the spatial term, as the reference) and the gradient
    double grad(const double *theta, double *gr) {
        map_theta(theta);
        double Jd = 0.0;
        if (r.check(sw2dg_assim_gradient(r.ctx, m_buf.data(), nullptr, nullptr, &Jd, gm.data(), nullptr,
                                         nullptr, nullptr, &n_steps),
                    "assim_gradient")) {
            error = true;
            return NAN;
        }
        if (n_steps >= r.max_steps) {
            reached_max(r, n_steps);
            error = true;
            return NAN;
        }
        std::vector<double> dphi;
        const std::vector<double> phi = phi_now(&dphi);
        const AdConfig &a = r.ad;
        std::vector<double> gsp;
        if (a.alpha_spatial > 0.0) {
            gsp.assign(phi.size(), 0.0);
            const auto &L = r.cs.g.int_L, &R = r.cs.g.int_R;
            for (size_t e = 0; e < L.size(); ++e) {
                const double d = a.alpha_spatial * (phi[L[e]] - phi[R[e]]);
                gsp[L[e]] += d;
                gsp[R[e]] -= d;
            }
        }
        double pmin = 2.0, pmax = -1.0, psum = 0.0, gl2 = 0.0, ginf = 0.0;
        long garg = -1, n05 = 0, n005 = 0, nup = 0;
        std::vector<double> gcell(free_idx.size());
        for (size_t j = 0; j < free_idx.size(); ++j) {
            const int64_t i = free_idx[j];
            const double ph = phi[i], dp = dphi[i];
            double dl1 = 0.0;
            if (a.alpha_l1 > 0.0) {
                const double diff = ph - phi_b[i];
                if (diff > 1e-12) dl1 = a.alpha_l1 * dp;
                else if (diff < -1e-12) dl1 = -a.alpha_l1 * dp;
            }
            double g = gm[i] + alpha * (ph - phi_b[i]) * dp + dl1;
            if (!gsp.empty()) g += gsp[i] * dp;
            if (a.alpha_phi_ref > 0.0 && !phi_ref.empty()) g += a.alpha_phi_ref * (ph - phi_ref[i]) * dp;
            gcell[j] = g;
            pmin = std::min(pmin, ph);
            pmax = std::max(pmax, ph);
            psum += ph;
            n05 += ph < 0.5;
            n005 += ph < 0.05;
            nup += ph > 1.049;
            gl2 += g * g;
            if (std::fabs(g) > ginf) ginf = std::fabs(g), garg = (long)i;
        }
        gl2 = std::sqrt(gl2);
        if (zon) {
            std::fill(gr, gr + size(), 0.0);
            for (size_t j = 0; j < free_idx.size(); ++j) gr[zone_of_free[j]] += gcell[j];
        } else {
            std::copy(gcell.begin(), gcell.end(), gr);
        }
        const Pen p = penalties(phi);
        const double J = Jd + p.ridge + p.l1 + p.ref;
        if (!J_init_set) J_init = J, J_init_set = true;
        ++grad_calls;
        const double rel = J_init != 0.0 ? (J - J_init) / J_init * 100.0 : 0.0;
        const double relp = (J_prev_set && J_prev != 0.0) ? (J - J_prev) / J_prev * 100.0 : 0.0;
        // the data gradient over all cells (its NaN/inf already zeroed by the library)
        double pl2 = 0.0, pmx = 0.0;
        long parg = 0;
        for (size_t i = 0; i < gm.size(); ++i) {
            pl2 += gm[i] * gm[i];
            if (std::fabs(gm[i]) > std::fabs(pmx)) pmx = gm[i], parg = (long)i;
        }
        pl2 = std::sqrt(pl2);
        double peak = 0.0;
        sw2dg_assim_adjoint_peak(r.ctx, &peak);
        const double mean = free_idx.empty() ? 0.0 : psum / (double)free_idx.size();
        {
            std::ostringstream m;
            m << "[Minim] grad #" << grad_calls << " : J_current = " << std::scientific
              << std::setprecision(6) << J << " | J_init = " << J_init << " | rel = "
              << (rel >= 0.0 ? '+' : '-') << std::fixed << std::setprecision(2) << std::fabs(rel) << " %";
            info(m.str());
        }
        if (!descent_dir.empty()) {
            if (!csv.is_open()) {
                csv.open(descent_dir + "/descent_log.csv");
                csv << "iter,J_total,J_data,ridge,rel_init_pct,rel_prev_pct,gradJm_L2,gradJm_Linf,"
                       "gradJm_argmax,gradJphys_L2,gradJphys_max,gradJphys_argmax,dh_peak,phi_min,"
                       "phi_max,phi_mean,n_phi_lt_0p5,n_phi_lt_0p05,n_phi_upper_clamp,adam_step,"
                       "sse_phi_ref\n";
            }
            csv << grad_calls << "," << std::scientific << std::setprecision(10) << J << "," << Jd << ","
                << p.ridge << "," << std::setprecision(6) << rel << "," << relp << "," << gl2 << ","
                << ginf << "," << garg << "," << pl2 << "," << pmx << "," << parg << "," << peak << ","
                << pmin << "," << pmax << "," << mean << "," << n05 << "," << n005 << "," << nup << ","
                << a.adam_step_size << "," << std::setprecision(10) << sse_ref(phi) << "\n";
            csv.flush();
            char name[64];
            std::snprintf(name, sizeof name, "/phi_iter/phi_iter_%04d.txt", grad_calls);
            std::vector<double> out = phi_b;
            for (int64_t i : free_idx) out[i] = phi[i];
            write_sp_map(descent_dir + name, out);
        }
        J_prev = J, J_prev_set = true;
        return J;
    }
};

// ---- minimisers of run_minimization (swSpE_minimize.cpp:748-1015) ----
struct Objective {
    Problem *p;
    bool failed = false;
};

int lbfgsb_obj(const double *x, double *f, double *g, void *user) {
    auto *o = static_cast<Objective *>(user);
    *f = o->p->grad(x, g);
    return o->p->error ? -1 : 0;
}

// projected gradient descent + Armijo (:748-831)
double run_gd(Problem &P, std::vector<double> &x, size_t max_it, double step0) {
    const size_t n = x.size();
    std::vector<double> g(n), v(n), best = x;
    double best_J = INFINITY, t = step0 > 0.0 ? step0 : 1e-6;
    const bool project = P.r.pmode == 0 || P.r.pmode == 1;
    for (size_t it = 0; it < max_it; ++it) {
        const double J = P.grad(x.data(), g.data());
        if (J < best_J) best_J = J, best = x;
        double gg = 0.0;
        for (double q : g) gg += q * q;
        if (!(gg > 0.0)) break;
        double tt = 4.0 * t;
        bool ok = false;
        for (int bt = 0; bt < 40; ++bt) {
            double pg = 0.0;
            for (size_t i = 0; i < n; ++i) {
                v[i] = x[i] - tt * g[i];
                if (project) v[i] = std::min(1.0, std::max(0.0, v[i]));
                pg += g[i] * (x[i] - v[i]);
            }
            const double Jn = P.func(v.data());
            if (std::isfinite(Jn) && Jn <= J - 1e-4 * pg && pg > 0.0) {
                ok = true;
                break;
            }
            tt *= 0.5;
        }
        if (!ok) break;
        x = v, t = tt;
    }
    x = best;
    return best_J;
}

// projected Adam (:930-1015)
double run_adam(Problem &P, std::vector<double> &x, size_t max_it, double step, bool safe) {
    const size_t n = x.size();
    std::vector<double> g(n), m(n, 0.0), v(n, 0.0), best = x;
    double best_J = INFINITY, gprev = -1.0;
    const bool project = P.r.pmode == 0 || P.r.pmode == 1;
    for (size_t it = 1; it <= max_it; ++it) {
        const double J = P.grad(x.data(), g.data());
        if (J < best_J) best_J = J, best = x;
        double ginf = 0.0;
        for (double q : g) ginf = std::max(ginf, std::fabs(q));
        if (gprev > 0.0 && ginf > 50.0 * gprev) {
            info("[Adam] gradient explosion (||g||_inf x" + std::to_string(ginf / gprev) + ")");
            if (safe) break;
        }
        gprev = ginf;
        const double b1 = 1.0 - std::pow(0.9, (double)it), b2 = 1.0 - std::pow(0.999, (double)it);
        for (size_t i = 0; i < n; ++i) {
            m[i] = 0.9 * m[i] + 0.1 * g[i];
            v[i] = 0.999 * v[i] + 0.001 * g[i] * g[i];
            x[i] -= step * (m[i] / b1) / (std::sqrt(v[i] / b2) + 1e-8);
            if (project) x[i] = std::min(1.0, std::max(0.0, x[i]));
        }
    }
    x = best;
    return best_J;
}

// dense BFGS, strong-Wolfe line search (the default branch, :913-930; PSQN
// is not in the reference tree: Nocedal & Wright algorithms 6.1, 3.5, 3.6)
double run_bfgs(Problem &P, std::vector<double> &x, size_t max_it) {
    const size_t n = x.size();
    std::vector<double> g(n), H(n * n, 0.0), d(n), xt(n), gt(n), s(n), y(n), Hy(n);
    for (size_t i = 0; i < n; ++i) H[i * n + i] = 1.0;
    double J = P.grad(x.data(), g.data());
    bool first = true;
    const double c1 = 1e-4, c2 = 0.9;
    for (size_t it = 0; it < max_it; ++it) {
        double dg = 0.0;
        for (size_t i = 0; i < n; ++i) {
            double a = 0.0;
            for (size_t k = 0; k < n; ++k) a -= H[i * n + k] * g[k];
            d[i] = a;
            dg += a * g[i];
        }
        if (!(dg < 0.0)) {
            std::fill(H.begin(), H.end(), 0.0);
            for (size_t i = 0; i < n; ++i) H[i * n + i] = 1.0, d[i] = -g[i];
            dg = 0.0;
            for (size_t i = 0; i < n; ++i) dg -= g[i] * g[i];
        }
        // bracketing + zoom
        int evals = 0;
        double t = 1.0, t_prev = 0.0, J_prev = J, Jt = J, lo = 0, hi = 0, J_lo = J;
        bool found = false, zoom = false;
        auto eval = [&](double st, double &dgt) {
            ++evals;
            for (size_t i = 0; i < n; ++i) xt[i] = x[i] + st * d[i];
            const double v = P.grad(xt.data(), gt.data());
            dgt = 0.0;
            for (size_t i = 0; i < n; ++i) dgt += gt[i] * d[i];
            return v;
        };
        while (evals < 40 && !found && !zoom) {
            double dgt;
            Jt = eval(t, dgt);
            if (!std::isfinite(Jt) || Jt > J + c1 * t * dg || (t_prev > 0.0 && Jt >= J_prev)) {
                lo = t_prev, J_lo = J_prev, hi = t, zoom = true;
                break;
            }
            if (std::fabs(dgt) <= -c2 * dg) {
                found = true;
                break;
            }
            if (dgt >= 0.0) {
                lo = t, J_lo = Jt, hi = t_prev, zoom = true;
                break;
            }
            t_prev = t, J_prev = Jt, t *= 2.0;
        }
        while (zoom && evals < 40 && !found) {
            t = 0.5 * (lo + hi);
            double dgt;
            Jt = eval(t, dgt);
            if (!std::isfinite(Jt) || Jt > J + c1 * t * dg || Jt >= J_lo) {
                hi = t;
            } else {
                if (std::fabs(dgt) <= -c2 * dg) {
                    found = true;
                    break;
                }
                if (dgt * (hi - lo) >= 0.0) hi = lo;
                lo = t, J_lo = Jt;
            }
        }
        if (!found) break;
        double sy = 0.0, yy = 0.0;
        for (size_t i = 0; i < n; ++i) s[i] = t * d[i], y[i] = gt[i] - g[i], sy += s[i] * y[i], yy += y[i] * y[i];
        if (sy > 0.0) {
            if (first) {
                std::fill(H.begin(), H.end(), 0.0);
                for (size_t i = 0; i < n; ++i) H[i * n + i] = sy / yy;
                first = false;
            }
            const double rho = 1.0 / sy;
            double yHy = 0.0;
            for (size_t i = 0; i < n; ++i) {
                double a = 0.0;
                for (size_t k = 0; k < n; ++k) a += H[i * n + k] * y[k];
                Hy[i] = a;
                yHy += y[i] * a;
            }
            for (size_t i = 0; i < n; ++i)
                for (size_t k = 0; k < n; ++k)
                    H[i * n + k] += -rho * (s[i] * Hy[k] + Hy[i] * s[k]) + (rho * rho * yHy + rho) * s[i] * s[k];
        }
        const bool done = std::fabs(Jt - J) <= 1e-8 * (std::fabs(J) + 1e-8);
        for (size_t i = 0; i < n; ++i) x[i] += s[i];
        J = Jt, g = gt;
        if (done) break;
    }
    return J;
}

int mode_minimize(Run &r, const std::string &phi_path, double alpha_reg) {
    const size_t nc = r.nc;
    std::vector<double> phi_start;
    std::string err;
    if (!load_phi_map(phi_path, nc, phi_start, err)) return fail(err);
    const AdConfig &a = r.ad;
    std::vector<int64_t> free_idx;
    if (a.mask_dry_cells) {  // native forward of the background: cells wet at the end are free
        info("identifiability mask on (mask_dry_cells): native forward of the background");
        if (r.check(sw2dg_set_cell_params(r.ctx, phi_start.data(), r.cs.beta.data(),
                                          r.cs.params().friction ? r.cs.fc1.data() : nullptr,
                                          r.cs.params().friction ? r.cs.fc2.data() : nullptr,
                                          r.cs.params().friction ? r.cs.falpha.data() : nullptr,
                                          r.cs.params().head_loss ? r.cs.hsing.data() : nullptr,
                                          r.cs.params().head_loss ? r.cs.hk1.data() : nullptr,
                                          r.cs.params().head_loss ? r.cs.hk2.data() : nullptr,
                                          r.cs.params().head_loss ? r.cs.halpha.data() : nullptr),
                    "set_cell_params"))
            return 1;
        int64_t n = 0;
        if (r.check(sw2dg_run(r.ctx, r.cs.params().t_max, INT64_MAX / 4, 0, &n), "run")) return 1;
        std::vector<double> h(nc), q(nc), s(nc);
        if (r.check(sw2dg_get_state(r.ctx, h.data(), q.data(), s.data()), "get_state")) return 1;
        for (size_t i = 0; i < nc; ++i)
            if (h[i] > 1e-4) free_idx.push_back((int64_t)i);
        if (r.check(sw2dg_set_state(r.ctx, r.cs.h.data(), r.cs.qx.data(), r.cs.qy.data(), r.cs.params().t0),
                    "set_state"))
            return 1;
        info(std::to_string(nc - free_idx.size()) + " dry cells fixed at the background, " +
             std::to_string(free_idx.size()) + " free");
    } else {
        for (size_t i = 0; i < nc; ++i) free_idx.push_back((int64_t)i);
    }
    if (!a.porosity_free_zone_map.empty()) {
        std::vector<double> zone;
        if (!load_phi_map((fs::path(r.cs.folder) / a.porosity_free_zone_map).string(), nc, zone, err)) {
            info("zone map unreadable (" + err + ") -> zonation ignored");
        } else {
            std::vector<int64_t> z;
            for (int64_t i : free_idx)
                if (zone[i] < 0.5) z.push_back(i);
            free_idx.swap(z);
        }
    }
    if (r.check(sw2dg_assim_set_phi_transform(r.ctx, r.pmode, a.lambda_tanh, free_idx.data(),
                                              (int64_t)free_idx.size(), a.mask_wall_gradient ? 1 : 0),
                "set_phi_transform"))
        return 1;
    Problem P(r, alpha_reg, phi_start, free_idx, a.zonation);
    P.free_mask.assign(nc, 0);
    for (int64_t i : free_idx) P.free_mask[i] = 1;
    if (!a.porosity_ref_map.empty()) {
        if (!load_phi_map((fs::path(r.cs.folder) / a.porosity_ref_map).string(), nc, P.phi_ref, err)) {
            info("porosity_ref_map unreadable: sse_phi_ref column off");
            P.phi_ref.clear();
        }
    }
    {
        const fs::path dd = fs::path(r.cs.folder) / "descent";
        std::error_code ec;
        fs::remove_all(dd, ec);
        fs::create_directories(dd / "phi_iter", ec);
        P.descent_dir = dd.string();
    }
    // theta0 = m_from_phi of the start (phi itself for none / clamp; the
    // logistic's inverse for tanh; the normalised map's own value otherwise)
    std::vector<double> theta(P.size());
    auto m_from_phi = [&](double phi) {
        if (r.pmode == 2) {
            const double p = std::min(1.0 - 1e-12, std::max(1e-12, phi));
            return std::log(p / (1.0 - p)) / a.lambda_tanh;
        }
        return phi;
    };
    if (a.zonation)
        for (size_t z = 0; z < P.size(); ++z) theta[z] = m_from_phi(P.zone_value[z]);
    else
        for (size_t j = 0; j < free_idx.size(); ++j) theta[j] = m_from_phi(phi_start[free_idx[j]]);
    const std::string algo = a.minimize_algo;
    info("minimisation: " + algo + ", " + std::to_string(free_idx.size()) + " free cells / " +
         std::to_string(nc) + ", max_it " + std::to_string(a.max_it) + ", stride " +
         std::to_string(a.checkpoint_stride) + ", alpha_reg " + std::to_string(alpha_reg));
    double J_final = 0.0;
    if (algo == "gd") {
        J_final = run_gd(P, theta, a.max_it, 1e-6);
    } else if (algo == "adam") {
        J_final = run_adam(P, theta, a.max_it, a.adam_step_size, a.adam_safe);
    } else if (algo == "lbfgsb") {
        // probe gradient at the start fixes the truncation window, frozen for
        // the solve so every line-search evaluation covers the same segments
        std::vector<double> gp(theta.size());
        P.grad(theta.data(), gp.data());
        int64_t fks = 0, fkseg = 0;
        sw2dg_assim_truncation(r.ctx, &fks, &fkseg);
        sw2dg_assim_set_first_segment(r.ctx, std::max<int64_t>(0, fkseg));
        info("L-BFGS-B: truncation window frozen at segment " + std::to_string(std::max<int64_t>(0, fkseg)));
        sw2dh_lbfgsb_param lp;
        sw2dh_lbfgsb_default(&lp);
        lp.m = a.lbfgsb_m;
        lp.epsilon = a.lbfgsb_epsilon;
        lp.epsilon_rel = a.lbfgsb_epsilon;
        lp.max_iterations = (int32_t)a.max_it;
        lp.max_linesearch = a.lbfgsb_max_linesearch;
        std::vector<double> lb(theta.size(), a.phi_lower_bound), ub(theta.size(), a.phi_upper_bound);
        Objective ob{&P};
        int32_t it = 0, ne = 0, why = 0;
        double pg = 0.0;
        const int st = sw2dh_lbfgsb((int64_t)theta.size(), theta.data(), lb.data(), ub.data(), lbfgsb_obj,
                                    &ob, &lp, &J_final, &it, &ne, &pg, &why);
        sw2dg_assim_set_first_segment(r.ctx, -1);
        if (st) info(std::string("L-BFGS-B stopped: ") + sw2dh_lbfgsb_error() + " (last iterate kept)");
        std::ostringstream m;
        m << "[L-BFGS-B] done: J_final=" << std::scientific << std::setprecision(10) << J_final
          << "  iterations=" << it << "  ||proj grad||_inf=" << pg << "  n_eval(grad)=" << P.grad_calls;
        info(m.str());
    } else {
        J_final = run_bfgs(P, theta, a.max_it);
    }
    if (P.error) return 1;
    // final map: fixed cells = background, free cells = phi_correct(theta)
    std::vector<double> m_final = phi_start;
    for (size_t j = 0; j < free_idx.size(); ++j) m_final[free_idx[j]] = theta[a.zonation ? P.zone_of_free[j] : j];
    std::vector<double> phi_fin;
    r.phi_of(m_final, free_idx, phi_fin);
    std::vector<double> out = phi_start;
    for (int64_t i : free_idx) out[i] = phi_fin[i];
    double reg = 0.0;
    for (size_t i = 0; i < nc; ++i) reg += (out[i] - phi_start[i]) * (out[i] - phi_start[i]);
    const double J_data = J_final - 0.5 * alpha_reg * reg;
    {
        std::ostringstream m;
        m << "[Minim] alpha=" << alpha_reg << "  J_data=" << std::scientific << std::setprecision(10) << J_data
          << "  ||phi-phi_b||^2=" << reg << "  J_total=" << J_final;
        info(m.str());
    }
    write_sp_map((fs::path(r.cs.folder) / "optimized_porosity.txt").string(), out);
    const std::string sweep = (fs::path(r.cs.folder) / "alpha_sweep.csv").string();
    const bool exists = fs::exists(sweep);
    std::ofstream sf(sweep, std::ios::app);
    if (!exists) sf << "alpha,J_data,reg_norm2,J_total,n_grad\n";
    sf << std::setprecision(12) << alpha_reg << "," << J_data << "," << reg << "," << J_final << ","
       << P.grad_calls << "\n";
    return 0;
}

// one gradient + Taylor test (run_taylor_test_checkpointed with one top cell)
int mode_taylor(Run &r, std::vector<double> phi) {
    const size_t N = r.nc;
    if (r.check(sw2dg_assim_set_phi_transform(r.ctx, r.pmode, r.ad.lambda_tanh, nullptr, 0,
                                              r.ad.mask_wall_gradient ? 1 : 0),
                "set_phi_transform"))
        return 1;
    std::vector<double> g(N);
    double J = 0.0;
    int64_t n = 0;
    if (r.check(sw2dg_assim_gradient(r.ctx, phi.data(), nullptr, nullptr, &J, g.data(), nullptr, nullptr,
                                     nullptr, &n),
                "assim_gradient"))
        return 1;
    if (reached_max(r, n)) return 1;
    double l2 = 0.0, gmax = 0.0;
    size_t imax = 0, nz = 0;
    for (size_t i = 0; i < N; ++i) {
        l2 += g[i] * g[i];
        nz += g[i] != 0.0;
        if (std::fabs(g[i]) > std::fabs(gmax)) gmax = g[i], imax = i;
    }
    l2 = std::sqrt(l2);
    {
        std::ostringstream m;
        m << "=== GRADIENT CHECKPOINTED === J = " << std::scientific << std::setprecision(10) << J
          << "  ||dJ/dPhi||_2 = " << l2 << "  non-zero " << nz << " / " << N << "  max dJ/dphi[" << imax
          << "] = " << gmax;
        info(m.str());
    }
    if (nz > 0) {
        std::vector<std::pair<double, size_t>> srt;
        for (size_t i = 0; i < N; ++i)
            if (g[i] != 0.0) srt.emplace_back(std::fabs(g[i]), i);
        std::sort(srt.rbegin(), srt.rend());
        std::vector<size_t> cells = {srt[0].second};
        if (srt.size() > 2) cells.push_back(srt[srt.size() / 2].second);
        double J0 = 0.0;
        if (r.check(sw2dg_assim_cost(r.ctx, phi.data(), nullptr, nullptr, &J0), "assim_cost")) return 1;
        std::ofstream csv("taylor_test_checkpointed.csv");
        csv << "cell_id,grad_enzyme,epsilon,J_traj_pert,fd_grad,ratio\n" << std::scientific
            << std::setprecision(10);
        std::vector<double> pp = phi;
        for (size_t c : cells) {
            const double gi = g[c], p0 = phi[c];
            const double sgn = p0 >= 1.0 - 1e-8 ? -1.0 : 1.0;
            double eps = 1e-2;
            bool ok = false;
            for (int k = 0; k < 7; ++k, eps *= 0.1) {
                pp[c] = p0 + sgn * eps;
                double Jp = 0.0;
                if (r.check(sw2dg_assim_cost(r.ctx, pp.data(), nullptr, nullptr, &Jp), "assim_cost")) return 1;
                const double fd = (Jp - J0) / (sgn * eps);
                const double ratio = std::fabs(gi) > 1e-40 ? fd / gi : 0.0;
                if (k > 0 && std::fabs(ratio - 1.0) < 0.05) ok = true;
                csv << c << "," << gi << "," << eps << "," << Jp << "," << fd << "," << ratio << "\n";
                pp[c] = p0;
            }
            info("Taylor cell " + std::to_string(c) + (ok ? ": GRADIENT CORRECT" : ": GRADIENT SUSPECT"));
        }
        const double eps0 = 0.01 * l2 / srt[0].first;
        double eps = eps0;
        for (int k = 0; k < 7; ++k, eps *= 0.1) {
            double dga = 0.0;
            for (size_t i = 0; i < N; ++i) {
                const double pn = std::min(1.0, std::max(0.0, phi[i] + eps * (g[i] / l2)));
                dga += g[i] * (pn - phi[i]);
                pp[i] = pn;
            }
            double Jp = 0.0;
            if (r.check(sw2dg_assim_cost(r.ctx, pp.data(), nullptr, nullptr, &Jp), "assim_cost")) return 1;
            const double dJ = Jp - J0;
            const double ratio = std::fabs(dga) > 1e-30 ? dJ / dga : 0.0;
            const double fd = eps > 1e-300 ? dJ / eps : 0.0;
            csv << "global," << l2 << "," << eps << "," << Jp << "," << fd << "," << ratio << "\n";
            pp = phi;
        }
        info("[Taylor] results in taylor_test_checkpointed.csv");
    }
    std::ofstream out((fs::path(r.cs.folder) / "gradient_porosity.csv").string());
    out << "cell_id,phi,gradient\n";
    for (size_t i = 0; i < N; ++i) out << i << "," << phi[i] << "," << g[i] << "\n";
    info("gradient written to gradient_porosity.csv");
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2 || argc > 5) {
        std::cerr << "sw2d_enzyme_gpu usage:\n"
                  << "  sw2d_enzyme_gpu <folder> [mode] [phi_init.csv] [alpha_reg]\n"
                  << "    mode : (none/forward) | taylor | gen-ref | gen-ref-e | minimize\n";
        return EXIT_FAILURE;
    }
    if (!std::strcmp(argv[1], "-h") || !std::strcmp(argv[1], "--help")) {
        std::cout << "sw2d_enzyme_gpu: porosity assimilation (4D-Var) on an MI355X GPU\n";
        return EXIT_SUCCESS;
    }
    if (!std::strcmp(argv[1], "-v") || !std::strcmp(argv[1], "--version")) {
        std::cout << sw2dg_version() << std::endl;
        return EXIT_SUCCESS;
    }
    const std::string folder = argv[1];
    if (!fs::exists(folder)) return fail("Folder " + folder + " does not exist");
    const std::string mode = argc >= 3 ? argv[2] : "";
    if (mode != "" && mode != "forward" && mode != "gen-ref-e" && mode != "taylor" && mode != "minimize" &&
        mode != "gen-ref")
        return fail("unknown mode " + mode);
    std::error_code ec;
    fs::create_directories(fs::path(folder) / "log", ec);
    g_log.open((fs::path(folder) / "log" / "console_Solver.log").string());
    {
        std::string cmd;
        for (int i = 0; i < argc; ++i) cmd += (i ? " " : "") + std::string(argv[i]);
        info("Command line:" + cmd);
    }
    if (mode == "gen-ref")
        return fail("gen-ref: the reference's classic generation stores no reference map (its native "
                    "time loop never writes one, so it ends with 'no reference generated'); use gen-ref-e");
    Run r;
    std::string err;
    if (!sw2d::load_case(folder, r.cs, err)) return fail(err);
    if (!read_ad(r.cs.cfg, r.ad, err)) return fail(err);
    r.nc = r.cs.g.n_cell();
    r.pmode = phi_mode(r.ad.phi_transform);
    if (r.ad.dt_map_ad <= 0.0) r.ad.dt_map_ad = r.cs.params().dt_map;  // resolved to dt_map when unset
    {
        auto it = r.cs.cfg.gpu.find("gpu_max_steps");
        if (it != r.cs.cfg.gpu.end()) r.max_steps = std::max<int64_t>(1, std::stoll(it->second));
    }
    if (sw2d::make_ctx(r.cs, false, &r.ctx, err)) return fail(err);
    const auto t0w = std::chrono::steady_clock::now();
    const sw2dg_params p = r.cs.params();
    // porosity override: ref map for gen-ref-e, background otherwise (minimize reads its own)
    std::vector<double> phi = r.cs.phi;
    {
        std::string ov = mode == "gen-ref-e" ? r.ad.porosity_ref_map
                                             : (mode != "minimize" ? r.ad.porosity_ebauche_map : "");
        if (!ov.empty() && !load_phi_map((fs::path(folder) / ov).string(), r.nc, phi, err))
            return fail(err);
    }
    int rc = 0;
    if (mode == "gen-ref-e") {
        if (r.ad.dt_map_ad <= kSmall) return fail("gen-ref-e: dt_map_ad invalid");
        rc = dump_forward(r, phi, r.ad.dt_map_ad, r.ad.t_state_assimilation, "reference", true);
    } else if (mode == "" || mode == "forward") {
        const double dt = p.dt_map > kSmall ? p.dt_map : p.dt_max;
        rc = dump_forward(r, phi, dt, p.t0, "output", false);
    } else {
        // observations, cost weights, window
        Observations o;
        const bool gauges = r.ad.obs_type == "gauges";
        if (gauges ? !load_gauges(r.cs, r.ad.t_state_assimilation, p.t_max, o, err)
                   : !load_frames(folder, r.nc, r.ad.t_state_assimilation, p.t_max, o, err))
            return fail(err);
        if (!gauges) {
            o.wh = cost_weight(r.ad.h_weight, folder, r.nc);
            o.wq = cost_weight(r.ad.q_weight, folder, r.nc);
        }
        info(std::to_string(o.t.size()) + " observation(s) loaded (" + (gauges ? "gauges" : "frames") + ")");
        double tw = window_start(r, {});
        if (r.ad.windowed_assimilation && r.ad.assimilation_window_steps > 0) {
            // the window counts steps of the recorded run: a first dry run fixes it
            if (setup_assim(r, o, -1e300)) return 1;
            int64_t n = 0;
            if (r.check(sw2dg_assim_record(r.ctx, phi.data(), nullptr, nullptr, &n), "assim_record")) return 1;
            std::vector<double> ts(n + 1), dts(n), bcs(n * std::max<size_t>(r.cs.bc.codes.size(), 1));
            std::vector<int32_t> tr(n);
            if (r.check(sw2dg_assim_get_record(r.ctx, ts.data(), dts.data(), bcs.data(), tr.data()),
                        "get_record"))
                return 1;
            tw = window_start(r, ts);
        }
        if (setup_assim(r, o, tw)) return 1;
        if (mode == "taylor") {
            rc = mode_taylor(r, phi);
        } else {
            std::string path = argc >= 4 ? std::string(argv[3])
                               : !r.ad.porosity_ebauche_map.empty()
                                   ? (fs::path(folder) / r.ad.porosity_ebauche_map).string()
                                   : (fs::path(folder) / "phi_init_ones.csv").string();
            const double alpha = argc >= 5 ? std::atof(argv[4]) : r.ad.alpha_reg;
            rc = mode_minimize(r, path, alpha);
        }
    }
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0w).count();
    info("sw2d_enzyme_gpu " + std::string(sw2dg_version()) + ": mode " + (mode.empty() ? "forward" : mode) +
         ", wall " + std::to_string(wall) + " s");
    return rc;
}
