// sw2d_case.cpp -- case-folder formats and synthetic meshes (see sw2d_case.hpp).
#include "sw2d_case.hpp"

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <random>
#include <set>
#include <sstream>
#include <unordered_map>

namespace sw2d {

namespace {

std::string lower(std::string s) {
    for (auto &ch : s) ch = (char)std::tolower((unsigned char)ch);
    return s;
}

std::string trim(const std::string &s) {
    const size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    const size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}

std::vector<std::string> split_ws(const std::string &s) {
    std::vector<std::string> out;
    std::istringstream is(s);
    std::string tok;
    while (is >> tok) out.push_back(tok);
    return out;
}

bool to_double(const std::string &s, double &v) {
    char *end = nullptr;
    v = std::strtod(s.c_str(), &end);
    return end && end != s.c_str() && *end == '\0';
}

// reads one line, stripping a trailing '\r' (swIO::readLine)
bool get_line(std::istream &is, std::string &line) {
    if (!std::getline(is, line)) return false;
    if (!line.empty() && line.back() == '\r') line.pop_back();
    return true;
}

template <class T>
bool rd(std::istream &is, T &v) {
    is.read(reinterpret_cast<char *>(&v), sizeof(T));
    return (bool)is;
}
template <class T>
void wr(std::ostream &os, const T &v) {
    os.write(reinterpret_cast<const char *>(&v), sizeof(T));
}

std::string fmt_e(double v, int prec) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.*e", prec, v);
    return buf;
}

}  // namespace

// ---------------------------------------------------------------------------
// geometry.sw2d v2
// ---------------------------------------------------------------------------
bool read_geometry(const std::string &path, Geometry &g, std::string &err) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        err = "The file does not exist " + path;
        return false;
    }
    uint64_t vlen = 0;
    if (!rd(f, vlen) || vlen > 64) {
        err = "bad geometry header";
        return false;
    }
    std::string ver(vlen, '\0');
    f.read(&ver[0], (std::streamsize)vlen);
    if (ver != "2") {
        err = "Unknown version number " + ver;
        return false;
    }
    uint64_t nc, nb, ni, nn, nbd;
    if (!rd(f, nc) || !rd(f, nb) || !rd(f, ni) || !rd(f, nn) || !rd(f, nbd)) {
        err = "truncated geometry header";
        return false;
    }
    g = Geometry();
    g.cell_sms.resize(nc);
    g.cell_area.resize(nc);
    g.cell_perim.resize(nc);
    g.cell_x.resize(nc);
    g.cell_y.resize(nc);
    g.cell_zb.resize(nc);
    g.cell_nodes.resize(nc);
    g.cell_bnd.resize(nc);
    g.cell_int.resize(nc);
    g.cell_neigh.resize(nc);
    for (uint64_t n = 0; n < nc; ++n) {
        uint64_t idx, sms, nnode, nbe, nie, nne;
        rd(f, idx);
        rd(f, sms);
        rd(f, nnode);
        rd(f, nbe);
        rd(f, nie);
        rd(f, nne);
        rd(f, g.cell_area[n]);
        rd(f, g.cell_perim[n]);
        rd(f, g.cell_x[n]);
        rd(f, g.cell_y[n]);
        if (!rd(f, g.cell_zb[n])) {
            err = "truncated cell section";
            return false;
        }
        if (idx != n) {
            err = "cell indices must be contiguous from 0";
            return false;
        }
        g.cell_sms[n] = sms;
        g.cell_nodes[n].resize(nnode);
        g.cell_bnd[n].resize(nbe);
        g.cell_int[n].resize(nie);
        g.cell_neigh[n].resize(nne);
    }
    g.node_sms.resize(nn);
    g.node_x.resize(nn);
    g.node_y.resize(nn);
    g.node_z.resize(nn);
    g.node_cells.resize(nn);
    g.node_nedge.resize(nn);
    for (uint64_t p = 0; p < nn; ++p) {
        uint64_t id, sms, ncell, nedge;
        rd(f, id);
        rd(f, sms);
        rd(f, g.node_x[p]);
        rd(f, g.node_y[p]);
        rd(f, g.node_z[p]);
        rd(f, ncell);
        if (!rd(f, nedge)) {
            err = "truncated node section";
            return false;
        }
        g.node_sms[p] = sms;
        g.node_cells[p].resize(ncell);
        g.node_nedge[p] = nedge;
    }
    auto resize_b = [&](size_t n) {
        g.bnd_n1.resize(n);
        g.bnd_n2.resize(n);
        g.bnd_cell.resize(n);
        g.bnd_xm.resize(n);
        g.bnd_ym.resize(n);
        g.bnd_zm.resize(n);
        g.bnd_dnorm.resize(n);
        g.bnd_len.resize(n);
        g.bnd_nx.resize(n);
        g.bnd_ny.resize(n);
        g.bnd_code.resize(n);
    };
    resize_b(nb);
    for (uint64_t i = 0; i < nb; ++i) {
        uint64_t idx;
        rd(f, idx);
        rd(f, g.bnd_n1[i]);
        rd(f, g.bnd_n2[i]);
        rd(f, g.bnd_xm[i]);
        rd(f, g.bnd_ym[i]);
        rd(f, g.bnd_zm[i]);
        rd(f, g.bnd_dnorm[i]);
        rd(f, g.bnd_len[i]);
        rd(f, g.bnd_nx[i]);
        rd(f, g.bnd_ny[i]);
        rd(f, g.bnd_code[i]);
        if (!rd(f, g.bnd_cell[i]) || g.bnd_cell[i] >= nc) {
            err = "bad boundary edge section";
            return false;
        }
    }
    g.int_n1.resize(ni);
    g.int_n2.resize(ni);
    g.int_L.resize(ni);
    g.int_R.resize(ni);
    g.int_xm.resize(ni);
    g.int_ym.resize(ni);
    g.int_zm.resize(ni);
    g.int_dnorm.resize(ni);
    g.int_len.resize(ni);
    g.int_nx.resize(ni);
    g.int_ny.resize(ni);
    for (uint64_t i = 0; i < ni; ++i) {
        uint64_t idx;
        rd(f, idx);
        rd(f, g.int_n1[i]);
        rd(f, g.int_n2[i]);
        rd(f, g.int_xm[i]);
        rd(f, g.int_ym[i]);
        rd(f, g.int_zm[i]);
        rd(f, g.int_dnorm[i]);
        rd(f, g.int_len[i]);
        rd(f, g.int_nx[i]);
        rd(f, g.int_ny[i]);
        rd(f, g.int_L[i]);
        if (!rd(f, g.int_R[i]) || g.int_L[i] >= nc || g.int_R[i] >= nc) {
            err = "bad internal edge section";
            return false;
        }
    }
    g.boundaries.resize(nbd);
    for (uint64_t k = 0; k < nbd; ++k) {
        uint64_t sz;
        rd(f, sz);
        g.boundaries[k].resize(sz);
        for (uint64_t i = 0; i < sz; ++i) rd(f, g.boundaries[k][i]);
    }
    g.boundary_codes.resize(nbd);
    for (uint64_t k = 0; k < nbd; ++k) rd(f, g.boundary_codes[k]);
    for (uint64_t n = 0; n < nc; ++n) {
        for (auto &v : g.cell_nodes[n]) rd(f, v);
        for (auto &v : g.cell_bnd[n]) rd(f, v);
        for (auto &v : g.cell_int[n]) rd(f, v);
        for (auto &v : g.cell_neigh[n]) rd(f, v);
    }
    if (!f) {
        err = "truncated cell links";
        return false;
    }
    // node -> cell links: the reference reader sizes them to 0 and reads
    // nothing; read them when present, tolerate their absence.
    for (uint64_t p = 0; p < nn; ++p)
        for (auto &v : g.node_cells[p])
            if (!rd(f, v)) v = 0;
    return true;
}

bool write_geometry(const std::string &path, const Geometry &g, std::string &err) {
    std::ofstream f(path, std::ios::binary);
    if (!f) {
        err = "Unable to open file " + path;
        return false;
    }
    const std::string ver = "2";
    wr(f, (uint64_t)ver.size());
    f.write(ver.data(), (std::streamsize)ver.size());
    wr(f, (uint64_t)g.n_cell());
    wr(f, (uint64_t)g.n_bnd());
    wr(f, (uint64_t)g.n_int());
    wr(f, (uint64_t)g.n_node());
    wr(f, (uint64_t)g.boundaries.size());
    for (size_t n = 0; n < g.n_cell(); ++n) {
        wr(f, (uint64_t)n);
        wr(f, g.cell_sms[n]);
        wr(f, (uint64_t)g.cell_nodes[n].size());
        wr(f, (uint64_t)g.cell_bnd[n].size());
        wr(f, (uint64_t)g.cell_int[n].size());
        wr(f, (uint64_t)g.cell_neigh[n].size());
        wr(f, g.cell_area[n]);
        wr(f, g.cell_perim[n]);
        wr(f, g.cell_x[n]);
        wr(f, g.cell_y[n]);
        wr(f, g.cell_zb[n]);
    }
    for (size_t p = 0; p < g.n_node(); ++p) {
        wr(f, (uint64_t)p);
        wr(f, g.node_sms[p]);
        wr(f, g.node_x[p]);
        wr(f, g.node_y[p]);
        wr(f, g.node_z[p]);
        wr(f, (uint64_t)g.node_cells[p].size());
        wr(f, g.node_nedge[p]);
    }
    for (size_t i = 0; i < g.n_bnd(); ++i) {
        wr(f, (uint64_t)i);
        wr(f, g.bnd_n1[i]);
        wr(f, g.bnd_n2[i]);
        wr(f, g.bnd_xm[i]);
        wr(f, g.bnd_ym[i]);
        wr(f, g.bnd_zm[i]);
        wr(f, g.bnd_dnorm[i]);
        wr(f, g.bnd_len[i]);
        wr(f, g.bnd_nx[i]);
        wr(f, g.bnd_ny[i]);
        wr(f, g.bnd_code[i]);
        wr(f, g.bnd_cell[i]);
    }
    for (size_t i = 0; i < g.n_int(); ++i) {
        wr(f, (uint64_t)i);
        wr(f, g.int_n1[i]);
        wr(f, g.int_n2[i]);
        wr(f, g.int_xm[i]);
        wr(f, g.int_ym[i]);
        wr(f, g.int_zm[i]);
        wr(f, g.int_dnorm[i]);
        wr(f, g.int_len[i]);
        wr(f, g.int_nx[i]);
        wr(f, g.int_ny[i]);
        wr(f, g.int_L[i]);
        wr(f, g.int_R[i]);
    }
    for (const auto &b : g.boundaries) {
        wr(f, (uint64_t)b.size());
        for (uint64_t e : b) wr(f, e);
    }
    for (int32_t c : g.boundary_codes) wr(f, c);
    for (size_t n = 0; n < g.n_cell(); ++n) {
        for (uint64_t v : g.cell_nodes[n]) wr(f, v);
        for (uint64_t v : g.cell_bnd[n]) wr(f, v);
        for (uint64_t v : g.cell_int[n]) wr(f, v);
        for (uint64_t v : g.cell_neigh[n]) wr(f, v);
    }
    for (size_t p = 0; p < g.n_node(); ++p)
        for (uint64_t v : g.node_cells[p]) wr(f, v);
    if (!f) {
        err = "write failed " + path;
        return false;
    }
    return true;
}

// Geometry from nodes + cells.  Formulae follow the reference geometry classes:
// area by the shoelace sum (swCell.cpp:85-101), centre and zb as vertex means
// (swCell.cpp:103-125), edge length / normal / orientation (swEdgeIntern.cpp:
// 61-112, swEdgeBoundary.cpp:60-107).
void build_geometry(const std::vector<double> &x, const std::vector<double> &y,
                    const std::vector<double> &z, const std::vector<std::vector<uint64_t>> &cells,
                    BcCodeFn bc_code, void *user, Geometry &g) {
    g = Geometry();
    const size_t nc = cells.size(), nn = x.size();
    g.node_x = x;
    g.node_y = y;
    g.node_z = z;
    g.node_sms.resize(nn);
    for (size_t p = 0; p < nn; ++p) g.node_sms[p] = p + 1;
    g.node_cells.assign(nn, {});
    g.node_nedge.assign(nn, 0);
    g.cell_sms.resize(nc);
    g.cell_area.resize(nc);
    g.cell_perim.resize(nc);
    g.cell_x.resize(nc);
    g.cell_y.resize(nc);
    g.cell_zb.resize(nc);
    g.cell_nodes = cells;
    g.cell_bnd.assign(nc, {});
    g.cell_int.assign(nc, {});
    g.cell_neigh.assign(nc, {});
    for (size_t c = 0; c < nc; ++c) {
        const auto &nd = cells[c];
        const size_t k = nd.size();
        double a = 0, xg = 0, yg = 0, zg = 0, per = 0;
        for (size_t i = 0; i < k; ++i) {
            const uint64_t p = nd[i], q = nd[(i + 1) % k];
            a += x[p] * y[q] - x[q] * y[p];
            per += std::sqrt(std::pow(x[p] - x[q], 2) + std::pow(y[p] - y[q], 2));
            xg += x[p];
            yg += y[p];
            zg += z[p];
            g.node_cells[p].push_back(c);
        }
        g.cell_sms[c] = c + 1;
        g.cell_area[c] = std::fabs(a) / 2;
        g.cell_perim[c] = per;
        g.cell_x[c] = xg / (double)k;
        g.cell_y[c] = yg / (double)k;
        g.cell_zb[c] = zg / (double)k;
    }
    // unique edges in order of discovery
    std::unordered_map<uint64_t, int64_t> emap;
    emap.reserve(nc * 3);
    struct E {
        uint64_t a, b;
        int64_t c0, c1;
    };
    std::vector<E> edges;
    edges.reserve(nc * 2);
    for (size_t c = 0; c < nc; ++c) {
        const auto &nd = cells[c];
        for (size_t i = 0; i < nd.size(); ++i) {
            uint64_t a = nd[i], b = nd[(i + 1) % nd.size()];
            const uint64_t key = std::min(a, b) * (uint64_t)nn + std::max(a, b);
            auto it = emap.find(key);
            if (it == emap.end()) {
                emap.emplace(key, (int64_t)edges.size());
                edges.push_back({a, b, (int64_t)c, -1});
            } else {
                edges[it->second].c1 = (int64_t)c;
            }
        }
    }
    std::map<int32_t, size_t> code_slot;
    for (const E &e : edges) {
        g.node_nedge[e.a]++;
        g.node_nedge[e.b]++;
        const double xa = x[e.a], ya = y[e.a], xb = x[e.b], yb = y[e.b];
        const double len = std::sqrt(std::pow(xa - xb, 2) + std::pow(ya - yb, 2));
        double nxv = (yb - ya) / len;
        double nyv = -(xb - xa) / len;
        uint64_t n1 = e.a, n2 = e.b;
        const double xm = 0.5 * (xa + xb), ym = 0.5 * (ya + yb), zm = 0.5 * (z[e.a] + z[e.b]);
        if (e.c1 < 0) {
            const size_t c = (size_t)e.c0;
            if (nxv * (g.cell_x[c] - xm) + nyv * (g.cell_y[c] - ym) < 0) {
                nxv = -nxv;
                nyv = -nyv;
                std::swap(n1, n2);
            }
            const size_t i = g.bnd_len.size();
            const int32_t code = bc_code ? bc_code(xm, ym, user) : 1;
            g.bnd_n1.push_back(n1);
            g.bnd_n2.push_back(n2);
            g.bnd_cell.push_back(c);
            g.bnd_xm.push_back(xm);
            g.bnd_ym.push_back(ym);
            g.bnd_zm.push_back(zm);
            g.bnd_dnorm.push_back(std::fabs(nxv * g.cell_x[c] + nyv * g.cell_y[c]));
            g.bnd_len.push_back(len);
            g.bnd_nx.push_back(nxv);
            g.bnd_ny.push_back(nyv);
            g.bnd_code.push_back(code);
            g.cell_bnd[c].push_back(i);
            auto it = code_slot.find(code);
            if (it == code_slot.end()) {
                code_slot.emplace(code, g.boundaries.size());
                g.boundaries.push_back({i});
                g.boundary_codes.push_back(code);
            } else {
                g.boundaries[it->second].push_back(i);
            }
        } else {
            const size_t L = (size_t)e.c0, R = (size_t)e.c1;
            if (nxv * (g.cell_x[R] - g.cell_x[L]) + nyv * (g.cell_y[R] - g.cell_y[L]) < 0) {
                nxv = -nxv;
                nyv = -nyv;
                std::swap(n1, n2);
            }
            const size_t i = g.int_len.size();
            g.int_n1.push_back(n1);
            g.int_n2.push_back(n2);
            g.int_L.push_back(L);
            g.int_R.push_back(R);
            g.int_xm.push_back(xm);
            g.int_ym.push_back(ym);
            g.int_zm.push_back(zm);
            g.int_dnorm.push_back(std::fabs(nxv * (g.cell_x[R] - g.cell_x[L]) +
                                            nyv * (g.cell_y[R] - g.cell_y[L])));
            g.int_len.push_back(len);
            g.int_nx.push_back(nxv);
            g.int_ny.push_back(nyv);
            g.cell_int[L].push_back(i);
            g.cell_int[R].push_back(i);
            g.cell_neigh[L].push_back(R);
            g.cell_neigh[R].push_back(L);
        }
    }
}

void set_bnd_codes(Geometry &g, const std::vector<int32_t> &codes) {
    g.bnd_code = codes;
    g.boundaries.clear();
    g.boundary_codes.clear();
    std::map<int32_t, size_t> slot;
    for (size_t i = 0; i < codes.size(); ++i) {
        auto it = slot.find(codes[i]);
        if (it == slot.end()) {
            slot.emplace(codes[i], g.boundaries.size());
            g.boundaries.push_back({(uint64_t)i});
            g.boundary_codes.push_back(codes[i]);
        } else {
            g.boundaries[it->second].push_back(i);
        }
    }
}

void set_node_z(Geometry &g, const std::vector<double> &z) {
    g.node_z = z;
    for (size_t c = 0; c < g.n_cell(); ++c) {
        double zg = 0;
        for (uint64_t p : g.cell_nodes[c]) zg += z[p];
        g.cell_zb[c] = zg / (double)g.cell_nodes[c].size();
    }
    for (size_t i = 0; i < g.n_bnd(); ++i) g.bnd_zm[i] = 0.5 * (z[g.bnd_n1[i]] + z[g.bnd_n2[i]]);
    for (size_t i = 0; i < g.n_int(); ++i) g.int_zm[i] = 0.5 * (z[g.int_n1[i]] + z[g.int_n2[i]]);
}

// ---------------------------------------------------------------------------
// maps
// ---------------------------------------------------------------------------
bool read_map(const std::string &path, size_t n_params, size_t n_records,
              std::vector<std::string> &names, std::vector<std::vector<double>> &cols,
              std::string &err) {
    std::ifstream f(path);
    if (!f) {
        err = "Unable to open file " + path;
        return false;
    }
    std::string line;
    get_line(f, line);
    auto tok = split_ws(line);
    if (tok.size() != 2) {
        err = path + ": line 1 is misformatted";
        return false;
    }
    const int unif = std::atoi(tok[1].c_str());
    if (unif != 0 && unif != 1) {
        err = path + ": Incorrect value for 'Unif' parameter";
        return false;
    }
    get_line(f, line);  // ==== Default Param
    get_line(f, line);
    names = split_ws(line);
    if (names.size() != n_params) {
        err = path + ": header expects " + std::to_string(n_params) + " parameters";
        return false;
    }
    cols.assign(n_params, {});
    if (unif == 1) {
        get_line(f, line);
        tok = split_ws(line);
        if (tok.size() != n_params) {
            err = path + ": uniform line is misformatted";
            return false;
        }
        for (size_t k = 0; k < n_params; ++k) {
            double v;
            if (!to_double(tok[k], v)) {
                err = path + ": format conversion";
                return false;
            }
            cols[k].assign(n_records, v);
        }
        return true;
    }
    get_line(f, line);  // default values
    get_line(f, line);  // ==== Distrib
    for (size_t k = 0; k < n_params; ++k) cols[k].reserve(n_records);
    for (size_t r = 0; r < n_records; ++r) {
        if (!get_line(f, line)) {
            err = path + ": Not enough values in file";
            return false;
        }
        tok = split_ws(line);
        if (tok.size() != n_params) {
            err = path + ": line " + std::to_string(r + 6) + " is misformatted";
            return false;
        }
        for (size_t k = 0; k < n_params; ++k) {
            double v;
            if (!to_double(tok[k], v)) {
                err = path + ": format conversion";
                return false;
            }
            cols[k].push_back(v);
        }
    }
    return true;
}

bool write_map_uniform(const std::string &path, const std::vector<std::string> &names,
                       const std::vector<double> &vals, std::string &err) {
    std::ofstream f(path);
    if (!f) {
        err = "Unable to open file " + path;
        return false;
    }
    f << "Unif 1\n==== Default Param\n";
    for (size_t k = 0; k < names.size(); ++k) f << (k ? "\t" : "") << names[k];
    f << "\n";
    for (size_t k = 0; k < vals.size(); ++k) f << (k ? "\t" : "") << fmt_e(vals[k], 17);
    f << "\n";
    return (bool)f;
}

bool write_map_distrib(const std::string &path, const std::vector<std::string> &names,
                       const std::vector<double> &defaults,
                       const std::vector<std::vector<double>> &cols, std::string &err) {
    std::ofstream f(path);
    if (!f) {
        err = "Unable to open file " + path;
        return false;
    }
    f << "Unif 0\n==== Default Param\n";
    for (size_t k = 0; k < names.size(); ++k) f << (k ? "\t" : "") << names[k];
    f << "\n";
    for (size_t k = 0; k < defaults.size(); ++k) f << (k ? "\t" : "") << fmt_e(defaults[k], 17);
    f << "\n==== Distrib\n";
    const size_t n = cols.empty() ? 0 : cols[0].size();
    std::string row;
    for (size_t r = 0; r < n; ++r) {
        row.clear();
        for (size_t k = 0; k < cols.size(); ++k) {
            if (k) row += '\t';
            row += fmt_e(cols[k][r], 17);
        }
        row += '\n';
        f << row;
    }
    return (bool)f;
}

// ---------------------------------------------------------------------------
// boundary time series
// ---------------------------------------------------------------------------
bool read_bc_series(const std::string &path, BcSeries &bc, std::string &err) {
    std::ifstream f(path);
    if (!f) {
        err = path + " cannot be opened";
        return false;
    }
    std::string line;
    get_line(f, line);
    get_line(f, line);
    get_line(f, line);
    const long ns = std::atol(trim(line).c_str());
    if (ns < 0) {
        err = path + ": bad series count";
        return false;
    }
    get_line(f, line);
    get_line(f, line);
    auto codes = split_ws(line);
    if ((long)codes.size() != ns) {
        err = path + ": incorrect number of specified boundary conditions";
        return false;
    }
    bc = BcSeries();
    for (auto &c : codes) bc.codes.push_back((int32_t)std::atol(c.c_str()));
    get_line(f, line);
    get_line(f, line);
    bc.types = split_ws(line);
    if ((long)bc.types.size() != ns) {
        err = path + ": incorrect number of boundary types";
        return false;
    }
    get_line(f, line);
    get_line(f, line);
    bc.v.assign(ns, {});
    while (get_line(f, line)) {
        if (line.empty()) continue;
        auto tok = split_ws(line);
        if (tok.empty()) continue;
        if ((long)tok.size() < ns + 1) {
            err = path + ": line is misformatted";
            return false;
        }
        double t;
        if (!to_double(tok[0], t)) {
            err = path + ": format conversion";
            return false;
        }
        bc.t.push_back(t);
        for (long j = 0; j < ns; ++j) {
            double v;
            if (!to_double(tok[j + 1], v)) {
                err = path + ": format conversion";
                return false;
            }
            bc.v[j].push_back(v);
        }
    }
    if (bc.t.empty()) {
        err = "The file " + path + " does not contain any boundary condition values";
        return false;
    }
    return true;
}

bool read_scalar_series(const std::string &path, std::vector<double> &t,
                        std::vector<std::vector<double>> &v, std::string &err) {
    std::ifstream f(path);
    if (!f) {
        err = path + " cannot be opened";
        return false;
    }
    std::string line;
    get_line(f, line);
    get_line(f, line);
    get_line(f, line);
    const long ns = std::atol(trim(line).c_str());
    if (ns < 0) {
        err = path + ": bad series count";
        return false;
    }
    get_line(f, line);
    get_line(f, line);
    t.clear();
    v.assign(ns, {});
    while (get_line(f, line)) {
        auto tok = split_ws(line);
        if (tok.empty()) continue;
        if ((long)tok.size() < ns + 1) {
            err = path + ": line is misformatted";
            return false;
        }
        double x;
        if (!to_double(tok[0], x)) {
            err = path + ": format conversion";
            return false;
        }
        t.push_back(x);
        for (long j = 0; j < ns; ++j) {
            if (!to_double(tok[j + 1], x)) {
                err = path + ": format conversion";
                return false;
            }
            v[j].push_back(x);
        }
    }
    if (t.empty()) {
        err = "The file " + path + " does not contain any values";
        return false;
    }
    return true;
}

bool read_wind_series(const std::string &path, std::vector<double> &t,
                      std::vector<std::vector<double>> &norm,
                      std::vector<std::vector<double>> &angle_deg, std::string &err) {
    std::ifstream f(path);
    if (!f) {
        err = path + " cannot be opened";
        return false;
    }
    std::string line;
    get_line(f, line);
    get_line(f, line);
    get_line(f, line);
    const long ns = std::atol(trim(line).c_str());
    if (ns < 0) {
        err = path + ": bad station count";
        return false;
    }
    get_line(f, line);
    get_line(f, line);
    t.clear();
    norm.assign(ns, {});
    angle_deg.assign(ns, {});
    while (get_line(f, line)) {
        auto tok = split_ws(line);
        if (tok.empty()) continue;
        if ((long)tok.size() < 2 * ns + 1) {
            err = path + ": line is misformatted";
            return false;
        }
        double v;
        if (!to_double(tok[0], v)) {
            err = path + ": format conversion";
            return false;
        }
        t.push_back(v);
        for (long j = 0; j < ns; ++j) {
            double a, b;
            if (!to_double(tok[2 * j + 1], a) || !to_double(tok[2 * j + 2], b)) {
                err = path + ": format conversion";
                return false;
            }
            norm[j].push_back(a);
            angle_deg[j].push_back(b);
        }
    }
    if (t.empty()) {
        err = "The file " + path + " does not contain any wind values";
        return false;
    }
    return true;
}

bool write_bc_series(const std::string &path, const BcSeries &bc, std::string &err) {
    std::ofstream f(path);
    if (!f) {
        err = "Unable to open file " + path;
        return false;
    }
    f << "# boundary time series\n# number of series\n" << bc.codes.size() << "\n# bc codes\n";
    for (size_t k = 0; k < bc.codes.size(); ++k) f << (k ? "\t" : "") << bc.codes[k];
    f << "\n# bc types\n";
    for (size_t k = 0; k < bc.types.size(); ++k) f << (k ? "\t" : "") << bc.types[k];
    f << "\n# values\n# time";
    for (size_t k = 0; k < bc.codes.size(); ++k) f << "\t" << bc.codes[k];
    f << "\n";
    for (size_t r = 0; r < bc.t.size(); ++r) {
        f << fmt_e(bc.t[r], 17);
        for (size_t k = 0; k < bc.v.size(); ++k) f << "\t" << fmt_e(bc.v[k][r], 17);
        f << "\n";
    }
    return (bool)f;
}

// ---------------------------------------------------------------------------
// input.solv
// ---------------------------------------------------------------------------
bool read_solver_config(const std::string &path, SolverConfig &cfg, std::string &err) {
    static const std::set<std::string> ad = {
        "adam_safe", "adam_step_size", "alpha_l1", "alpha_phi_ref", "alpha_reg", "alpha_spatial",
        "assimilation_window_duration", "assimilation_window_steps", "checkpoint_stride",
        "cost_normalize", "dt_map_ad", "h_weight", "lambda_tanh", "lbfgsb_epsilon", "lbfgsb_m",
        "lbfgsb_max_linesearch", "mask_dry_cells", "mask_wall_gradient", "max_it",
        "minimize_algo", "n_assimilation_windows", "obs_type", "phi_lower_bound",
        "phi_transform", "phi_upper_bound", "porosity_ebauche_map", "porosity_free_zone_map",
        "porosity_ref_map", "q_weight", "t_state_assimilation", "windowed_assimilation",
        "zonation"};
    static const std::set<std::string> native = {
        "cfl_max", "contact_restitution", "contact_ext_margin", "ddp_chi_phi_max",
        "ddp_chi_theta_max", "ddp_ztab0_log", "div_corr", "dt_map", "dt_max", "dt_time_series",
        "fr_max", "gravity", "h_min", "hotstart_time", "log_write", "model_building",
        "model_friction", "model_hyperb", "model_infiltration", "model_logjam",
        "model_porosity_rainfall", "model_runoff", "model_singular_head_loss",
        "model_transport_reaction", "model_transport_diffusion", "model_wind",
        "model_momentum_diffusion", "model_momentum_dissipation", "momentum_diffusion_coefficient",
        "negative_depth_detection", "n_decimal", "n_it_div_max", "n_por_tab_max", "num_scheme",
        "output_format", "phi_min", "q_section", "read_wall_from_file", "store_flag", "t_0",
        "t_max", "topo_discretization", "transport_diffusion_coefficient", "transport_hydro",
        "v_max", "momentum_diff_max_iter", "momentum_diff_convergence_eps", "h_min_ddp"};
    std::ifstream f(path);
    if (!f) {
        err = "Unable to open: " + path;
        return false;
    }
    std::string line;
    size_t ln = 0;
    while (get_line(f, line)) {
        ++ln;
        if (line.empty() || line[0] == '#' || line[0] == '=') continue;
        std::string content = line;
        auto tok = split_ws(content);
        if (tok.empty()) continue;
        const std::string key = lower(tok[0]);
        const size_t hash = content.find('#');
        if (hash != std::string::npos) tok = split_ws(content.substr(0, hash));
        if (tok.size() != 2) {
            err = path + ": line " + std::to_string(ln) + " is misformatted";
            return false;
        }
        if (native.count(key)) {
            cfg.native[key] = tok[1];
        } else if (ad.count(key)) {
            cfg.assim[key] = tok[1];
        } else if (key.rfind("gpu_", 0) == 0) {
            cfg.gpu[key] = tok[1];
        } else {
            err = path + ": parameter not recognized: " + key;
            return false;
        }
    }
    auto num = [&](const char *k, double &v) {
        auto it = cfg.native.find(k);
        if (it != cfg.native.end()) v = std::strtod(it->second.c_str(), nullptr);
    };
    auto str = [&](const char *k, std::string &v) {
        auto it = cfg.native.find(k);
        if (it != cfg.native.end()) v = lower(it->second);
    };
    num("cfl_max", cfg.cfl_max);
    num("dt_map", cfg.dt_map);
    num("dt_max", cfg.dt_max);
    num("dt_time_series", cfg.dt_time_series);
    num("fr_max", cfg.fr_max);
    num("h_min", cfg.h_min);
    num("phi_min", cfg.phi_min);
    num("t_0", cfg.t0);
    num("t_max", cfg.t_max);
    num("v_max", cfg.v_max);
    double nd = cfg.n_decimal;
    num("n_decimal", nd);
    cfg.n_decimal = (int)nd;
    str("model_hyperb", cfg.model_hyp);
    str("model_friction", cfg.model_friction);
    str("model_singular_head_loss", cfg.model_singular_head_loss);
    str("num_scheme", cfg.num_scheme);
    str("topo_discretization", cfg.topo);
    str("store_flag", cfg.store_flag);
    str("negative_depth_detection", cfg.negative_depth_detection);
    str("div_corr", cfg.div_corr);
    str("model_runoff", cfg.model_runoff);
    str("model_infiltration", cfg.model_infiltration);
    str("model_wind", cfg.model_wind);
    str("model_momentum_diffusion", cfg.model_momentum_diffusion);
    str("model_porosity_rainfall", cfg.model_porosity_rainfall);
    return true;
}

// ---------------------------------------------------------------------------
// case folder
// ---------------------------------------------------------------------------
sw2dg_params Case::params() const {
    sw2dg_params p{};
    p.scheme = cfg.num_scheme == "muscl_evr" ? SW2DG_MUSCL_EVR : SW2DG_GODUNOV;
    p.topo = cfg.topo == "blending" ? SW2DG_TOPO_BLENDING : SW2DG_TOPO_DISCONTINUOUS;
    p.friction = cfg.model_friction == "manning"     ? SW2DG_FRIC_MANNING
                 : cfg.model_friction == "strickler" ? SW2DG_FRIC_STRICKLER
                 : cfg.model_friction == "chezy"     ? SW2DG_FRIC_CHEZY
                                                     : SW2DG_FRIC_NONE;
    p.head_loss = cfg.model_singular_head_loss != "none";
    p.neg_depth_detection = cfg.negative_depth_detection == "active";
    p.h_min = cfg.h_min;
    p.phi_min = cfg.phi_min;
    p.cfl_max = cfg.cfl_max;
    p.dt_max = cfg.dt_max;
    p.dt_map = cfg.dt_map;
    p.dt_time_series = cfg.dt_time_series;
    p.t0 = cfg.t0;
    p.t_max = cfg.t_max;
    p.v_max = cfg.v_max;
    p.fr_max = cfg.fr_max;
    return p;
}

sw2dg_mesh Case::mesh() const {
    sw2dg_mesh m{};
    m.n_cell = (int64_t)g.n_cell();
    m.n_bnd = (int64_t)g.n_bnd();
    m.n_int = (int64_t)g.n_int();
    m.cell_area = g.cell_area.data();
    m.cell_x = g.cell_x.data();
    m.cell_y = g.cell_y.data();
    m.cell_zb = g.cell_zb.data();
    m.bnd_cell = bcell64.data();
    m.bnd_len = g.bnd_len.data();
    m.bnd_nx = g.bnd_nx.data();
    m.bnd_ny = g.bnd_ny.data();
    m.bnd_series = bnd_series.data();
    m.int_L = L64.data();
    m.int_R = R64.data();
    m.int_len = g.int_len.data();
    m.int_nx = g.int_nx.data();
    m.int_ny = g.int_ny.data();
    m.int_xm = g.int_xm.data();
    m.int_ym = g.int_ym.data();
    return m;
}

std::vector<int32_t> Case::series_types() const {
    std::vector<int32_t> t;
    for (const auto &s : bc.types) {
        const std::string l = lower(s);
        t.push_back(l == "fr" ? SW2DG_BC_FR : l == "h" ? SW2DG_BC_H : l == "z" ? SW2DG_BC_Z
                    : l == "q" ? SW2DG_BC_Q : -1);
    }
    return t;
}

std::vector<double> Case::series_values() const {
    std::vector<double> v;
    for (const auto &col : bc.v) v.insert(v.end(), col.begin(), col.end());
    return v;
}

bool read_probes(const std::string &path, Probes &p, std::string &err) {
    p = Probes();
    std::ifstream f(path);
    if (!f) return true;  // no probe file: no probes
    std::string line;
    std::getline(f, line);  // header
    while (std::getline(f, line)) {
        line.erase(std::remove(line.begin(), line.end(), '\r'), line.end());
        if (line.empty()) break;
        // tokens between runs of white space; a leading run gives an empty first
        // token, a trailing run none (std::sregex_token_iterator with "\\s+")
        std::vector<std::string> tok;
        size_t i = 0;
        const auto ws = [](char ch) { return std::isspace((unsigned char)ch) != 0; };
        if (ws(line[0])) tok.emplace_back();
        while (i < line.size()) {
            while (i < line.size() && ws(line[i])) ++i;
            size_t j = i;
            while (j < line.size() && !ws(line[j])) ++j;
            if (j > i) tok.emplace_back(line.substr(i, j - i));
            i = j;
        }
        if (tok.size() != 3) {
            err = path + ": line is misformatted: " + line;
            return false;
        }
        char *end = nullptr;
        const double x = std::strtod(tok[0].c_str(), &end);
        const bool okx = end != tok[0].c_str();
        const double y = std::strtod(tok[1].c_str(), &end);
        if (!okx || end == tok[1].c_str()) {
            err = path + ": line is misformatted: format conversion: " + line;
            return false;
        }
        p.names.push_back(tok[2]);
        p.x.push_back(x);
        p.y.push_back(y);
    }
    return true;
}

int64_t locate_cell(const Geometry &g, double x, double y) {
    const double pi = 4.0 * std::atan(1.0);
    for (size_t c = 0; c < g.n_cell(); ++c) {
        const auto &nd = g.cell_nodes[c];
        const size_t nn = nd.size();
        if (nn == 0) continue;
        // outside the vertex bounding box the angle sum is 0 up to rounding,
        // so skipping such cells cannot change which cell is found first
        double x0 = g.node_x[nd[0]], x1 = x0, y0 = g.node_y[nd[0]], y1 = y0;
        for (size_t p = 1; p < nn; ++p) {
            x0 = std::min(x0, g.node_x[nd[p]]), x1 = std::max(x1, g.node_x[nd[p]]);
            y0 = std::min(y0, g.node_y[nd[p]]), y1 = std::max(y1, g.node_y[nd[p]]);
        }
        if (x < x0 || x > x1 || y < y0 || y > y1) continue;
        double theta = 0.0;
        for (size_t p = 0; p < nn; ++p) {
            const double xa = g.node_x[nd[p]], ya = g.node_y[nd[p]];
            const double xb = g.node_x[nd[(p + 1) % nn]], yb = g.node_y[nd[(p + 1) % nn]];
            const double cr = (yb - y) * (xa - x) - (xb - x) * (ya - y);
            const double dp = (xa - x) * (xb - x) + (ya - y) * (yb - y);
            const double na = std::max(std::sqrt((xa - x) * (xa - x) + (ya - y) * (ya - y)), 1e-16);
            const double nb = std::max(std::sqrt((xb - x) * (xb - x) + (yb - y) * (yb - y)), 1e-16);
            const double s = cr / na / nb, cs = dp / na / nb;
            double dth;
            if (cs > 0.0)
                dth = std::asin(std::min(s, 1.0));
            else if (s < 0.0)
                dth = -pi - std::asin(std::min(s, 1.0));
            else
                dth = pi - std::asin(std::min(s, 1.0));
            theta = theta + dth;
        }
        if (std::fabs(theta) > 6.0) return (int64_t)c;
    }
    return -1;
}

namespace {
std::string sci10(double v) {
    char b[64];
    std::snprintf(b, sizeof b, "%.10e", v);
    return b;
}
}  // namespace

std::string probe_header(const Case &c) {
    std::string s;
    for (size_t i = 0; i < c.probes.names.size(); ++i)
        s += c.probes.names[i] + "\t" + sci10(c.probes.x[i]) + "\t" + sci10(c.probes.y[i]) + "\n";
    s += "time\t";
    for (const auto &n : c.probes.names) s += n + "_zb\t" + n + "_h\t" + n + "_q\t" + n + "_r\t";
    return s + "\n";
}

std::string probe_line(const Case &c, double t, const std::vector<double> &h,
                       const std::vector<double> &qx, const std::vector<double> &qy) {
    std::string s = sci10(t) + "\t";
    for (int64_t k : c.probe_cell)
        s += sci10(c.g.cell_zb[k]) + "\t" + sci10(h[k]) + "\t" + sci10(qx[k]) + "\t" + sci10(qy[k]) +
             "\t";
    return s + "\n";
}

bool load_case(const std::string &folder, Case &c, std::string &err) {
    namespace fs = std::filesystem;
    c = Case();
    c.folder = folder;
    const std::string in = (fs::path(folder) / "input").string() + "/";
    if (!read_solver_config(in + "input.solv", c.cfg, err)) return false;
    if (c.cfg.model_hyp != "sp") {
        err = "only model_hyperb sp is handled by this solver (got " + c.cfg.model_hyp + ")";
        return false;
    }
    if (c.cfg.div_corr != "none") {
        err = "option not handled by the GPU solver: " + c.cfg.div_corr;
        return false;
    }
    // swSolverParam::setModelMomentumDiffusion (swSolverParam.cpp:1615-1635) accepts
    // any value and swSp::execute diffuses for anything but "none" (swSp.cpp:537)
    c.momentum_diffusion = c.cfg.model_momentum_diffusion != "none" ? 1 : 0;
    // runoff / infiltration models (swSolverParam.cpp:1486-1530, swSpSourceRunoffInfiltration.cpp)
    const std::string &ro = c.cfg.model_runoff, &inf = c.cfg.model_infiltration;
    if (ro == "scs" || inf == "h_proportional") {
        err = "model '" + (ro == "scs" ? ro : inf) + "' available only for industrial use (CEREG)";
        return false;
    }
    c.runoff_model = ro == "none" ? SW2DG_RUNOFF_NONE : ro == "full_rainfall" ? SW2DG_RUNOFF_FULL
                   : ro == "infiltration_coefficient" ? SW2DG_RUNOFF_INFIL_COEFF
                   : ro == "runoff_coefficient" ? SW2DG_RUNOFF_RUNOFF_COEFF : -1;
    if (c.runoff_model < 0) {
        err = "unallowed value for parameter 'model_runoff' " + ro;
        return false;
    }
    if (inf != "none" && inf != "horton") {
        err = "Infiltration model not handled = " + inf;
        return false;
    }
    c.infil_model = inf == "horton" ? SW2DG_INFIL_HORTON : SW2DG_INFIL_NONE;
    const std::string &pr = c.cfg.model_porosity_rainfall;
    if (pr != "everywhere" && pr != "water_only") {
        err = "unallowed value for parameter 'model_porosity_rainfall' " + pr;
        return false;
    }
    c.rain_water_only = pr == "water_only" ? 1 : 0;
    if (c.cfg.model_wind != "none" && c.cfg.model_wind != "smithbanke") {
        err = "Wind model unknown : " + c.cfg.model_wind;  // swSpSourceWind.cpp:37-41
        return false;
    }
    if (!read_geometry(in + "geometry.sw2d", c.g, err)) return false;
    const size_t nc = c.g.n_cell();
    std::vector<std::string> names;
    std::vector<std::vector<double>> cols;
    // initial conditions (swSp.cpp:1119-1224)
    if (!read_map(in + "initial_conditions_map.txt", 3, nc, names, cols, err)) return false;
    c.h.assign(nc, 0.0);
    c.qx.assign(nc, 0.0);
    c.qy.assign(nc, 0.0);
    auto col = [&](const char *n) -> int {
        for (size_t k = 0; k < names.size(); ++k)
            if (names[k] == n) return (int)k;
        return -1;
    };
    int k;
    bool found = false;
    if ((k = col("h")) >= 0) {
        c.h = cols[k];
        found = true;
    }
    if ((k = col("z")) >= 0) {
        for (size_t i = 0; i < nc; ++i) c.h[i] = std::max(cols[k][i] - c.g.cell_zb[i], 0.0);
        found = true;
    }
    if (!found) {
        err = "Could find neither h nor z in initial conditions";
        return false;
    }
    found = false;
    if ((k = col("q")) >= 0) {
        c.qx = cols[k];
        found = true;
    }
    if ((k = col("u")) >= 0) {
        for (size_t i = 0; i < nc; ++i) c.qx[i] = c.h[i] * cols[k][i];
        found = true;
    }
    if (!found) {
        err = "Could find neither q nor u in initial conditions";
        return false;
    }
    found = false;
    if ((k = col("r")) >= 0) {
        c.qy = cols[k];
        found = true;
    }
    if ((k = col("v")) >= 0) {
        for (size_t i = 0; i < nc; ++i) c.qy[i] = c.h[i] * cols[k][i];
        found = true;
    }
    if (!found) {
        err = "Could find neither r nor v in initial conditions";
        return false;
    }
    // boundary series (swParent.cpp:192-328)
    if (!read_bc_series(in + "hydro_boundary_time_series.txt", c.bc, err)) return false;
    c.bnd_series.resize(c.g.n_bnd());
    for (size_t e = 0; e < c.g.n_bnd(); ++e) {
        int s = -1;
        for (size_t j = 0; j < c.bc.codes.size(); ++j)
            if (c.bc.codes[j] == c.g.bnd_code[e]) {
                s = (int)j;
                break;
            }
        if (s < 0) {
            err = "Looking for bc code " + std::to_string(c.g.bnd_code[e]) + " not found";
            return false;
        }
        c.bnd_series[e] = s;
    }
    for (int32_t t : c.series_types())
        if (t < 0) {
            err = "BC type not handled";
            return false;
        }
    // Boussinesq and porosity (swSp.cpp:998-1021, 1250-1273)
    if (!read_map(in + "Boussinesq_map.txt", 1, nc, names, cols, err)) return false;
    c.beta = cols[0];
    if (!read_map(in + "SP_cell_porosity_map.txt", 1, nc, names, cols, err)) return false;
    c.phi = cols[0];
    // friction (swSp.cpp:1024-1072)
    const std::string fm = c.cfg.model_friction;
    if (fm != "none") {
        std::string file = fm == "chezy"       ? "friction_chezy_map.txt"
                           : fm == "manning"   ? "friction_manning_map.txt"
                           : fm == "strickler" ? "friction_strickler_map.txt"
                                               : "";
        if (file.empty()) {
            err = "Friction model not handled = " + fm;
            return false;
        }
        if (!read_map(in + file, 3, nc, names, cols, err)) return false;
        c.fc1 = cols[0];
        c.fc2 = cols[1];
        c.falpha = cols[2];
    }
    if (c.cfg.model_singular_head_loss != "none") {
        if (!read_map(in + "singular_head_loss_map.txt", 4, nc, names, cols, err)) return false;
        c.hsing = cols[0];
        c.hk1 = cols[1];
        c.hk2 = cols[2];
        c.halpha = cols[3];
    }
    // 1-based station codes of a Unif/Distrib map (swParent.cpp:516-541, 571-595)
    auto station_codes = [&](const std::string &file, const char *what,
                             std::vector<int32_t> &code) -> bool {
        if (!read_map(in + file, 1, nc, names, cols, err)) return false;
        code.assign(nc, 0);
        for (size_t i = 0; i < nc; ++i) {
            double ip;
            const double x = cols[0][i];
            if (std::abs(std::modf(x, &ip)) > 1e-16 || x < 0) {
                err = std::string("Non-integer/negative ") + what + " code read for cell " +
                      std::to_string(i);
                return false;
            }
            if (x < 1) {
                err = std::string(what) + " station code is zero or smaller in cell " +
                      std::to_string(i);
                return false;
            }
            code[i] = (int32_t)x;
        }
        return true;
    };
    if (c.runoff_model != SW2DG_RUNOFF_NONE) {
        // swSp::setRunoffParameters (swSp.cpp:1276-1320), setPrecipitationStationCodes /
        // setPrecipitationTimeSeriesTable (swParent.cpp:516-569), checks of swSp.cpp:150-215
        if (c.runoff_model != SW2DG_RUNOFF_FULL) {
            const char *f = c.runoff_model == SW2DG_RUNOFF_INFIL_COEFF ? "infiltration_coefficient_map.txt"
                                                                        : "runoff_coefficient_map.txt";
            if (!read_map(in + f, 1, nc, names, cols, err)) return false;
            c.runoff_coeff = cols[0];
            for (size_t i = 0; i < nc; ++i)
                if (c.runoff_coeff[i] < 0.0 || c.runoff_coeff[i] > 1.0) {
                    err = "Infiltrataion / runoff coefficient < 0 or > 1 in cell " + std::to_string(i);
                    return false;
                }
        }
        if (!station_codes("precipitation_station_codes_map.txt", "Precipitation", c.rain_code))
            return false;
        std::vector<std::vector<double>> rate;
        if (!read_scalar_series(in + "precipitation_time_series.txt", c.rain_t, rate, err)) return false;
        for (size_t k = 1; k < c.rain_t.size(); ++k)
            if (c.rain_t[k - 1] > c.rain_t[k]) {
                err = "Precipitation time series - non-increasing times for records " +
                      std::to_string(k - 1) + " " + std::to_string(k);
                return false;
            }
        c.rain_stations = (int32_t)rate.size();
        for (size_t i = 0; i < nc; ++i)
            if (c.rain_code[i] > c.rain_stations) {
                err = "Precipitation station code " + std::to_string(c.rain_code[i]) + " of cell " +
                      std::to_string(i) + " exceeds the number of stations";
                return false;
            }
        for (const auto &r : rate) c.rain_v.insert(c.rain_v.end(), r.begin(), r.end());
    }
    if (c.infil_model != SW2DG_INFIL_NONE) {
        // swSp::setInfiltrationParameters (swSp.cpp:1075-1115), checks of swSp.cpp:99-124
        if (!read_map(in + "infiltration_horton_parameter_map.txt", 1, nc, names, cols, err)) return false;
        c.infil_rate = cols[0];
        for (size_t i = 0; i < nc; ++i)
            if (c.infil_rate[i] < 0.) {
                err = "Infiltration model = horton parameter is negative in cell " + std::to_string(i);
                return false;
            }
    }
    if (c.cfg.model_wind != "none") {
        // station codes: swParent::setWindStationCodes (swParent.cpp:571-595)
        // and the checks of swSp::init (swSp.cpp:292-333)
        if (!station_codes("wind_station_codes_map.txt", "Wind", c.wind_code)) return false;
        std::vector<std::vector<double>> norm, angle;
        if (!read_wind_series(in + "wind_time_series.txt", c.wind_t, norm, angle, err))
            return false;
        for (size_t k = 1; k < c.wind_t.size(); ++k)
            if (c.wind_t[k - 1] > c.wind_t[k]) {
                err = "Wind time series - non-increasing times for records " + std::to_string(k - 1) +
                      " " + std::to_string(k);
                return false;
            }
        c.wind_stations = (int32_t)norm.size();
        for (size_t i = 0; i < nc; ++i)
            if (c.wind_code[i] > c.wind_stations) {
                err = "Wind station code " + std::to_string(c.wind_code[i]) + " of cell " +
                      std::to_string(i) + " exceeds the number of stations";
                return false;
            }
        // (norm, angle clockwise from North) -> (x, y): swSourceWind.cpp:9-38
        const size_t nt = c.wind_t.size();
        c.wind_x.assign(norm.size() * nt, 0.0);
        c.wind_y.assign(norm.size() * nt, 0.0);
        for (size_t s = 0; s < norm.size(); ++s)
            for (size_t r = 0; r < nt; ++r) {
                const double a = atan(1.0) / 45. * (90. - angle[s][r]);
                c.wind_x[s * nt + r] = -norm[s][r] * cos(a);
                c.wind_y[s * nt + r] = -norm[s][r] * sin(a);
            }
    }
    // probes (swSolverParam::readProbes, then their cells in swSp::init, swSp.cpp:754-771)
    if (!read_probes(in + "probe_locations.txt", c.probes, err)) return false;
    for (size_t i = 0; i < c.probes.names.size(); ++i) {
        const int64_t cell = locate_cell(c.g, c.probes.x[i], c.probes.y[i]);
        if (cell < 0) {
            err = "Error while looking for probe coordinates. Probe could be out of the mesh (" +
                  c.probes.names[i] + ")";
            return false;
        }
        c.probe_cell.push_back(cell);
    }
    c.bcell64.assign(c.g.bnd_cell.begin(), c.g.bnd_cell.end());
    c.L64.assign(c.g.int_L.begin(), c.g.int_L.end());
    c.R64.assign(c.g.int_R.begin(), c.g.int_R.end());
    return true;
}

std::string result_file_name(double t_time) {
    // swIO::computeFileNameFromTime (swIO.cpp:51-86), same integer arithmetic
    size_t nb_d = int(t_time / 86400.);
    size_t nb_d_e3 = int(nb_d / 1000.);
    size_t nb_d_e2 = int((nb_d - 1000. * nb_d_e3) / 100.);
    size_t nb_d_e1 = int((nb_d - 1000. * nb_d_e3 - 100. * nb_d_e2) / 10.);
    size_t nb_d_e0 = int(nb_d - 1000. * nb_d_e3 - 100. * nb_d_e2 - 10. * nb_d_e1);
    size_t nb_h = int((t_time - float(nb_d) * 86400.) / 3600.);
    size_t nb_h_e1 = int(nb_h / 10.);
    size_t nb_h_e0 = int(nb_h - 10. * nb_h_e1);
    size_t nb_m = int((t_time - nb_d * 86400. - nb_h * 3600.) / 60.);
    size_t nb_m_e1 = int(nb_m / 10.);
    size_t nb_m_e0 = int(nb_m - 10. * nb_m_e1);
    size_t nb_s = int(t_time - 86400. * nb_d - 3600. * nb_h - 60. * nb_m);
    size_t nb_s_e1 = int(nb_s / 10.);
    size_t nb_s_e0 = int(nb_s - 10. * nb_s_e1);
    size_t nb_mill = int((t_time - int(t_time)) * 1000. + 0.499999);
    size_t nb_mill_e2 = int(nb_mill / 100.);
    size_t nb_mill_e1 = int((nb_mill - 100 * nb_mill_e2) / 10.);
    size_t nb_mill_e0 = int(nb_mill - 100 * nb_mill_e2 - 10 * nb_mill_e1);
    return "ResultHydro_p" + std::to_string(nb_d_e3) + std::to_string(nb_d_e2) +
           std::to_string(nb_d_e1) + std::to_string(nb_d_e0) + "d" + std::to_string(nb_h_e1) +
           std::to_string(nb_h_e0) + "h" + std::to_string(nb_m_e1) + std::to_string(nb_m_e0) +
           "m" + std::to_string(nb_s_e1) + std::to_string(nb_s_e0) + "s" +
           std::to_string(nb_mill_e2) + std::to_string(nb_mill_e1) + std::to_string(nb_mill_e0) +
           ".txt";
}

bool write_result_map(const std::string &path, const Case &c, const std::vector<double> &h,
                      const std::vector<double> &qx, const std::vector<double> &qy,
                      int n_decimal, const std::string &mode, std::string &err) {
    std::ofstream f(path);
    if (!f) {
        err = "Unable to open file " + path;
        return false;
    }
    std::vector<std::string> names = {"x", "y", "zb", "water_depth", "unit_discharge_x",
                                      "unit_discharge_y", "unit_discharge_norm"};
    if (mode == "all") {
        for (const char *n : {"surface_elevation", "velocity_x", "velocity_y", "velocity_norm",
                              "froude", "specific_head", "total_head", "phi_by_water_depth",
                              "phi_by_unit_discharge_x", "phi_by_unit_discharge_y",
                              "phi_by_unit_discharge_norm"})
            names.push_back(n);
    }
    const size_t nb_char = (size_t)n_decimal + 7;
    for (const auto &n : names) {
        f << n;
        if (n.size() < nb_char) f << std::string(nb_char - n.size(), ' ');
        f << "\t";
    }
    f << "\n";
    std::string row;
    const double g = 9.81, small = 1e-16;
    for (size_t i = 0; i < c.g.n_cell(); ++i) {
        const double qn = std::sqrt(qx[i] * qx[i] + qy[i] * qy[i]);
        std::vector<double> v = {c.g.cell_x[i], c.g.cell_y[i], c.g.cell_zb[i], h[i], qx[i], qy[i], qn};
        if (mode == "all") {
            const double hh = h[i];
            const double vn = qn / std::max(small, hh);
            const double sh = hh + 0.5 * vn * vn / g;
            const double ph = c.phi[i];
            v.insert(v.end(), {hh + c.g.cell_zb[i], qx[i] / std::max(hh, small),
                               qy[i] / std::max(hh, small), vn, vn / std::max(std::sqrt(g * hh), small),
                               sh, sh + c.g.cell_zb[i], ph * hh, ph * qx[i], ph * qy[i], ph * qn});
        }
        row.clear();
        for (double x : v) {
            row += fmt_e(x, n_decimal);
            row += '\t';
        }
        row += '\n';
        f << row;
    }
    return (bool)f;
}

// ---------------------------------------------------------------------------
// generators
// ---------------------------------------------------------------------------
namespace {
int32_t rect_code(double, double, void *) { return 1; }
}  // namespace

void make_rect_tri_mesh(int nx, int ny, double Lx, double Ly, const std::vector<double> *node_z,
                        Geometry &g) {
    std::vector<double> x, y, z;
    const double dx = Lx / nx, dy = Ly / ny;
    for (int j = 0; j <= ny; ++j)
        for (int i = 0; i <= nx; ++i) {
            x.push_back(i * dx);
            y.push_back(j * dy);
            z.push_back(node_z ? (*node_z)[(size_t)j * (nx + 1) + i] : 0.0);
        }
    std::vector<std::vector<uint64_t>> cells;
    cells.reserve((size_t)2 * nx * ny);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const uint64_t a = (uint64_t)j * (nx + 1) + i, b = a + 1, d = a + (nx + 1), c = d + 1;
            cells.push_back({a, b, c});  // split along the a-c diagonal
            cells.push_back({a, c, d});
        }
    build_geometry(x, y, z, cells, rect_code, nullptr, g);
}

namespace {
inline double orient(double ax, double ay, double bx, double by, double cx, double cy) {
    return (bx - ax) * (cy - ay) - (by - ay) * (cx - ax);
}
// > 0 when d is strictly inside the circumcircle of the CCW triangle abc
inline double incircle(double ax, double ay, double bx, double by, double cx, double cy, double dx,
                       double dy) {
    const double adx = ax - dx, ady = ay - dy, bdx = bx - dx, bdy = by - dy, cdx = cx - dx,
                 cdy = cy - dy;
    const double ad = adx * adx + ady * ady, bd = bdx * bdx + bdy * bdy, cd = cdx * cdx + cdy * cdy;
    return adx * (bdy * cd - bd * cdy) - ady * (bdx * cd - bd * cdx) + ad * (bdx * cdy - bdy * cdx);
}
}  // namespace

void make_jittered_delaunay(int px, int py, double Lx, double Ly, double jitter, uint64_t seed,
                            ZFn zfun, void *zuser, Geometry &g, int64_t *n_flips) {
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    const double dx = Lx / px, dy = Ly / py;
    std::vector<double> x, y, z;
    for (int j = 0; j <= py; ++j)
        for (int i = 0; i <= px; ++i) {
            double xx = i * dx, yy = j * dy;
            const double jx = U(rng) * jitter * dx, jy = U(rng) * jitter * dy;
            const bool bx = (i == 0 || i == px), by = (j == 0 || j == py);
            if (!bx) xx += jx;  // boundary nodes only slide along their side
            if (!by) yy += jy;
            x.push_back(xx);
            y.push_back(yy);
        }
    const size_t nn = x.size();
    z.resize(nn);
    for (size_t p = 0; p < nn; ++p) z[p] = zfun ? zfun(x[p], y[p], zuser) : 0.0;
    // initial triangulation: each lattice quad split along its Delaunay diagonal
    std::vector<std::array<int64_t, 3>> tri;
    tri.reserve((size_t)2 * px * py);
    for (int j = 0; j < py; ++j)
        for (int i = 0; i < px; ++i) {
            const int64_t a = (int64_t)j * (px + 1) + i, b = a + 1, d = a + (px + 1), c = d + 1;
            if (incircle(x[a], y[a], x[b], y[b], x[c], y[c], x[d], y[d]) > 0) {
                tri.push_back({a, b, d});
                tri.push_back({b, c, d});
            } else {
                tri.push_back({a, b, c});
                tri.push_back({a, c, d});
            }
        }
    // neighbour table: nb[t][k] = triangle across the edge opposite vertex k
    const size_t nt = tri.size();
    std::vector<std::array<int64_t, 3>> nb(nt, {-1, -1, -1});
    {
        std::unordered_map<uint64_t, int64_t> half;
        half.reserve(nt * 3);
        for (size_t t = 0; t < nt; ++t)
            for (int k = 0; k < 3; ++k) {
                const uint64_t u = (uint64_t)tri[t][(k + 1) % 3], v = (uint64_t)tri[t][(k + 2) % 3];
                const uint64_t key = std::min(u, v) * nn + std::max(u, v);
                auto it = half.find(key);
                if (it == half.end()) {
                    half.emplace(key, (int64_t)(t * 3 + k));
                } else {
                    const int64_t t2 = it->second / 3, k2 = it->second % 3;
                    nb[t][k] = t2;
                    nb[t2][k2] = (int64_t)t;
                }
            }
    }
    // Lawson flips until every internal edge is locally Delaunay
    std::vector<int64_t> stack;
    stack.reserve(nt * 3);
    for (size_t t = 0; t < nt; ++t)
        for (int k = 0; k < 3; ++k) stack.push_back((int64_t)(t * 3 + k));
    int64_t flips = 0;
    auto idx_of = [&](int64_t t, int64_t vtx) {
        for (int k = 0; k < 3; ++k)
            if (tri[t][k] == vtx) return k;
        return -1;
    };
    while (!stack.empty()) {
        const int64_t he = stack.back();
        stack.pop_back();
        const int64_t t = he / 3, k = he % 3;
        const int64_t u = nb[t][k];
        if (u < 0) continue;
        const int64_t a = tri[t][k], b = tri[t][(k + 1) % 3], c = tri[t][(k + 2) % 3];
        // opposite vertex in u: the one not in {b, c}
        int ku = -1;
        for (int q = 0; q < 3; ++q)
            if (tri[u][q] != b && tri[u][q] != c) ku = q;
        if (ku < 0) continue;
        const int64_t d = tri[u][ku];
        if (incircle(x[a], y[a], x[b], y[b], x[c], y[c], x[d], y[d]) <= 0) continue;
        // the quad a-b-d-c must be convex for the flip (always true when d is in the circle)
        if (orient(x[a], y[a], x[b], y[b], x[d], y[d]) <= 0 ||
            orient(x[a], y[a], x[d], y[d], x[c], y[c]) <= 0)
            continue;
        // neighbours before the flip
        const int64_t n_ab = nb[t][(k + 2) % 3];  // across edge a-b (opposite c)
        const int64_t n_ca = nb[t][(k + 1) % 3];  // across edge c-a (opposite b)
        const int ku_b = idx_of(u, b), ku_c = idx_of(u, c);
        const int64_t n_cd = nb[u][ku_b];         // across edge c-d (opposite b)
        const int64_t n_db = nb[u][ku_c];         // across edge d-b (opposite c)
        // new triangles t = (a, b, d), u = (a, d, c)
        tri[t] = {a, b, d};
        tri[u] = {a, d, c};
        nb[t] = {n_db, u, n_ab};  // opp a: edge b-d ; opp b: edge d-a (u) ; opp d: edge a-b
        nb[u] = {n_cd, n_ca, t};  // opp a: edge d-c ; opp d: edge c-a ; opp c: edge a-d (t)
        auto relink = [&](int64_t nbr, int64_t from, int64_t to) {
            if (nbr < 0) return;
            for (int q = 0; q < 3; ++q)
                if (nb[nbr][q] == from) {
                    nb[nbr][q] = to;
                    return;
                }
        };
        relink(n_db, u, t);
        relink(n_ca, t, u);
        ++flips;
        stack.push_back(t * 3 + 0);
        stack.push_back(t * 3 + 2);
        stack.push_back(u * 3 + 0);
        stack.push_back(u * 3 + 1);
    }
    if (n_flips) *n_flips = flips;
    std::vector<std::vector<uint64_t>> cells(nt);
    for (size_t t = 0; t < nt; ++t)
        cells[t] = {(uint64_t)tri[t][0], (uint64_t)tri[t][1], (uint64_t)tri[t][2]};
    build_geometry(x, y, z, cells, rect_code, nullptr, g);
}

}  // namespace sw2d
