// lbfgsb.cpp -- box-constrained limited-memory BFGS (L-BFGS-B), the method
// the reference's minimiser drives through LBFGSpp's LBFGSBSolver
// (assim/swSpE_minimize.cpp:832-912, call at :873-898).  LBFGSpp is not in
// the reference tree; this is written from the published method:
//   * R. H. Byrd, P. Lu, J. Nocedal, C. Zhu, "A limited memory algorithm for
//     bound constrained optimization", SIAM J. Sci. Comput. 16 (1995):
//     compact representation B = theta I - W M W^T, W = [Y, theta S]
//     (section 3), generalized Cauchy point along the projected gradient
//     path (algorithm CP, section 4), subspace minimisation over the free
//     variables by the direct primal method (section 5.1);
//   * J. L. Morales, J. Nocedal, "Remark on Algorithm 778", ACM TOMS 38
//     (2011): the subspace step is projected onto the box and kept when it
//     is a descent direction, else the step is cut back to the box;
//   * J. J. More, D. J. Thuente, "Line search algorithms with guaranteed
//     sufficient decrease", ACM TOMS 20 (1994): the line search (sufficient
//     decrease ftol, curvature wolfe, safeguarded cubic / quadratic steps).
// Conventions of the reference's library kept so that a user switching over
// gets the same kind of iteration: the first direction (no correction pairs)
// is normalised to unit length and tried at step min(1, step to the box)
// (first_step 1 takes the Fortran code's full step to the Cauchy point);
// later directions at step 1; convergence when the projected gradient's max
// norm is <= max(epsilon, epsilon_rel ||x||_2), or when the objective
// changes by <= delta max(|f|, |f_prev|, 1) over `past` iterations; every
// evaluation is a value and a gradient.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/sw2d_host.h"

namespace {

using Vec = std::vector<double>;
constexpr double kInf = std::numeric_limits<double>::infinity();

double dot(const double *a, const double *b, size_t n) {
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

// LU factorisation with partial pivoting of a small dense row-major matrix.
bool lu_factor(std::vector<double> &A, int n, std::vector<int> &piv) {
    piv.resize(n);
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = std::fabs(A[k * n + k]);
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(A[i * n + k]) > best) best = std::fabs(A[i * n + k]), p = i;
        if (!(best > 0.0)) return false;
        piv[k] = p;
        if (p != k)
            for (int j = 0; j < n; ++j) std::swap(A[k * n + j], A[p * n + j]);
        for (int i = k + 1; i < n; ++i) {
            const double f = A[i * n + k] / A[k * n + k];
            A[i * n + k] = f;
            for (int j = k + 1; j < n; ++j) A[i * n + j] -= f * A[k * n + j];
        }
    }
    return true;
}

void lu_solve(const std::vector<double> &A, int n, const std::vector<int> &piv, double *b) {
    for (int k = 0; k < n; ++k)
        if (piv[k] != k) std::swap(b[k], b[piv[k]]);
    for (int i = 1; i < n; ++i)
        for (int j = 0; j < i; ++j) b[i] -= A[i * n + j] * b[j];
    for (int i = n - 1; i >= 0; --i) {
        for (int j = i + 1; j < n; ++j) b[i] -= A[i * n + j] * b[j];
        b[i] /= A[i * n + i];
    }
}

// Compact limited-memory BFGS matrix B = theta I - W M W^T (Byrd, Nocedal,
// Schnabel 1994), W = [Y, theta S], M^{-1} = [[-D, L^T], [L, theta S^T S]],
// D = diag(s_i . y_i), L_ij = s_i . y_j (i > j).  Pairs oldest first.
struct Bfgs {
    size_t n = 0;
    int m = 6;
    std::vector<Vec> S, Y;
    double theta = 1.0;
    std::vector<double> Minv;  // LU factors of M^{-1} (2k x 2k)
    std::vector<int> piv;

    int k() const { return (int)S.size(); }
    void reset() {
        S.clear();
        Y.clear();
        theta = 1.0;
        Minv.clear();
    }
    bool add(const Vec &s, const Vec &y) {
        if ((int)S.size() == m) {
            S.erase(S.begin());
            Y.erase(Y.begin());
        }
        S.push_back(s);
        Y.push_back(y);
        theta = dot(y.data(), y.data(), n) / dot(s.data(), y.data(), n);
        const int kk = k(), t = 2 * kk;
        Minv.assign((size_t)t * t, 0.0);
        for (int i = 0; i < kk; ++i) {
            Minv[i * t + i] = -dot(S[i].data(), Y[i].data(), n);
            for (int j = 0; j < i; ++j) {
                const double l = dot(S[i].data(), Y[j].data(), n);  // L_ij, i > j
                Minv[(kk + i) * t + j] = l;                         // lower-left L
                Minv[j * t + kk + i] = l;                           // upper-right L^T
            }
            for (int j = 0; j <= i; ++j) {
                const double ss = theta * dot(S[i].data(), S[j].data(), n);
                Minv[(kk + i) * t + kk + j] = ss;
                Minv[(kk + j) * t + kk + i] = ss;
            }
        }
        if (!lu_factor(Minv, t, piv)) {
            reset();
            return false;
        }
        return true;
    }
    // out[2k] = W^T v
    void Wt(const double *v, double *out) const {
        const int kk = k();
        for (int j = 0; j < kk; ++j) {
            out[j] = dot(Y[j].data(), v, n);
            out[kk + j] = theta * dot(S[j].data(), v, n);
        }
    }
    // out[2k] = row i of W
    void Wrow(size_t i, double *out) const {
        const int kk = k();
        for (int j = 0; j < kk; ++j) {
            out[j] = Y[j][i];
            out[kk + j] = theta * S[j][i];
        }
    }
    // out[2k] = M v
    void Mv(const double *v, double *out) const {
        const int t = 2 * k();
        std::copy(v, v + t, out);
        if (t) lu_solve(Minv, t, piv, out);
    }
    // out[n] += a * W v
    void add_Wv(const double *v, double a, double *out) const {
        const int kk = k();
        for (int j = 0; j < kk; ++j) {
            const double cy = a * v[j], cs = a * theta * v[kk + j];
            const double *y = Y[j].data(), *s = S[j].data();
            for (size_t i = 0; i < n; ++i) out[i] += cy * y[i] + cs * s[i];
        }
    }
};

double proj_grad_norm(const Vec &x, const Vec &g, const double *lb, const double *ub) {
    double r = 0.0;
    for (size_t i = 0; i < x.size(); ++i) {
        double p;
        if (g[i] < 0.0)
            p = std::max(x[i] - ub[i], g[i]);
        else
            p = std::min(x[i] - lb[i], g[i]);
        r = std::max(r, std::fabs(p));
    }
    return r;
}

// Generalized Cauchy point (algorithm CP).  Outputs xcp, c = W^T (xcp - x)
// and the free set at xcp.
void cauchy(const Bfgs &B, const Vec &x, const Vec &g, const double *lb, const double *ub,
            Vec &xcp, Vec &c, std::vector<char> &free_at) {
    const size_t n = x.size();
    const int t2 = 2 * B.k();
    xcp = x;
    c.assign(t2, 0.0);
    free_at.assign(n, 1);
    Vec d(n, 0.0), tb(n, kInf);
    std::vector<std::pair<double, size_t>> brk;
    brk.reserve(n);
    for (size_t i = 0; i < n; ++i) {
        const double gi = g[i];
        if ((x[i] - lb[i] <= 0.0 && gi >= 0.0) || (ub[i] - x[i] <= 0.0 && gi <= 0.0)) {
            tb[i] = 0.0;  // at a bound, the gradient pushes outwards: fixed
            free_at[i] = 0;
            continue;
        }
        if (gi < 0.0)
            tb[i] = (x[i] - ub[i]) / gi;
        else if (gi > 0.0)
            tb[i] = (x[i] - lb[i]) / gi;
        d[i] = -gi;
        if (tb[i] < kInf) brk.emplace_back(tb[i], i);
    }
    std::sort(brk.begin(), brk.end());
    Vec p(t2, 0.0), wb(t2), Mc(t2), Mp(t2), Mw(t2);
    B.Wt(d.data(), p.data());
    double f1 = -dot(d.data(), d.data(), n);
    if (f1 >= 0.0) return;  // projected gradient is zero
    B.Mv(p.data(), Mp.data());
    double f2 = -B.theta * f1 - dot(p.data(), Mp.data(), t2);
    const double f2_org = f2;
    f2 = std::max(f2, std::numeric_limits<double>::epsilon() * f2_org);
    double dt_min = -f1 / f2;
    double t_old = 0.0;
    size_t ib = 0;
    double tcur = brk.empty() ? kInf : brk[0].first;
    double dt = tcur;
    std::vector<char> passed(n, 0);
    while (ib < brk.size() && dt_min >= dt) {
        const size_t b = brk[ib].second;
        const double gb = g[b];
        xcp[b] = d[b] > 0.0 ? ub[b] : lb[b];
        const double zb = xcp[b] - x[b];
        for (int j = 0; j < t2; ++j) c[j] += dt * p[j];
        B.Wrow(b, wb.data());
        B.Mv(c.data(), Mc.data());
        B.Mv(p.data(), Mp.data());
        B.Mv(wb.data(), Mw.data());
        f1 += dt * f2 + gb * gb + B.theta * gb * zb - gb * dot(wb.data(), Mc.data(), t2);
        f2 -= B.theta * gb * gb + 2.0 * gb * dot(wb.data(), Mp.data(), t2) +
              gb * gb * dot(wb.data(), Mw.data(), t2);
        f2 = std::max(f2, std::numeric_limits<double>::epsilon() * f2_org);
        for (int j = 0; j < t2; ++j) p[j] += gb * wb[j];
        d[b] = 0.0;
        passed[b] = 1;
        free_at[b] = 0;
        dt_min = -f1 / f2;
        t_old = tcur;
        ++ib;
        tcur = ib < brk.size() ? brk[ib].first : kInf;
        dt = tcur - t_old;
    }
    dt_min = std::max(dt_min, 0.0);
    t_old += dt_min;
    for (size_t i = 0; i < n; ++i)
        if (!passed[i] && tb[i] > 0.0) xcp[i] = x[i] + t_old * d[i];
    for (int j = 0; j < t2; ++j) c[j] += dt_min * p[j];
}

// Subspace minimisation from the Cauchy point (direct primal method with the
// Morales-Nocedal projection).  Returns the point xbar.
void subspace(const Bfgs &B, const Vec &x, const Vec &g, const double *lb, const double *ub,
              const Vec &xcp, const Vec &c, const std::vector<char> &free_at, Vec &xbar) {
    const size_t n = x.size();
    const int kk = B.k(), t2 = 2 * kk;
    xbar = xcp;
    std::vector<size_t> Z;
    for (size_t i = 0; i < n; ++i)
        if (free_at[i]) Z.push_back(i);
    if (Z.empty() || kk == 0) return;
    // reduced gradient r = Z^T (g + theta (xcp - x) - W M c)
    Vec Mc(t2), r(Z.size());
    B.Mv(c.data(), Mc.data());
    Vec WMc(n, 0.0);
    B.add_Wv(Mc.data(), 1.0, WMc.data());
    for (size_t a = 0; a < Z.size(); ++a) {
        const size_t i = Z[a];
        r[a] = g[i] + B.theta * (xcp[i] - x[i]) - WMc[i];
    }
    // A = W^T Z (2k x |Z|): v = A r, N = I - (1/theta) M A A^T
    Vec v(t2, 0.0), wr(t2);
    std::vector<double> AAt((size_t)t2 * t2, 0.0);
    for (size_t a = 0; a < Z.size(); ++a) {
        B.Wrow(Z[a], wr.data());
        for (int j = 0; j < t2; ++j) {
            v[j] += wr[j] * r[a];
            for (int l = 0; l < t2; ++l) AAt[j * t2 + l] += wr[j] * wr[l];
        }
    }
    Vec Mv_(t2);
    B.Mv(v.data(), Mv_.data());
    std::vector<double> N((size_t)t2 * t2, 0.0), col(t2), Mcol(t2);
    for (int l = 0; l < t2; ++l) {
        for (int j = 0; j < t2; ++j) col[j] = AAt[j * t2 + l];
        B.Mv(col.data(), Mcol.data());
        for (int j = 0; j < t2; ++j) N[j * t2 + l] = (j == l ? 1.0 : 0.0) - Mcol[j] / B.theta;
    }
    std::vector<int> piv;
    if (!lu_factor(N, t2, piv)) return;  // keep the Cauchy point
    lu_solve(N, t2, piv, Mv_.data());
    // du = -(1/theta) r - (1/theta^2) A^T v
    Vec du(Z.size());
    for (size_t a = 0; a < Z.size(); ++a) {
        B.Wrow(Z[a], wr.data());
        du[a] = -r[a] / B.theta - dot(wr.data(), Mv_.data(), t2) / (B.theta * B.theta);
    }
    // projected step, kept when it is a descent direction from x
    Vec xp = xcp;
    for (size_t a = 0; a < Z.size(); ++a) {
        const size_t i = Z[a];
        xp[i] = std::min(ub[i], std::max(lb[i], xcp[i] + du[a]));
    }
    double dd = 0.0;
    for (size_t i = 0; i < n; ++i) dd += (xp[i] - x[i]) * g[i];
    if (dd < 0.0) {
        xbar = xp;
        return;
    }
    // otherwise the longest feasible fraction of the step
    double alpha = 1.0;
    for (size_t a = 0; a < Z.size(); ++a) {
        const size_t i = Z[a];
        if (du[a] < 0.0)
            alpha = std::min(alpha, (lb[i] - xcp[i]) / du[a]);
        else if (du[a] > 0.0)
            alpha = std::min(alpha, (ub[i] - xcp[i]) / du[a]);
    }
    alpha = std::max(alpha, 0.0);
    for (size_t a = 0; a < Z.size(); ++a) xbar[Z[a]] = xcp[Z[a]] + alpha * du[a];
}

// One step of the More-Thuente interval update (their section 4, cases 1-4).
void mt_step(double &stx, double &fx, double &dx, double &sty, double &fy, double &dy, double &stp,
             double fp, double dp, bool &brackt, double stpmin, double stpmax) {
    const double sgnd = dp * (dx / std::fabs(dx));
    double stpf;
    if (fp > fx) {  // case 1: higher value, the minimum is bracketed
        const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
        const double s = std::max({std::fabs(theta), std::fabs(dx), std::fabs(dp)});
        double gamma = s * std::sqrt(std::max(0.0, (theta / s) * (theta / s) - (dx / s) * (dp / s)));
        if (stp < stx) gamma = -gamma;
        const double p = (gamma - dx) + theta, q = ((gamma - dx) + gamma) + dp;
        const double stpc = stx + (p / q) * (stp - stx);
        const double stpq = stx + ((dx / ((fx - fp) / (stp - stx) + dx)) / 2.0) * (stp - stx);
        stpf = std::fabs(stpc - stx) < std::fabs(stpq - stx) ? stpc : stpc + (stpq - stpc) / 2.0;
        brackt = true;
    } else if (sgnd < 0.0) {  // case 2: lower value, derivatives of opposite sign
        const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
        const double s = std::max({std::fabs(theta), std::fabs(dx), std::fabs(dp)});
        double gamma = s * std::sqrt(std::max(0.0, (theta / s) * (theta / s) - (dx / s) * (dp / s)));
        if (stp > stx) gamma = -gamma;
        const double p = (gamma - dp) + theta, q = ((gamma - dp) + gamma) + dx;
        const double stpc = stp + (p / q) * (stx - stp);
        const double stpq = stp + (dp / (dp - dx)) * (stx - stp);
        stpf = std::fabs(stpc - stp) > std::fabs(stpq - stp) ? stpc : stpq;
        brackt = true;
    } else if (std::fabs(dp) < std::fabs(dx)) {  // case 3: derivative magnitude decreases
        const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
        const double s = std::max({std::fabs(theta), std::fabs(dx), std::fabs(dp)});
        double gamma = s * std::sqrt(std::max(0.0, (theta / s) * (theta / s) - (dx / s) * (dp / s)));
        if (stp > stx) gamma = -gamma;
        const double p = (gamma - dp) + theta, q = (gamma + (dx - dp)) + gamma;
        const double r = p / q;
        double stpc;
        if (r < 0.0 && gamma != 0.0)
            stpc = stp + r * (stx - stp);
        else if (stp > stx)
            stpc = stpmax;
        else
            stpc = stpmin;
        const double stpq = stp + (dp / (dp - dx)) * (stx - stp);
        if (brackt) {
            stpf = std::fabs(stpc - stp) < std::fabs(stpq - stp) ? stpc : stpq;
            if (stp > stx)
                stpf = std::min(stp + 0.66 * (sty - stp), stpf);
            else
                stpf = std::max(stp + 0.66 * (sty - stp), stpf);
        } else {
            stpf = std::fabs(stpc - stp) > std::fabs(stpq - stp) ? stpc : stpq;
            stpf = std::max(stpmin, std::min(stpmax, stpf));
        }
    } else {  // case 4: lower value, same sign, derivative does not decrease
        if (brackt) {
            const double theta = 3.0 * (fp - fy) / (sty - stp) + dy + dp;
            const double s = std::max({std::fabs(theta), std::fabs(dy), std::fabs(dp)});
            double gamma =
                s * std::sqrt(std::max(0.0, (theta / s) * (theta / s) - (dy / s) * (dp / s)));
            if (stp > sty) gamma = -gamma;
            const double p = (gamma - dp) + theta, q = ((gamma - dp) + gamma) + dy;
            stpf = stp + (p / q) * (sty - stp);
        } else {
            stpf = stp > stx ? stpmax : stpmin;
        }
    }
    if (fp > fx) {
        sty = stp, fy = fp, dy = dp;
    } else {
        if (sgnd < 0.0) sty = stx, fy = fx, dy = dx;
        stx = stp, fx = fp, dx = dp;
    }
    stp = stpf;
}

struct Eval {
    sw2dh_objective fg;
    void *user;
    int32_t n_eval = 0;
    bool stop = false;
    int call(const Vec &x, double &f, Vec &g) {
        ++n_eval;
        const int rc = fg(x.data(), &f, g.data(), user);
        if (rc > 0) stop = true;
        if (!std::isfinite(f)) f = kInf;
        return rc;
    }
};

// More-Thuente search along d from x (value f0, directional derivative dg0 <
// 0).  On return: 0 = strong Wolfe point found, 1 = the best point of the
// search lowers f (accepted), 2 = no decrease, 3 = stop requested by the
// objective (the evaluated point is in xn / fn / gn), -1 = objective error.
int line_search(Eval &ev, const Vec &x, double f0, const Vec &g0, const Vec &d, double dg0,
                double stp, double stpmax, const sw2dh_lbfgsb_param &P, Vec &xn, double &fn,
                Vec &gn, const double *lb, const double *ub) {
    const size_t n = x.size();
    const double ftol = P.ftol, gtol = P.wolfe, xtol = 0.1, stpmin = 0.0;
    const double xtrapl = 1.1, xtrapu = 4.0;
    bool brackt = false;
    int stage = 1;
    const double gtest = ftol * dg0;
    double width = stpmax - stpmin, width1 = 2.0 * width;
    double stx = 0.0, fx = f0, gx = dg0, sty = 0.0, fy = f0, gy = dg0;
    double stmin = 0.0, stmax = stp + xtrapu * stp;
    // best point seen (lowest value)
    double best_f = f0;
    Vec best_x = x, best_g = g0;
    Vec xt(n), gt(n);
    for (int it = 0; it < P.max_linesearch; ++it) {
        for (size_t i = 0; i < n; ++i)
            xt[i] = std::min(ub[i], std::max(lb[i], x[i] + stp * d[i]));
        double f;
        const int rc = ev.call(xt, f, gt);
        if (rc < 0) return -1;
        if (f < best_f) best_f = f, best_x = xt, best_g = gt;
        if (ev.stop) {
            xn = xt, fn = f, gn = gt;
            return 3;
        }
        if (!std::isfinite(f)) {  // step back towards the best bracket end
            brackt = true;
            sty = stp, fy = kInf, gy = 0.0;
            stp = stx + 0.5 * (stp - stx);
            stmin = std::min(stx, sty), stmax = std::max(stx, sty);
            continue;
        }
        const double dg = dot(gt.data(), d.data(), n);
        const double ftest = f0 + stp * gtest;
        if (stage == 1 && f <= ftest && dg >= 0.0) stage = 2;
        if (f <= ftest && std::fabs(dg) <= gtol * (-dg0)) {  // strong Wolfe
            xn = xt, fn = f, gn = gt;
            return 0;
        }
        const bool warn = (brackt && (stp <= stmin || stp >= stmax)) ||
                          (brackt && stmax - stmin <= xtol * stmax) ||
                          (stp == stpmax && f <= ftest && dg <= gtest) ||
                          (stp == stpmin && (f > ftest || dg >= gtest));
        if (warn) break;
        if (stage == 1 && f <= fx && f > ftest) {  // modified function psi
            double fm = f - stp * gtest, fxm = fx - stx * gtest, fym = fy - sty * gtest;
            double gm = dg - gtest, gxm = gx - gtest, gym = gy - gtest;
            mt_step(stx, fxm, gxm, sty, fym, gym, stp, fm, gm, brackt, stmin, stmax);
            fx = fxm + stx * gtest, fy = fym + sty * gtest;
            gx = gxm + gtest, gy = gym + gtest;
        } else {
            mt_step(stx, fx, gx, sty, fy, gy, stp, f, dg, brackt, stmin, stmax);
        }
        if (brackt) {
            if (std::fabs(sty - stx) >= 0.66 * width1) stp = stx + 0.5 * (sty - stx);
            width1 = width;
            width = std::fabs(sty - stx);
            stmin = std::min(stx, sty), stmax = std::max(stx, sty);
        } else {
            stmin = stp + xtrapl * (stp - stx);
            stmax = stp + xtrapu * (stp - stx);
        }
        stp = std::max(stpmin, std::min(stpmax, stp));
        if ((brackt && (stp <= stmin || stp >= stmax)) ||
            (brackt && stmax - stmin <= xtol * stmax))
            stp = stx;
        if (!(stp > 0.0)) break;
    }
    if (best_f < f0) {
        xn = best_x, fn = best_f, gn = best_g;
        return 1;
    }
    return 2;
}

thread_local std::string g_lb_err;

}  // namespace

extern "C" void sw2dh_lbfgsb_default(sw2dh_lbfgsb_param *p) {
    if (!p) return;
    p->m = 6;
    p->epsilon = 1e-5;
    p->epsilon_rel = 1e-5;
    p->past = 1;
    p->delta = 1e-10;
    p->max_iterations = 0;
    p->max_linesearch = 20;
    p->ftol = 1e-4;
    p->wolfe = 0.9;
    p->max_step = 1e20;
    p->first_step = 0;
}

extern "C" const char *sw2dh_lbfgsb_error(void) { return g_lb_err.c_str(); }

extern "C" int sw2dh_lbfgsb(int64_t n, double *x, const double *lb, const double *ub,
                            sw2dh_objective fg, void *user, const sw2dh_lbfgsb_param *param,
                            double *fx_out, int32_t *n_iter, int32_t *n_eval, double *pgnorm_out,
                            int32_t *stop_reason) {
    g_lb_err.clear();
    if (n <= 0 || !x || !fg) {
        g_lb_err = "sw2dh_lbfgsb: need n > 0, x and an objective";
        return 1;
    }
    sw2dh_lbfgsb_param P;
    sw2dh_lbfgsb_default(&P);
    if (param) P = *param;
    if (P.m < 1 || P.max_linesearch < 1 || !(P.ftol > 0.0 && P.ftol < P.wolfe && P.wolfe < 1.0)) {
        g_lb_err = "sw2dh_lbfgsb: need m >= 1, max_linesearch >= 1, 0 < ftol < wolfe < 1";
        return 1;
    }
    const size_t N = (size_t)n;
    Vec L(N, -kInf), U(N, kInf);
    if (lb) std::copy(lb, lb + N, L.begin());
    if (ub) std::copy(ub, ub + N, U.begin());
    for (size_t i = 0; i < N; ++i)
        if (!(L[i] <= U[i])) {
            g_lb_err = "sw2dh_lbfgsb: lower bound above upper bound at " + std::to_string(i);
            return 1;
        }
    Vec X(x, x + N), G(N);
    for (size_t i = 0; i < N; ++i) X[i] = std::min(U[i], std::max(L[i], X[i]));
    Eval ev{fg, user};
    double f = 0.0;
    int32_t k = 0;
    auto finish = [&](int32_t why) {
        std::copy(X.begin(), X.end(), x);
        if (fx_out) *fx_out = f;
        if (n_iter) *n_iter = k;
        if (n_eval) *n_eval = ev.n_eval;
        if (pgnorm_out) *pgnorm_out = proj_grad_norm(X, G, L.data(), U.data());
        if (stop_reason) *stop_reason = why;
        return 0;
    };
    if (ev.call(X, f, G) < 0) {
        g_lb_err = "sw2dh_lbfgsb: the objective failed at the starting point";
        return 2;
    }
    if (ev.stop) return finish(SW2DH_LBFGSB_STOPPED);
    if (!std::isfinite(f)) {
        g_lb_err = "sw2dh_lbfgsb: non-finite objective at the starting point";
        return 2;
    }
    if (proj_grad_norm(X, G, L.data(), U.data()) <=
        std::max(P.epsilon, P.epsilon_rel * std::sqrt(dot(X.data(), X.data(), N))))
        return finish(SW2DH_LBFGSB_CONVERGED_GRAD);
    Bfgs B;
    B.n = N;
    B.m = P.m;
    std::vector<double> f_hist;
    f_hist.push_back(f);
    Vec xcp, c, xbar, d(N), Xn(N), Gn(N);
    std::vector<char> free_at;
    bool retried = false;
    for (;;) {
        cauchy(B, X, G, L.data(), U.data(), xcp, c, free_at);
        subspace(B, X, G, L.data(), U.data(), xcp, c, free_at, xbar);
        for (size_t i = 0; i < N; ++i) d[i] = xbar[i] - X[i];
        double dg = dot(G.data(), d.data(), N);
        if (!(dg < 0.0) && B.k() > 0) {  // pathological subspace step: restart from the GCP
            B.reset();
            cauchy(B, X, G, L.data(), U.data(), xcp, c, free_at);
            for (size_t i = 0; i < N; ++i) d[i] = xcp[i] - X[i];
            dg = dot(G.data(), d.data(), N);
        }
        if (!(dg < 0.0)) return finish(SW2DH_LBFGSB_CONVERGED_GRAD);
        if (B.k() == 0 && P.first_step == 0) {  // LBFGSpp's first step: unit-length direction
            const double dn = std::sqrt(dot(d.data(), d.data(), N));
            for (size_t i = 0; i < N; ++i) d[i] /= dn;
            dg /= dn;
        }
        double stpmax = P.max_step;
        for (size_t i = 0; i < N; ++i) {
            if (d[i] < 0.0 && L[i] > -kInf) stpmax = std::min(stpmax, (L[i] - X[i]) / d[i]);
            if (d[i] > 0.0 && U[i] < kInf) stpmax = std::min(stpmax, (U[i] - X[i]) / d[i]);
        }
        stpmax = std::max(stpmax, 0.0);
        const double stp = std::min(1.0, stpmax);
        double fn = kInf;
        const int ls = stp > 0.0 ? line_search(ev, X, f, G, d, dg, stp, stpmax, P, Xn, fn, Gn,
                                               L.data(), U.data())
                                 : 2;
        if (ls < 0) {
            g_lb_err = "sw2dh_lbfgsb: the objective failed";
            return 2;
        }
        if (ls == 3) {
            if (fn < f || !std::isfinite(f)) X = Xn, f = fn, G = Gn;
            ++k;
            return finish(SW2DH_LBFGSB_STOPPED);
        }
        if (ls == 2) {  // no decrease: drop the memory once and retry from the GCP
            if (B.k() > 0 && !retried) {
                B.reset();
                retried = true;
                continue;
            }
            return finish(SW2DH_LBFGSB_LINESEARCH_FAILED);
        }
        retried = false;
        ++k;
        Vec s(N), y(N);
        for (size_t i = 0; i < N; ++i) s[i] = Xn[i] - X[i], y[i] = Gn[i] - G[i];
        X = Xn, G = Gn;
        f = fn;
        f_hist.push_back(f);
        const double pg = proj_grad_norm(X, G, L.data(), U.data());
        if (pg <= std::max(P.epsilon, P.epsilon_rel * std::sqrt(dot(X.data(), X.data(), N))))
            return finish(SW2DH_LBFGSB_CONVERGED_GRAD);
        if (P.past > 0 && (int)f_hist.size() > P.past) {
            const double fp = f_hist[f_hist.size() - 1 - P.past];
            if (std::fabs(fp - f) <= P.delta * std::max({std::fabs(f), std::fabs(fp), 1.0}))
                return finish(SW2DH_LBFGSB_CONVERGED_F);
        }
        if (P.max_iterations > 0 && k >= P.max_iterations) return finish(SW2DH_LBFGSB_MAX_ITER);
        const double sy = dot(s.data(), y.data(), N);
        if (sy > std::numeric_limits<double>::epsilon() * dot(y.data(), y.data(), N)) B.add(s, y);
    }
}
