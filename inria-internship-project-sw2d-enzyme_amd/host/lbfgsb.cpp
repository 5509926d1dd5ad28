// lbfgsb.cpp -- box-constrained limited-memory BFGS (L-BFGS-B), the method
// the reference's minimiser drives through LBFGSpp's LBFGSBSolver
// (assim/swSpE_minimize.cpp:832-912, call at :873-898).  LBFGSpp (a
// header-only library found through LBFGSPP_DIR, assim/CMakeLists.txt:24-27,
// no version pinned) is not in the reference tree; this is written from the
// published methods that library implements:
//   * R. H. Byrd, P. Lu, J. Nocedal, C. Zhu, "A limited memory algorithm for
//     bound constrained optimization", SIAM J. Sci. Comput. 16 (1995):
//     compact representation B = theta I - W M W^T, W = [Y, theta S]
//     (section 3), generalized Cauchy point along the projected gradient
//     path (algorithm CP, section 4);
//   * C. Voglis, I. E. Lagaris, "BOXCQP: an algorithm for bound constrained
//     convex quadratic problems" (2004): the subspace step minimises the
//     model over the variables free at the Cauchy point exactly inside the
//     box (primal-dual active sets), LBFGSpp's replacement for the direct
//     primal method of section 5.1;
//   * J. J. More, D. J. Thuente, "Line search algorithms with guaranteed
//     sufficient decrease", ACM TOMS 20 (1994): the line search (sufficient
//     decrease ftol, curvature wolfe, safeguarded cubic / quadratic steps).
// Conventions of the reference's library kept so that a user switching over
// gets the same iteration: the first direction (no correction pairs) is
// normalised to unit length and tried at step min(1, step to the box)
// (first_step 1 takes the Fortran code's full step to the Cauchy point);
// later directions at step 1; convergence when the projected gradient's max
// norm is <= max(epsilon, epsilon_rel ||x||_2), or when the objective
// changes by <= delta max(|f|, |f_prev|, 1) over `past` iterations; a line
// search without sufficient decrease ends the run (restart 1: the Fortran
// code's reset of the memory instead); every evaluation is a value and a
// gradient.
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

// a . b summed in the order of a 2-wide vector reduction with two packet
// accumulators (the reduction order of the reference library's vector
// arithmetic on x86-64): lanes (0,1) and (2,3) accumulate every fourth pair,
// are added, a leftover pair joins, the two lanes are added, then the last
// element.  Used for the search direction's norm and its slopes, whose last
// bits decide the trial points.
double dot2(const double *a, const double *b, size_t n) {
    if (n < 2) return n ? a[0] * b[0] : 0.0;
    const size_t end2 = (n / 4) * 4, end = (n / 2) * 2;
    double p0 = a[0] * b[0], p1 = a[1] * b[1];
    if (end > 2) {
        double q0 = a[2] * b[2], q1 = a[3] * b[3];
        for (size_t i = 4; i < end2; i += 4) {
            p0 += a[i] * b[i], p1 += a[i + 1] * b[i + 1];
            q0 += a[i + 2] * b[i + 2], q1 += a[i + 3] * b[i + 3];
        }
        p0 += q0, p1 += q1;
        if (end > end2) p0 += a[end2] * b[end2], p1 += a[end2 + 1] * b[end2 + 1];
    }
    double r = p0 + p1;
    for (size_t i = end; i < n; ++i) r += a[i] * b[i];
    return r;
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

// Generalized Cauchy point (algorithm CP).  Outputs xcp, c = W^T (xcp - x),
// the variables free at xcp and those that reached a bound on the way
// (`passed`).  A variable is held from the start only when its breakpoint is
// 0 (at a bound with the gradient pointing out of the box, or lb == ub); a
// zero gradient component leaves it free whatever its position.
void cauchy(const Bfgs &B, const Vec &x, const Vec &g, const double *lb, const double *ub,
            Vec &xcp, Vec &c, std::vector<char> &free_at, std::vector<char> &passed) {
    const size_t n = x.size();
    const int t2 = 2 * B.k();
    xcp = x;
    c.assign(t2, 0.0);
    free_at.assign(n, 1);
    passed.assign(n, 0);
    Vec d(n, 0.0), tb(n, kInf);
    std::vector<std::pair<double, size_t>> brk;
    brk.reserve(n);
    for (size_t i = 0; i < n; ++i) {
        const double gi = g[i];
        if (lb[i] == ub[i])
            tb[i] = 0.0;
        else if (gi < 0.0)
            tb[i] = (x[i] - ub[i]) / gi;
        else if (gi > 0.0)
            tb[i] = (x[i] - lb[i]) / gi;
        if (tb[i] == 0.0) {
            free_at[i] = 0;
            continue;
        }
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

// The quadratic model restricted to the variables free at the Cauchy point,
// H = B_FF = theta I - W_F M W_F^T: products and solves on subsets of F
// through the rows of W (Sherman-Morrison-Woodbury, a 2k x 2k system).
struct FreeModel {
    const Bfgs &B;
    int t2;
    std::vector<double> WF;  // |F| x 2k, row a = row F[a] of W
    FreeModel(const Bfgs &b, const std::vector<size_t> &F) : B(b), t2(2 * b.k()) {
        WF.resize(F.size() * (size_t)t2);
        for (size_t a = 0; a < F.size(); ++a) B.Wrow(F[a], &WF[a * t2]);
    }
    const double *w(size_t a) const { return &WF[a * t2]; }
    // out[2k] = sum over a in S of W_F[a] v[a]  (v indexed like S)
    void Wt(const std::vector<size_t> &S, const double *v, double *out) const {
        std::fill(out, out + t2, 0.0);
        for (size_t q = 0; q < S.size(); ++q)
            for (int j = 0; j < t2; ++j) out[j] += w(S[q])[j] * v[q];
    }
    // out[q] = -(W_F[S[q]] . M u)   (the off-diagonal part of H applied to W^T v = u)
    void minus_WMu(const std::vector<size_t> &S, const double *u, double *out) const {
        Vec Mu(t2);
        B.Mv(u, Mu.data());
        for (size_t q = 0; q < S.size(); ++q) out[q] = -dot(w(S[q]), Mu.data(), t2);
    }
    // out = H_SS^{-1} rhs; false if the small system is singular
    bool solve(const std::vector<size_t> &S, const double *rhs, double *out) const {
        const double th = B.theta;
        if (t2 == 0) {
            for (size_t q = 0; q < S.size(); ++q) out[q] = rhs[q] / th;
            return true;
        }
        Vec v(t2), Mv_(t2), col(t2), Mcol(t2);
        std::vector<double> AtA((size_t)t2 * t2, 0.0), N((size_t)t2 * t2);
        Wt(S, rhs, v.data());
        for (size_t q = 0; q < S.size(); ++q) {
            const double *r = w(S[q]);
            for (int j = 0; j < t2; ++j)
                for (int l = 0; l < t2; ++l) AtA[j * t2 + l] += r[j] * r[l];
        }
        B.Mv(v.data(), Mv_.data());
        for (int l = 0; l < t2; ++l) {
            for (int j = 0; j < t2; ++j) col[j] = AtA[j * t2 + l];
            B.Mv(col.data(), Mcol.data());
            for (int j = 0; j < t2; ++j) N[j * t2 + l] = (j == l ? 1.0 : 0.0) - Mcol[j] / th;
        }
        std::vector<int> piv;
        if (!lu_factor(N, t2, piv)) return false;
        lu_solve(N, t2, piv, Mv_.data());
        for (size_t q = 0; q < S.size(); ++q)
            out[q] = rhs[q] / th + dot(w(S[q]), Mv_.data(), t2) / (th * th);
        return true;
    }
};

// Subspace minimisation from the Cauchy point, the reference library's
// method: with the variables that are not free held at their xcp values,
// the model m(x0 + y) = c.y + y^T H y / 2 over the free ones (y = x_F - x0_F,
// c = g_F + B_FA (xcp - x0)_A) is minimised exactly inside the box
// l = lb - x0 <= y <= u = ub - x0 by the primal-dual active-set iteration
// BOXCQP of C. Voglis, I. E. Lagaris, "BOXCQP: an algorithm for bound
// constrained convex quadratic problems" (2004): start from the unconstrained
// minimiser; each pass puts below-l (or at-l with multiplier >= 0) variables
// at l, above-u ones at u, solves the rest, and recomputes the multipliers
// lambda_L = (H y + c)_L, mu_U = -(H y + c)_U; done when they are >= 0 and
// the solved variables lie in the box.  After maxit passes without that, the
// projected last iterate, else the projected unconstrained minimiser, when it
// is a descent direction, else the unconstrained minimiser.  Output: the
// direction d = xsm - x0 (d = xcp - x0 outside F).
void subspace(const Bfgs &B, const Vec &x, const Vec &g, const double *lb, const double *ub,
              const Vec &xcp, const std::vector<char> &free_at, const std::vector<char> &passed,
              int maxit, Vec &d) {
    const size_t n = x.size();
    for (size_t i = 0; i < n; ++i) d[i] = xcp[i] - x[i];
    std::vector<size_t> F;
    for (size_t i = 0; i < n; ++i)
        if (free_at[i]) F.push_back(i);
    const size_t nf = F.size();
    if (nf == 0) return;
    const FreeModel H(B, F);
    const int t2 = H.t2;
    const double th = B.theta;
    // c = g_F + B_FA d_A, B_FA d_A = -W_F M W_A^T d_A over the variables that became active
    Vec cvec(nf), l(nf), u(nf), y(nf), tmp(nf);
    std::vector<size_t> all(nf);
    std::iota(all.begin(), all.end(), 0);
    {
        Vec wd(t2, 0.0), wr(t2);
        for (size_t i = 0; i < n; ++i)
            if (passed[i]) {
                B.Wrow(i, wr.data());
                for (int j = 0; j < t2; ++j) wd[j] += wr[j] * d[i];
            }
        H.minus_WMu(all, wd.data(), cvec.data());
    }
    for (size_t a = 0; a < nf; ++a) {
        const size_t i = F[a];
        cvec[a] += g[i];
        l[a] = lb[i] - x[i];
        u[a] = ub[i] - x[i];
        tmp[a] = -cvec[a];
    }
    if (!H.solve(all, tmp.data(), y.data())) return;  // singular model: the Cauchy point
    bool inside = true;
    for (size_t a = 0; a < nf && inside; ++a) inside = (y[a] >= l[a] && y[a] <= u[a]);
    auto assign = [&](const Vec &yy) {
        for (size_t a = 0; a < nf; ++a) d[F[a]] = yy[a];
    };
    if (inside) {
        assign(y);
        return;
    }
    const Vec yfb = y;
    Vec lam(nf, 0.0), mu(nf, 0.0), Fy(t2), sub(nf), vals(nf);
    std::vector<size_t> Ls, Us, Ps;
    int it;
    for (it = 0; it < maxit; ++it) {
        Ls.clear(), Us.clear(), Ps.clear();
        for (size_t a = 0; a < nf; ++a) {
            if (y[a] < l[a] || (y[a] == l[a] && lam[a] >= 0.0)) {
                Ls.push_back(a), y[a] = l[a], mu[a] = 0.0;
            } else if (y[a] > u[a] || (y[a] == u[a] && mu[a] >= 0.0)) {
                Us.push_back(a), y[a] = u[a], lam[a] = 0.0;
            } else {
                Ps.push_back(a), lam[a] = 0.0, mu[a] = 0.0;
            }
        }
        if (!Ps.empty()) {
            // y_P = -H_PP^{-1} (c_P + H_PL l_L + H_PU u_U)
            Vec wq(t2, 0.0), wtmp(t2);
            for (size_t q = 0; q < Ls.size(); ++q) vals[q] = l[Ls[q]];
            H.Wt(Ls, vals.data(), wtmp.data());
            for (int j = 0; j < t2; ++j) wq[j] += wtmp[j];
            for (size_t q = 0; q < Us.size(); ++q) vals[q] = u[Us[q]];
            H.Wt(Us, vals.data(), wtmp.data());
            for (int j = 0; j < t2; ++j) wq[j] += wtmp[j];
            H.minus_WMu(Ps, wq.data(), sub.data());
            for (size_t q = 0; q < Ps.size(); ++q) sub[q] = -(sub[q] + cvec[Ps[q]]);
            if (!H.solve(Ps, sub.data(), tmp.data())) break;
            for (size_t q = 0; q < Ps.size(); ++q) y[Ps[q]] = tmp[q];
        }
        if (!Ls.empty() || !Us.empty()) {
            H.Wt(all, y.data(), Fy.data());
            H.minus_WMu(Ls, Fy.data(), sub.data());
            for (size_t q = 0; q < Ls.size(); ++q) {
                const size_t a = Ls[q];
                lam[a] = sub[q] + cvec[a] + th * y[a];
            }
            H.minus_WMu(Us, Fy.data(), sub.data());
            for (size_t q = 0; q < Us.size(); ++q) {
                const size_t a = Us[q];
                mu[a] = -(sub[q] + cvec[a] + th * y[a]);
            }
        }
        bool ok = true;
        for (size_t a : Ls) ok = ok && lam[a] >= 0.0;
        for (size_t a : Us) ok = ok && mu[a] >= 0.0;
        for (size_t a : Ps) ok = ok && y[a] >= l[a] && y[a] <= u[a];
        if (ok) break;
    }
    if (it < maxit) {
        assign(y);
        return;
    }
    const double eps = std::numeric_limits<double>::epsilon();
    auto descent = [&](const Vec &yy) {
        assign(yy);
        return dot2(d.data(), g.data(), n) <= -eps;
    };
    Vec yp(nf);
    for (size_t a = 0; a < nf; ++a) yp[a] = std::min(u[a], std::max(l[a], y[a]));
    if (descent(yp)) return;
    for (size_t a = 0; a < nf; ++a) yp[a] = std::min(u[a], std::max(l[a], yfb[a]));
    if (descent(yp)) return;
    assign(yfb);
}

// Moré-Thuente line search in the form of the reference's library: the
// auxiliary function psi(a) = f(a) - f(0) - a ftol f'(0) brackets the step;
// a trial with psi above the best one shrinks the interval (case 1), a lower
// psi with psi' still of the initial sign extrapolates by 1.1 x the last
// move (case 2), a sign change interpolates (case 3); interpolation picks
// among the cubic and quadratic models of J. J. More, D. J. Thuente, ACM
// TOMS 20 (1994), section 4.  Every trial is a value and a gradient.
double quad_min_f(double a, double b, double fa, double ga, double fb) {
    const double ba = b - a;
    const double w = 0.5 * ba * ga / (fa - fb + ba * ga);
    return a + w * ba;
}

double quad_min_g(double a, double b, double ga, double gb) {
    const double w = ga / (ga - gb);
    return a + w * (b - a);
}

double cubic_min(double a, double b, double fa, double fb, double ga, double gb, bool &exists) {
    const double apb = a + b, ba = b - a, ba2 = ba * ba, fba = fb - fa, gba = gb - ga;
    const double z3 = (ga + gb) * ba - 2.0 * fba;
    const double z2 = 0.5 * (gba * ba2 - 3.0 * apb * z3);
    const double z1 = fba * ba2 - apb * z2 - (a * apb + b * b) * z3;
    const double eps = std::numeric_limits<double>::epsilon();
    if (std::fabs(z3) < eps * std::fabs(z2) || std::fabs(z3) < eps * std::fabs(z1)) {
        exists = z2 * ba > 0.0;
        return exists ? (-0.5 * z1 / z2) : b;
    }
    const double u = z2 / (3.0 * z3), v = z1 / z2;
    const double vu = v / u;
    exists = vu <= 1.0;
    if (!exists) return b;
    double r1, r2;
    if (std::fabs(u) >= std::fabs(v)) {
        const double w = 1.0 + std::sqrt(1.0 - vu);
        r1 = -u * w;
        r2 = -v / w;
    } else {
        const double sq = std::sqrt(std::fabs(u)) * std::sqrt(std::fabs(v)) * std::sqrt(1.0 - u / v);
        r1 = -u - sq;
        r2 = -u + sq;
    }
    return (z3 * ba > 0.0) ? std::max(r1, r2) : std::min(r1, r2);
}

double step_selection(double al, double au, double at, double fl, double fu, double ft, double gl,
                      double gu, double gt) {
    if (al == au) return al;
    if (!std::isfinite(ft) || !std::isfinite(gt)) return (al + at) / 2.0;
    bool ac_ok;
    const double ac = cubic_min(al, at, fl, ft, gl, gt, ac_ok);
    const double aq = quad_min_f(al, at, fl, gl, ft);
    if (ft > fl) {
        if (!ac_ok) return aq;
        return (std::fabs(ac - al) < std::fabs(aq - al)) ? ac : ((aq + ac) / 2.0);
    }
    const double as = quad_min_g(al, at, gl, gt);
    if (gt * gl < 0.0) return (std::fabs(ac - at) >= std::fabs(as - at)) ? ac : as;
    const double deltal = 1.1, deltau = 0.66;
    if (std::fabs(gt) < std::fabs(gl)) {
        const double res =
            (ac_ok && (ac - at) * (at - al) > 0.0 && std::fabs(ac - at) < std::fabs(as - at)) ? ac : as;
        return (at > al) ? std::min(at + deltau * (au - at), res) : std::max(at + deltau * (au - at), res);
    }
    if (!std::isfinite(au) || !std::isfinite(fu) || !std::isfinite(gu)) return at + deltal * (at - al);
    bool ae_ok;
    const double ae = cubic_min(at, au, ft, fu, gt, gu, ae_ok);
    return (at > al) ? std::min(at + deltau * (au - at), ae) : std::max(at + deltau * (au - at), ae);
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

// Search along d from xp (value fx0, slope dg0 < 0), first trial at `step`,
// never beyond stpmax.  Trial points are xp + step d as computed (the caller
// clamps the accepted one).  On return xn / fn / gn hold the accepted point
// and `step` its step: 0 = strong Wolfe met or the largest step kept, 1 =
// the best point of an exhausted search (sufficient decrease), 2 = no step
// of sufficient decrease, 3 = stop requested by the objective (xn is that
// evaluation), 4 = the step fell below the minimum, -1 = objective error.
int line_search(Eval &ev, const Vec &xp, double fx0, const Vec &gp, const Vec &d, double dg0,
                double &step, double stpmax, const sw2dh_lbfgsb_param &P, Vec &xn, double &fn,
                Vec &gn) {
    const size_t n = xp.size();
    const double min_step = 1e-20;
    const double test_decr = P.ftol * dg0, test_curv = -P.wolfe * dg0;
    double I_lo = 0.0, I_hi = kInf, fI_lo = 0.0, fI_hi = kInf;
    double gI_lo = (1.0 - P.ftol) * dg0, gI_hi = kInf;
    Vec x_lo = xp, g_lo = gp, x(n), g(n);
    double fx_lo = fx0;
    auto eval_at = [&](double stp, double &fx, double &dg) -> int {
        for (size_t i = 0; i < n; ++i) x[i] = xp[i] + stp * d[i];
        const int rc = ev.call(x, fx, g);
        dg = dot2(g.data(), d.data(), n);
        return rc;
    };
    double fx, dg;
    if (eval_at(step, fx, dg) < 0) return -1;
    if (ev.stop) {
        xn = x, fn = fx, gn = g;
        return 3;
    }
    if (fx <= fx0 + step * test_decr && std::fabs(dg) <= test_curv) {
        xn = x, fn = fx, gn = g;
        return 0;
    }
    const double delta = 1.1;
    int it;
    for (it = 0; it < P.max_linesearch; ++it) {
        const double ft = fx - fx0 - step * test_decr;
        const double gt = dg - P.ftol * dg0;
        double new_step;
        if (ft > fI_lo) {  // case 1: the interval shrinks to [I_lo, step]
            new_step = step_selection(I_lo, I_hi, step, fI_lo, fI_hi, ft, gI_lo, gI_hi, gt);
            if (new_step <= min_step) new_step = (I_lo + step) / 2.0;
            I_hi = step, fI_hi = ft, gI_hi = gt;
        } else if (gt * (I_lo - step) > 0.0) {  // case 2: extrapolate
            new_step = std::min(stpmax, step + delta * (step - I_lo));
            I_lo = step, fI_lo = ft, gI_lo = gt;
            x_lo.swap(x), g_lo.swap(g), fx_lo = fx;
        } else {  // case 3: the minimum lies between the old and the new lower end
            new_step = step_selection(I_lo, I_hi, step, fI_lo, fI_hi, ft, gI_lo, gI_hi, gt);
            I_hi = I_lo, fI_hi = fI_lo, gI_hi = gI_lo;
            I_lo = step, fI_lo = ft, gI_lo = gt;
            x_lo.swap(x), g_lo.swap(g), fx_lo = fx;
        }
        if (step == stpmax && new_step >= stpmax) {  // the largest step stays: keep the best
            xn = x_lo, fn = fx_lo, gn = g_lo;
            step = I_lo;
            return 0;
        }
        // a collapsed bracket proposes the best point again: its value is known,
        // and the search would only repeat it until the evaluations run out
        if (new_step == I_lo) {
            if (!(I_lo > 0.0)) return 2;
            xn = x_lo, fn = fx_lo, gn = g_lo;
            step = I_lo;
            return 1;
        }
        step = new_step;
        if (!(step >= min_step)) return 4;
        if (eval_at(step, fx, dg) < 0) return -1;
        if (ev.stop) {
            xn = x, fn = fx, gn = g;
            return 3;
        }
        if (fx <= fx0 + step * test_decr && std::fabs(dg) <= test_curv) {
            xn = x, fn = fx, gn = g;
            return 0;
        }
        if (step >= stpmax && fx - fx0 - step * test_decr <= fI_lo) {
            xn = x, fn = fx, gn = g;
            return 0;
        }
    }
    // exhausted: the last trial if it is the best, else the best end point
    if (fx - fx0 - step * test_decr <= fI_lo) {
        xn = x, fn = fx, gn = g;
        return 1;
    }
    if (I_lo > 0.0) {
        xn = x_lo, fn = fx_lo, gn = g_lo;
        step = I_lo;
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
    p->max_submin = 10;
    p->restart = 0;
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
    if (P.m < 1 || P.max_linesearch < 1 || P.max_submin < 1 ||
        !(P.ftol > 0.0 && P.ftol < P.wolfe && P.wolfe < 1.0)) {
        g_lb_err = "sw2dh_lbfgsb: need m >= 1, max_linesearch >= 1, max_submin >= 1, "
                   "0 < ftol < wolfe < 1";
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
    Vec xcp, c, d(N), Xn(N), Gn(N);
    std::vector<char> free_at, passed;
    auto max_step_to_box = [&](const Vec &dir) {
        double st = kInf;
        for (size_t i = 0; i < N; ++i) {
            if (dir[i] > 0.0) st = std::min(st, (U[i] - X[i]) / dir[i]);
            else if (dir[i] < 0.0) st = std::min(st, (L[i] - X[i]) / dir[i]);
        }
        return st;
    };
    // first direction: towards the Cauchy point of the identity model
    cauchy(B, X, G, L.data(), U.data(), xcp, c, free_at, passed);
    for (size_t i = 0; i < N; ++i) d[i] = xcp[i] - X[i];
    if (P.first_step == 0) {  // unit length (the reference library's rule)
        const double dn = std::sqrt(dot2(d.data(), d.data(), N));
        if (dn > 0.0)
            for (size_t i = 0; i < N; ++i) d[i] /= dn;
    }
    for (;;) {
        double dg = dot2(G.data(), d.data(), N);
        double stpmax = max_step_to_box(d);
        if (stpmax <= 1e-20 || !(dg < 0.0)) {  // pathological step: the Cauchy direction, no memory
            for (size_t i = 0; i < N; ++i) d[i] = xcp[i] - X[i];
            B.reset();
            dg = dot2(G.data(), d.data(), N);
            stpmax = max_step_to_box(d);
        }
        if (!(dg < 0.0)) return finish(SW2DH_LBFGSB_CONVERGED_GRAD);
        stpmax = std::min(P.max_step, stpmax);
        if (!(stpmax > 0.0)) return finish(SW2DH_LBFGSB_LINESEARCH_FAILED);
        double step = std::min(1.0, stpmax);
        double fn = kInf;
        const int ls = line_search(ev, X, f, G, d, dg, step, stpmax, P, Xn, fn, Gn);
        if (ls < 0) {
            g_lb_err = "sw2dh_lbfgsb: the objective failed";
            return 2;
        }
        if (ls == 3) {
            if (fn < f) {
                X = Xn, f = fn, G = Gn;
                for (size_t i = 0; i < N; ++i) X[i] = std::min(U[i], std::max(L[i], X[i]));
            }
            ++k;
            return finish(SW2DH_LBFGSB_STOPPED);
        }
        // no step of sufficient decrease: the reference's library gives up here
        // and the reference keeps the last iterate (swSpE_minimize.cpp:895-902);
        // with `restart` the correction pairs go and the search starts again
        // along the Cauchy direction of the identity model, once per failure
        if (ls == 2 || ls == 4) {
            if (!P.restart || B.k() == 0) return finish(SW2DH_LBFGSB_LINESEARCH_FAILED);
            B.reset();
            cauchy(B, X, G, L.data(), U.data(), xcp, c, free_at, passed);
            for (size_t i = 0; i < N; ++i) d[i] = xcp[i] - X[i];
            if (P.first_step == 0) {
                const double dn = std::sqrt(dot2(d.data(), d.data(), N));
                if (dn > 0.0)
                    for (size_t i = 0; i < N; ++i) d[i] /= dn;
            }
            continue;
        }
        ++k;
        Vec sv(N), yv(N);
        for (size_t i = 0; i < N; ++i) sv[i] = Xn[i] - X[i], yv[i] = Gn[i] - G[i];
        X = Xn, G = Gn;
        f = fn;
        f_hist.push_back(f);
        const double pg = proj_grad_norm(X, G, L.data(), U.data());
        if (pg <= std::max(P.epsilon, P.epsilon_rel * std::sqrt(dot2(X.data(), X.data(), N))))
            return finish(SW2DH_LBFGSB_CONVERGED_GRAD);
        if (P.past > 0 && (int)f_hist.size() > P.past) {
            const double fp = f_hist[f_hist.size() - 1 - P.past];
            if (std::fabs(fp - f) <= P.delta * std::max({std::fabs(f), std::fabs(fp), 1.0}))
                return finish(SW2DH_LBFGSB_CONVERGED_F);
        }
        if (P.max_iterations > 0 && k >= P.max_iterations) return finish(SW2DH_LBFGSB_MAX_ITER);
        if (dot2(sv.data(), yv.data(), N) >
            std::numeric_limits<double>::epsilon() * dot2(yv.data(), yv.data(), N))
            B.add(sv, yv);
        for (size_t i = 0; i < N; ++i) X[i] = std::min(U[i], std::max(L[i], X[i]));
        cauchy(B, X, G, L.data(), U.data(), xcp, c, free_at, passed);
        subspace(B, X, G, L.data(), U.data(), xcp, free_at, passed, P.max_submin, d);
    }
}
