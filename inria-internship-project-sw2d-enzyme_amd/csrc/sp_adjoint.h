// sp_adjoint.h -- the reference's differentiable ("_E") single-porosity
// physics and hand-written reverse-mode derivatives (vector-Jacobian
// products) of each local function.  Shared by the HIP adjoint kernels and a
// host build used by the per-function derivative tests.
//
// Forward semantics (the function the reference differentiates):
//   hypGodunovComputeLRStates_E  assim/swSpE_hyp.cpp:507-573
//   hypBoundaryFlux_E            assim/swSpE_hyp.cpp:143-239
//   hypHllc_E                    assim/swSpE_hyp.cpp:725-832
//   hypComputeBalance_E          assim/swSpE_hyp.cpp:336-394
//   sourceFrictionStrickler_E    assim/swSpE_friction.cpp:19-76
//   sourceSingularHeadLoss_E     assim/swSpE_friction.cpp:79-126
//   hypMusclEvrComputeLRStates_E assim/swSpE_muscl.cpp:22-187
//   computeMusclSlopes           swSpHypMusclEvr.cpp:11-115 (free-surface slopes)
// Derivative conventions: branch decisions (wet/dry regimes, min/max
// selections, clamps) are taken from the primal values and held fixed; the
// derivative of sqrt at 0 is taken as 0.
#pragma once

#include "sp_physics.h"

namespace spe {

using sp::kG;
using sp::kSmall;

constexpr double kPhiEps = 1e-3;  // smoothed min of the porosities (swSpE_hyp.cpp:794)

SPD double dsqrt_arg(double s, double ds) { return (s > 0.0) ? ds * 0.5 / s : 0.0; }

// ---------------------------------------------------------------------------
// Godunov Riemann state on one side (swSpE_hyp.cpp:532-565)
// ---------------------------------------------------------------------------
struct Side {
    double z, q, r, lm, lp;
};

SPD Side lr_fwd(double h, double qx, double qy, double zb, double b, double nx, double ny,
                double h_min) {
    Side s;
    s.z = h + zb;
    s.q = qx * nx + qy * ny;
    s.r = -qx * ny + qy * nx;
    const double u = (h > h_min) ? s.q / h : 0.0;
    const double c = sqrt(kG * h);
    s.lm = sp::lambda_minus(b, c, u);
    s.lp = sp::lambda_plus(b, c, u);
    return s;
}

// the intermediates of lr_fwd its adjoint needs (normal discharge q,
// velocity u, celerity c = sqrt(g h), s = sqrt((b - 1) b u^2 + c^2), 1/h)
struct LrAux {
    double q, u, c, s, inv_h;
};

// lr_fwd keeping its intermediates: the same operations, so the returned Side
// is bit-identical to lr_fwd's (lambda_minus / lambda_plus share one root)
SPD Side lr_fwd_x(double h, double qx, double qy, double zb, double b, double nx, double ny,
                  double h_min, LrAux &x) {
    Side s;
    s.z = h + zb;
    s.q = qx * nx + qy * ny;
    s.r = -qx * ny + qy * nx;
    const bool wet = h > h_min;
    const double u = wet ? s.q / h : 0.0;
    const double c = sqrt(kG * h);
    const double root = sqrt((b - 1.0) * b * u * u + c * c);
    s.lm = b * u - root;
    s.lp = b * u + root;
    x.q = s.q;
    x.u = u;
    x.c = c;
    x.s = root;
    x.inv_h = wet ? 1.0 / h : 0.0;
    return s;
}

// ---------------------------------------------------------------------------
// MUSCL-EVR _E (swSpE_muscl.cpp:22-187) with the Kurganov-Petrova
// desingularised velocity (swSpDesingVelocity.h:28-39)
// ---------------------------------------------------------------------------
constexpr double kHDesing = 5e-3;

SPD double desing(double q, double h) {
    const double h2 = h * h;
    const double h4 = h2 * h2;
    const double e4 = kHDesing * kHDesing * kHDesing * kHDesing;
    const double den = sqrt(h4 + ((h4 < e4) ? e4 : h4));
    return 1.4142135623730951 * h * q / den;
}

// boundary edge: first order, desingularised velocity
SPD Side lr_bnd_muscl(double h, double qx, double qy, double zb, double b, double nx, double ny) {
    Side s;
    s.z = h + zb;
    s.q = qx * nx + qy * ny;
    s.r = -qx * ny + qy * nx;
    const double u = desing(s.q, h);
    const double c = sqrt(kG * h);
    s.lm = sp::lambda_minus(b, c, u);
    s.lp = sp::lambda_plus(b, c, u);
    return s;
}

// internal edge, one side: characteristic-offset reconstruction of the free
// surface with the cell's limited slope (sx, sy), dxm / dym = edge midpoint
// minus cell centre, dt = the dt the reconstruction reads
SPD Side recon_muscl(int side, double h, double qx, double qy, double zb, double b, double sx,
                     double sy, double nx, double ny, double dxm, double dym, double dt) {
    const double zc = h + zb;
    const double q_local = qx * nx + qy * ny;
    const double r_local = qy * nx - qx * ny;
    const double c_local = sqrt(kG * ((kSmall < h) ? h : kSmall));
    const double u_local = desing(q_local, h);
    const double v_local = desing(r_local, h);
    double l1 = sp::lambda_minus(b, c_local, u_local);
    double l2 = sp::lambda_plus(b, c_local, u_local);
    if (side == 0) {
        l1 = (0.0 < l1) ? l1 : 0.0;
        l2 = (0.0 < l2) ? l2 : 0.0;
    } else {
        l1 = (l1 < 0.0) ? l1 : 0.0;
        l2 = (l2 < 0.0) ? l2 : 0.0;
    }
    const double dx1 = dxm - 0.5 * l1 * dt * nx;
    const double dy1 = dym - 0.5 * l1 * dt * ny;
    const double z1 = zc + dx1 * sx + dy1 * sy;
    const double h1 = (0.0 < z1 - zb) ? (z1 - zb) : 0.0;
    const double dx2 = dxm - 0.5 * l2 * dt * nx;
    const double dy2 = dym - 0.5 * l2 * dt * ny;
    const double z2 = zc + dx2 * sx + dy2 * sy;
    const double h2 = (0.0 < z2 - zb) ? (z2 - zb) : 0.0;
    const double hr = 0.5 * (h1 + h2);
    Side s;
    s.z = 0.5 * (h1 + h2) + zb;
    s.q = 0.5 * (h1 + h2) * u_local;
    s.r = 0.5 * (h1 + h2) * v_local;
    const double u = desing(s.q, hr);
    const double c = sqrt(kG * hr);
    s.lm = sp::lambda_minus(b, c, u);
    s.lp = sp::lambda_plus(b, c, u);
    return s;
}

// d desing(q, h) / dq and / dh; the max picks its branch from the primal h
SPD void desing_d(double q, double h, double &dq, double &dh) {
    const double h2 = h * h;
    const double h4 = h2 * h2;
    const double e4 = kHDesing * kHDesing * kHDesing * kHDesing;
    const bool big = !(h4 < e4);
    const double D2 = h4 + (big ? h4 : e4);
    const double rD = 1.0 / sqrt(D2);
    // d(h q / D)/dh = q / D (1 - h dD2/dh / (2 D2)), dD2/dh = (big ? 8 : 4) h^3
    dq = 1.4142135623730951 * h * rD;
    dh = 1.4142135623730951 * q * rD * (1.0 - (big ? 4.0 : 2.0) * h4 / D2);
}

// adjoint of (lm, lp) = b u -/+ sqrt((b - 1) b u^2 + c^2): adds to (du, dc)
SPD void lam_vjp(double b, double c, double u, double dlm, double dlp, double &du, double &dc) {
    const double root = sqrt((b - 1.0) * b * u * u + c * c);
    const double dA = dsqrt_arg(root, dlp - dlm);
    du += b * (dlm + dlp) + dA * (b - 1.0) * b * 2.0 * u;
    dc += dA * 2.0 * c;
}

// adjoint of lr_bnd_muscl: (dz, dq, dr, dlp) -> accumulate into (dh, dqx, dqy)
SPD void lr_bnd_muscl_vjp(double h, double qx, double qy, double b, double nx, double ny, double dz,
                          double dq, double dr, double dlp, double &dh, double &dqx, double &dqy) {
    const double q = qx * nx + qy * ny;
    const double u = desing(q, h);
    const double c = sqrt(kG * h);
    double du = 0.0, dc = 0.0, gq, gh;
    lam_vjp(b, c, u, 0.0, dlp, du, dc);
    desing_d(q, h, gq, gh);
    const double dqq = dq + du * gq;
    dh += dz + dsqrt_arg(c, dc * kG) + du * gh;
    dqx += dqq * nx - dr * ny;
    dqy += dqq * ny + dr * nx;
}

// adjoint of recon_muscl: (dz, dq, dr, dlm, dlp) of the side -> accumulate
// into the cell's (dh, dqx, dqy) (its free surface h + zb included) and its
// limited slope (dsx, dsy).  Clipped eigenvalues and dry offsets pass no
// derivative.
SPD void recon_muscl_vjp(int side, double h, double qx, double qy, double zb, double b, double sx,
                         double sy, double nx, double ny, double dxm, double dym, double dt,
                         double dz, double dq, double dr, double dlm, double dlp, double &dh,
                         double &dqx, double &dqy, double &dsx, double &dsy) {
    const double zc = h + zb;
    const double ql = qx * nx + qy * ny;
    const double rl = qy * nx - qx * ny;
    const double cl = sqrt(kG * ((kSmall < h) ? h : kSmall));
    const double ul = desing(ql, h);
    const double vl = desing(rl, h);
    double l1 = sp::lambda_minus(b, cl, ul);
    double l2 = sp::lambda_plus(b, cl, ul);
    const bool p1 = side ? (l1 < 0.0) : (0.0 < l1);
    const bool p2 = side ? (l2 < 0.0) : (0.0 < l2);
    if (!p1) l1 = 0.0;
    if (!p2) l2 = 0.0;
    const double dx1 = dxm - 0.5 * l1 * dt * nx, dy1 = dym - 0.5 * l1 * dt * ny;
    const double dx2 = dxm - 0.5 * l2 * dt * nx, dy2 = dym - 0.5 * l2 * dt * ny;
    const double z1 = zc + dx1 * sx + dy1 * sy;
    const double z2 = zc + dx2 * sx + dy2 * sy;
    const bool w1 = 0.0 < z1 - zb, w2 = 0.0 < z2 - zb;
    const double hr = 0.5 * ((w1 ? (z1 - zb) : 0.0) + (w2 ? (z2 - zb) : 0.0));
    const double qo = hr * ul;
    const double cc = sqrt(kG * hr);
    const double uu = desing(qo, hr);
    // eigenvalues of the reconstructed state
    double duu = 0.0, dcc = 0.0, gq, gh;
    lam_vjp(b, cc, uu, dlm, dlp, duu, dcc);
    desing_d(qo, hr, gq, gh);
    const double dqo = dq + duu * gq;
    const double dhr = dz + dsqrt_arg(cc, dcc * kG) + duu * gh + dqo * ul + dr * vl;
    double dul = dqo * hr;
    const double dvl = dr * hr;
    // the two characteristic-offset free surfaces
    const double dz1 = w1 ? 0.5 * dhr : 0.0;
    const double dz2 = w2 ? 0.5 * dhr : 0.0;
    dsx += dz1 * dx1 + dz2 * dx2;
    dsy += dz1 * dy1 + dz2 * dy2;
    const double proj = -0.5 * dt * (nx * sx + ny * sy);
    // eigenvalues of the cell state
    double dcl = 0.0;
    lam_vjp(b, cl, ul, p1 ? proj * dz1 : 0.0, p2 ? proj * dz2 : 0.0, dul, dcl);
    double dhh = dz1 + dz2 + ((kSmall < h) ? dsqrt_arg(cl, dcl * kG) : 0.0);
    desing_d(ql, h, gq, gh);
    const double dql = dul * gq;
    dhh += dul * gh;
    desing_d(rl, h, gq, gh);
    const double drl = dvl * gq;
    dhh += dvl * gh;
    dh += dhh;
    dqx += dql * nx - drl * ny;
    dqy += dql * ny + drl * nx;
}

// ---------------------------------------------------------------------------
// Free-surface slope of one cell (computeMusclSlopes, swSpHypMusclEvr.cpp:
// 11-115): inverse-distance weighted gradient over the neighbours, then the
// in-place limiter visits of the cell's internal edges in ascending edge
// order (a cell's slope changes only at its own edges' visits, so the
// per-cell replay is the reference's global loop restricted to the cell).
// Nb gives, for incidence k = 0 .. deg-1 in that order: zj(k), xj(k), yj(k)
// of the neighbour, xm(k), ym(k) of the edge midpoint and side(k) (1: the
// cell is the edge's right cell).
// ---------------------------------------------------------------------------
// one neighbour's term of the raw slope (dx, dy = neighbour centre minus
// the cell's)
SPD void slope_raw_term(double zi, double zj, double dx, double dy, double &sx, double &sy,
                        double &wsum) {
    const double dist = sqrt(dx * dx + dy * dy);
    const double w = 1.0 / dist;
    sx += (zj - zi) * dx / dist / dist * w;
    sy += (zj - zi) * dy / dist / dist * w;
    wsum += w;
}

// limiter visit of one incidence (s = the cell's side, ax / ay = edge
// midpoint minus the cell's centre): scales (sx, sy) when the
// reconstruction at the midpoint leaves [min, max] of the two free
// surfaces.  Returns the cell whose surface bounds it (0 this cell, 1 the
// neighbour; -1 no scaling), with D = fr - zi and the scale sf.
SPD int slope_visit_term(double zi, double zj, int s, double ax, double ay, double &sx, double &sy,
                         double &D, double &sf) {
    const double zL = s ? zj : zi;
    const double zR = s ? zi : zj;
    const double fmx = sp::smax(zL, zR);
    const double fmn = sp::smin(zL, zR);
    const double fr = zi + ax * sx + ay * sy;
    int src = -1;
    if (fr < fmn) {
        sf = (fmn - zi) / (fr - zi);
        sx *= sf;
        sy *= sf;
        D = fr - zi;
        src = ((zR < zL) == (s == 0)) ? 1 : 0;  // fmn is zR when zR < zL
    }
    if (fr > fmx) {
        sf = (fmx - zi) / (fr - zi);
        sx *= sf;
        sy *= sf;
        D = fr - zi;
        src = ((zL < zR) == (s == 0)) ? 1 : 0;  // fmx is zR when zL < zR
    }
    return src;
}

template <class Nb>
SPD void slope_raw(const Nb &nb, int deg, double zi, double xi, double yi, double &sx, double &sy,
                   double &wsum) {
    sx = 0.0;
    sy = 0.0;
    wsum = 0.0;
    if (deg <= 1) return;
    for (int k = 0; k < deg; ++k) slope_raw_term(zi, nb.zj(k), nb.xj(k) - xi, nb.yj(k) - yi, sx, sy, wsum);
    sx /= wsum;
    sy /= wsum;
}

template <class Nb>
SPD int slope_visit(const Nb &nb, int k, double zi, double xi, double yi, double &sx, double &sy,
                    double &D, double &sf) {
    return slope_visit_term(zi, nb.zj(k), nb.side(k), nb.xm(k) - xi, nb.ym(k) - yi, sx, sy, D, sf);
}

template <class Nb>
SPD void slope_fwd(const Nb &nb, int deg, double zi, double xi, double yi, double &sx, double &sy) {
    double wsum, D, sf;
    slope_raw(nb, deg, zi, xi, yi, sx, sy, wsum);
    for (int k = 0; k < deg; ++k) (void)slope_visit(nb, k, zi, xi, yi, sx, sy, D, sf);
}

// adjoint of slope_fwd: (gx, gy) = adjoint of the limited slope -> adds the
// adjoint of this cell's free surface to dzi and gives that of the neighbour
// across incidence k through out.set(k, v) then out.add(k, v).  The slope
// each visit saw is replayed from the raw slope (deg (deg + 1) / 2 visits).
// Each scaling sf = (z_f - zi) / D with D = a . slope (independent of zi).
template <class Nb, class Out>
SPD void slope_vjp(const Nb &nb, int deg, double zi, double xi, double yi, double gx, double gy,
                   double &dzi, Out &out) {
    double sx0, sy0, wsum;
    slope_raw(nb, deg, zi, xi, yi, sx0, sy0, wsum);
    for (int k = deg - 1; k >= 0; --k) {
        double sx = sx0, sy = sy0, D = 1.0, sf = 1.0;
        for (int q = 0; q < k; ++q) (void)slope_visit(nb, q, zi, xi, yi, sx, sy, D, sf);
        const double px = sx, py = sy;
        const int src = slope_visit(nb, k, zi, xi, yi, sx, sy, D, sf);
        double zn = 0.0;
        if (src >= 0) {
            const double gs = gx * px + gy * py;
            const double c = gs * sf / D;
            gx = sf * gx - c * (nb.xm(k) - xi);
            gy = sf * gy - c * (nb.ym(k) - yi);
            if (src == 1) {  // bounded by the neighbour's surface (else d/dzi cancels)
                zn = gs / D;
                dzi -= zn;
            }
        }
        out.set(k, zn);
    }
    if (deg <= 1) return;
    for (int k = 0; k < deg; ++k) {
        const double dx = nb.xj(k) - xi;
        const double dy = nb.yj(k) - yi;
        const double dist = sqrt(dx * dx + dy * dy);
        const double w = 1.0 / dist;
        const double a = (gx * dx + gy * dy) / dist / dist * w / wsum;
        out.add(k, a);
        dzi -= a;
    }
}

// slope_vjp for cells of at most MAXD incidences, with the incidences' data
// and the slope each visit saw held in registers (one forward replay); each
// neighbour's term is passed once, through out.put(k, v).  Returns false,
// doing nothing, when deg > MAXD.
template <int MAXD, class Nb, class Out>
SPD bool slope_vjp_reg(const Nb &nb, int deg, double zi, double xi, double yi, double gx, double gy,
                       double &dzi, const Out &out) {
    if (deg > MAXD) return false;
    double zj[MAXD], dx[MAXD], dy[MAXD], ax[MAXD], ay[MAXD], px[MAXD], py[MAXD], D[MAXD], sf[MAXD];
    int side[MAXD], src[MAXD];
#pragma unroll
    for (int k = 0; k < MAXD; ++k) {
        if (k < deg) {
            zj[k] = nb.zj(k);
            dx[k] = nb.xj(k) - xi;
            dy[k] = nb.yj(k) - yi;
            ax[k] = nb.xm(k) - xi;
            ay[k] = nb.ym(k) - yi;
            side[k] = nb.side(k);
        }
    }
    double sx = 0.0, sy = 0.0, wsum = 0.0;
    if (deg > 1) {
#pragma unroll
        for (int k = 0; k < MAXD; ++k)
            if (k < deg) slope_raw_term(zi, zj[k], dx[k], dy[k], sx, sy, wsum);
        sx /= wsum;
        sy /= wsum;
    }
#pragma unroll
    for (int k = 0; k < MAXD; ++k) {
        if (k < deg) {
            px[k] = sx;
            py[k] = sy;
            D[k] = 1.0;
            sf[k] = 1.0;
            src[k] = slope_visit_term(zi, zj[k], side[k], ax[k], ay[k], sx, sy, D[k], sf[k]);
        }
    }
    double zn[MAXD];
#pragma unroll
    for (int k = MAXD - 1; k >= 0; --k) {
        if (k < deg) {
            zn[k] = 0.0;
            if (src[k] >= 0) {
                const double gs = gx * px[k] + gy * py[k];
                const double c = gs * sf[k] / D[k];
                gx = sf[k] * gx - c * ax[k];
                gy = sf[k] * gy - c * ay[k];
                if (src[k] == 1) {
                    zn[k] = gs / D[k];
                    dzi -= zn[k];
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < MAXD; ++k) {
        if (k < deg) {
            if (deg > 1) {
                const double dist = sqrt(dx[k] * dx[k] + dy[k] * dy[k]);
                const double w = 1.0 / dist;
                const double a = (gx * dx[k] + gy * dy[k]) / dist / dist * w / wsum;
                zn[k] += a;
                dzi -= a;
            }
            out.put(k, zn[k]);
        }
    }
    return true;
}

// adjoint of lr_fwd from its intermediates: (dz, dq, dr, dlm, dlp) ->
// accumulate into (dh, dqx, dqy)
SPD void lr_vjp_x(const LrAux &x, double b, double nx, double ny, double dz, double dq, double dr,
                  double dlm, double dlp, double &dh, double &dqx, double &dqy) {
    double du = b * (dlm + dlp);
    const double ds = dlp - dlm;
    const double dA = dsqrt_arg(x.s, ds);
    du += dA * (b - 1.0) * b * 2.0 * x.u;
    const double dc = dA * 2.0 * x.c;
    double dhh = dz + dsqrt_arg(x.c, dc) * kG;
    double dqq = dq;
    if (x.inv_h != 0.0) {  // h > h_min
        dqq += du * x.inv_h;
        dhh += -du * x.q * (x.inv_h * x.inv_h);
    }
    dh += dhh;
    dqx += dqq * nx - dr * ny;
    dqy += dqq * ny + dr * nx;
}

// adjoint of lr_fwd: (dz, dq, dr, dlm, dlp) -> accumulate into (dh, dqx, dqy)
SPD void lr_vjp(double h, double qx, double qy, double b, double nx, double ny, double h_min,
                double dz, double dq, double dr, double dlm, double dlp, double &dh, double &dqx,
                double &dqy) {
    LrAux x;
    (void)lr_fwd_x(h, qx, qy, 0.0, b, nx, ny, h_min, x);
    lr_vjp_x(x, b, nx, ny, dz, dq, dr, dlm, dlp, dh, dqx, dqy);
}

// ---------------------------------------------------------------------------
// HLLC with the smoothed porosity min (swSpE_hyp.cpp:725-832).
// in: zL qL rL zR qR rR lmL lmR lpL lpR phiL phiR ; out: F1 F2 F3 sL sR
// ---------------------------------------------------------------------------
enum Regime { R_NOWHERE = 0, R_BOTH = 1, R_LHS = 2, R_RHS = 3 };

SPD int hllc_regime(double h_min, double phi_min, double phiL, double phiR, double zbL, double zbR,
                    double zL, double zR, double hL, double hR) {
    bool both = false, lhs = false, rhs = false, nowhere = false;
    if (((hL <= h_min) && (hR <= h_min)) || ((phiL <= phi_min) && (phiR <= phi_min))) {
        nowhere = true;
    } else {
        both = true;
        if ((phiR <= phi_min) || ((phiR > phi_min) && (hR <= h_min) && (zL <= zbR + h_min))) {
            lhs = true;
            both = false;
        }
        if ((phiL <= phi_min) || ((phiL > phi_min) && (hL <= h_min) && (zR <= zbL + h_min))) {
            rhs = true;
            both = false;
        }
    }
    // the reference's if-else chain: both, then rhs, then lhs, else nothing
    if (both) return R_BOTH;
    if (rhs) return R_RHS;
    if (lhs) return R_LHS;
    return R_NOWHERE;
}

SPD void hllc_fwd(double h_min, double phi_min, double bL, double bR, double phiL, double phiR,
                  double zbL, double zbR, double zL, double qL, double rL, double zR, double qR,
                  double rR, double lmL, double lmR, double lpL, double lpR, double *f) {
    const double hL = (zL - zbL > 0.0) ? (zL - zbL) : 0.0;
    const double hR = (zR - zbR > 0.0) ? (zR - zbR) : 0.0;
    const int reg = hllc_regime(h_min, phi_min, phiL, phiR, zbL, zbR, zL, zR, hL, hR);
    double f1 = 0.0, f2 = 0.0, f3 = 0.0, sl = 0.0, sr = 0.0;
    if (reg == R_BOTH) {
        const double hLs = (hL > kSmall) ? hL : kSmall;
        const double hRs = (hR > kSmall) ? hR : kSmall;
        const double uL = qL / hLs, uR = qR / hRs;
        double L1 = (lmL < lmR) ? lmL : lmR;
        if (L1 > 0.0) L1 = 0.0;
        double L2 = (lpL > lpR) ? lpL : lpR;
        if (L2 < 0.0) L2 = 0.0;
        double den = L2 - L1;
        if (den < kSmall) den = kSmall;
        const double d = phiL - phiR;
        const double phiLR = 0.5 * (phiL + phiR - sqrt(d * d + kPhiEps * kPhiEps));
        f1 = phiLR * (L2 * qL - L1 * qR + L1 * L2 * (zR - zL)) / den;
        const double fact = 0.5 * phiLR * (hL + hR);
        double s2 = (phiR * hR * hR - phiL * hL * hL) * 0.5 + fact * (zL - zR);
        s2 = s2 * kG / den;
        sl = -L1 * s2;
        sr = L2 * s2;
        const double FL = sp::specific_force(bL, phiL, hL, uL);
        const double FR = sp::specific_force(bR, phiR, hR, uR);
        f2 = (L2 * (FL + sl) - L1 * (FR - sr) + L1 * L2 * (phiR * qR - phiL * qL)) / den;
        const double vL = rL / hLs, vR = rR / hRs;
        f3 = (f1 >= 0.0) ? (bL * f1 * vL) : (bR * f1 * vR);
    } else if (reg == R_RHS) {
        sr = phiR * 0.5 * kG * hR * hR;
    } else if (reg == R_LHS) {
        sl = -phiL * 0.5 * kG * hL * hL;
    }
    f[0] = f1;
    f[1] = f2;
    f[2] = f3;
    f[3] = sl;
    f[4] = sr;
}

// d[0..11] accumulates the adjoints of zL qL rL zR qR rR lmL lmR lpL lpR phiL phiR
SPD void hllc_vjp(double h_min, double phi_min, double bL, double bR, double phiL, double phiR,
                  double zbL, double zbR, double zL, double qL, double rL, double zR, double qR,
                  double rR, double lmL, double lmR, double lpL, double lpR, const double *df,
                  double *d) {
    const double hL = (zL - zbL > 0.0) ? (zL - zbL) : 0.0;
    const double hR = (zR - zbR > 0.0) ? (zR - zbR) : 0.0;
    const int reg = hllc_regime(h_min, phi_min, phiL, phiR, zbL, zbR, zL, zR, hL, hR);
    double dhL = 0.0, dhR = 0.0;
    if (reg == R_BOTH) {
        const double hLs = (hL > kSmall) ? hL : kSmall;
        const double hRs = (hR > kSmall) ? hR : kSmall;
        const double ihL = 1.0 / hLs, ihR = 1.0 / hRs;
        const double uL = qL * ihL, uR = qR * ihR;
        int s1 = (lmL < lmR) ? 1 : 2;
        double L1 = (s1 == 1) ? lmL : lmR;
        if (L1 > 0.0) {
            L1 = 0.0;
            s1 = 0;
        }
        int s2s = (lpL > lpR) ? 1 : 2;
        double L2 = (s2s == 1) ? lpL : lpR;
        if (L2 < 0.0) {
            L2 = 0.0;
            s2s = 0;
        }
        const double den0 = L2 - L1;
        const bool den_free = !(den0 < kSmall);
        const double den = den_free ? den0 : kSmall;
        const double iden = 1.0 / den, iden2 = iden * iden;
        const double dphi = phiL - phiR;
        const double sq = sqrt(dphi * dphi + kPhiEps * kPhiEps);
        const double phiLR = 0.5 * (phiL + phiR - sq);
        const double A1 = L2 * qL - L1 * qR + L1 * L2 * (zR - zL);
        const double N1 = phiLR * A1;
        const double f1 = N1 * iden;  // sign of N1 / den
        const double fact = 0.5 * phiLR * (hL + hR);
        const double s2a = (phiR * hR * hR - phiL * hL * hL) * 0.5 + fact * (zL - zR);
        const double s2b = s2a * kG;
        const double s2 = s2b * iden;
        const double sl = -L1 * s2, sr = L2 * s2;
        const double FL = sp::specific_force(bL, phiL, hL, uL);
        const double FR = sp::specific_force(bR, phiR, hR, uR);
        const double P = phiR * qR - phiL * qL;
        const double A2 = L2 * (FL + sl) - L1 * (FR - sr) + L1 * L2 * P;
        const double vL = rL * ihL, vR = rR * ihR;
        double df1 = df[0];
        const double df2 = df[1], df3 = df[2];
        double dsl = df[3], dsr = df[4];
        double dhLs = 0.0, dhRs = 0.0, dL1 = 0.0, dL2 = 0.0, dden = 0.0, dphiLR = 0.0;
        double duL = 0.0, duR = 0.0;
        // f3
        if (f1 >= 0.0) {
            df1 += df3 * bL * vL;
            const double dvL = df3 * bL * f1;
            d[2] += dvL * ihL;
            dhLs += -dvL * rL * (ihL * ihL);
        } else {
            df1 += df3 * bR * vR;
            const double dvR = df3 * bR * f1;
            d[5] += dvR * ihR;
            dhRs += -dvR * rR * (ihR * ihR);
        }
        // f2 = A2 / den
        const double dA2 = df2 * iden;
        dden += -df2 * A2 * iden2;
        dL2 += dA2 * (FL + sl) + dA2 * L1 * P;
        const double dFL = dA2 * L2;
        dsl += dA2 * L2;
        dL1 += -dA2 * (FR - sr) + dA2 * L2 * P;
        const double dFR = -dA2 * L1;
        dsr += dA2 * L1;
        const double dP = dA2 * L1 * L2;
        d[11] += dP * qR;
        d[4] += dP * phiR;
        d[10] += -dP * qL;
        d[1] += -dP * phiL;
        // specific forces
        d[10] += dFL * (bL * hL * uL * uL + 0.5 * kG * hL * hL);
        dhL += dFL * phiL * (bL * uL * uL + kG * hL);
        duL += dFL * phiL * bL * hL * 2.0 * uL;
        d[11] += dFR * (bR * hR * uR * uR + 0.5 * kG * hR * hR);
        dhR += dFR * phiR * (bR * uR * uR + kG * hR);
        duR += dFR * phiR * bR * hR * 2.0 * uR;
        // sl = -L1 s2, sr = L2 s2
        dL1 += -dsl * s2;
        dL2 += dsr * s2;
        const double ds2 = -dsl * L1 + dsr * L2;
        const double ds2a = ds2 * kG * iden;
        dden += -ds2 * s2b * iden2;
        d[11] += ds2a * 0.5 * hR * hR;
        dhR += ds2a * phiR * hR;
        d[10] += -ds2a * 0.5 * hL * hL;
        dhL += -ds2a * phiL * hL;
        const double dfact = ds2a * (zL - zR);
        d[0] += ds2a * fact;
        d[3] += -ds2a * fact;
        dphiLR += dfact * 0.5 * (hL + hR);
        dhL += dfact * 0.5 * phiLR;
        dhR += dfact * 0.5 * phiLR;
        // f1 = phiLR A1 / den
        const double dN1 = df1 * iden;
        dden += -df1 * N1 * iden2;
        dphiLR += dN1 * A1;
        const double dA1 = dN1 * phiLR;
        dL2 += dA1 * (qL + L1 * (zR - zL));
        d[1] += dA1 * L2;
        dL1 += dA1 * (-qR + L2 * (zR - zL));
        d[4] += -dA1 * L1;
        d[3] += dA1 * L1 * L2;
        d[0] += -dA1 * L1 * L2;
        // smoothed min
        const double dsq = -0.5 * dphiLR;
        d[10] += 0.5 * dphiLR;
        d[11] += 0.5 * dphiLR;
        const double dd = (sq > 0.0) ? dsq * dphi / sq : 0.0;
        d[10] += dd;
        d[11] -= dd;
        // den
        if (den_free) {
            dL2 += dden;
            dL1 -= dden;
        }
        // selected by value, not by index, so d stays in registers
        d[6] += (s1 == 1) ? dL1 : 0.0;
        d[7] += (s1 == 2) ? dL1 : 0.0;
        d[8] += (s2s == 1) ? dL2 : 0.0;
        d[9] += (s2s == 2) ? dL2 : 0.0;
        // velocities
        d[1] += duL * ihL;
        dhLs += -duL * qL * (ihL * ihL);
        d[4] += duR * ihR;
        dhRs += -duR * qR * (ihR * ihR);
        if (hL > kSmall) dhL += dhLs;
        if (hR > kSmall) dhR += dhRs;
    } else if (reg == R_RHS) {
        d[11] += df[4] * 0.5 * kG * hR * hR;
        dhR += df[4] * phiR * kG * hR;
    } else if (reg == R_LHS) {
        d[10] += -df[3] * 0.5 * kG * hL * hL;
        dhL += -df[3] * phiL * kG * hL;
    }
    if (zL - zbL > 0.0) d[0] += dhL;
    if (zR - zbR > 0.0) d[3] += dhR;
}

// ---------------------------------------------------------------------------
// boundary flux (swSpE_hyp.cpp:143-239); in: zR qR rR lpR phiR bc ; out F1 F2 F3
// ---------------------------------------------------------------------------
SPD void bnd_fwd(double h_min, double phi_min, double bR, double phiR, double zbR, double zR,
                 double qR, double rR, double lpR, int type, double bc, double *f) {
    const double hR = zR - zbR;
    const bool dry = (hR < h_min);
    const double q1 = dry ? 0.0 : qR, r1 = dry ? 0.0 : rR;
    const double hRe = dry ? 0.0 : hR;
    const double hden = (hRe > kSmall) ? hRe : kSmall;
    const double cR = sqrt(kG * hden);
    const double uR = q1 / hden;
    const double L1 = (lpR > 0.0) ? lpR : 0.0;
    const double f2R = sp::specific_force(bR, phiR, hRe, uR);
    double f1 = 0.0, f2 = f2R, f3;
    if (type == sp::BC_Q) {
        f1 = bc;
        if (bc > kSmall) {
            const double ps = (phiR > phi_min) ? phiR : phi_min;
            const double hc = sw_exp((1.0 / 3.0) * sw_log((bc / ps) / kG));
            if (hc >= hRe) {
                const double hcs = (hc > kSmall) ? hc : kSmall;
                f1 = phiR * 1.5 * sqrt(kG * hcs) * hc;
            }
        }
        f2 = f2R + (f1 - q1 * phiR) * L1;
    } else if (type == sp::BC_FR) {
        const double fr = (bc < 0.0) ? 0.0 : ((bc > 1.0) ? 1.0 : bc);
        double c = (uR - 2.0 * cR) / (-fr - 2.0);
        c = (c > 0.0) ? c : 0.0;
        const double u = -fr * c;
        const double h = c * c / kG;
        f1 = phiR * u * h;
        f2 = f2R + (f1 - q1 * phiR) * L1;
    } else if (type == sp::BC_H || type == sp::BC_Z) {
        const double hb = (type == sp::BC_Z) ? ((bc - zbR > 0.0) ? bc - zbR : 0.0)
                                             : ((bc > 0.0) ? bc : 0.0);
        if (hRe > h_min) {
            f1 = phiR * (q1 + L1 * (hb - hRe));
            f2 = f2R + (f1 - q1 * phiR) * L1;
        } else {
            const double c = sqrt(kG * ((hb > kSmall) ? hb : kSmall));
            f1 = phiR * hb * c;
            f2 = sp::specific_force(bR, phiR, hb, c);
        }
    }
    if (f1 > 0.0) {
        f3 = 0.0;
    } else {
        const double vR = r1 / ((hRe > kSmall) ? hRe : kSmall);
        f3 = f1 * vR;
    }
    f[0] = f1;
    f[1] = f2;
    f[2] = f3;
}

// d[0..5] accumulates the adjoints of zR qR rR lpR phiR bc
SPD void bnd_vjp(double h_min, double phi_min, double bR, double phiR, double zbR, double zR,
                 double qR, double rR, double lpR, int type, double bc, const double *df, double *d) {
    const double hR = zR - zbR;
    const bool dry = (hR < h_min);
    const double q1 = dry ? 0.0 : qR, r1 = dry ? 0.0 : rR;
    const double hRe = dry ? 0.0 : hR;
    const double hden = (hRe > kSmall) ? hRe : kSmall;
    const double cR = sqrt(kG * hden);
    const double uR = q1 / hden;
    const double L1 = (lpR > 0.0) ? lpR : 0.0;
    double f[3];
    bnd_fwd(h_min, phi_min, bR, phiR, zbR, zR, qR, rR, lpR, type, bc, f);
    const double f1 = f[0];
    double df1 = df[0];
    const double df2 = df[1], df3 = df[2];
    double dq1 = 0.0, dr1 = 0.0, dhRe = 0.0, duR = 0.0, dcR = 0.0, dL1 = 0.0, dphi = 0.0,
           dbc = 0.0, df2R = 0.0;
    // f3
    if (!(f1 > 0.0)) {
        const double hd = (hRe > kSmall) ? hRe : kSmall;
        df1 += df3 * r1 / hd;
        dr1 += df3 * f1 / hd;
        if (hRe > kSmall) dhRe += -df3 * f1 * r1 / (hd * hd);
    }
    if (type == sp::BC_Q || type == sp::BC_FR ||
        ((type == sp::BC_H || type == sp::BC_Z) && hRe > h_min)) {
        // f2 = f2R + (f1 - q1 phiR) L1
        df2R += df2;
        df1 += df2 * L1;
        dq1 += -df2 * L1 * phiR;
        dphi += -df2 * L1 * q1;
        dL1 += df2 * (f1 - q1 * phiR);
    }
    if (type == sp::BC_Q) {
        if (bc > kSmall) {
            const double ps = (phiR > phi_min) ? phiR : phi_min;
            const double x = (bc / ps) / kG;
            const double hc = sw_exp((1.0 / 3.0) * sw_log(x));
            if (hc >= hRe) {
                const double hcs = (hc > kSmall) ? hc : kSmall;
                const double sg = sqrt(kG * hcs);
                dphi += df1 * 1.5 * sg * hc;
                double dhc = df1 * phiR * 1.5 * sg;
                if (hc > kSmall) dhc += dsqrt_arg(sg, df1 * phiR * 1.5 * hc) * kG;
                const double dx = dhc * hc * (1.0 / 3.0) / x;
                dbc += dx / (ps * kG);
                if (phiR > phi_min) dphi += -dx * bc / (ps * ps * kG);
            } else {
                dbc += df1;
            }
        } else {
            dbc += df1;
        }
    } else if (type == sp::BC_FR) {
        const double fr = (bc < 0.0) ? 0.0 : ((bc > 1.0) ? 1.0 : bc);
        const double B = -fr - 2.0;
        const double A = uR - 2.0 * cR;
        const double c0 = A / B;
        const double c = (c0 > 0.0) ? c0 : 0.0;
        const double u = -fr * c;
        const double h = c * c / kG;
        dphi += df1 * u * h;
        const double du = df1 * phiR * h;
        const double dh = df1 * phiR * u;
        double dc = dh * 2.0 * c / kG - du * fr;
        double dfr = -du * c;
        if (c0 > 0.0) {
            const double dA = dc / B;
            const double dB = -dc * A / (B * B);
            duR += dA;
            dcR += -2.0 * dA;
            dfr += -dB;
        }
        if (!(bc < 0.0) && !(bc > 1.0)) dbc += dfr;
    } else if (type == sp::BC_H || type == sp::BC_Z) {
        const double hb = (type == sp::BC_Z) ? ((bc - zbR > 0.0) ? bc - zbR : 0.0)
                                             : ((bc > 0.0) ? bc : 0.0);
        double dhb = 0.0;
        if (hRe > h_min) {
            dphi += df1 * (q1 + L1 * (hb - hRe));
            dq1 += df1 * phiR;
            dL1 += df1 * phiR * (hb - hRe);
            dhb += df1 * phiR * L1;
            dhRe += -df1 * phiR * L1;
        } else {
            const double c = sqrt(kG * ((hb > kSmall) ? hb : kSmall));
            dphi += df1 * hb * c + df2 * (bR * hb * c * c + 0.5 * kG * hb * hb);
            dhb += df1 * phiR * c + df2 * phiR * (bR * c * c + kG * hb);
            const double dc = df1 * phiR * hb + df2 * phiR * bR * hb * 2.0 * c;
            if (hb > kSmall) dhb += dsqrt_arg(c, dc) * kG;
        }
        if (type == sp::BC_Z ? (bc - zbR > 0.0) : (bc > 0.0)) dbc += dhb;
    }
    // f2R = phiR (bR hRe uR^2 + g hRe^2 / 2)
    dphi += df2R * (bR * hRe * uR * uR + 0.5 * kG * hRe * hRe);
    dhRe += df2R * phiR * (bR * uR * uR + kG * hRe);
    duR += df2R * phiR * bR * hRe * 2.0 * uR;
    // L1 = max(lpR, 0)
    if (lpR > 0.0) d[3] += dL1;
    // uR = q1 / hden ; cR = sqrt(g hden)
    dq1 += duR / hden;
    double dhden = -duR * q1 / (hden * hden) + dsqrt_arg(cR, dcR) * kG;
    if (hRe > kSmall) dhRe += dhden;
    if (!dry) {
        d[1] += dq1;
        d[2] += dr1;
        d[0] += dhRe;
    }
    d[4] += dphi;
    d[5] += dbc;
}

// ---------------------------------------------------------------------------
// semi-implicit Strickler friction (swSpE_friction.cpp:19-76): (h,qx,qy,K1,K2) -> (qx,qy)
// ---------------------------------------------------------------------------
SPD void fric_fwd(double h_min, double h, double &qx, double &qy, double ca, double sa, double K1in,
                  double K2in, double dt) {
    const double hthr = (h_min > kSmall) ? h_min : kSmall;
    if (h > hthr) {
        const double hs = h + kSmall;
        const double u = qx / hs, v = qy / hs;
        double ul = u * ca + v * sa;
        double vl = -u * sa + v * ca;
        const double un = sqrt(u * u + v * v + 1e-16);
        const double K1 = (K1in > 1e-10) ? K1in : 20.0;
        const double K2 = (K2in > 1e-10) ? K2in : 20.0;
        const double h43 = h * sw_cbrt(h + 1e-16);
        const double d1 = 1.0 + (dt * kG * un) / (K1 * K1 * h43 + 1e-10);
        const double d2 = 1.0 + (dt * kG * un) / (K2 * K2 * h43 + 1e-10);
        ul /= d1;
        vl /= d2;
        qx = (ul * ca - vl * sa) * h;
        qy = (ul * sa + vl * ca) * h;
    } else {
        qx = 0.0;
        qy = 0.0;
    }
}

// adjoint: dqx, dqy (of the outputs) in/out become the adjoints of the inputs;
// dh accumulates; dK1, dK2 accumulate.
SPD void fric_vjp(double h_min, double h, double qx, double qy, double ca, double sa, double K1in,
                  double K2in, double dt, double &dh, double &dqx, double &dqy, double &dK1,
                  double &dK2) {
    const double hthr = (h_min > kSmall) ? h_min : kSmall;
    if (!(h > hthr)) {
        dqx = 0.0;
        dqy = 0.0;
        return;
    }
    const double hs = h + kSmall;
    const double ihs = 1.0 / hs;
    const double u = qx * ihs, v = qy * ihs;
    const double ul = u * ca + v * sa;
    const double vl = -u * sa + v * ca;
    const double un = sqrt(u * u + v * v + 1e-16);
    const double K1 = (K1in > 1e-10) ? K1in : 20.0;
    const double K2 = (K2in > 1e-10) ? K2in : 20.0;
    const double cb = sw_cbrt(h + 1e-16);
    const double h43 = h * cb;
    const double D1 = K1 * K1 * h43 + 1e-10, D2 = K2 * K2 * h43 + 1e-10;
    const double iD1 = 1.0 / D1, iD2 = 1.0 / D2;
    const double gdt = dt * kG;
    const double d1 = 1.0 + gdt * un * iD1;
    const double d2 = 1.0 + gdt * un * iD2;
    const double id1 = 1.0 / d1, id2 = 1.0 / d2;
    const double ul2 = ul * id1, vl2 = vl * id2;
    const double T1 = ul2 * ca - vl2 * sa, T2 = ul2 * sa + vl2 * ca;
    dh += dqx * T1 + dqy * T2;
    const double dT1 = dqx * h, dT2 = dqy * h;
    const double dul2 = dT1 * ca + dT2 * sa;
    const double dvl2 = -dT1 * sa + dT2 * ca;
    const double dul = dul2 * id1, dvl = dvl2 * id2;
    const double dd1 = -dul2 * ul * (id1 * id1), dd2 = -dvl2 * vl * (id2 * id2);
    const double dun = (dd1 * iD1 + dd2 * iD2) * gdt;
    const double dD1 = -dd1 * gdt * un * (iD1 * iD1), dD2 = -dd2 * gdt * un * (iD2 * iD2);
    if (K1in > 1e-10) dK1 += dD1 * 2.0 * K1 * h43;
    if (K2in > 1e-10) dK2 += dD2 * 2.0 * K2 * h43;
    const double dh43 = dD1 * K1 * K1 + dD2 * K2 * K2;
    dh += dh43 * (cb + h * cb / (3.0 * (h + 1e-16)));
    const double iun = 1.0 / un;
    double du = dun * u * iun + dul * ca - dvl * sa;
    double dv = dun * v * iun + dul * sa + dvl * ca;
    dqx = du * ihs;
    dqy = dv * ihs;
    dh += -(du * qx + dv * qy) * (ihs * ihs);
}

// ---------------------------------------------------------------------------
// singular head loss (swSpE_friction.cpp:79-126): (h,qx,qy) -> (qx,qy)
// ---------------------------------------------------------------------------
SPD void hl_fwd(double h_min, double h, double &qx, double &qy, double ca, double sa, double hsing,
                double k1, double k2, double dt) {
    const double hthr = (h_min > kSmall) ? h_min : kSmall;
    double ax = qx * ca + qy * sa;
    double ay = -qx * sa + qy * ca;
    if (h > hthr) {
        double Ax = 0.0, Ay = 0.0;
        if (h > hsing) {
            const double hs = h + kSmall;
            const double qn = sqrt(ax * ax + ay * ay + 1e-16);
            Ax = qn * k1 / hs;
            Ay = qn * k2 / hs;
        }
        ax = ax * sw_exp(-Ax * dt);
        ay = ay * sw_exp(-Ay * dt);
    } else {
        ax = 0.0;
        ay = 0.0;
    }
    qx = ax * ca - ay * sa;
    qy = ax * sa + ay * ca;
}

SPD void hl_vjp(double h_min, double h, double qx, double qy, double ca, double sa, double hsing,
                double k1, double k2, double dt, double &dh, double &dqx, double &dqy) {
    const double hthr = (h_min > kSmall) ? h_min : kSmall;
    const double dbx = dqx * ca + dqy * sa;
    const double dby = -dqx * sa + dqy * ca;
    double dax = 0.0, day = 0.0;
    if (h > hthr) {
        const double ax = qx * ca + qy * sa;
        const double ay = -qx * sa + qy * ca;
        double Ax = 0.0, Ay = 0.0, qn = 0.0, hs = 0.0;
        if (h > hsing) {
            hs = h + kSmall;
            qn = sqrt(ax * ax + ay * ay + 1e-16);
            Ax = qn * k1 / hs;
            Ay = qn * k2 / hs;
        }
        const double ex = sw_exp(-Ax * dt), ey = sw_exp(-Ay * dt);
        dax = dbx * ex;
        day = dby * ey;
        const double dAx = -dbx * ax * ex * dt, dAy = -dby * ay * ey * dt;
        if (h > hsing) {
            const double dqn = dAx * k1 / hs + dAy * k2 / hs;
            dh += -(dAx * qn * k1 + dAy * qn * k2) / (hs * hs);
            dax += dqn * ax / qn;
            day += dqn * ay / qn;
        }
    }
    dqx = dax * ca - day * sa;
    dqy = dax * sa + day * ca;
}

// Balance contributions of one internal-edge side (swSpE_hyp.cpp:356-383) and
// their adjoint.  side 0 = left cell (factors -c,-c nx,+c ny,+c nx,...).
SPD void side_contrib(int side, double co, double nx, double ny, const double *F, double dt,
                      double &dh, double &dqx, double &dqy) {
    const double s = F[3 + side];
    if (side == 0) {
        dh = (-co) * (F[0] * dt);
        dqx = ((-co * nx) * F[1] + (co * nx) * s + (co * ny) * F[2]) * dt;
        dqy = ((-co * ny) * F[1] + (co * ny) * s + (-co * nx) * F[2]) * dt;
    } else {
        dh = (co) * (F[0] * dt);
        dqx = ((co * nx) * F[1] + (co * nx) * s + (-co * ny) * F[2]) * dt;
        dqy = ((co * ny) * F[1] + (co * ny) * s + (co * nx) * F[2]) * dt;
    }
}

// adjoint of side_contrib w.r.t. F (accumulated into dF[5]) and the factor co
SPD double side_contrib_vjp(int side, double co, double nx, double ny, const double *F, double dt,
                            double ah, double ax, double ay, double *dF) {
    const double sg = side == 0 ? -1.0 : 1.0;
    dF[0] += ah * sg * co * dt;
    dF[1] += dt * (ax * sg * co * nx + ay * sg * co * ny);
    dF[2] += dt * (ax * (-sg) * co * ny + ay * sg * co * nx);
    dF[3 + side] += dt * (ax * co * nx + ay * co * ny);
    const double s = F[3 + side];
    return ah * sg * F[0] * dt +
           dt * (ax * (sg * nx * F[1] + nx * s - sg * ny * F[2]) +
                 ay * (sg * ny * F[1] + ny * s + sg * nx * F[2]));
}

// coeff = w / area / phi (floored as computeEdgeFactors_E) and d coeff / d phi
SPD double edge_coeff(double w, double area, double phi, double phi_min) {
    const double pc = (phi > phi_min) ? phi : phi_min;
    return (phi > phi_min) ? (w / area / pc) : 0.0;
}
SPD double edge_coeff_dphi(double w, double area, double phi, double phi_min) {
    return (phi > phi_min) ? -(w / area / phi) / phi : 0.0;
}

}  // namespace spe
