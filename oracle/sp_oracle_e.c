/*
 * sp_oracle_e.c -- CPU oracle (test infrastructure only) of the reference's
 * differentiable "_E" forward, the function its 4D-Var layer differentiates:
 *   assim/swSpE_hyp.cpp      hypGodunovComputeLRStates_E (507-573),
 *                            hypComputeTimeStep_E (578-635),
 *                            hypBoundaryFlux_E (143-239), hypHllc_E (725-832),
 *                            hypComputeFluxes_E (396-482),
 *                            hypComputeBalance_E (336-394),
 *                            computeEdgeFactors_E (277-333)
 *   assim/swSpE_friction.cpp sourceFrictionStrickler_E (19-76),
 *                            sourceSingularHeadLoss_E (79-126)
 *   assim/swSpE.h            interpolateTimeSeries_E (138-153)
 *   assim/swSpE_checkpoint.cpp build_checkpoint_records (97-455, the dry run
 *                            that fixes dt / BC / observation triggers) and
 *                            ckpt_segment_E (584-715, the frozen replay).
 * Godunov reconstruction only.  Costs: gauge misfit (swSpE_gauges.cpp:166-179)
 * and full-frame misfit (swSpE_ad.cpp:273-308), optional normalisation by the
 * number of observations.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sp_oracle.h"

#define SMALL 1e-16
#define G 9.81

typedef struct {
    const orc_case *c;
    double *h, *qx, *qy;
    double *fac_b, *fac_L, *fac_R;   /* coeff = w/area/phi (E floors phi inside) */
    double *zrb, *qrb, *rrb, *lpb, *lmaxb;
    double *zri, *qri, *rri, *lmi, *lpi, *lmaxi;
    double *F1b, *F2b, *F3b, *F1i, *F2i, *F3i, *sLi, *sRi;
    double *swept;
    double *fcos, *fsin, *hcos, *hsin;
    const double *phi, *K1, *K2;
} esolver;

static double lam_m(double b, double c, double u) { return b * u - sqrt((b - 1.0) * b * u * u + c * c); }
static double lam_p(double b, double c, double u) { return b * u + sqrt((b - 1.0) * b * u * u + c * c); }
static double sforce(double b, double phi, double h, double u) {
    return phi * (b * h * u * u + 0.5 * G * h * h);
}

/* interpolateTimeSeries_E -- assim/swSpE.h:138-153 */
double orc_e_interp(const double *tt, const double *vv, int n, double t) {
    if (n == 0) return 0.0;
    if (n == 1 || t <= tt[0] - 1.0e-12) return vv[0];
    if (t >= tt[n - 1]) return vv[n - 1];
    int it = 0;
    while (it < n - 1 && t > tt[it + 1]) it++;
    double t1 = tt[it], t2 = tt[it + 1], v1 = vv[it], v2 = vv[it + 1];
    double den = t2 - t1;
    if (den < 1e-12) return v1;
    return v1 + (v2 - v1) * (t - t1) / den;
}

/* computeEdgeFactors_E -- swSpE_hyp.cpp:277-333 */
static void factors(esolver *s) {
    const orc_case *c = s->c;
    for (int64_t i = 0; i < c->n_bnd; i++) {
        int64_t k = c->bnd_cell[i];
        double ph = (s->phi[k] > c->phi_min) ? s->phi[k] : c->phi_min;
        s->fac_b[i] = (s->phi[k] > c->phi_min) ? (c->bnd_len[i] / c->cell_area[k] / ph) : 0.0;
    }
    for (int64_t i = 0; i < c->n_int; i++) {
        int64_t L = c->int_L[i], R = c->int_R[i];
        double pL = (s->phi[L] > c->phi_min) ? s->phi[L] : c->phi_min;
        double pR = (s->phi[R] > c->phi_min) ? s->phi[R] : c->phi_min;
        s->fac_L[i] = (s->phi[L] > c->phi_min) ? (c->int_len[i] / c->cell_area[L] / pL) : 0.0;
        s->fac_R[i] = (s->phi[R] > c->phi_min) ? (c->int_len[i] / c->cell_area[R] / pR) : 0.0;
    }
}

/* hypGodunovComputeLRStates_E -- swSpE_hyp.cpp:507-573 */
static void lr(esolver *s) {
    const orc_case *c = s->c;
    const double hmin = c->h_min;
    for (int64_t e = 0; e < c->n_bnd; e++) {
        double nx = c->bnd_nx[e], ny = c->bnd_ny[e];
        int64_t id = c->bnd_cell[e];
        double b = c->beta[id], q = s->qx[id], r = s->qy[id], hR = s->h[id];
        s->zrb[e] = s->h[id] + c->cell_zb[id];
        s->qrb[e] = q * nx + r * ny;
        s->rrb[e] = -q * ny + r * nx;
        double uR = (hR > hmin) ? s->qrb[e] / hR : 0.0;
        double cR = sqrt(G * hR);
        double lm = lam_m(b, cR, uR), lp = lam_p(b, cR, uR);
        s->lpb[e] = lp;
        s->lmaxb[e] = (fabs(lp) < fabs(lm)) ? fabs(lm) : fabs(lp);
    }
    for (int64_t e = 0; e < c->n_int; e++) {
        double nx = c->int_nx[e], ny = c->int_ny[e];
        for (int sd = 0; sd < 2; sd++) {
            int64_t id = sd ? c->int_R[e] : c->int_L[e];
            double b = c->beta[id], q = s->qx[id], r = s->qy[id], hh = s->h[id];
            s->zri[2 * e + sd] = s->h[id] + c->cell_zb[id];
            s->qri[2 * e + sd] = q * nx + r * ny;
            s->rri[2 * e + sd] = -q * ny + r * nx;
            double u = (hh > hmin) ? s->qri[2 * e + sd] / hh : 0.0;
            double cc = sqrt(G * hh);
            s->lmi[2 * e + sd] = lam_m(b, cc, u);
            s->lpi[2 * e + sd] = lam_p(b, cc, u);
        }
        double m = fabs(s->lmi[2 * e]), v;
        if (m < (v = fabs(s->lpi[2 * e]))) m = v;
        if (m < (v = fabs(s->lmi[2 * e + 1]))) m = v;
        if (m < (v = fabs(s->lpi[2 * e + 1]))) m = v;
        s->lmaxi[e] = m;
    }
}

/* hypComputeTimeStep_E -- swSpE_hyp.cpp:578-635 (without the t_max clamp and
 * floor, applied by the caller as in build_checkpoint_records) */
static double cfl_dt(esolver *s) {
    const orc_case *c = s->c;
    double dt = c->dt_max;
    for (int64_t i = 0; i < c->n_cell; i++) s->swept[i] = 0.0;
    for (int64_t e = 0; e < c->n_bnd; e++) {
        int64_t id = c->bnd_cell[e];
        if (s->h[id] > c->h_min) s->swept[id] += s->lmaxb[e] * c->bnd_len[e];
    }
    for (int64_t e = 0; e < c->n_int; e++) {
        int64_t L = c->int_L[e], R = c->int_R[e];
        if (s->h[L] > c->h_min) s->swept[L] += s->lmaxi[e] * c->int_len[e];
        if (s->h[R] > c->h_min) s->swept[R] += s->lmaxi[e] * c->int_len[e];
    }
    for (int64_t i = 0; i < c->n_cell; i++) {
        double den = (s->swept[i] < SMALL) ? SMALL : s->swept[i];
        double x = c->cfl_max * (c->cell_area[i] / den);
        if (x < dt) dt = x;
    }
    return dt;
}

/* hypBoundaryFlux_E -- swSpE_hyp.cpp:143-239 */
void orc_e_bnd_flux(double h_min, double phi_min, double bR, double phiR, double zbR, double zR,
                    double qR, double rR, double lpR, int type, double bc, double *f) {
    double hR = zR - zbR;
    int dry = (hR < h_min);
    double q1 = dry ? 0.0 : qR, r1 = dry ? 0.0 : rR;
    double hRe = dry ? 0.0 : hR;
    double hden = (hRe > SMALL) ? hRe : SMALL;
    double cR = sqrt(G * hden);
    double uR = q1 / hden;
    double L1 = (lpR > 0.0) ? lpR : 0.0;
    double f2R = sforce(bR, phiR, hRe, uR);
    double f1 = 0.0, f2 = f2R, f3;
    if (type == ORC_BC_Q) {
        f1 = bc;
        if (bc > SMALL) {
            double ps = (phiR > phi_min) ? phiR : phi_min;
            double hc = exp((1.0 / 3.0) * log((bc / ps) / G));
            if (hc >= hRe) {
                double hcs = (hc > SMALL) ? hc : SMALL;
                f1 = phiR * 1.5 * sqrt(G * hcs) * hc;
            }
        }
        f2 = f2R + (f1 - q1 * phiR) * L1;
    } else if (type == ORC_BC_FR) {
        double fr = (bc < 0.0) ? 0.0 : ((bc > 1.0) ? 1.0 : bc);
        double cc = (uR - 2.0 * cR) / (-fr - 2.0);
        cc = (cc > 0.0) ? cc : 0.0;
        double u = -fr * cc;
        double h = cc * cc / G;
        f1 = phiR * u * h;
        f2 = f2R + (f1 - q1 * phiR) * L1;
    } else if (type == ORC_BC_H || type == ORC_BC_Z) {
        double hb = (type == ORC_BC_Z) ? ((bc - zbR > 0.0) ? bc - zbR : 0.0) : ((bc > 0.0) ? bc : 0.0);
        if (hRe > h_min) {
            f1 = phiR * (q1 + L1 * (hb - hRe));
            f2 = f2R + (f1 - q1 * phiR) * L1;
        } else {
            double cc = sqrt(G * ((hb > SMALL) ? hb : SMALL));
            f1 = phiR * hb * cc;
            f2 = sforce(bR, phiR, hb, cc);
        }
    }
    if (f1 > 0.0) {
        f3 = 0.0;
    } else {
        double vR = r1 / ((hRe > SMALL) ? hRe : SMALL);
        f3 = f1 * vR;
    }
    f[0] = f1;
    f[1] = f2;
    f[2] = f3;
}

/* hypHllc_E -- swSpE_hyp.cpp:725-832 */
void orc_e_hllc(double h_min, double phi_min, double bL, double bR, double phiL, double phiR,
                double zbL, double zbR, double zL, double qL, double rL, double zR, double qR,
                double rR, double lmL, double lmR, double lpL, double lpR, double *f) {
    double hL = (zL - zbL > 0.0) ? (zL - zbL) : 0.0;
    double hR = (zR - zbR > 0.0) ? (zR - zbR) : 0.0;
    int both = 0, lhs = 0, rhs = 0; /* all zero: no flow through the edge */
    if (!(((hL <= h_min) && (hR <= h_min)) || ((phiL <= phi_min) && (phiR <= phi_min)))) {
        both = 1;
        if ((phiR <= phi_min) || ((phiR > phi_min) && (hR <= h_min) && (zL <= zbR + h_min))) {
            lhs = 1;
            both = 0;
        }
        if ((phiL <= phi_min) || ((phiL > phi_min) && (hL <= h_min) && (zR <= zbL + h_min))) {
            rhs = 1;
            both = 0;
        }
    }
    double f1 = 0, f2 = 0, f3 = 0, sl = 0, sr = 0;
    if (both) {
        double hLs = (hL > SMALL) ? hL : SMALL, hRs = (hR > SMALL) ? hR : SMALL;
        double uL = qL / hLs, uR = qR / hRs;
        double L1 = (lmL < lmR) ? lmL : lmR;
        if (L1 > 0.0) L1 = 0.0;
        double L2 = (lpL > lpR) ? lpL : lpR;
        if (L2 < 0.0) L2 = 0.0;
        double den = L2 - L1;
        if (den < SMALL) den = SMALL;
        double d = phiL - phiR;
        double phiLR = 0.5 * (phiL + phiR - sqrt(d * d + 1e-3 * 1e-3));
        f1 = phiLR * (L2 * qL - L1 * qR + L1 * L2 * (zR - zL)) / den;
        double fact = 0.5 * phiLR * (hL + hR);
        double s2 = (phiR * hR * hR - phiL * hL * hL) * 0.5 + fact * (zL - zR);
        s2 = s2 * G / den;
        sl = -L1 * s2;
        sr = L2 * s2;
        double FL = sforce(bL, phiL, hL, uL), FR = sforce(bR, phiR, hR, uR);
        f2 = (L2 * (FL + sl) - L1 * (FR - sr) + L1 * L2 * (phiR * qR - phiL * qL)) / den;
        double vL = rL / hLs, vR = rR / hRs;
        f3 = (f1 >= 0.0) ? (bL * f1 * vL) : (bR * f1 * vR);
    } else if (rhs) {
        sr = phiR * 0.5 * G * hR * hR;
    } else if (lhs) {
        sl = -phiL * 0.5 * G * hL * hL;
    }
    f[0] = f1;
    f[1] = f2;
    f[2] = f3;
    f[3] = sl;
    f[4] = sr;
}

static void fluxes(esolver *s, const double *bc_step) {
    const orc_case *c = s->c;
    for (int64_t e = 0; e < c->n_bnd; e++) {
        int64_t id = c->bnd_cell[e];
        int k = c->bnd_series[e];
        double f[3];
        orc_e_bnd_flux(c->h_min, c->phi_min, c->beta[id], s->phi[id], c->cell_zb[id], s->zrb[e],
                       s->qrb[e], s->rrb[e], s->lpb[e], c->series_type[k], bc_step[k], f);
        s->F1b[e] = f[0];
        s->F2b[e] = f[1];
        s->F3b[e] = f[2];
    }
    for (int64_t e = 0; e < c->n_int; e++) {
        int64_t L = c->int_L[e], R = c->int_R[e];
        double f[5];
        orc_e_hllc(c->h_min, c->phi_min, c->beta[L], c->beta[R], s->phi[L], s->phi[R], c->cell_zb[L],
                   c->cell_zb[R], s->zri[2 * e], s->qri[2 * e], s->rri[2 * e], s->zri[2 * e + 1],
                   s->qri[2 * e + 1], s->rri[2 * e + 1], s->lmi[2 * e], s->lmi[2 * e + 1],
                   s->lpi[2 * e], s->lpi[2 * e + 1], f);
        s->F1i[e] = f[0];
        s->F2i[e] = f[1];
        s->F3i[e] = f[2];
        s->sLi[e] = f[3];
        s->sRi[e] = f[4];
    }
}

/* hypComputeBalance_E -- swSpE_hyp.cpp:336-394 (signed factors of
 * computeEdgeFactors_E written out) */
static void balance(esolver *s, double dt) {
    const orc_case *c = s->c;
    double *h = s->h, *qx = s->qx, *qy = s->qy;
    for (int64_t e = 0; e < c->n_bnd; e++) {
        int64_t id = c->bnd_cell[e];
        double co = s->fac_b[e], nx = c->bnd_nx[e], ny = c->bnd_ny[e];
        h[id] += (+co) * s->F1b[e] * dt;
        qx[id] += ((+co * nx) * s->F2b[e] + (+co * nx) * 0.0 + (-co * ny) * s->F3b[e]) * dt;
        qy[id] += ((+co * ny) * s->F2b[e] + (+co * ny) * 0.0 + (+co * nx) * s->F3b[e]) * dt;
    }
    for (int64_t e = 0; e < c->n_int; e++) {
        int64_t L = c->int_L[e], R = c->int_R[e];
        double cl = s->fac_L[e], cr = s->fac_R[e], nx = c->int_nx[e], ny = c->int_ny[e];
        double dh = s->F1i[e] * dt;
        h[L] += (-cl) * dh;
        h[R] += (+cr) * dh;
        qx[L] += ((-cl * nx) * s->F2i[e] + (+cl * nx) * s->sLi[e] + (+cl * ny) * s->F3i[e]) * dt;
        qx[R] += ((+cr * nx) * s->F2i[e] + (+cr * nx) * s->sRi[e] + (-cr * ny) * s->F3i[e]) * dt;
        qy[L] += ((-cl * ny) * s->F2i[e] + (+cl * ny) * s->sLi[e] + (-cl * nx) * s->F3i[e]) * dt;
        qy[R] += ((+cr * ny) * s->F2i[e] + (+cr * ny) * s->sRi[e] + (+cr * nx) * s->F3i[e]) * dt;
    }
    for (int64_t n = 0; n < c->n_cell; ++n) {
        if (fabs(h[n]) < SMALL) h[n] = 0.0;
        if (fabs(qx[n]) < SMALL) qx[n] = 0.0;
        if (fabs(qy[n]) < SMALL) qy[n] = 0.0;
    }
}

static void sources(esolver *s, double dt) {
    const orc_case *c = s->c;
    const double hthr = (c->h_min > SMALL) ? c->h_min : SMALL;
    if (c->friction != ORC_FRIC_NONE) {
        /* sourceFrictionStrickler_E -- swSpE_friction.cpp:19-76 */
        for (int64_t i = 0; i < c->n_cell; i++) {
            double h = s->h[i];
            if (h > hthr) {
                double hs = h + SMALL;
                double u = s->qx[i] / hs, v = s->qy[i] / hs;
                double ca = s->fcos[i], sa = s->fsin[i];
                double ul = u * ca + v * sa, vl = -u * sa + v * ca;
                double un = sqrt(u * u + v * v + 1e-16);
                double K1 = (s->K1[i] > 1e-10) ? s->K1[i] : 20.0;
                double K2 = (s->K2[i] > 1e-10) ? s->K2[i] : 20.0;
                double h43 = h * cbrt(h + 1e-16);
                double d1 = 1.0 + (dt * G * un) / (K1 * K1 * h43 + 1e-10);
                double d2 = 1.0 + (dt * G * un) / (K2 * K2 * h43 + 1e-10);
                ul /= d1;
                vl /= d2;
                s->qx[i] = (ul * ca - vl * sa) * h;
                s->qy[i] = (ul * sa + vl * ca) * h;
            } else {
                s->qx[i] = 0.0;
                s->qy[i] = 0.0;
            }
        }
    }
    if (c->head_loss) {
        /* sourceSingularHeadLoss_E -- swSpE_friction.cpp:79-126 */
        for (int64_t i = 0; i < c->n_cell; i++) {
            double ca = s->hcos[i], sa = s->hsin[i];
            double qx = s->qx[i] * ca + s->qy[i] * sa;
            double qy = -s->qx[i] * sa + s->qy[i] * ca;
            double h = s->h[i];
            if (h > hthr) {
                double ax = 0.0, ay = 0.0;
                if (h > c->hl_hsing[i]) {
                    double hs = h + SMALL;
                    double qn = sqrt(qx * qx + qy * qy + 1e-16);
                    ax = qn * c->hl_c1[i] / hs;
                    ay = qn * c->hl_c2[i] / hs;
                }
                qx = qx * exp(-ax * dt);
                qy = qy * exp(-ay * dt);
            } else {
                qx = 0.0;
                qy = 0.0;
            }
            s->qx[i] = qx * ca - qy * sa;
            s->qy[i] = qx * sa + qy * ca;
        }
    }
}

/* one _E step with a given dt and per-series BC values (ckpt_segment_E body) */
static void e_step(esolver *s, double dt, const double *bc_step) {
    const orc_case *c = s->c;
    lr(s);
    fluxes(s, bc_step);
    balance(s, dt);
    for (int64_t n = 0; n < c->n_cell; ++n) {
        if (s->h[n] < 0.) {
            s->h[n] = 0.;
            s->qx[n] = 0.;
            s->qy[n] = 0.;
        }
    }
    sources(s, dt);
}

static void e_init(esolver *s, const orc_case *c, const double *phi, const double *K1,
                   const double *K2, const double *h0, const double *qx0, const double *qy0) {
    int64_t nc = c->n_cell, nb = c->n_bnd ? c->n_bnd : 1, ni = c->n_int ? c->n_int : 1;
    memset(s, 0, sizeof(*s));
    s->c = c;
    s->phi = phi;
    s->K1 = K1;
    s->K2 = K2;
#define A(p, n) (p) = (double *)calloc((size_t)(n), sizeof(double))
    A(s->h, nc); A(s->qx, nc); A(s->qy, nc); A(s->swept, nc);
    A(s->fcos, nc); A(s->fsin, nc); A(s->hcos, nc); A(s->hsin, nc);
    A(s->fac_b, nb); A(s->fac_L, ni); A(s->fac_R, ni);
    A(s->zrb, nb); A(s->qrb, nb); A(s->rrb, nb); A(s->lpb, nb); A(s->lmaxb, nb);
    A(s->zri, 2 * ni); A(s->qri, 2 * ni); A(s->rri, 2 * ni); A(s->lmi, 2 * ni); A(s->lpi, 2 * ni);
    A(s->lmaxi, ni);
    A(s->F1b, nb); A(s->F2b, nb); A(s->F3b, nb);
    A(s->F1i, ni); A(s->F2i, ni); A(s->F3i, ni); A(s->sLi, ni); A(s->sRi, ni);
#undef A
    memcpy(s->h, h0, sizeof(double) * (size_t)nc);
    memcpy(s->qx, qx0, sizeof(double) * (size_t)nc);
    memcpy(s->qy, qy0, sizeof(double) * (size_t)nc);
    for (int64_t i = 0; i < nc; i++) {
        double a = (c->friction != ORC_FRIC_NONE && c->fric_alpha_deg) ? c->fric_alpha_deg[i] * atan(1.0) / 45.0 : 0.0;
        s->fcos[i] = cos(a);
        s->fsin[i] = sin(a);
        double b = (c->head_loss && c->hl_alpha_deg) ? c->hl_alpha_deg[i] * atan(1.0) / 45.0 : 0.0;
        s->hcos[i] = cos(b);
        s->hsin[i] = sin(b);
    }
    factors(s);
}

static void e_free(esolver *s) {
    double *p[] = {s->h, s->qx, s->qy, s->swept, s->fcos, s->fsin, s->hcos, s->hsin, s->fac_b,
                   s->fac_L, s->fac_R, s->zrb, s->qrb, s->rrb, s->lpb, s->lmaxb, s->zri, s->qri,
                   s->rri, s->lmi, s->lpi, s->lmaxi, s->F1b, s->F2b, s->F3b, s->F1i, s->F2i,
                   s->F3i, s->sLi, s->sRi};
    for (size_t i = 0; i < sizeof(p) / sizeof(p[0]); i++) free(p[i]);
}

/* Dry run (build_checkpoint_records, swSpE_checkpoint.cpp:210-401): steps from
 * t0 until t_max, dt = min(CFL_E, next observation - t, dt_max), clamped to
 * t_max and floored at 1e-6; records dt, the per-series BC values at the step
 * start and the observation index triggered by the step (-1 if none).
 * Returns the number of steps (<= max_steps). */
int64_t orc_e_record(const orc_case *c, const double *phi, const double *K1, const double *K2,
                     const double *h0, const double *qx0, const double *qy0, int32_t n_obs,
                     const double *obs_t, int64_t max_steps, double *dt_seq, double *bc_seq,
                     int32_t *obs_trig) {
    esolver s;
    e_init(&s, c, phi, K1, K2, h0, qx0, qy0);
    double t = c->t0;
    int32_t oi = 0;
    int64_t n = 0;
    double *bc = (double *)calloc((size_t)(c->n_series ? c->n_series : 1), sizeof(double));
    while (t < c->t_max && n < max_steps) {
        for (int k = 0; k < c->n_series; k++)
            bc[k] = orc_e_interp(c->series_t, c->series_v + (int64_t)k * c->n_times, c->n_times, t);
        double target = c->dt_max;
        if (oi < n_obs) {
            double to = obs_t[oi] - t;
            if (to - 10e-6 > 0.0 && to < target) target = to;
        }
        lr(&s);
        double dtc = cfl_dt(&s);
        if ((t + dtc) > c->t_max) dtc = c->t_max - t;
        if (dtc < 1e-6) dtc = 1e-6;
        double dt = (dtc < target) ? dtc : target;
        if ((t + dt) > c->t_max) dt = c->t_max - t;
        if (dt < 1e-6) dt = 1e-6;
        fluxes(&s, bc);
        balance(&s, dt);
        for (int64_t i = 0; i < c->n_cell; ++i)
            if (s.h[i] < 0.) { s.h[i] = 0.; s.qx[i] = 0.; s.qy[i] = 0.; }
        sources(&s, dt);
        t += dt;
        while (oi < n_obs && t > obs_t[oi] + 1e-6) oi++;
        int32_t trig = -1;
        if (oi < n_obs && fabs(t - obs_t[oi]) < 1e-6) trig = oi++;
        dt_seq[n] = dt;
        for (int k = 0; k < c->n_series; k++) bc_seq[n * c->n_series + k] = bc[k];
        obs_trig[n] = trig;
        n++;
    }
    free(bc);
    e_free(&s);
    return n;
}

/* Frozen replay (ckpt_segment_E over the whole trajectory) and cost.
 * Gauge cost when n_gauges > 0: sum_g w_g (h[cell_g] - obs[k*n_gauges + g])^2;
 * otherwise no cost.  Divided by n_obs when normalize != 0.  The final state
 * is returned in h/qx/qy when they are not NULL. */
double orc_e_replay(const orc_case *c, const double *phi, const double *K1, const double *K2,
                    const double *h0, const double *qx0, const double *qy0, int64_t n_steps,
                    const double *dt_seq, const double *bc_seq, const int32_t *obs_trig,
                    int32_t n_obs, int32_t n_gauges, const int64_t *gauge_cell,
                    const double *gauge_w, const double *obs, int32_t normalize, double *h_out,
                    double *qx_out, double *qy_out) {
    esolver s;
    e_init(&s, c, phi, K1, K2, h0, qx0, qy0);
    double J = 0.0;
    for (int64_t n = 0; n < n_steps; n++) {
        e_step(&s, dt_seq[n], bc_seq + n * c->n_series);
        int32_t oi = obs_trig[n];
        if (oi >= 0 && oi < n_obs && n_gauges > 0) {
            double Jo = 0.0;
            for (int g = 0; g < n_gauges; g++) {
                double d = s.h[gauge_cell[g]] - obs[(int64_t)oi * n_gauges + g];
                Jo += gauge_w[g] * d * d;
            }
            if (normalize && n_obs > 0) Jo /= (double)n_obs;
            J += Jo;
        }
    }
    size_t b = sizeof(double) * (size_t)c->n_cell;
    if (h_out) memcpy(h_out, s.h, b);
    if (qx_out) memcpy(qx_out, s.qx, b);
    if (qy_out) memcpy(qy_out, s.qy, b);
    e_free(&s);
    return J;
}
