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
 *                            ckpt_segment_E (584-715, the frozen replay)
 *   assim/swSpE_muscl.cpp    hypMusclEvrComputeLRStates_E (22-187) with the
 *                            slopes of computeMusclSlopes
 *                            (swSpHypMusclEvr.cpp:11-115)
 * Godunov and MUSCL-EVR reconstructions; the reverse sweep below covers both
 * (the slope limiter included).  Costs: gauge misfit (swSpE_gauges.cpp:166-179)
 * and full-frame misfit (swSpE_ad.cpp:273-308), optional normalisation by the
 * number of observations.
 */
#include "orc_omp.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sp_oracle.h"
#include "orc_math.h"

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
    /* MUSCL-EVR (_E): free surface, limited slopes, neighbour lists and the
     * dt its reconstruction reads (the dry run's previous step, the replay's
     * current step) */
    double *zc, *dzdx, *dzdy;
    int64_t *nb_ptr, *nb_idx;
    double dt_recon;
    orc_inc inc; /* per-cell incidences (the OpenMP build's gathers) */
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
    ORC_PAR
    for (int64_t i = 0; i < c->n_bnd; i++) {
        int64_t k = c->bnd_cell[i];
        double ph = (s->phi[k] > c->phi_min) ? s->phi[k] : c->phi_min;
        s->fac_b[i] = (s->phi[k] > c->phi_min) ? (c->bnd_len[i] / c->cell_area[k] / ph) : 0.0;
    }
    ORC_PAR
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
    ORC_PAR
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
    ORC_PAR
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

/* Kurganov-Petrova desingularised velocity -- assim/swSpDesingVelocity.h:28-39 */
#define H_DESING 5e-3
static double desing(double q, double h, double he) {
    double h2 = h * h;
    double h4 = h2 * h2;
    double e4 = he * he * he * he;
    double den = sqrt(h4 + ((h4 > e4) ? h4 : e4));
    return 1.4142135623730951 * h * q / den;
}

/* computeMusclSlopes on the free surface -- swSpHypMusclEvr.cpp:11-115 */
static void e_slopes(esolver *s) {
    const orc_case *c = s->c;
    ORC_PAR
    for (int64_t i = 0; i < c->n_cell; i++) {
        s->dzdx[i] = 0.0;
        s->dzdy[i] = 0.0;
        double wsum = 0.0;
        if (s->nb_ptr[i + 1] - s->nb_ptr[i] > 1) {
            for (int64_t k = s->nb_ptr[i]; k < s->nb_ptr[i + 1]; k++) {
                int64_t j = s->nb_idx[k];
                double dx = c->cell_x[j] - c->cell_x[i];
                double dy = c->cell_y[j] - c->cell_y[i];
                double dist = sqrt(dx * dx + dy * dy);
                double weight = 1.0 / dist;
                s->dzdx[i] += (s->zc[j] - s->zc[i]) * dx / dist / dist * weight;
                s->dzdy[i] += (s->zc[j] - s->zc[i]) * dy / dist / dist * weight;
                wsum += weight;
            }
            s->dzdx[i] /= wsum;
            s->dzdy[i] /= wsum;
        }
    }
#ifdef ORC_OMP
    /* the limiter of cell k only changes its own slope, in its edges' order */
    ORC_PAR
    for (int64_t k = 0; k < c->n_cell; k++) {
        for (int64_t p = s->inc.iptr[k]; p < s->inc.iptr[k + 1]; p++) {
            const int64_t e = s->inc.iinc[p] >> 1;
            int64_t L = c->int_L[e], R = c->int_R[e];
            double fmx = (s->zc[L] < s->zc[R]) ? s->zc[R] : s->zc[L];
            double fmn = (s->zc[R] < s->zc[L]) ? s->zc[R] : s->zc[L];
            double fr = s->zc[k] + (c->int_xm[e] - c->cell_x[k]) * s->dzdx[k] +
                        (c->int_ym[e] - c->cell_y[k]) * s->dzdy[k];
            if (fr < fmn) {
                double sf = (fmn - s->zc[k]) / (fr - s->zc[k]);
                s->dzdx[k] *= sf;
                s->dzdy[k] *= sf;
            }
            if (fr > fmx) {
                double sf = (fmx - s->zc[k]) / (fr - s->zc[k]);
                s->dzdx[k] *= sf;
                s->dzdy[k] *= sf;
            }
        }
    }
    return;
#endif
    for (int64_t e = 0; e < c->n_int; e++) {
        int64_t L = c->int_L[e], R = c->int_R[e];
        double fmx = (s->zc[L] < s->zc[R]) ? s->zc[R] : s->zc[L];
        double fmn = (s->zc[R] < s->zc[L]) ? s->zc[R] : s->zc[L];
        for (int side = 0; side < 2; side++) {
            int64_t k = side ? R : L;
            double fr = s->zc[k] + (c->int_xm[e] - c->cell_x[k]) * s->dzdx[k] +
                        (c->int_ym[e] - c->cell_y[k]) * s->dzdy[k];
            if (fr < fmn) {
                double sf = (fmn - s->zc[k]) / (fr - s->zc[k]);
                s->dzdx[k] *= sf;
                s->dzdy[k] *= sf;
            }
            if (fr > fmx) {
                double sf = (fmx - s->zc[k]) / (fr - s->zc[k]);
                s->dzdx[k] *= sf;
                s->dzdy[k] *= sf;
            }
        }
    }
}

/* hypMusclEvrComputeLRStates_E -- assim/swSpE_muscl.cpp:22-187 */
static void lr_muscl(esolver *s) {
    const orc_case *c = s->c;
    const double dt = s->dt_recon;
    ORC_PAR
    for (int64_t i = 0; i < c->n_cell; i++) s->zc[i] = s->h[i] + c->cell_zb[i];
    e_slopes(s);
    ORC_PAR
    for (int64_t e = 0; e < c->n_bnd; e++) {
        double nx = c->bnd_nx[e], ny = c->bnd_ny[e];
        int64_t id = c->bnd_cell[e];
        double b = c->beta[id], q = s->qx[id], r = s->qy[id], hR = s->h[id];
        s->zrb[e] = s->h[id] + c->cell_zb[id];
        s->qrb[e] = q * nx + r * ny;
        s->rrb[e] = -q * ny + r * nx;
        double uR = desing(s->qrb[e], hR, H_DESING);
        double cR = sqrt(G * hR);
        double lm = lam_m(b, cR, uR), lp = lam_p(b, cR, uR);
        s->lpb[e] = lp;
        s->lmaxb[e] = (fabs(lm) < fabs(lp)) ? fabs(lp) : fabs(lm);
    }
    ORC_PAR
    for (int64_t e = 0; e < c->n_int; e++) {
        double nx = c->int_nx[e], ny = c->int_ny[e];
        for (int sd = 0; sd < 2; sd++) {
            int64_t id = sd ? c->int_R[e] : c->int_L[e];
            double b = c->beta[id], zb = c->cell_zb[id];
            double h_local = s->h[id];
            double q_local = s->qx[id] * nx + s->qy[id] * ny;
            double r_local = s->qy[id] * nx - s->qx[id] * ny;
            double c_local = sqrt(G * ((SMALL < h_local) ? h_local : SMALL));
            double u_local = desing(q_local, h_local, H_DESING);
            double v_local = desing(r_local, h_local, H_DESING);
            double l1 = lam_m(b, c_local, u_local), l2 = lam_p(b, c_local, u_local);
            if (sd == 0) {
                l1 = (0.0 < l1) ? l1 : 0.0;
                l2 = (0.0 < l2) ? l2 : 0.0;
            } else {
                l1 = (l1 < 0.0) ? l1 : 0.0;
                l2 = (l2 < 0.0) ? l2 : 0.0;
            }
            double dx1 = c->int_xm[e] - c->cell_x[id] - 0.5 * l1 * dt * nx;
            double dy1 = c->int_ym[e] - c->cell_y[id] - 0.5 * l1 * dt * ny;
            double z1 = s->zc[id] + dx1 * s->dzdx[id] + dy1 * s->dzdy[id];
            double h1 = (0.0 < z1 - zb) ? (z1 - zb) : 0.0;
            double dx2 = c->int_xm[e] - c->cell_x[id] - 0.5 * l2 * dt * nx;
            double dy2 = c->int_ym[e] - c->cell_y[id] - 0.5 * l2 * dt * ny;
            double z2 = s->zc[id] + dx2 * s->dzdx[id] + dy2 * s->dzdy[id];
            double h2 = (0.0 < z2 - zb) ? (z2 - zb) : 0.0;
            double hr = 0.5 * (h1 + h2);
            s->zri[2 * e + sd] = 0.5 * (h1 + h2) + zb;
            s->qri[2 * e + sd] = 0.5 * (h1 + h2) * u_local;
            s->rri[2 * e + sd] = 0.5 * (h1 + h2) * v_local;
            double uu = desing(s->qri[2 * e + sd], hr, H_DESING);
            double cc = sqrt(G * hr);
            s->lmi[2 * e + sd] = lam_m(b, cc, uu);
            s->lpi[2 * e + sd] = lam_p(b, cc, uu);
        }
        double m1 = fabs(s->lmi[2 * e]), v = fabs(s->lpi[2 * e]);
        if (m1 < v) m1 = v;
        double m2 = fabs(s->lmi[2 * e + 1]);
        v = fabs(s->lpi[2 * e + 1]);
        if (m2 < v) m2 = v;
        s->lmaxi[e] = (m1 < m2) ? m2 : m1;
    }
}

/* hypComputeLRStates_E -- swSpE_hyp.cpp:487-498: Godunov or MUSCL-EVR */
static void lr_any(esolver *s) {
    if (s->c->scheme == ORC_MUSCL_EVR)
        lr_muscl(s);
    else
        lr(s);
}

/* hypComputeTimeStep_E -- swSpE_hyp.cpp:578-635 (without the t_max clamp and
 * floor, applied by the caller as in build_checkpoint_records) */
static double cfl_dt(esolver *s) {
    const orc_case *c = s->c;
    double dt = c->dt_max;
#ifdef ORC_OMP
    /* per cell: its boundary edges, then its internal edges, each ascending */
#pragma omp parallel
    {
        double dtl = dt;
#pragma omp for schedule(static)
        for (int64_t i = 0; i < c->n_cell; i++) {
            double sw = 0.0;
            if (s->h[i] > c->h_min) {
                for (int64_t p = s->inc.bptr[i]; p < s->inc.bptr[i + 1]; p++) {
                    const int64_t e = s->inc.binc[p];
                    sw += s->lmaxb[e] * c->bnd_len[e];
                }
                for (int64_t p = s->inc.iptr[i]; p < s->inc.iptr[i + 1]; p++) {
                    const int64_t e = s->inc.iinc[p] >> 1;
                    sw += s->lmaxi[e] * c->int_len[e];
                }
            }
            s->swept[i] = sw;
            double den = (sw < SMALL) ? SMALL : sw;
            double x = c->cfl_max * (c->cell_area[i] / den);
            if (x < dtl) dtl = x;
        }
#pragma omp critical
        if (dtl < dt) dt = dtl;
    }
    return dt;
#endif
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
    ORC_PAR
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
    ORC_PAR
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
#ifdef ORC_OMP
/* the balance of the cells of U (h | qx | qy, the cell's own additions in
 * the serial order), without the clean-up */
static void balance_gather(const esolver *s, double dt, double *h, double *qx, double *qy) {
    const orc_case *c = s->c;
    ORC_PAR
    for (int64_t i = 0; i < c->n_cell; i++) {
        for (int64_t p = s->inc.bptr[i]; p < s->inc.bptr[i + 1]; p++) {
            const int64_t e = s->inc.binc[p];
            double co = s->fac_b[e], nx = c->bnd_nx[e], ny = c->bnd_ny[e];
            h[i] += (+co) * s->F1b[e] * dt;
            qx[i] += ((+co * nx) * s->F2b[e] + (+co * nx) * 0.0 + (-co * ny) * s->F3b[e]) * dt;
            qy[i] += ((+co * ny) * s->F2b[e] + (+co * ny) * 0.0 + (+co * nx) * s->F3b[e]) * dt;
        }
        for (int64_t p = s->inc.iptr[i]; p < s->inc.iptr[i + 1]; p++) {
            const int64_t e = s->inc.iinc[p] >> 1;
            double cl = s->fac_L[e], cr = s->fac_R[e], nx = c->int_nx[e], ny = c->int_ny[e];
            double dh = s->F1i[e] * dt;
            if (s->inc.iinc[p] & 1) {
                h[i] += (+cr) * dh;
                qx[i] += ((+cr * nx) * s->F2i[e] + (+cr * nx) * s->sRi[e] + (-cr * ny) * s->F3i[e]) * dt;
                qy[i] += ((+cr * ny) * s->F2i[e] + (+cr * ny) * s->sRi[e] + (+cr * nx) * s->F3i[e]) * dt;
            } else {
                h[i] += (-cl) * dh;
                qx[i] += ((-cl * nx) * s->F2i[e] + (+cl * nx) * s->sLi[e] + (+cl * ny) * s->F3i[e]) * dt;
                qy[i] += ((-cl * ny) * s->F2i[e] + (+cl * ny) * s->sLi[e] + (-cl * nx) * s->F3i[e]) * dt;
            }
        }
    }
}
#endif

static void balance(esolver *s, double dt) {
    const orc_case *c = s->c;
    double *h = s->h, *qx = s->qx, *qy = s->qy;
#ifdef ORC_OMP
    balance_gather(s, dt, h, qx, qy);
    ORC_PAR
    for (int64_t n = 0; n < c->n_cell; ++n) {
        if (fabs(h[n]) < SMALL) h[n] = 0.0;
        if (fabs(qx[n]) < SMALL) qx[n] = 0.0;
        if (fabs(qy[n]) < SMALL) qy[n] = 0.0;
    }
    return;
#endif
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
        ORC_PAR
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
        ORC_PAR
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
    s->dt_recon = dt; /* ckpt_segment_E sets solver->dt first (swSpE_checkpoint.cpp:658-659) */
    lr_any(s);
    fluxes(s, bc_step);
    balance(s, dt);
    ORC_PAR
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
    A(s->zc, nc); A(s->dzdx, nc); A(s->dzdy, nc);
#undef A
    /* neighbour lists in ascending internal-edge order -- swParent.cpp:476-486 */
    s->nb_ptr = (int64_t *)calloc((size_t)nc + 1, sizeof(int64_t));
    for (int64_t e = 0; e < c->n_int; e++) {
        s->nb_ptr[c->int_L[e] + 1]++;
        s->nb_ptr[c->int_R[e] + 1]++;
    }
    for (int64_t i = 0; i < nc; i++) s->nb_ptr[i + 1] += s->nb_ptr[i];
    s->nb_idx = (int64_t *)calloc((size_t)(2 * ni + 1), sizeof(int64_t));
    {
        int64_t *fill = (int64_t *)calloc((size_t)nc, sizeof(int64_t));
        for (int64_t e = 0; e < c->n_int; e++) {
            int64_t L = c->int_L[e], R = c->int_R[e];
            s->nb_idx[s->nb_ptr[L] + fill[L]++] = R;
            s->nb_idx[s->nb_ptr[R] + fill[R]++] = L;
        }
        free(fill);
    }
#ifdef ORC_OMP
    orc_inc_build(&s->inc, nc, c->n_int, c->n_bnd, c->int_L, c->int_R, c->bnd_cell);
#endif
    s->dt_recon = 0.0; /* the solver's dt member before any step (swSp.h:155) */
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
                   s->F3i, s->sLi, s->sRi, s->zc, s->dzdx, s->dzdy};
    for (size_t i = 0; i < sizeof(p) / sizeof(p[0]); i++) free(p[i]);
    free(s->nb_ptr);
    free(s->nb_idx);
#ifdef ORC_OMP
    orc_inc_free(&s->inc);
#endif
}

/* Dry run (build_checkpoint_records, swSpE_checkpoint.cpp:210-401): steps from
 * t0 until t_max, dt = min(CFL_E, next observation - t, dt_max), clamped to
 * t_max and floored at 1e-6; records dt, the per-series BC values at the step
 * start and the observation index triggered by the step (-1 if none).
 * Returns the number of steps (<= max_steps). */
int64_t orc_e_record_states(const orc_case *c, const double *phi, const double *K1,
                            const double *K2, const double *h0, const double *qx0,
                            const double *qy0, int32_t n_obs, const double *obs_t,
                            int64_t max_steps, double *dt_seq, double *bc_seq, int32_t *obs_trig,
                            double *states);

int64_t orc_e_record(const orc_case *c, const double *phi, const double *K1, const double *K2,
                     const double *h0, const double *qx0, const double *qy0, int32_t n_obs,
                     const double *obs_t, int64_t max_steps, double *dt_seq, double *bc_seq,
                     int32_t *obs_trig) {
    return orc_e_record_states(c, phi, K1, K2, h0, qx0, qy0, n_obs, obs_t, max_steps, dt_seq,
                               bc_seq, obs_trig, NULL);
}

/* As orc_e_record; states (optional, (max_steps + 1) x 3 n_cell): the dry
 * run's state [h | qx | qy] after every step, the initial state first. */
int64_t orc_e_record_states(const orc_case *c, const double *phi, const double *K1,
                            const double *K2, const double *h0, const double *qx0,
                            const double *qy0, int32_t n_obs, const double *obs_t,
                            int64_t max_steps, double *dt_seq, double *bc_seq, int32_t *obs_trig,
                            double *states) {
    esolver s;
    e_init(&s, c, phi, K1, K2, h0, qx0, qy0);
    const int64_t ncs = c->n_cell;
    if (states) {
        memcpy(states, h0, sizeof(double) * ncs);
        memcpy(states + ncs, qx0, sizeof(double) * ncs);
        memcpy(states + 2 * ncs, qy0, sizeof(double) * ncs);
    }
    double t = c->t0;
    int32_t oi = 0;
    int64_t n = 0;
    double *bc = (double *)calloc((size_t)(c->n_series ? c->n_series : 1), sizeof(double));
    while (t < c->t_max && n < max_steps) {
        ORC_AUDIT_STEP(n);
        for (int k = 0; k < c->n_series; k++)
            bc[k] = orc_e_interp(c->series_t, c->series_v + (int64_t)k * c->n_times, c->n_times, t);
        double target = c->dt_max;
        if (oi < n_obs) {
            double to = obs_t[oi] - t;
            if (to - 10e-6 > 0.0 && to < target) target = to;
        }
        /* the reconstruction reads the solver's dt, still the previous step's
         * here: build_checkpoint_records sets it after the CFL step
         * (swSpE_checkpoint.cpp:244-251) */
        lr_any(&s);
        double dtc = cfl_dt(&s);
        if ((t + dtc) > c->t_max) dtc = c->t_max - t;
        if (dtc < 1e-6) dtc = 1e-6;
        double dt = (dtc < target) ? dtc : target;
        if ((t + dt) > c->t_max) dt = c->t_max - t;
        if (dt < 1e-6) dt = 1e-6;
        s.dt_recon = dt;
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
        if (states) {
            double *St = states + 3 * ncs * n;
            memcpy(St, s.h, sizeof(double) * ncs);
            memcpy(St + ncs, s.qx, sizeof(double) * ncs);
            memcpy(St + 2 * ncs, s.qy, sizeof(double) * ncs);
        }
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
        ORC_AUDIT_STEP(n);
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

/* The dry run's states on a given schedule: as orc_e_replay, but the
 * reconstruction reads the previous step's dt (0 before the first), as
 * build_checkpoint_records does (swSpE_checkpoint.cpp:244-251); on the dry
 * run's own schedule these are its states.  states: (n_steps + 1) x 3
 * n_cell [h | qx | qy], the initial state first. */
void orc_e_dry_states(const orc_case *c, const double *phi, const double *K1, const double *K2,
                      const double *h0, const double *qx0, const double *qy0, int64_t n_steps,
                      const double *dt_seq, const double *bc_seq, double *states) {
    esolver s;
    e_init(&s, c, phi, K1, K2, h0, qx0, qy0);
    const int64_t nc = c->n_cell;
    const size_t b = sizeof(double) * (size_t)nc;
    memcpy(states, h0, b);
    memcpy(states + nc, qx0, b);
    memcpy(states + 2 * nc, qy0, b);
    double prev = 0.0;
    for (int64_t n = 0; n < n_steps; n++) {
        ORC_AUDIT_STEP(n);
        const double dt = dt_seq[n];
        s.dt_recon = prev;
        lr_any(&s);
        s.dt_recon = dt;
        fluxes(&s, bc_seq + n * c->n_series);
        balance(&s, dt);
        for (int64_t i = 0; i < nc; ++i)
            if (s.h[i] < 0.) { s.h[i] = 0.; s.qx[i] = 0.; s.qy[i] = 0.; }
        sources(&s, dt);
        prev = dt;
        double *St = states + 3 * nc * (n + 1);
        memcpy(St, s.h, b);
        memcpy(St + nc, s.qx, b);
        memcpy(St + 2 * nc, s.qy, b);
    }
    e_free(&s);
}

/* ------------------------------------------------------------------------
 * Reverse sweep of the frozen replay: the derivative that the reference
 * obtains from Enzyme (get_spatial_gradient_checkpointed,
 * swSpE_checkpoint.cpp:722-1281, reversing ckpt_segment_E :584-715), written
 * out by hand for the functions above, in the same loop order.  All states
 * of the trajectory are kept (no checkpointing: test sizes).
 * Conventions at non-differentiable points: the derivative of sqrt at 0 is
 * taken as 0; min/max/clamps pass the derivative to the branch the forward
 * took; the round-off clean-up and the negative-depth clamp zero it.
 * ------------------------------------------------------------------------ */

static double dsq(double x) { return (x > 0.0) ? 0.5 / sqrt(x) : 0.0; } /* d sqrt(x)/dx */

/* lam_m / lam_p adjoint: returns d lam / du and d lam / dc */
static void lam_grad(double b, double c, double u, int plus, double *du, double *dc) {
    double a = (b - 1.0) * b * u * u + c * c;
    double ds = dsq(a);
    double sg = plus ? 1.0 : -1.0;
    *du = b + sg * ds * 2.0 * (b - 1.0) * b * u;
    *dc = sg * ds * 2.0 * c;
}

/* adjoint of one side of hypGodunovComputeLRStates_E: (z, q, r, lm, lp) of a
 * cell state rotated to the edge normal -> accumulates into (ah, aqx, aqy) */
static void lr_adj(double h, double qx, double qy, double b, double nx, double ny, double hmin,
                   double az, double aq, double ar, double alm, double alp, double *ah,
                   double *aqx, double *aqy) {
    double q = qx * nx + qy * ny;
    double u = (h > hmin) ? q / h : 0.0;
    double c = sqrt(G * h);
    double dum, dcm, dup, dcp;
    lam_grad(b, c, u, 0, &dum, &dcm);
    lam_grad(b, c, u, 1, &dup, &dcp);
    double au = alm * dum + alp * dup;
    double ac = alm * dcm + alp * dcp;
    double ahh = az + ac * G * dsq(G * h);
    if (h > hmin) {
        aq += au / h;
        ahh += -au * q / (h * h);
    }
    *ah += ahh;
    *aqx += aq * nx - ar * ny;
    *aqy += aq * ny + ar * nx;
}

/* d desing(q, h) / dq and / dh (swSpDesingVelocity.h:28-39): the max picks
 * its branch from the primal values */
static void desing_grad(double q, double h, double he, double *dq, double *dh) {
    double h2 = h * h, h4 = h2 * h2, e4 = he * he * he * he;
    int big = h4 > e4;
    double D2 = h4 + (big ? h4 : e4), D = sqrt(D2);
    double dD2 = (big ? 8.0 : 4.0) * h2 * h; /* d(h4 + max(h4, e4))/dh */
    *dq = 1.4142135623730951 * h / D;
    *dh = 1.4142135623730951 * q / D - 0.5 * 1.4142135623730951 * h * q * dD2 / (D2 * D);
}

/* adjoint of a boundary side of hypMusclEvrComputeLRStates_E
 * (swSpE_muscl.cpp:48-64, desingularised velocity): (z, q, r, lp) ->
 * accumulates into (ah, aqx, aqy) */
static void lr_bnd_muscl_adj(double h, double qx, double qy, double b, double nx, double ny,
                             double az, double aq, double ar, double alp, double *ah, double *aqx,
                             double *aqy) {
    double q = qx * nx + qy * ny;
    double u = desing(q, h, H_DESING);
    double c = sqrt(G * h);
    double dup, dcp, dq, dh;
    lam_grad(b, c, u, 1, &dup, &dcp);
    desing_grad(q, h, H_DESING, &dq, &dh);
    double au = alp * dup;
    double ahh = az + alp * dcp * G * dsq(G * h) + au * dh;
    aq += au * dq;
    *ah += ahh;
    *aqx += aq * nx - ar * ny;
    *aqy += aq * ny + ar * nx;
}

/* adjoint of one side (sd 0 = L, 1 = R) of the internal-edge reconstruction
 * of hypMusclEvrComputeLRStates_E (swSpE_muscl.cpp:66-180): (z, q, r, lm,
 * lp) of the side -> accumulates into the cell's (ah, aqx, aqy) (its free
 * surface z = h + zb included) and its limited slope (asx, asy) */
static void recon_adj(int sd, double h, double qx, double qy, double zb, double b, double sx,
                      double sy, double nx, double ny, double dxm, double dym, double dt, double az,
                      double aq, double ar, double alm, double alp, double *ah, double *aqx,
                      double *aqy, double *asx, double *asy) {
    /* forward intermediates */
    double zc = h + zb;
    double ql = qx * nx + qy * ny, rl = qy * nx - qx * ny;
    double cl = sqrt(G * ((SMALL < h) ? h : SMALL));
    double ul = desing(ql, h, H_DESING), vl = desing(rl, h, H_DESING);
    double l1 = lam_m(b, cl, ul), l2 = lam_p(b, cl, ul);
    int p1 = sd ? (l1 < 0.0) : (0.0 < l1), p2 = sd ? (l2 < 0.0) : (0.0 < l2);
    if (!p1) l1 = 0.0;
    if (!p2) l2 = 0.0;
    double dx1 = dxm - 0.5 * l1 * dt * nx, dy1 = dym - 0.5 * l1 * dt * ny;
    double dx2 = dxm - 0.5 * l2 * dt * nx, dy2 = dym - 0.5 * l2 * dt * ny;
    double z1 = zc + dx1 * sx + dy1 * sy, z2 = zc + dx2 * sx + dy2 * sy;
    int w1 = 0.0 < z1 - zb, w2 = 0.0 < z2 - zb;
    double hr = 0.5 * ((w1 ? z1 - zb : 0.0) + (w2 ? z2 - zb : 0.0));
    double qo = hr * ul;
    double cc = sqrt(G * hr);
    double uu = desing(qo, hr, H_DESING);
    /* reverse: eigenvalues of the reconstructed state */
    double dum, dcm, dup, dcp, dq, dh;
    lam_grad(b, cc, uu, 0, &dum, &dcm);
    lam_grad(b, cc, uu, 1, &dup, &dcp);
    double auu = alm * dum + alp * dup, acc = alm * dcm + alp * dcp;
    desing_grad(qo, hr, H_DESING, &dq, &dh);
    double aqo = aq + auu * dq;
    double ahr = az + acc * G * dsq(G * hr) + auu * dh + aqo * ul + ar * vl;
    double aul = aqo * hr, avl = ar * hr;
    /* the two characteristic-offset free surfaces */
    double az1 = w1 ? 0.5 * ahr : 0.0, az2 = w2 ? 0.5 * ahr : 0.0;
    *asx += az1 * dx1 + az2 * dx2;
    *asy += az1 * dy1 + az2 * dy2;
    double al1 = p1 ? -0.5 * dt * (nx * sx + ny * sy) * az1 : 0.0;
    double al2 = p2 ? -0.5 * dt * (nx * sx + ny * sy) * az2 : 0.0;
    /* eigenvalues of the cell state */
    lam_grad(b, cl, ul, 0, &dum, &dcm);
    lam_grad(b, cl, ul, 1, &dup, &dcp);
    aul += al1 * dum + al2 * dup;
    double acl = al1 * dcm + al2 * dcp;
    double ahh = az1 + az2 + ((SMALL < h) ? acl * G * dsq(G * h) : 0.0);
    desing_grad(ql, h, H_DESING, &dq, &dh);
    double aql = aul * dq;
    ahh += aul * dh;
    desing_grad(rl, h, H_DESING, &dq, &dh);
    double arl = avl * dq;
    ahh += avl * dh;
    *ah += ahh;
    *aqx += aql * nx - arl * ny;
    *aqy += aql * ny + arl * nx;
}

/* adjoint of e_slopes (computeMusclSlopes, swSpHypMusclEvr.cpp:11-115):
 * (asx, asy) = adjoint of the limited slopes -> adds the adjoint of the free
 * surface zc to azc.  The limiter visits are replayed forward (keeping the
 * slope each visit saw), then reversed in the opposite edge / side order. */
static void e_slopes_adj(esolver *s, double *asx, double *asy, double *azc) {
    const orc_case *c = s->c;
    const int64_t nc = c->n_cell, ni = c->n_int;
    double *sx = (double *)calloc((size_t)nc, sizeof(double));
    double *sy = (double *)calloc((size_t)nc, sizeof(double));
    double *ws = (double *)calloc((size_t)nc, sizeof(double));
    double *tx = (double *)calloc((size_t)(2 * ni + 1), sizeof(double));
    double *ty = (double *)calloc((size_t)(2 * ni + 1), sizeof(double));
    const double *z = s->zc;
    for (int64_t i = 0; i < nc; i++) {
        if (s->nb_ptr[i + 1] - s->nb_ptr[i] <= 1) continue;
        for (int64_t k = s->nb_ptr[i]; k < s->nb_ptr[i + 1]; k++) {
            int64_t j = s->nb_idx[k];
            double dx = c->cell_x[j] - c->cell_x[i], dy = c->cell_y[j] - c->cell_y[i];
            double dist = sqrt(dx * dx + dy * dy), w = 1.0 / dist;
            sx[i] += (z[j] - z[i]) * dx / dist / dist * w;
            sy[i] += (z[j] - z[i]) * dy / dist / dist * w;
            ws[i] += w;
        }
        sx[i] /= ws[i];
        sy[i] /= ws[i];
    }
    for (int64_t e = 0; e < ni; e++) {
        int64_t L = c->int_L[e], R = c->int_R[e];
        double fmx = (z[L] < z[R]) ? z[R] : z[L], fmn = (z[R] < z[L]) ? z[R] : z[L];
        for (int sd = 0; sd < 2; sd++) {
            int64_t k = sd ? R : L;
            tx[2 * e + sd] = sx[k];
            ty[2 * e + sd] = sy[k];
            double fr = z[k] + (c->int_xm[e] - c->cell_x[k]) * sx[k] + (c->int_ym[e] - c->cell_y[k]) * sy[k];
            double f = (fr < fmn) ? fmn : (fr > fmx) ? fmx : fr;
            if (f != fr) {
                double sf = (f - z[k]) / (fr - z[k]);
                sx[k] *= sf;
                sy[k] *= sf;
            }
        }
    }
    for (int64_t e = ni - 1; e >= 0; e--) {
        int64_t L = c->int_L[e], R = c->int_R[e];
        for (int sd = 1; sd >= 0; sd--) {
            int64_t k = sd ? R : L;
            double ax = c->int_xm[e] - c->cell_x[k], ay = c->int_ym[e] - c->cell_y[k];
            double px = tx[2 * e + sd], py = ty[2 * e + sd];
            double fr = z[k] + ax * px + ay * py;
            int64_t fsrc; /* the cell whose free surface bounds the reconstruction */
            if (fr < ((z[R] < z[L]) ? z[R] : z[L]))
                fsrc = (z[R] < z[L]) ? R : L;
            else if (fr > ((z[L] < z[R]) ? z[R] : z[L]))
                fsrc = (z[L] < z[R]) ? R : L;
            else
                continue;
            double D = fr - z[k], sf = (z[fsrc] - z[k]) / D;
            double gs = asx[k] * px + asy[k] * py; /* adjoint of sf, times D / (f - z_k) */
            asx[k] = sf * asx[k] - gs * sf / D * ax;
            asy[k] = sf * asy[k] - gs * sf / D * ay;
            /* sf = (z_f - z_k) / D with D = a . slope independent of z_k */
            azc[fsrc] += gs / D;
            azc[k] -= gs / D;
        }
    }
    for (int64_t i = 0; i < nc; i++) {
        if (s->nb_ptr[i + 1] - s->nb_ptr[i] <= 1) continue;
        for (int64_t k = s->nb_ptr[i]; k < s->nb_ptr[i + 1]; k++) {
            int64_t j = s->nb_idx[k];
            double dx = c->cell_x[j] - c->cell_x[i], dy = c->cell_y[j] - c->cell_y[i];
            double dist = sqrt(dx * dx + dy * dy), w = 1.0 / dist;
            double a = (asx[i] * dx + asy[i] * dy) / dist / dist * w / ws[i];
            azc[j] += a;
            azc[i] -= a;
        }
    }
    free(sx);
    free(sy);
    free(ws);
    free(tx);
    free(ty);
}

/* adjoint of hypHllc_E: df[5] -> d[12] = d(zL qL rL zR qR rR lmL lmR lpL lpR phiL phiR) */
static void hllc_adj(double h_min, double phi_min, double bL, double bR, double phiL, double phiR,
                     double zbL, double zbR, double zL, double qL, double rL, double zR, double qR,
                     double rR, double lmL, double lmR, double lpL, double lpR, const double *df,
                     double *d) {
    double hL = (zL - zbL > 0.0) ? (zL - zbL) : 0.0;
    double hR = (zR - zbR > 0.0) ? (zR - zbR) : 0.0;
    int both = 0, lhs = 0, rhs = 0;
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
    double ahL = 0, ahR = 0;
    for (int k = 0; k < 12; k++) d[k] = 0.0;
    if (both) {
        double hLs = (hL > SMALL) ? hL : SMALL, hRs = (hR > SMALL) ? hR : SMALL;
        double uL = qL / hLs, uR = qR / hRs;
        double L1 = (lmL < lmR) ? lmL : lmR;
        int L1z = (L1 > 0.0);
        if (L1z) L1 = 0.0;
        double L2 = (lpL > lpR) ? lpL : lpR;
        int L2z = (L2 < 0.0);
        if (L2z) L2 = 0.0;
        double den = L2 - L1;
        int denf = (den < SMALL);
        if (denf) den = SMALL;
        double dd = phiL - phiR;
        double sq = sqrt(dd * dd + 1e-3 * 1e-3);
        double phiLR = 0.5 * (phiL + phiR - sq);
        double N1 = L2 * qL - L1 * qR + L1 * L2 * (zR - zL);
        double f1 = phiLR * N1 / den;
        double fact = 0.5 * phiLR * (hL + hR);
        double s2a = (phiR * hR * hR - phiL * hL * hL) * 0.5 + fact * (zL - zR);
        double s2 = s2a * G / den;
        double sl = -L1 * s2, sr = L2 * s2;
        double FLk = bL * hL * uL * uL + 0.5 * G * hL * hL, FRk = bR * hR * uR * uR + 0.5 * G * hR * hR;
        double FL = phiL * FLk, FR = phiR * FRk;
        double P = phiR * qR - phiL * qL;
        double num2 = L2 * (FL + sl) - L1 * (FR - sr) + L1 * L2 * P;
        double f2 = num2 / den;
        double vL = rL / hLs, vR = rR / hRs;
        double af1 = df[0], af2 = df[1], af3 = df[2], asl = df[3], asr = df[4];
        double ahLs = 0, ahRs = 0, auL = 0, auR = 0, aL1 = 0, aL2 = 0, aden = 0, aphiLR = 0;
        /* f3 */
        if (f1 >= 0.0) {
            af1 += af3 * bL * vL;
            double avL = af3 * bL * f1;
            d[2] += avL / hLs;
            ahLs += -avL * rL / (hLs * hLs);
        } else {
            af1 += af3 * bR * vR;
            double avR = af3 * bR * f1;
            d[5] += avR / hRs;
            ahRs += -avR * rR / (hRs * hRs);
        }
        /* f2 */
        double an2 = af2 / den;
        aden += -af2 * f2 / den;
        aL2 += an2 * (FL + sl) + an2 * L1 * P;
        double aFL = an2 * L2;
        asl += an2 * L2;
        aL1 += -an2 * (FR - sr) + an2 * L2 * P;
        double aFR = -an2 * L1;
        asr += an2 * L1;
        double aP = an2 * L1 * L2;
        d[11] += aP * qR;
        d[4] += aP * phiR;
        d[10] += -aP * qL;
        d[1] += -aP * phiL;
        /* FL, FR */
        d[10] += aFL * FLk;
        ahL += aFL * phiL * (bL * uL * uL + G * hL);
        auL += aFL * phiL * 2.0 * bL * hL * uL;
        d[11] += aFR * FRk;
        ahR += aFR * phiR * (bR * uR * uR + G * hR);
        auR += aFR * phiR * 2.0 * bR * hR * uR;
        /* sl, sr */
        aL1 += -asl * s2;
        double as2 = -asl * L1;
        aL2 += asr * s2;
        as2 += asr * L2;
        /* s2 */
        double as2a = as2 * G / den;
        aden += -as2 * s2 / den;
        d[11] += as2a * 0.5 * hR * hR;
        ahR += as2a * phiR * hR;
        d[10] += -as2a * 0.5 * hL * hL;
        ahL += -as2a * phiL * hL;
        double afact = as2a * (zL - zR);
        d[0] += as2a * fact;
        d[3] += -as2a * fact;
        aphiLR += afact * 0.5 * (hL + hR);
        ahL += afact * 0.5 * phiLR;
        ahR += afact * 0.5 * phiLR;
        /* f1 */
        aphiLR += af1 * N1 / den;
        double aN1 = af1 * phiLR / den;
        aden += -af1 * f1 / den;
        aL2 += aN1 * (qL + L1 * (zR - zL));
        d[1] += aN1 * L2;
        aL1 += aN1 * (-qR + L2 * (zR - zL));
        d[4] += -aN1 * L1;
        d[3] += aN1 * L1 * L2;
        d[0] += -aN1 * L1 * L2;
        /* den, phiLR, u, h */
        if (!denf) {
            aL2 += aden;
            aL1 += -aden;
        }
        d[10] += aphiLR * 0.5 * (1.0 - dd / sq);
        d[11] += aphiLR * 0.5 * (1.0 + dd / sq);
        d[1] += auL / hLs;
        ahLs += -auL * qL / (hLs * hLs);
        d[4] += auR / hRs;
        ahRs += -auR * qR / (hRs * hRs);
        if (hL > SMALL) ahL += ahLs;
        if (hR > SMALL) ahR += ahRs;
        if (!L2z) {
            if (lpL > lpR) d[8] += aL2;
            else d[9] += aL2;
        }
        if (!L1z) {
            if (lmL < lmR) d[6] += aL1;
            else d[7] += aL1;
        }
    } else if (rhs) {
        d[11] += df[4] * 0.5 * G * hR * hR;
        ahR += df[4] * phiR * G * hR;
    } else if (lhs) {
        d[10] += -df[3] * 0.5 * G * hL * hL;
        ahL += -df[3] * phiL * G * hL;
    }
    if (zL - zbL > 0.0) d[0] += ahL;
    if (zR - zbR > 0.0) d[3] += ahR;
}

/* adjoint of hypBoundaryFlux_E: df[3] -> d[6] = d(zR qR rR lpR phiR bc) */
static void bnd_adj(double h_min, double phi_min, double bR, double phiR, double zbR, double zR,
                    double qR, double rR, double lpR, int type, double bc, const double *df,
                    double *d) {
    for (int k = 0; k < 6; k++) d[k] = 0.0;
    double hR = zR - zbR;
    int dry = (hR < h_min);
    double q1 = dry ? 0.0 : qR, r1 = dry ? 0.0 : rR;
    double hRe = dry ? 0.0 : hR;
    int hden_f = !(hRe > SMALL);
    double hden = hden_f ? SMALL : hRe;
    double cR = sqrt(G * hden);
    double uR = q1 / hden;
    double L1 = (lpR > 0.0) ? lpR : 0.0;
    double f2Rk = bR * hRe * uR * uR + 0.5 * G * hRe * hRe;
    /* forward values needed for the reverse */
    double f1 = 0.0;
    double aq1 = 0, ar1 = 0, ahRe = 0, auR = 0, acR = 0, aL1 = 0, aphi = 0, abc = 0;
    double af1 = df[0], af2 = df[1], af3 = df[2];
    /* recompute f1 by type, keeping what the reverse needs */
    int q_crit = 0;
    double ps = 0, hc = 0, hcs = 0;
    double fr = 0, cc = 0, cc0 = 0, uu = 0, hh = 0;
    double hb = 0, ccw = 0;
    int hz_wet = 0, fr_in = 0, cc_pos = 0, hb_pos = 0, hbs_f = 0;
    if (type == ORC_BC_Q) {
        f1 = bc;
        if (bc > SMALL) {
            ps = (phiR > phi_min) ? phiR : phi_min;
            hc = exp((1.0 / 3.0) * log((bc / ps) / G));
            if (hc >= hRe) {
                q_crit = 1;
                hcs = (hc > SMALL) ? hc : SMALL;
                f1 = phiR * 1.5 * sqrt(G * hcs) * hc;
            }
        }
    } else if (type == ORC_BC_FR) {
        fr_in = (bc >= 0.0 && bc <= 1.0);
        fr = (bc < 0.0) ? 0.0 : ((bc > 1.0) ? 1.0 : bc);
        cc0 = (uR - 2.0 * cR) / (-fr - 2.0);
        cc_pos = (cc0 > 0.0);
        cc = cc_pos ? cc0 : 0.0;
        uu = -fr * cc;
        hh = cc * cc / G;
        f1 = phiR * uu * hh;
    } else if (type == ORC_BC_H || type == ORC_BC_Z) {
        hb_pos = (type == ORC_BC_Z) ? (bc - zbR > 0.0) : (bc > 0.0);
        hb = hb_pos ? ((type == ORC_BC_Z) ? bc - zbR : bc) : 0.0;
        hz_wet = (hRe > h_min);
        if (hz_wet) {
            f1 = phiR * (q1 + L1 * (hb - hRe));
        } else {
            hbs_f = !(hb > SMALL);
            ccw = sqrt(G * (hbs_f ? SMALL : hb));
            f1 = phiR * hb * ccw;
        }
    }
    /* f3 */
    if (!(f1 > 0.0)) {
        double hv = (hRe > SMALL) ? hRe : SMALL;
        double vR = r1 / hv;
        af1 += af3 * vR;
        double avR = af3 * f1;
        ar1 += avR / hv;
        if (hRe > SMALL) ahRe += -avR * r1 / (hv * hv);
    }
    /* f2 and f1 by type */
    int f2_std = 1;
    if (type == ORC_BC_H || type == ORC_BC_Z) f2_std = hz_wet;
    if (type != ORC_BC_Q && type != ORC_BC_FR && type != ORC_BC_H && type != ORC_BC_Z) f2_std = 2;
    double af2R = 0;
    if (f2_std == 1) {
        /* f2 = f2R + (f1 - q1 phiR) L1 */
        af2R += af2;
        af1 += af2 * L1;
        aq1 += -af2 * phiR * L1;
        aphi += -af2 * q1 * L1;
        aL1 += af2 * (f1 - q1 * phiR);
    } else if (f2_std == 0) {
        /* f2 = sforce(bR, phiR, hb, ccw) */
        aphi += af2 * (bR * hb * ccw * ccw + 0.5 * G * hb * hb);
        double ahb = af2 * phiR * (bR * ccw * ccw + G * hb);
        double accw = af2 * phiR * 2.0 * bR * hb * ccw;
        /* f1 = phiR hb ccw */
        aphi += af1 * hb * ccw;
        ahb += af1 * phiR * ccw;
        accw += af1 * phiR * hb;
        af1 = 0.0;
        if (!hbs_f) ahb += accw * G * dsq(G * hb);
        if (hb_pos) abc += ahb;
    } else {
        af2R += af2;
    }
    if (type == ORC_BC_Q) {
        if (q_crit) {
            /* f1 = phiR 1.5 sqrt(G hcs) hc */
            double sg = sqrt(G * hcs);
            aphi += af1 * 1.5 * sg * hc;
            double ahc = af1 * phiR * 1.5 * sg;
            if (hc > SMALL) ahc += af1 * phiR * 1.5 * hc * G * dsq(G * hcs);
            /* hc = ((bc/ps)/G)^(1/3): d hc / d bc = hc / (3 bc), d hc / d ps = -hc / (3 ps) */
            abc += ahc * hc / (3.0 * bc);
            if (phiR > phi_min) aphi += -ahc * hc / (3.0 * ps);
        } else {
            abc += af1;
        }
    } else if (type == ORC_BC_FR) {
        /* f1 = phiR uu hh, uu = -fr cc, hh = cc^2/G */
        aphi += af1 * uu * hh;
        double auu = af1 * phiR * hh, ahh = af1 * phiR * uu;
        double acc = -auu * fr + ahh * 2.0 * cc / G;
        double afr = -auu * cc;
        if (cc_pos) {
            /* cc0 = (uR - 2 cR) / (-fr - 2) */
            double den = -fr - 2.0;
            auR += acc / den;
            acR += -2.0 * acc / den;
            afr += acc * cc0 / den; /* d cc0 / d fr = cc0 / den (den' = -1) */
        }
        if (fr_in) abc += afr;
    } else if (type == ORC_BC_H || type == ORC_BC_Z) {
        if (hz_wet) {
            /* f1 = phiR (q1 + L1 (hb - hRe)) */
            aphi += af1 * (q1 + L1 * (hb - hRe));
            aq1 += af1 * phiR;
            aL1 += af1 * phiR * (hb - hRe);
            double ahb = af1 * phiR * L1;
            ahRe += -af1 * phiR * L1;
            if (hb_pos) abc += ahb;
        }
    }
    /* f2R = phiR (bR hRe uR^2 + 0.5 G hRe^2) */
    aphi += af2R * f2Rk;
    ahRe += af2R * phiR * (bR * uR * uR + G * hRe);
    auR += af2R * phiR * 2.0 * bR * hRe * uR;
    /* uR = q1 / hden, cR = sqrt(G hden) */
    aq1 += auR / hden;
    double ahden = -auR * q1 / (hden * hden) + acR * G * dsq(G * hden);
    if (!hden_f) ahRe += ahden;
    /* L1 = max(lpR, 0) */
    if (lpR > 0.0) d[3] += aL1;
    if (!dry) {
        d[0] += ahRe; /* hR = zR - zbR */
        d[1] += aq1;
        d[2] += ar1;
    }
    d[4] += aphi;
    d[5] += abc;
}

/* adjoint of the semi-implicit Strickler friction (sourceFrictionStrickler_E) */
static void fric_adj(double hthr, double h, double qx, double qy, double ca, double sa,
                     double K1in, double K2in, double dt, double *ah, double *aqx, double *aqy,
                     double *aK1, double *aK2) {
    if (!(h > hthr)) {
        *aqx = 0.0;
        *aqy = 0.0;
        return;
    }
    double hs = h + SMALL;
    double u = qx / hs, v = qy / hs;
    double ul = u * ca + v * sa, vl = -u * sa + v * ca;
    double un = sqrt(u * u + v * v + 1e-16);
    int k1f = !(K1in > 1e-10), k2f = !(K2in > 1e-10);
    double K1 = k1f ? 20.0 : K1in, K2 = k2f ? 20.0 : K2in;
    double cb = cbrt(h + 1e-16);
    double h43 = h * cb;
    double D1 = K1 * K1 * h43 + 1e-10, D2 = K2 * K2 * h43 + 1e-10;
    double d1 = 1.0 + (dt * G * un) / D1, d2 = 1.0 + (dt * G * un) / D2;
    double ul2 = ul / d1, vl2 = vl / d2;
    double bx = *aqx, by = *aqy;
    double aul2 = (bx * ca + by * sa) * h;
    double avl2 = (-bx * sa + by * ca) * h;
    double ahh = bx * (ul2 * ca - vl2 * sa) + by * (ul2 * sa + vl2 * ca);
    double aul = aul2 / d1, ad1 = -aul2 * ul2 / d1;
    double avl = avl2 / d2, ad2 = -avl2 * vl2 / d2;
    double aun = ad1 * dt * G / D1 + ad2 * dt * G / D2;
    double aD1 = -ad1 * dt * G * un / (D1 * D1), aD2 = -ad2 * dt * G * un / (D2 * D2);
    if (!k1f) *aK1 += aD1 * 2.0 * K1 * h43;
    if (!k2f) *aK2 += aD2 * 2.0 * K2 * h43;
    double ah43 = aD1 * K1 * K1 + aD2 * K2 * K2;
    ahh += ah43 * (cb + h * cb / (3.0 * (h + 1e-16)));
    double au = aun * u / un + aul * ca - avl * sa;
    double av = aun * v / un + aul * sa + avl * ca;
    *aqx = au / hs;
    *aqy = av / hs;
    ahh += -(au * u + av * v) / hs;
    *ah += ahh;
}

/* adjoint of sourceSingularHeadLoss_E */
static void hl_adj(double hthr, double h, double qx, double qy, double ca, double sa, double hsing,
                   double c1, double c2, double dt, double *ah, double *aqx, double *aqy) {
    if (!(h > hthr)) {
        *aqx = 0.0;
        *aqy = 0.0;
        return;
    }
    double xr = qx * ca + qy * sa, yr = -qx * sa + qy * ca;
    double ax = 0, ay = 0, hs = h + SMALL, qn = 1.0;
    int act = (h > hsing);
    if (act) {
        qn = sqrt(xr * xr + yr * yr + 1e-16);
        ax = qn * c1 / hs;
        ay = qn * c2 / hs;
    }
    double ex = exp(-ax * dt), ey = exp(-ay * dt);
    double bx = *aqx, by = *aqy;
    double axr2 = bx * ca + by * sa, ayr2 = -bx * sa + by * ca;
    double axr = axr2 * ex, ayr = ayr2 * ey;
    if (act) {
        double aax = -axr2 * xr * dt * ex, aay = -ayr2 * yr * dt * ey;
        double aqn = aax * c1 / hs + aay * c2 / hs;
        *ah += -(aax * ax + aay * ay) / hs;
        axr += aqn * xr / qn;
        ayr += aqn * yr / qn;
    }
    *aqx = axr * ca - ayr * sa;
    *aqy = axr * sa + ayr * ca;
}

static void fac_dphi(const orc_case *c, const double *phi, int64_t cell, double w, double *out) {
    *out = (phi[cell] > c->phi_min) ? -(w / c->cell_area[cell] / phi[cell]) / phi[cell] : 0.0;
}

/* Gradient of the frozen-replay gauge cost (orc_e_replay) with respect to
 * porosity, the two Strickler coefficients per cell and the boundary value of
 * every step and series.  Returns J. */
/* Thin-film cut of the reverse sweep (no reference counterpart; 0 = the
 * exact adjoint): the adjoint of U^{n+1} is dropped at cells whose depth
 * h^{n+1} is below the cut, before the reverse of step n.  Restates the
 * library's sw2dg_assim_set_adjoint_cut for parity. */
static double g_adj_cut = 0.0;
void orc_e_set_adjoint_cut(double h_cut) { g_adj_cut = h_cut; }

/* Adjoint explosion guard (get_spatial_gradient_checkpointed,
 * assim/swSpE_checkpoint.cpp:1035-1117, untiled): segments of `stride`
 * steps; after the reverse of segment k, ||lambda_h|| of its start state
 * (before that state's own observation seed) against the first non-zero
 * one; watermark = the gradient after milestone segments (k % 50 == 0,
 * first, last, segments with an observation); above ratio x ref the
 * gradient returns to the watermark, boundary-value gradients of earlier
 * steps are zeroed and the cost keeps only observations from the
 * watermark segment on.  ratio <= 0 or stride <= 0: off.
 * orc_e_truncation() gives the first kept step of the last gradient. */
static double g_expl_ratio = 0.0;
static int64_t g_expl_stride = 0, g_first_kept = 0;
void orc_e_set_explosion_guard(double ratio, int64_t stride) {
    g_expl_ratio = ratio;
    g_expl_stride = stride;
}
int64_t orc_e_truncation(void) { return g_first_kept; }

/* Checkpoint restarts of the replay under the reverse sweep
 * (get_spatial_gradient_checkpointed, swSpE_checkpoint.cpp:985-1030): each
 * segment of `stride` steps is replayed from the dry run's state at its
 * first step (states, as orc_e_dry_states writes them, kept by the caller),
 * and the adjoint of that start state is handed to the previous segment's
 * end state.  The cost sums the segments' observations.  Where the dry run
 * and the replay differ (MUSCL-EVR: the reconstruction's dt) the result
 * depends on the stride; otherwise it is the plain reverse sweep.  stride
 * <= 0 or states NULL: off. */
static int64_t g_ck_stride = 0;
static const double *g_ck_states = NULL;

/* Planted mistakes, for the tests that show the checks fail on them (0 =
 * none): 1 the slopes' adjoint is dropped (the limited slope held fixed),
 * 2 the reverse reconstructs with the previous step's dt (the dry run's). */
static int g_plant = 0;
void orc_e_set_plant(int code) { g_plant = code; }
void orc_e_set_checkpoints(int64_t stride, const double *states) {
    g_ck_stride = stride;
    g_ck_states = states;
}

static double e_gradient(const orc_case *c, const double *phi, const double *K1, const double *K2,
                         const double *h0, const double *qx0, const double *qy0, int64_t n_steps,
                         const double *dt_seq, const double *bc_seq, const int32_t *obs_trig,
                         int32_t n_obs, int32_t n_gauges, const int64_t *gauge_cell,
                         const double *gauge_w, const double *obs, const double *frames,
                         const double *w_h, const double *w_q, int32_t normalize, double *g_phi,
                         double *g_K1, double *g_K2, double *g_bc) {
    const int64_t nc = c->n_cell;
    const int ns = c->n_series;
    const double hthr = (c->h_min > SMALL) ? c->h_min : SMALL;
    const double norm = (normalize && n_obs > 0) ? (double)n_obs : 1.0;
    esolver s;
    e_init(&s, c, phi, K1, K2, h0, qx0, qy0);
    /* forward: keep every state */
    double *U = (double *)malloc(sizeof(double) * 3 * (size_t)nc * (size_t)(n_steps + 1));
    memcpy(U, h0, sizeof(double) * nc);
    memcpy(U + nc, qx0, sizeof(double) * nc);
    memcpy(U + 2 * nc, qy0, sizeof(double) * nc);
    double J = 0.0;
    double *Jstep = (double *)calloc((size_t)(n_steps > 0 ? n_steps : 1), sizeof(double));
    g_first_kept = 0;
    const int restart = g_ck_stride > 0 && g_ck_states != NULL;
    for (int64_t n = 0; n < n_steps; n++) {
        if (restart && n > 0 && n % g_ck_stride == 0) { /* segment start: the dry run's state */
            const double *ck = g_ck_states + 3 * (size_t)nc * (size_t)n;
            memcpy(s.h, ck, sizeof(double) * nc);
            memcpy(s.qx, ck + nc, sizeof(double) * nc);
            memcpy(s.qy, ck + 2 * nc, sizeof(double) * nc);
        }
        e_step(&s, dt_seq[n], bc_seq + n * ns);
        double *Un = U + 3 * (size_t)nc * (size_t)(n + 1);
        memcpy(Un, s.h, sizeof(double) * nc);
        memcpy(Un + nc, s.qx, sizeof(double) * nc);
        memcpy(Un + 2 * nc, s.qy, sizeof(double) * nc);
        int32_t oi = obs_trig[n];
        if (oi >= 0 && oi < n_obs && n_gauges > 0) {
            double Jo = 0.0;
            for (int g = 0; g < n_gauges; g++) {
                double d = s.h[gauge_cell[g]] - obs[(int64_t)oi * n_gauges + g];
                Jo += gauge_w[g] * d * d;
            }
            if (normalize && n_obs > 0) Jo /= (double)n_obs;
            J += Jo;
            Jstep[n] += Jo;
        }
        if (oi >= 0 && oi < n_obs && frames) {
            const double *fh = frames + (int64_t)oi * 3 * nc, *fqx = fh + nc, *fqy = fh + 2 * nc;
            double Jo = 0.0;
            for (int64_t i = 0; i < nc; i++) {
                double dh = s.h[i] - fh[i], dqx = s.qx[i] - fqx[i], dqy = s.qy[i] - fqy[i];
                double wh = w_h ? w_h[i] : 1.0, wq = w_q ? w_q[i] : 1.0;
                Jo += wh * (dh * dh) + wq * ((dqx * dqx) + (dqy * dqy));
            }
            if (normalize && n_obs > 0) Jo /= (double)n_obs;
            J += Jo;
            Jstep[n] += Jo;
        }
    }
    const int guard = g_expl_ratio > 0.0 && g_expl_stride > 0;
    const int64_t S = g_expl_stride > 0 ? g_expl_stride : 1;
    const int64_t K = (n_steps + S - 1) / S;
    double ref = -1.0;
    int64_t wm_seg = K - 1;
    double *wm = guard ? (double *)calloc(3 * (size_t)nc, sizeof(double)) : NULL;
    /* reverse */
    double *lam = (double *)calloc(3 * (size_t)nc, sizeof(double));  /* adjoint of U^{n+1} */
    double *lams = (double *)calloc(3 * (size_t)nc, sizeof(double)); /* adjoint after balance */
    double *bh = (double *)malloc(sizeof(double) * 3 * (size_t)nc);
    const int muscl = c->scheme == ORC_MUSCL_EVR;
    double *asx = (double *)calloc((size_t)nc, sizeof(double)); /* adjoint of the limited slopes */
    double *asy = (double *)calloc((size_t)nc, sizeof(double));
#ifdef ORC_OMP
    double *wb = (double *)calloc(6 * (size_t)(c->n_bnd + 1), sizeof(double));   /* per boundary edge */
    double *wi = (double *)calloc(14 * (size_t)(c->n_int + 1), sizeof(double));  /* per internal edge */
#endif
    for (int64_t i = 0; i < nc; i++) g_phi[i] = g_K1[i] = g_K2[i] = 0.0;
    for (int64_t k = 0; k < n_steps * (ns > 0 ? ns : 1); k++) g_bc[k] = 0.0;
    for (int64_t n = n_steps - 1; n >= 0; n--) {
        const double dt = dt_seq[n];
        const double *bcn = bc_seq + n * ns;
        /* start state of step n (a segment start: the checkpoint), end state
         * of step n (the replayed one, which its observation sees) */
        const double *Un = (restart && n > 0 && n % g_ck_stride == 0)
                               ? g_ck_states + 3 * (size_t)nc * (size_t)n
                               : U + 3 * (size_t)nc * (size_t)n;
        const double *Un1 = U + 3 * (size_t)nc * (size_t)(n + 1);
        int32_t oi = obs_trig[n];
        if (oi >= 0 && oi < n_obs && n_gauges > 0)
            for (int g = 0; g < n_gauges; g++) {
                int64_t gc = gauge_cell[g];
                lam[gc] += 2.0 * gauge_w[g] * (Un1[gc] - obs[(int64_t)oi * n_gauges + g]) / norm;
            }
        if (oi >= 0 && oi < n_obs && frames) {
            const double *fh = frames + (int64_t)oi * 3 * nc, *fqx = fh + nc, *fqy = fh + 2 * nc;
            for (int64_t i = 0; i < nc; i++) {
                double wh = w_h ? w_h[i] : 1.0, wq = w_q ? w_q[i] : 1.0;
                lam[i] += 2.0 * wh * (Un1[i] - fh[i]) / norm;
                lam[nc + i] += 2.0 * wq * (Un1[nc + i] - fqx[i]) / norm;
                lam[2 * nc + i] += 2.0 * wq * (Un1[2 * nc + i] - fqy[i]) / norm;
            }
        }
        if (g_adj_cut > 0.0)
            for (int64_t i = 0; i < nc; i++)
                if (Un1[i] < g_adj_cut) lam[i] = lam[nc + i] = lam[2 * nc + i] = 0.0;
        /* recompute the step's intermediates (the replay reconstructs with
         * the step's own dt) */
        memcpy(s.h, Un, sizeof(double) * nc);
        memcpy(s.qx, Un + nc, sizeof(double) * nc);
        memcpy(s.qy, Un + 2 * nc, sizeof(double) * nc);
        s.dt_recon = (g_plant == 2) ? (n > 0 ? dt_seq[n - 1] : 0.0) : dt;
        lr_any(&s);
        fluxes(&s, bcn);
#ifdef ORC_OMP
        memcpy(bh, Un, sizeof(double) * 3 * nc);
        balance_gather(&s, dt, bh, bh + nc, bh + 2 * nc);
#else
        memcpy(bh, Un, sizeof(double) * 3 * nc);
        {
            double *h = bh, *qx = bh + nc, *qy = bh + 2 * nc;
            for (int64_t e = 0; e < c->n_bnd; e++) {
                int64_t id = c->bnd_cell[e];
                double co = s.fac_b[e], nx = c->bnd_nx[e], ny = c->bnd_ny[e];
                h[id] += (+co) * s.F1b[e] * dt;
                qx[id] += ((+co * nx) * s.F2b[e] + (+co * nx) * 0.0 + (-co * ny) * s.F3b[e]) * dt;
                qy[id] += ((+co * ny) * s.F2b[e] + (+co * ny) * 0.0 + (+co * nx) * s.F3b[e]) * dt;
            }
            for (int64_t e = 0; e < c->n_int; e++) {
                int64_t L = c->int_L[e], R = c->int_R[e];
                double cl = s.fac_L[e], cr = s.fac_R[e], nx = c->int_nx[e], ny = c->int_ny[e];
                double dh = s.F1i[e] * dt;
                h[L] += (-cl) * dh;
                h[R] += (+cr) * dh;
                qx[L] += ((-cl * nx) * s.F2i[e] + (+cl * nx) * s.sLi[e] + (+cl * ny) * s.F3i[e]) * dt;
                qx[R] += ((+cr * nx) * s.F2i[e] + (+cr * nx) * s.sRi[e] + (-cr * ny) * s.F3i[e]) * dt;
                qy[L] += ((-cl * ny) * s.F2i[e] + (+cl * ny) * s.sLi[e] + (-cl * nx) * s.F3i[e]) * dt;
                qy[R] += ((+cr * ny) * s.F2i[e] + (+cr * ny) * s.sRi[e] + (+cr * nx) * s.F3i[e]) * dt;
            }
        }
#endif
        /* per cell: sources, clamp and clean-up in reverse */
        ORC_PAR
        for (int64_t i = 0; i < nc; i++) {
            double h = bh[i], qx = bh[nc + i], qy = bh[2 * nc + i];
            int kh = !(fabs(h) < SMALL), kx = !(fabs(qx) < SMALL), ky = !(fabs(qy) < SMALL);
            if (!kh) h = 0.0;
            if (!kx) qx = 0.0;
            if (!ky) qy = 0.0;
            int clamped = (h < 0.);
            if (clamped) h = qx = qy = 0.0;
            double fx = qx, fy = qy; /* state after friction */
            if (c->friction != ORC_FRIC_NONE) {
                if (h > hthr) {
                    double hs = h + SMALL;
                    double u = qx / hs, v = qy / hs;
                    double ca = s.fcos[i], sa = s.fsin[i];
                    double ul = u * ca + v * sa, vl = -u * sa + v * ca;
                    double un = sqrt(u * u + v * v + 1e-16);
                    double k1 = (K1[i] > 1e-10) ? K1[i] : 20.0, k2 = (K2[i] > 1e-10) ? K2[i] : 20.0;
                    double h43 = h * cbrt(h + 1e-16);
                    double d1 = 1.0 + (dt * G * un) / (k1 * k1 * h43 + 1e-10);
                    double d2 = 1.0 + (dt * G * un) / (k2 * k2 * h43 + 1e-10);
                    ul /= d1;
                    vl /= d2;
                    fx = (ul * ca - vl * sa) * h;
                    fy = (ul * sa + vl * ca) * h;
                } else {
                    fx = fy = 0.0;
                }
            }
            double ah = lam[i], ax = lam[nc + i], ay = lam[2 * nc + i];
            if (c->head_loss)
                hl_adj(hthr, h, fx, fy, s.hcos[i], s.hsin[i], c->hl_hsing[i], c->hl_c1[i],
                       c->hl_c2[i], dt, &ah, &ax, &ay);
            if (c->friction != ORC_FRIC_NONE)
                fric_adj(hthr, h, qx, qy, s.fcos[i], s.fsin[i], K1[i], K2[i], dt, &ah, &ax, &ay,
                         &g_K1[i], &g_K2[i]);
            if (clamped) ah = ax = ay = 0.0;
            if (!kh) ah = 0.0;
            if (!kx) ax = 0.0;
            if (!ky) ay = 0.0;
            lams[i] = ah;
            lams[nc + i] = ax;
            lams[2 * nc + i] = ay;
        }
#ifdef ORC_OMP
        /* every edge's reverse reads lams (fixed here) only: the terms of each
         * edge side go to per-edge slots, then each cell adds its own in
         * ascending edge order, boundary edges first (porosity and
         * boundary-value sums in the serial order; the adjoint state sums
         * each edge's terms before adding them) */
        memcpy(lam, lams, sizeof(double) * 3 * nc);
        ORC_PAR
        for (int64_t e = 0; e < c->n_bnd; e++) {
            int64_t id = c->bnd_cell[e];
            double co = s.fac_b[e], nx = c->bnd_nx[e], ny = c->bnd_ny[e];
            double lh = lams[id], lx = lams[nc + id], ly = lams[2 * nc + id];
            double df[3] = {lh * co * dt, (lx * co * nx + ly * co * ny) * dt,
                            (-lx * co * ny + ly * co * nx) * dt};
            double dco = lh * s.F1b[e] * dt +
                         (lx * (nx * s.F2b[e] - ny * s.F3b[e]) + ly * (ny * s.F2b[e] + nx * s.F3b[e])) * dt;
            double dfp;
            fac_dphi(c, phi, id, c->bnd_len[e], &dfp);
            double *tb = wb + 6 * e;
            tb[0] = dco * dfp;
            int k = c->bnd_series[e];
            double d[6];
            bnd_adj(c->h_min, c->phi_min, c->beta[id], phi[id], c->cell_zb[id], s.zrb[e], s.qrb[e],
                    s.rrb[e], s.lpb[e], c->series_type[k], bcn[k], df, d);
            tb[1] = d[4];
            tb[2] = d[5];
            tb[3] = tb[4] = tb[5] = 0.0;
            if (muscl)
                lr_bnd_muscl_adj(Un[id], Un[nc + id], Un[2 * nc + id], c->beta[id], nx, ny, d[0],
                                 d[1], d[2], d[3], &tb[3], &tb[4], &tb[5]);
            else
                lr_adj(Un[id], Un[nc + id], Un[2 * nc + id], c->beta[id], nx, ny, c->h_min, d[0],
                       d[1], d[2], 0.0, d[3], &tb[3], &tb[4], &tb[5]);
        }
        for (int64_t e = 0; e < c->n_bnd; e++) g_bc[n * ns + c->bnd_series[e]] += wb[6 * e + 2];
        ORC_PAR
        for (int64_t e = 0; e < c->n_int; e++) {
            int64_t L = c->int_L[e], R = c->int_R[e];
            double cl = s.fac_L[e], cr = s.fac_R[e], nx = c->int_nx[e], ny = c->int_ny[e];
            double lhL = lams[L], lxL = lams[nc + L], lyL = lams[2 * nc + L];
            double lhR = lams[R], lxR = lams[nc + R], lyR = lams[2 * nc + R];
            double F1 = s.F1i[e], F2 = s.F2i[e], F3 = s.F3i[e], sL = s.sLi[e], sR = s.sRi[e];
            double df[5];
            df[0] = (-lhL * cl + lhR * cr) * dt;
            df[1] = (-lxL * cl * nx + lxR * cr * nx - lyL * cl * ny + lyR * cr * ny) * dt;
            df[2] = (lxL * cl * ny - lxR * cr * ny - lyL * cl * nx + lyR * cr * nx) * dt;
            df[3] = (lxL * cl * nx + lyL * cl * ny) * dt;
            df[4] = (lxR * cr * nx + lyR * cr * ny) * dt;
            double dcl = (-lhL * F1 + lxL * (-nx * F2 + nx * sL + ny * F3) + lyL * (-ny * F2 + ny * sL - nx * F3)) * dt;
            double dcr = (lhR * F1 + lxR * (nx * F2 + nx * sR - ny * F3) + lyR * (ny * F2 + ny * sR + nx * F3)) * dt;
            double dpl, dpr;
            fac_dphi(c, phi, L, c->int_len[e], &dpl);
            fac_dphi(c, phi, R, c->int_len[e], &dpr);
            double *tL = wi + 14 * e, *tR = tL + 7;
            for (int q = 0; q < 14; q++) tL[q] = 0.0;
            tL[0] = dcl * dpl;
            tR[0] = dcr * dpr;
            double d[12];
            hllc_adj(c->h_min, c->phi_min, c->beta[L], c->beta[R], phi[L], phi[R], c->cell_zb[L],
                     c->cell_zb[R], s.zri[2 * e], s.qri[2 * e], s.rri[2 * e], s.zri[2 * e + 1],
                     s.qri[2 * e + 1], s.rri[2 * e + 1], s.lmi[2 * e], s.lmi[2 * e + 1], s.lpi[2 * e],
                     s.lpi[2 * e + 1], df, d);
            tL[1] = d[10];
            tR[1] = d[11];
            if (muscl) {
                recon_adj(0, Un[L], Un[nc + L], Un[2 * nc + L], c->cell_zb[L], c->beta[L], s.dzdx[L],
                          s.dzdy[L], nx, ny, c->int_xm[e] - c->cell_x[L], c->int_ym[e] - c->cell_y[L],
                          s.dt_recon, d[0], d[1], d[2], d[6], d[8], &tL[2], &tL[3], &tL[4], &tL[5], &tL[6]);
                recon_adj(1, Un[R], Un[nc + R], Un[2 * nc + R], c->cell_zb[R], c->beta[R], s.dzdx[R],
                          s.dzdy[R], nx, ny, c->int_xm[e] - c->cell_x[R], c->int_ym[e] - c->cell_y[R],
                          s.dt_recon, d[3], d[4], d[5], d[7], d[9], &tR[2], &tR[3], &tR[4], &tR[5], &tR[6]);
            } else {
                lr_adj(Un[L], Un[nc + L], Un[2 * nc + L], c->beta[L], nx, ny, c->h_min, d[0], d[1],
                       d[2], d[6], d[8], &tL[2], &tL[3], &tL[4]);
                lr_adj(Un[R], Un[nc + R], Un[2 * nc + R], c->beta[R], nx, ny, c->h_min, d[3], d[4],
                       d[5], d[7], d[9], &tR[2], &tR[3], &tR[4]);
            }
        }
        ORC_PAR
        for (int64_t i = 0; i < nc; i++) {
            for (int64_t p = s.inc.bptr[i]; p < s.inc.bptr[i + 1]; p++) {
                const double *tb = wb + 6 * s.inc.binc[p];
                g_phi[i] += tb[0];
                g_phi[i] += tb[1];
                lam[i] += tb[3];
                lam[nc + i] += tb[4];
                lam[2 * nc + i] += tb[5];
            }
            for (int64_t p = s.inc.iptr[i]; p < s.inc.iptr[i + 1]; p++) {
                const double *t = wi + 14 * (s.inc.iinc[p] >> 1) + 7 * (s.inc.iinc[p] & 1);
                g_phi[i] += t[0];
                g_phi[i] += t[1];
                lam[i] += t[2];
                lam[nc + i] += t[3];
                lam[2 * nc + i] += t[4];
                asx[i] += t[5];
                asy[i] += t[6];
            }
        }
#else
        /* balance: identity part, then every edge in reverse */
        memcpy(lam, lams, sizeof(double) * 3 * nc);
        for (int64_t e = 0; e < c->n_bnd; e++) {
            int64_t id = c->bnd_cell[e];
            double co = s.fac_b[e], nx = c->bnd_nx[e], ny = c->bnd_ny[e];
            double lh = lams[id], lx = lams[nc + id], ly = lams[2 * nc + id];
            double df[3] = {lh * co * dt, (lx * co * nx + ly * co * ny) * dt,
                            (-lx * co * ny + ly * co * nx) * dt};
            double dco = lh * s.F1b[e] * dt +
                         (lx * (nx * s.F2b[e] - ny * s.F3b[e]) + ly * (ny * s.F2b[e] + nx * s.F3b[e])) * dt;
            double dfp;
            fac_dphi(c, phi, id, c->bnd_len[e], &dfp);
            g_phi[id] += dco * dfp;
            int k = c->bnd_series[e];
            double d[6];
            bnd_adj(c->h_min, c->phi_min, c->beta[id], phi[id], c->cell_zb[id], s.zrb[e], s.qrb[e],
                    s.rrb[e], s.lpb[e], c->series_type[k], bcn[k], df, d);
            g_phi[id] += d[4];
            g_bc[n * ns + k] += d[5];
            if (muscl)
                lr_bnd_muscl_adj(Un[id], Un[nc + id], Un[2 * nc + id], c->beta[id], nx, ny, d[0],
                                 d[1], d[2], d[3], &lam[id], &lam[nc + id], &lam[2 * nc + id]);
            else
                lr_adj(Un[id], Un[nc + id], Un[2 * nc + id], c->beta[id], nx, ny, c->h_min, d[0],
                       d[1], d[2], 0.0, d[3], &lam[id], &lam[nc + id], &lam[2 * nc + id]);
        }
        for (int64_t e = 0; e < c->n_int; e++) {
            int64_t L = c->int_L[e], R = c->int_R[e];
            double cl = s.fac_L[e], cr = s.fac_R[e], nx = c->int_nx[e], ny = c->int_ny[e];
            double lhL = lams[L], lxL = lams[nc + L], lyL = lams[2 * nc + L];
            double lhR = lams[R], lxR = lams[nc + R], lyR = lams[2 * nc + R];
            double F1 = s.F1i[e], F2 = s.F2i[e], F3 = s.F3i[e], sL = s.sLi[e], sR = s.sRi[e];
            double df[5];
            df[0] = (-lhL * cl + lhR * cr) * dt;
            df[1] = (-lxL * cl * nx + lxR * cr * nx - lyL * cl * ny + lyR * cr * ny) * dt;
            df[2] = (lxL * cl * ny - lxR * cr * ny - lyL * cl * nx + lyR * cr * nx) * dt;
            df[3] = (lxL * cl * nx + lyL * cl * ny) * dt;
            df[4] = (lxR * cr * nx + lyR * cr * ny) * dt;
            double dcl = (-lhL * F1 + lxL * (-nx * F2 + nx * sL + ny * F3) + lyL * (-ny * F2 + ny * sL - nx * F3)) * dt;
            double dcr = (lhR * F1 + lxR * (nx * F2 + nx * sR - ny * F3) + lyR * (ny * F2 + ny * sR + nx * F3)) * dt;
            double dpl, dpr;
            fac_dphi(c, phi, L, c->int_len[e], &dpl);
            fac_dphi(c, phi, R, c->int_len[e], &dpr);
            g_phi[L] += dcl * dpl;
            g_phi[R] += dcr * dpr;
            double d[12];
            hllc_adj(c->h_min, c->phi_min, c->beta[L], c->beta[R], phi[L], phi[R], c->cell_zb[L],
                     c->cell_zb[R], s.zri[2 * e], s.qri[2 * e], s.rri[2 * e], s.zri[2 * e + 1],
                     s.qri[2 * e + 1], s.rri[2 * e + 1], s.lmi[2 * e], s.lmi[2 * e + 1], s.lpi[2 * e],
                     s.lpi[2 * e + 1], df, d);
            g_phi[L] += d[10];
            g_phi[R] += d[11];
            if (muscl) {
                recon_adj(0, Un[L], Un[nc + L], Un[2 * nc + L], c->cell_zb[L], c->beta[L], s.dzdx[L],
                          s.dzdy[L], nx, ny, c->int_xm[e] - c->cell_x[L], c->int_ym[e] - c->cell_y[L],
                          s.dt_recon, d[0], d[1], d[2], d[6], d[8], &lam[L], &lam[nc + L], &lam[2 * nc + L],
                          &asx[L], &asy[L]);
                recon_adj(1, Un[R], Un[nc + R], Un[2 * nc + R], c->cell_zb[R], c->beta[R], s.dzdx[R],
                          s.dzdy[R], nx, ny, c->int_xm[e] - c->cell_x[R], c->int_ym[e] - c->cell_y[R],
                          s.dt_recon, d[3], d[4], d[5], d[7], d[9], &lam[R], &lam[nc + R], &lam[2 * nc + R],
                          &asx[R], &asy[R]);
            } else {
                lr_adj(Un[L], Un[nc + L], Un[2 * nc + L], c->beta[L], nx, ny, c->h_min, d[0], d[1],
                       d[2], d[6], d[8], &lam[L], &lam[nc + L], &lam[2 * nc + L]);
                lr_adj(Un[R], Un[nc + R], Un[2 * nc + R], c->beta[R], nx, ny, c->h_min, d[3], d[4],
                       d[5], d[7], d[9], &lam[R], &lam[nc + R], &lam[2 * nc + R]);
            }
        }
#endif
        if (muscl) { /* slopes: free surface zc = h + zb of the cell and its neighbours */
            if (g_plant != 1) e_slopes_adj(&s, asx, asy, lam);
            memset(asx, 0, sizeof(double) * nc);
            memset(asy, 0, sizeof(double) * nc);
        }
        if (guard && n % S == 0) { /* segment k = n / S reversed: lam = adjoint of U^n */
            const int64_t k = n / S;
            double q = 0.0;
            for (int64_t i = 0; i < nc; i++) q += lam[i] * lam[i];
            const double nrm = sqrt(q);
            if (ref < 0.0 && nrm > 0.0) ref = nrm;
            int has_obs = 0;
            for (int64_t m = n; m < n + S && m < n_steps; m++) has_obs |= obs_trig[m] >= 0;
            if (k % 50 == 0 || k == K - 1 || k == 0 || has_obs) {
                memcpy(wm, g_phi, sizeof(double) * nc);
                memcpy(wm + nc, g_K1, sizeof(double) * nc);
                memcpy(wm + 2 * nc, g_K2, sizeof(double) * nc);
                wm_seg = k;
            }
            if (ref > 0.0 && nrm > g_expl_ratio * ref) {
                memcpy(g_phi, wm, sizeof(double) * nc);
                memcpy(g_K1, wm + nc, sizeof(double) * nc);
                memcpy(g_K2, wm + 2 * nc, sizeof(double) * nc);
                g_first_kept = wm_seg * S;
                for (int64_t m = 0; m < g_first_kept * (ns > 0 ? ns : 1); m++) g_bc[m] = 0.0;
                J = 0.0;
                for (int64_t m = g_first_kept; m < n_steps; m++)
                    if (obs_trig[m] >= 0) J += Jstep[m];
                break;
            }
        }
    }
    free(wm);
#ifdef ORC_OMP
    free(wb);
    free(wi);
#endif
    free(asx);
    free(asy);
    free(Jstep);
    free(U);
    free(lam);
    free(lams);
    free(bh);
    e_free(&s);
    return J;
}

/* Frozen replay with full-frame observations (compute_cost_frame_ptr,
 * assim/swSpE_ad.cpp:273-308): frames [n_obs][h | qx | qy] (3 n_cell each,
 * reference cell order), per-cell weights w_h / w_q (NULL = 1); each frame's
 * misfit divided by n_obs when normalize != 0, summed in time order. */
double orc_e_replay_frames(const orc_case *c, const double *phi, const double *K1, const double *K2,
                           const double *h0, const double *qx0, const double *qy0, int64_t n_steps,
                           const double *dt_seq, const double *bc_seq, const int32_t *obs_trig,
                           int32_t n_obs, const double *frames, const double *w_h,
                           const double *w_q, int32_t normalize) {
    esolver s;
    e_init(&s, c, phi, K1, K2, h0, qx0, qy0);
    const int64_t nc = c->n_cell;
    double J = 0.0;
    for (int64_t n = 0; n < n_steps; n++) {
        ORC_AUDIT_STEP(n);
        e_step(&s, dt_seq[n], bc_seq + n * c->n_series);
        int32_t oi = obs_trig[n];
        if (oi >= 0 && oi < n_obs) {
            const double *fh = frames + (int64_t)oi * 3 * nc, *fqx = fh + nc, *fqy = fh + 2 * nc;
            double Jo = 0.0;
            for (int64_t i = 0; i < nc; i++) {
                double dh = s.h[i] - fh[i], dqx = s.qx[i] - fqx[i], dqy = s.qy[i] - fqy[i];
                double wh = w_h ? w_h[i] : 1.0, wq = w_q ? w_q[i] : 1.0;
                Jo += wh * (dh * dh) + wq * ((dqx * dqx) + (dqy * dqy));
            }
            if (normalize && n_obs > 0) Jo /= (double)n_obs;
            J += Jo;
        }
    }
    e_free(&s);
    return J;
}

double orc_e_gradient(const orc_case *c, const double *phi, const double *K1, const double *K2,
                      const double *h0, const double *qx0, const double *qy0, int64_t n_steps,
                      const double *dt_seq, const double *bc_seq, const int32_t *obs_trig,
                      int32_t n_obs, int32_t n_gauges, const int64_t *gauge_cell,
                      const double *gauge_w, const double *obs, int32_t normalize, double *g_phi,
                      double *g_K1, double *g_K2, double *g_bc) {
    return e_gradient(c, phi, K1, K2, h0, qx0, qy0, n_steps, dt_seq, bc_seq, obs_trig, n_obs,
                      n_gauges, gauge_cell, gauge_w, obs, NULL, NULL, NULL, normalize, g_phi, g_K1,
                      g_K2, g_bc);
}

/* The same with full-frame observations (frames [n_obs][h | qx | qy]). */
double orc_e_gradient_frames(const orc_case *c, const double *phi, const double *K1,
                             const double *K2, const double *h0, const double *qx0,
                             const double *qy0, int64_t n_steps, const double *dt_seq,
                             const double *bc_seq, const int32_t *obs_trig, int32_t n_obs,
                             const double *frames, const double *w_h, const double *w_q,
                             int32_t normalize, double *g_phi, double *g_K1, double *g_K2,
                             double *g_bc) {
    return e_gradient(c, phi, K1, K2, h0, qx0, qy0, n_steps, dt_seq, bc_seq, obs_trig, n_obs, 0,
                      NULL, NULL, NULL, frames, w_h, w_q, normalize, g_phi, g_K1, g_K2, g_bc);
}
