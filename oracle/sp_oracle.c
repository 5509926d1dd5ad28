/*
 * sp_oracle.c -- CPU oracle (test infrastructure only) for the SW2D SP
 * forward step, native ("N") semantics.  See sp_oracle.h.
 *
 * Every function restates one reference routine; the citation is given at
 * the top of each.  Expression trees are kept in the reference's
 * evaluation order (left-to-right C/C++ semantics); build with
 * -ffp-contract=off so no multiply-add is fused.
 */
#include "sp_oracle.h"
#include "orc_math.h"
#include "orc_omp.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define SMALL_REAL 1e-16 /* src/common/param.h:2 */
#define GRAVITY 9.81     /* src/common/param.h:3 */
#define RHO_AIR 1.2      /* src/common/param.h:4 */
#define RHO_WATER 1.0e3  /* src/common/param.h:5 */
#define MOM_DIFF 1e-6    /* swSp.h:244, not read from the input */

/* std::max / std::min semantics: max returns a unless a < b, min returns a unless b < a */
static inline double smax(double a, double b) { return (a < b) ? b : a; }
static inline double smin(double a, double b) { return (b < a) ? b : a; }

struct orc_solver {
    orc_case c;
    int64_t n_audit; /* steps taken, for the math audit */
    /* owned copies of the per-cell parameters the step needs */
    double *h, *qx, *qy;
    /* edge factors (swSp.cpp:338-444) */
    double *bf1, *bf2F2, *bf2F3, *bf2S2, *bf3F2, *bf3F3, *bf3S2;
    double *fL1, *fR1, *f2F2L, *f2F2R, *f2F3L, *f2F3R, *f2S2L, *f2S2R;
    double *f3F2L, *f3F2R, *f3F3L, *f3F3R, *f3S2L, *f3S2R;
    /* Riemann states */
    double *zrb, *qrb, *rrb, *lpb, *lmb, *lmaxb;
    double *zri, *qri, *rri, *lmi, *lpi, *lmaxi; /* [2*n_int] except lmaxi */
    /* fluxes (persistent: dry-dry internal edges keep their previous value) */
    double *F1b, *F2b, *F3b, *S2Rb;
    double *F1i, *F2i, *F3i, *S2Li, *S2Ri;
    /* MUSCL work */
    double *zc, *dzdx, *dzdy;
    int64_t *nb_ptr, *nb_idx; /* neighbour list in ascending internal-edge order */
    double *swept;
    /* friction / head loss derived */
    double *fcos, *fsin, *hl_alpha;
    /* BC cursors (swBoundaryConditionTimeSeries state) */
    double *bc_vinf, *bc_vsup, *bc_tinf, *bc_tsup, *bc_val;
    int64_t *bc_rank;
    int *bc_atend;
    double *wvx, *wvy;  /* wind velocity per station at t + dt/2 */
    double *rain_now;   /* rain rate per station at t + dt/2 */
    double *pfac, *pcoef, *pflux;  /* momentum diffusion per internal edge */
    double t, dt;
    orc_inc inc; /* per-cell incidences (the OpenMP build's gathers) */
};

/* swSp::lambdaMinus / lambdaPlus -- swSp.cpp:972-995 */
static inline double lambda_minus(double b, double c, double u) {
    return b * u - sqrt((b - 1.0) * b * u * u + c * c);
}
static inline double lambda_plus(double b, double c, double u) {
    return b * u + sqrt((b - 1.0) * b * u * u + c * c);
}
/* swSp::specificForce -- swSp.cpp:1388-1399 */
static inline double specific_force(double b, double phi, double h, double u) {
    return phi * (b * h * u * u + 0.5 * GRAVITY * h * h);
}
/* swSp::hypHCritical -- swSpHyp.cpp:651-661 */
static inline double h_critical(double phi, double q) {
    return exp((1. / 3.) * log((q / smax(phi, SMALL_REAL)) / GRAVITY));
}

/* swSp::hypBoundaryFlux -- swSpHyp.cpp:13-135.  out = {F1, F2, F3, s2R} */
static int bnd_flux(double h_min, double bR, double phiR, double zbR, double zR, double qR,
                    double rR, double lambda_plus_R, int bc_type, double bc_val, double *out) {
    double hR = zR - zbR;
    double q1 = qR, r1 = rR;
    if (hR < h_min) {
        q1 = 0.;
        r1 = 0.;
        hR = 0.;
    }
    double cR = sqrt(GRAVITY * smax(hR, SMALL_REAL));
    double uR = q1 / smax(hR, SMALL_REAL);
    double L1 = smax(0.0, lambda_plus_R);
    double flux2R = specific_force(bR, phiR, hR, uR);
    double flux1 = 0.0, flux2 = 0.0, flux3 = 0.0;
    if (bc_type == ORC_BC_FR) {
        if ((bc_val < 0.) | (bc_val > 1.)) return ORC_ERR_BC;
        double c = (uR - 2. * cR) / (-bc_val - 2.);
        c = smax(c, 0.);
        double u = -bc_val * c;
        double h = c * c / GRAVITY;
        flux1 = phiR * u * h;
        flux2 = flux2R + (flux1 - q1 * phiR) * L1;
    } else if (bc_type == ORC_BC_H || bc_type == ORC_BC_Z) {
        double h_bc = (bc_type == ORC_BC_H) ? smax(bc_val, 0.) : smax(bc_val - zbR, 0.);
        if (hR > h_min) {
            flux1 = phiR * (q1 + L1 * (h_bc - hR));
            flux2 = flux2R + (flux1 - q1 * phiR) * L1;
        } else {
            double c = sqrt(GRAVITY * h_bc);
            flux1 = phiR * h_bc * c;
            flux2 = specific_force(bR, phiR, h_bc, c);
        }
    } else if (bc_type == ORC_BC_Q) {
        flux1 = bc_val;
        double hc = h_critical(phiR, bc_val);
        if (hc >= hR) flux1 = phiR * 1.5 * sqrt(GRAVITY * hc) * hc;
        flux2 = flux2R + (flux1 - q1 * phiR) * L1;
    } else {
        return ORC_ERR_BC;
    }
    if (flux1 > 0) {
        flux3 = 0;
    } else {
        double vR = r1 / smax(hR, SMALL_REAL);
        flux3 = flux1 * vR;
    }
    out[0] = flux1;
    out[1] = flux2;
    out[2] = flux3;
    out[3] = 0.0; /* source_right is never assigned by the reference: stays 0 */
    return ORC_OK;
}

/* swSp::hypHllc (swSpHyp.cpp:664-821) and swSp::hypHllcBlending (:824-998).
 * out = {F1, F2, F3, s2L, s2R} */
static int hllc(int blending, double h_min, double phi_min, double bL, double bR, double phiL,
                double phiR, double zbL, double zbR, double zL, double qL, double rL, double zR,
                double qR, double rR, double lmL, double lmR, double lpL, double lpR,
                double *out) {
    double hL, hR;
    if (blending) {
        hL = zL - zbL;
        hR = zR - zbR;
    } else {
        hL = smax(0.0, zL - zbL);
        hR = smax(0.0, zR - zbR);
    }
    int both = 0, lhs = 0, rhs = 0, nowhere = 0;
    if (((hL <= h_min) && (hR <= h_min)) || ((phiL <= phi_min) && (phiR <= phi_min))) {
        nowhere = 1;
    } else {
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
    if (both + lhs + rhs + nowhere != 1) return ORC_ERR_FLOW_CONFIG;
    double f1 = 0., f2 = 0., f3 = 0., sl = 0., sr = 0.;
    if (both) {
        double uL, uR;
        if (blending) {
            uL = qL / smax(hL, SMALL_REAL);
            uR = qR / smax(hR, SMALL_REAL);
        } else {
            uL = qL / smax(SMALL_REAL, hL);
            uR = qR / smax(SMALL_REAL, hR);
        }
        double L1 = smin(lmL, smin(lmR, 0.));
        double L2 = smax(lpL, smax(lpR, 0.));
        double s2;
        if (!blending) {
            double phiLR = smin(phiL, phiR);
            f1 = phiLR * (L2 * qL - L1 * qR + L1 * L2 * (zR - zL)) / smax(L2 - L1, SMALL_REAL);
            double fact = 0.5 * smin(phiL, phiR) * (hL + hR);
            s2 = (phiR * hR * hR - phiL * hL * hL) * 0.5 + fact * (zL - zR);
        } else {
            double alpha_h = 0.0;
            if (fabs(zbL - zbR) > SMALL_REAL) alpha_h = smax(0.0, smin(1.0, (zL - zR) / (zbL - zbR)));
            double phiLR = smin(phiL, phiR);
            double F1h = phiLR * (L2 * qL - L1 * qR + L1 * L2 * (hR - hL)) / smax(L2 - L1, SMALL_REAL);
            double F1z = phiLR * (L2 * qL - L1 * qR + L1 * L2 * (zR - zL)) / smax(L2 - L1, SMALL_REAL);
            f1 = alpha_h * F1h + (1.0 - alpha_h) * F1z;
            double fact = 0.5 * smin(phiL, phiR) * (hL + hR);
            double s2z = (phiR * hR * hR - phiL * hL * hL) * 0.5 + fact * (zL - zR);
            double s2h;
            if (zbL > zbR)
                s2h = phiR * hR * (zbL - zbR);
            else
                s2h = phiL * hL * (zbL - zbR);
            s2 = alpha_h * s2h + (1.0 - alpha_h) * s2z;
        }
        s2 = s2 * GRAVITY / smax(L2 - L1, SMALL_REAL);
        sl = -L1 * s2;
        sr = L2 * s2;
        double FL = specific_force(bL, phiL, hL, uL);
        double FR = specific_force(bR, phiR, hR, uR);
        f2 = (L2 * (FL + sl) - L1 * (FR - sr) + L1 * L2 * (phiR * qR - phiL * qL)) /
             smax(L2 - L1, SMALL_REAL);
        double vL = rL / smax(hL, SMALL_REAL);
        double vR = rR / smax(hR, SMALL_REAL);
        if (f1 >= 0)
            f3 = bL * f1 * vL;
        else
            f3 = bR * f1 * vR;
    } else if (rhs) {
        sr = phiR * 0.5 * GRAVITY * hR * hR;
    } else if (lhs) {
        sl = -phiL * 0.5 * GRAVITY * hL * hL;
    }
    out[0] = f1;
    out[1] = f2;
    out[2] = f3;
    out[3] = sl;
    out[4] = sr;
    return ORC_OK;
}

/* swSource::frictionManning1D / Strickler1D / Chezy1D -- swSourceFriction.cpp:7-79 */
static inline double friction_factor(int model, double coeff, double h, double u_norm, double dt) {
    double factor = 1.0;
    if (model == ORC_FRIC_MANNING) {
        if (h > SMALL_REAL) {
            double arg = -GRAVITY * u_norm * coeff * coeff * exp(-4.0 / 3.0 * log(h));
            factor = exp(arg * dt);
        } else {
            factor = 0.0;
        }
    } else if (model == ORC_FRIC_STRICKLER) {
        if ((h > SMALL_REAL) && (coeff > SMALL_REAL)) {
            double arg = -GRAVITY * u_norm / coeff / coeff * exp(-4.0 / 3.0 * log(h));
            factor = exp(arg * dt);
        } else {
            factor = 0.0;
        }
    } else if (model == ORC_FRIC_CHEZY) {
        if ((h > SMALL_REAL) && (coeff > SMALL_REAL)) {
            double arg = -GRAVITY * u_norm / coeff / coeff / h;
            factor = exp(arg * dt);
        } else {
            factor = 0.0;
        }
    }
    return factor;
}

/* swBoundaryConditionTimeSeries::initialise -- swBoundaryConditionTimeSeries.cpp:17-58 */
static void bc_init(orc_solver *s) {
    const orc_case *c = &s->c;
    for (int k = 0; k < c->n_series; ++k) {
        const double *v = c->series_v + (int64_t)k * c->n_times;
        s->bc_vinf[k] = v[0];
        s->bc_tinf[k] = c->series_t[0];
        if (c->n_times > 1) {
            s->bc_vsup[k] = v[1];
            s->bc_tsup[k] = c->series_t[1];
            s->bc_atend[k] = 0;
        } else {
            s->bc_vsup[k] = s->bc_vinf[k];
            s->bc_tsup[k] = s->bc_tinf[k];
            s->bc_atend[k] = 1;
        }
        s->bc_rank[k] = 1;
        s->bc_val[k] = 0.0;
    }
}

/* swBoundaryConditionTimeSeries::computeValBC -- swBoundaryConditionTimeSeries.cpp:61-96 */
static void bc_compute(orc_solver *s, double time) {
    const orc_case *c = &s->c;
    for (int k = 0; k < c->n_series; ++k) {
        const double *v = c->series_v + (int64_t)k * c->n_times;
        while ((time > s->bc_tsup[k]) && (s->bc_atend[k] == 0)) {
            s->bc_vinf[k] = s->bc_vsup[k];
            s->bc_tinf[k] = s->bc_tsup[k];
            s->bc_tsup[k] = c->series_t[s->bc_rank[k]];
            s->bc_vsup[k] = v[s->bc_rank[k]];
            ++s->bc_rank[k];
            if (s->bc_rank[k] == c->n_times) s->bc_atend[k] = 1;
        }
        if (time > s->bc_tsup[k]) {
            s->bc_val[k] = s->bc_vsup[k];
        } else if (time < s->bc_tinf[k]) {
            s->bc_val[k] = s->bc_vinf[k];
        } else {
            double delta_t = (time - s->bc_tinf[k]) / (s->bc_tsup[k] - s->bc_tinf[k]);
            s->bc_val[k] = s->bc_vinf[k] + (s->bc_vsup[k] - s->bc_vinf[k]) * delta_t;
        }
    }
}

/* swSp::computeEdgeFactors -- swSp.cpp:338-444 */
static void edge_factors(orc_solver *s) {
    const orc_case *c = &s->c;
    for (int64_t i = 0; i < c->n_bnd; i++) {
        double w = c->bnd_len[i];
        int64_t cell = c->bnd_cell[i];
        double area = c->cell_area[cell];
        double nx = c->bnd_nx[i], ny = c->bnd_ny[i];
        double coeff = 0.;
        if (c->phi[cell] > c->phi_min) coeff = w / area / c->phi[cell];
        s->bf1[i] = +coeff;
        s->bf2F2[i] = +coeff * nx;
        s->bf2F3[i] = -coeff * ny;
        s->bf2S2[i] = +coeff * nx;
        s->bf3F2[i] = +coeff * ny;
        s->bf3F3[i] = +coeff * nx;
        s->bf3S2[i] = +coeff * ny;
    }
    for (int64_t i = 0; i < c->n_int; i++) {
        double w = c->int_len[i];
        int64_t L = c->int_L[i], R = c->int_R[i];
        double nx = c->int_nx[i], ny = c->int_ny[i];
        double cl = 0.0, cr = 0.0;
        if (c->phi[L] > c->phi_min) cl = w / c->cell_area[L] / c->phi[L];
        if (c->phi[R] > c->phi_min) cr = w / c->cell_area[R] / c->phi[R];
        s->fL1[i] = -cl;
        s->fR1[i] = +cr;
        s->f2F2L[i] = -cl * nx;
        s->f2F2R[i] = +cr * nx;
        s->f2F3L[i] = +cl * ny;
        s->f2F3R[i] = -cr * ny;
        s->f2S2L[i] = +cl * nx;
        s->f2S2R[i] = +cr * nx;
        s->f3F2L[i] = -cl * ny;
        s->f3F2R[i] = +cr * ny;
        s->f3F3L[i] = -cl * nx;
        s->f3F3R[i] = +cr * nx;
        s->f3S2L[i] = +cl * ny;
        s->f3S2R[i] = +cr * ny;
    }
}

/* Godunov LR states -- swSpHypGodunov.cpp:9-84 (boundary part shared with MUSCL,
 * swSpHypMusclEvr.cpp:144-162) */
static void lr_bnd(orc_solver *s) {
    const orc_case *c = &s->c;
    ORC_PAR
    for (int64_t e = 0; e < c->n_bnd; e++) {
        double nx = c->bnd_nx[e], ny = c->bnd_ny[e];
        int64_t id = c->bnd_cell[e];
        double bR = c->beta[id];
        double q = s->qx[id], r = s->qy[id];
        double hrb = s->h[id];
        s->zrb[e] = s->h[id] + c->cell_zb[id];
        s->qrb[e] = q * nx + r * ny;
        s->rrb[e] = -q * ny + r * nx;
        double hR = hrb;
        double uR = s->qrb[e] / smax(hR, SMALL_REAL);
        double cR = sqrt(GRAVITY * hR);
        s->lmb[e] = lambda_minus(bR, cR, uR);
        s->lpb[e] = lambda_plus(bR, cR, uR);
        s->lmaxb[e] = smax(fabs(s->lmb[e]), fabs(s->lpb[e]));
    }
}

static void lr_godunov(orc_solver *s) {
    const orc_case *c = &s->c;
    lr_bnd(s);
    ORC_PAR
    for (int64_t e = 0; e < c->n_int; e++) {
        double nx = c->int_nx[e], ny = c->int_ny[e];
        for (int side = 0; side < 2; side++) {
            int64_t id = side ? c->int_R[e] : c->int_L[e];
            double b = c->beta[id];
            double q = s->qx[id], r = s->qy[id];
            double hh = s->h[id];
            s->zri[2 * e + side] = s->h[id] + c->cell_zb[id];
            s->qri[2 * e + side] = q * nx + r * ny;
            s->rri[2 * e + side] = -q * ny + r * nx;
            double u = s->qri[2 * e + side] / smax(hh, SMALL_REAL);
            double cc = sqrt(GRAVITY * hh);
            s->lmi[2 * e + side] = lambda_minus(b, cc, u);
            s->lpi[2 * e + side] = lambda_plus(b, cc, u);
        }
        /* std::max({...}) keeps the first maximum */
        double m = fabs(s->lmi[2 * e]);
        double v = fabs(s->lpi[2 * e]);
        if (m < v) m = v;
        v = fabs(s->lmi[2 * e + 1]);
        if (m < v) m = v;
        v = fabs(s->lpi[2 * e + 1]);
        if (m < v) m = v;
        s->lmaxi[e] = m;
    }
}

/* swSp::computeMusclSlopes -- swSpHypMusclEvr.cpp:11-115 */
static void muscl_slopes(orc_solver *s) {
    const orc_case *c = &s->c;
    ORC_PAR
    for (int64_t i = 0; i < c->n_cell; i++) {
        s->dzdx[i] = 0.0;
        s->dzdy[i] = 0.0;
        double wsum = 0.0;
        int64_t nn = s->nb_ptr[i + 1] - s->nb_ptr[i];
        if (nn > 1) {
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
            double fmx = smax(s->zc[L], s->zc[R]);
            double fmn = smin(s->zc[L], s->zc[R]);
            double fr = s->zc[k] + (c->int_xm[e] - c->cell_x[k]) * s->dzdx[k] +
                        (c->int_ym[e] - c->cell_y[k]) * s->dzdy[k];
            if (fr < fmn) {
                double dfr = fr - s->zc[k];
                double dft = fmn - s->zc[k];
                double sf = dft / dfr;
                s->dzdx[k] *= sf;
                s->dzdy[k] *= sf;
            }
            if (fr > fmx) {
                double dfr = fr - s->zc[k];
                double dft = fmx - s->zc[k];
                double sf = dft / dfr;
                s->dzdx[k] *= sf;
                s->dzdy[k] *= sf;
            }
        }
    }
    return;
#endif
    for (int64_t e = 0; e < c->n_int; e++) {
        int64_t L = c->int_L[e], R = c->int_R[e];
        double fmx = smax(s->zc[L], s->zc[R]);
        double fmn = smin(s->zc[L], s->zc[R]);
        for (int side = 0; side < 2; side++) {
            int64_t k = side ? R : L;
            double fr = s->zc[k] + (c->int_xm[e] - c->cell_x[k]) * s->dzdx[k] +
                        (c->int_ym[e] - c->cell_y[k]) * s->dzdy[k];
            if (fr < fmn) {
                double dfr = fr - s->zc[k];
                double dft = fmn - s->zc[k];
                double sf = dft / dfr;
                s->dzdx[k] *= sf;
                s->dzdy[k] *= sf;
            }
            if (fr > fmx) {
                double dfr = fr - s->zc[k];
                double dft = fmx - s->zc[k];
                double sf = dft / dfr;
                s->dzdx[k] *= sf;
                s->dzdy[k] *= sf;
            }
        }
    }
}

/* swSp::hypMusclEvrComputeLRStates -- swSpHypMusclEvr.cpp:119-277; dt is the
 * previous step's dt (the solver member, 0 before the first step, swSp.h:155) */
static void lr_muscl(orc_solver *s) {
    const orc_case *c = &s->c;
    ORC_PAR
    for (int64_t i = 0; i < c->n_cell; i++) s->zc[i] = s->h[i] + c->cell_zb[i];
    muscl_slopes(s);
    lr_bnd(s);
    double dt = s->dt;
    ORC_PAR
    for (int64_t e = 0; e < c->n_int; e++) {
        double nx = c->int_nx[e], ny = c->int_ny[e];
        for (int side = 0; side < 2; side++) {
            int64_t id = side ? c->int_R[e] : c->int_L[e];
            double b = c->beta[id];
            double zb = c->cell_zb[id];
            double h_local = s->h[id];
            double q_local, r_local;
            if (side == 0) {
                double q = s->qx[id], r = s->qy[id];
                q_local = q * nx + r * ny;
                r_local = r * nx - q * ny;
            } else {
                q_local = s->qx[id] * nx + s->qy[id] * ny;
                r_local = s->qy[id] * nx - s->qx[id] * ny;
            }
            double c_local = sqrt(GRAVITY * smax(SMALL_REAL, h_local));
            double u_local = q_local / smax(SMALL_REAL, h_local);
            double v_local = r_local / smax(SMALL_REAL, h_local);
            double l1, l2;
            if (side == 0) {
                l1 = smax(0.0, lambda_minus(b, c_local, u_local));
                l2 = smax(0.0, lambda_plus(b, c_local, u_local));
            } else {
                l1 = smin(0.0, lambda_minus(b, c_local, u_local));
                l2 = smin(0.0, lambda_plus(b, c_local, u_local));
            }
            double dx1 = c->int_xm[e] - c->cell_x[id] - 0.5 * l1 * dt * nx;
            double dy1 = c->int_ym[e] - c->cell_y[id] - 0.5 * l1 * dt * ny;
            double z1 = s->zc[id] + dx1 * s->dzdx[id] + dy1 * s->dzdy[id];
            double h1 = smax(0.0, z1 - zb);
            double dx2 = c->int_xm[e] - c->cell_x[id] - 0.5 * l2 * dt * nx;
            double dy2 = c->int_ym[e] - c->cell_y[id] - 0.5 * l2 * dt * ny;
            double z2 = s->zc[id] + dx2 * s->dzdx[id] + dy2 * s->dzdy[id];
            double h2 = smax(0.0, z2 - zb);
            double hr = 0.5 * (h1 + h2);
            s->zri[2 * e + side] = 0.5 * (h1 + h2) + zb;
            s->qri[2 * e + side] = 0.5 * (h1 + h2) * u_local;
            s->rri[2 * e + side] = 0.5 * (h1 + h2) * v_local;
            double u = s->qri[2 * e + side] / smax(hr, SMALL_REAL);
            double cc = sqrt(GRAVITY * hr);
            s->lmi[2 * e + side] = lambda_minus(b, cc, u);
            s->lpi[2 * e + side] = lambda_plus(b, cc, u);
        }
        double m = fabs(s->lmi[2 * e]);
        double v = fabs(s->lpi[2 * e]);
        if (m < v) m = v;
        v = fabs(s->lmi[2 * e + 1]);
        if (m < v) m = v;
        v = fabs(s->lpi[2 * e + 1]);
        if (m < v) m = v;
        s->lmaxi[e] = m;
    }
}

/* swSp::hypComputeTimeStep -- swSpHyp.cpp:430-512 (native: lambda carried L->R) */
static double time_step(orc_solver *s) {
    const orc_case *c = &s->c;
    double dt = c->dt_max;
#ifdef ORC_OMP
    /* per cell: its boundary edges, then its internal edges, each ascending */
#pragma omp parallel
    {
        double dtl = dt;
#pragma omp for schedule(static)
        for (int64_t i = 0; i < c->n_cell; i++) {
            double sw = 0.;
            for (int64_t p = s->inc.bptr[i]; p < s->inc.bptr[i + 1]; p++) {
                const int64_t e = s->inc.binc[p];
                if (s->h[i] > c->h_min) sw += s->lmaxb[e] * c->bnd_len[e];
            }
            for (int64_t p = s->inc.iptr[i]; p < s->inc.iptr[i + 1]; p++) {
                const int64_t e = s->inc.iinc[p] >> 1;
                const int side = (int)(s->inc.iinc[p] & 1);
                double lambda = 0.;
                if (s->h[c->int_L[e]] > c->h_min) lambda = s->lmaxi[e];
                if (side && s->h[c->int_R[e]] > c->h_min) lambda = s->lmaxi[e];
                sw += lambda * c->int_len[e];
            }
            s->swept[i] = sw;
            double x = c->cfl_max * (c->cell_area[i] / smax(sw, SMALL_REAL));
            if (x < dtl) dtl = x;
        }
#pragma omp critical
        if (dtl < dt) dt = dtl;
    }
    if (c->dt_map > SMALL_REAL) {
        int now = (int)(s->t / c->dt_map);
        int later = (int)((s->t + dt) / c->dt_map);
        if (later > now) dt = later * c->dt_map - s->t;
    }
    if ((s->t + dt) > c->t_max) dt = c->t_max - s->t;
    return dt;
#endif
    for (int64_t i = 0; i < c->n_cell; i++) s->swept[i] = 0.;
    for (int64_t e = 0; e < c->n_bnd; e++) {
        int64_t id = c->bnd_cell[e];
        if (s->h[id] > c->h_min) s->swept[id] += s->lmaxb[e] * c->bnd_len[e];
    }
    for (int64_t e = 0; e < c->n_int; e++) {
        double lambda = 0.;
        int64_t L = c->int_L[e], R = c->int_R[e];
        if (s->h[L] > c->h_min) lambda = s->lmaxi[e];
        s->swept[L] += lambda * c->int_len[e];
        if (s->h[R] > c->h_min) lambda = s->lmaxi[e];
        s->swept[R] += lambda * c->int_len[e];
    }
    for (int64_t i = 0; i < c->n_cell; i++) {
        double x = c->cfl_max * (c->cell_area[i] / smax(s->swept[i], SMALL_REAL));
        if (x < dt) dt = x;
    }
    if (c->dt_map > SMALL_REAL) {
        int now = (int)(s->t / c->dt_map);
        int later = (int)((s->t + dt) / c->dt_map);
        if (later > now) dt = later * c->dt_map - s->t;
    }
    if ((s->t + dt) > c->t_max) dt = c->t_max - s->t;
    return dt;
}

/* swSp::hypComputeFluxes -- swSpHyp.cpp:276-402 */
static int fluxes(orc_solver *s) {
    const orc_case *c = &s->c;
    bc_compute(s, s->t);
#ifdef ORC_OMP
    int err = ORC_OK;
    ORC_PAR
    for (int64_t e = 0; e < c->n_bnd; e++) {
        int64_t id = c->bnd_cell[e];
        int k = c->bnd_series[e];
        double out[4];
        int st = bnd_flux(c->h_min, c->beta[id], c->phi[id], c->cell_zb[id], s->zrb[e], s->qrb[e],
                          s->rrb[e], s->lpb[e], c->series_type[k], s->bc_val[k], out);
        if (st) {
#pragma omp critical
            err = err ? err : st;
        }
        s->F1b[e] = out[0];
        s->F2b[e] = out[1];
        s->F3b[e] = out[2];
        s->S2Rb[e] = out[3];
    }
    if (err) return err;
    ORC_PAR
    for (int64_t e = 0; e < c->n_int; e++) {
        int64_t L = c->int_L[e], R = c->int_R[e];
        double zbL = c->cell_zb[L], zbR = c->cell_zb[R];
        double zL = s->zri[2 * e], zR = s->zri[2 * e + 1];
        double hL = zL - zbL, hR = zR - zbR;
        if ((hL > c->h_min) || (hR > c->h_min)) {
            double out[5];
            int st = hllc(c->topo == ORC_TOPO_BLENDING, c->h_min, c->phi_min, c->beta[L], c->beta[R],
                          c->phi[L], c->phi[R], zbL, zbR, zL, s->qri[2 * e], s->rri[2 * e], zR,
                          s->qri[2 * e + 1], s->rri[2 * e + 1], s->lmi[2 * e], s->lmi[2 * e + 1],
                          s->lpi[2 * e], s->lpi[2 * e + 1], out);
            if (!st && isnan(out[0])) st = ORC_ERR_NAN;
            if (st) {
#pragma omp critical
                err = err ? err : st;
            }
            s->F1i[e] = out[0];
            s->F2i[e] = out[1];
            s->F3i[e] = out[2];
            s->S2Li[e] = out[3];
            s->S2Ri[e] = out[4];
        }
    }
    return err;
#endif
    for (int64_t e = 0; e < c->n_bnd; e++) {
        int64_t id = c->bnd_cell[e];
        int k = c->bnd_series[e];
        double out[4];
        int st = bnd_flux(c->h_min, c->beta[id], c->phi[id], c->cell_zb[id], s->zrb[e], s->qrb[e],
                          s->rrb[e], s->lpb[e], c->series_type[k], s->bc_val[k], out);
        if (st) return st;
        s->F1b[e] = out[0];
        s->F2b[e] = out[1];
        s->F3b[e] = out[2];
        s->S2Rb[e] = out[3];
    }
    for (int64_t e = 0; e < c->n_int; e++) {
        int64_t L = c->int_L[e], R = c->int_R[e];
        double zbL = c->cell_zb[L], zbR = c->cell_zb[R];
        double zL = s->zri[2 * e], zR = s->zri[2 * e + 1];
        double hL = zL - zbL, hR = zR - zbR;
        if ((hL > c->h_min) || (hR > c->h_min)) {
            double out[5];
            int st = hllc(c->topo == ORC_TOPO_BLENDING, c->h_min, c->phi_min, c->beta[L], c->beta[R],
                          c->phi[L], c->phi[R], zbL, zbR, zL, s->qri[2 * e], s->rri[2 * e], zR,
                          s->qri[2 * e + 1], s->rri[2 * e + 1], s->lmi[2 * e], s->lmi[2 * e + 1],
                          s->lpi[2 * e], s->lpi[2 * e + 1], out);
            if (st) return st;
            if (isnan(out[0])) return ORC_ERR_NAN;
            s->F1i[e] = out[0];
            s->F2i[e] = out[1];
            s->F3i[e] = out[2];
            s->S2Li[e] = out[3];
            s->S2Ri[e] = out[4];
        }
    }
    return ORC_OK;
}

/* swSp::hypComputeBalance -- swSpHyp.cpp:214-273 */
static void balance(orc_solver *s, double dt) {
    const orc_case *c = &s->c;
    double *h = s->h, *qx = s->qx, *qy = s->qy;
#ifdef ORC_OMP
    ORC_PAR
    for (int64_t i = 0; i < c->n_cell; i++) {
        for (int64_t p = s->inc.bptr[i]; p < s->inc.bptr[i + 1]; p++) {
            const int64_t e = s->inc.binc[p];
            h[i] += s->bf1[e] * s->F1b[e] * dt;
            qx[i] += (s->bf2F2[e] * s->F2b[e] + s->bf2S2[e] * s->S2Rb[e] + s->bf2F3[e] * s->F3b[e]) * dt;
            qy[i] += (s->bf3F2[e] * s->F2b[e] + s->bf3S2[e] * s->S2Rb[e] + s->bf3F3[e] * s->F3b[e]) * dt;
        }
        for (int64_t p = s->inc.iptr[i]; p < s->inc.iptr[i + 1]; p++) {
            const int64_t e = s->inc.iinc[p] >> 1;
            if (s->inc.iinc[p] & 1) {
                h[i] += s->fR1[e] * s->F1i[e] * dt;
                qx[i] += (s->f2F2R[e] * s->F2i[e] + s->f2S2R[e] * s->S2Ri[e] + s->f2F3R[e] * s->F3i[e]) * dt;
                qy[i] += (s->f3F2R[e] * s->F2i[e] + s->f3S2R[e] * s->S2Ri[e] + s->f3F3R[e] * s->F3i[e]) * dt;
            } else {
                h[i] += s->fL1[e] * s->F1i[e] * dt;
                qx[i] += (s->f2F2L[e] * s->F2i[e] + s->f2S2L[e] * s->S2Li[e] + s->f2F3L[e] * s->F3i[e]) * dt;
                qy[i] += (s->f3F2L[e] * s->F2i[e] + s->f3S2L[e] * s->S2Li[e] + s->f3F3L[e] * s->F3i[e]) * dt;
            }
        }
        if (fabs(h[i]) < SMALL_REAL) h[i] = 0.;
        if (fabs(qx[i]) < SMALL_REAL) qx[i] = 0.;
        if (fabs(qy[i]) < SMALL_REAL) qy[i] = 0.;
    }
    return;
#endif
    for (int64_t e = 0; e < c->n_bnd; e++) {
        int64_t id = c->bnd_cell[e];
        h[id] += s->bf1[e] * s->F1b[e] * dt;
        qx[id] += (s->bf2F2[e] * s->F2b[e] + s->bf2S2[e] * s->S2Rb[e] + s->bf2F3[e] * s->F3b[e]) * dt;
        qy[id] += (s->bf3F2[e] * s->F2b[e] + s->bf3S2[e] * s->S2Rb[e] + s->bf3F3[e] * s->F3b[e]) * dt;
    }
    for (int64_t e = 0; e < c->n_int; e++) {
        int64_t L = c->int_L[e], R = c->int_R[e];
        h[L] += s->fL1[e] * s->F1i[e] * dt;
        h[R] += s->fR1[e] * s->F1i[e] * dt;
        qx[L] += (s->f2F2L[e] * s->F2i[e] + s->f2S2L[e] * s->S2Li[e] + s->f2F3L[e] * s->F3i[e]) * dt;
        qx[R] += (s->f2F2R[e] * s->F2i[e] + s->f2S2R[e] * s->S2Ri[e] + s->f2F3R[e] * s->F3i[e]) * dt;
        qy[L] += (s->f3F2L[e] * s->F2i[e] + s->f3S2L[e] * s->S2Li[e] + s->f3F3L[e] * s->F3i[e]) * dt;
        qy[R] += (s->f3F2R[e] * s->F2i[e] + s->f3S2R[e] * s->S2Ri[e] + s->f3F3R[e] * s->F3i[e]) * dt;
    }
    for (int64_t n = 0; n < c->n_cell; ++n) {
        if (fabs(h[n]) < SMALL_REAL) h[n] = 0.;
        if (fabs(qx[n]) < SMALL_REAL) qx[n] = 0.;
        if (fabs(qy[n]) < SMALL_REAL) qy[n] = 0.;
    }
}

/* swSp::sourceFriction{Manning,Strickler,Chezy} -- swSpSourceFriction.cpp:39-171 */
static void source_friction(orc_solver *s, double dt) {
    const orc_case *c = &s->c;
    ORC_PAR
    for (int64_t i = 0; i < c->n_cell; i++) {
        double h = s->h[i];
        if (h > smax(c->h_min, SMALL_REAL)) {
            double u = s->qx[i] / h;
            double v = s->qy[i] / h;
            double ca = s->fcos[i], sa = s->fsin[i];
            double ul = u * ca + v * sa;
            double vl = -u * sa + v * ca;
            double un = sqrt(u * u + v * v);
            double f1 = friction_factor(c->friction, c->fric_c1[i], h, un, dt);
            double f2 = friction_factor(c->friction, c->fric_c2[i], h, un, dt);
            ul *= f1;
            vl *= f2;
            s->qx[i] = (ul * ca - vl * sa) * h;
            s->qy[i] = (ul * sa + vl * ca) * h;
        } else {
            s->qx[i] = 0.;
            s->qy[i] = 0.;
        }
    }
}

/* swSp::sourceSingularHeadLoss -- swSpSourceFriction.cpp:174-218 */
static void source_head_loss(orc_solver *s, double dt) {
    const orc_case *c = &s->c;
    for (int64_t i = 0; i < c->n_cell; i++) {
        double alpha = s->hl_alpha[i];
        double ca = cos(alpha), sa = sin(alpha);
        double qx = s->qx[i] * ca + s->qy[i] * sa;
        double qy = -s->qx[i] * sa + s->qy[i] * ca;
        double h = s->h[i];
        if (h > smax(c->h_min, SMALL_REAL)) {
            double ax = 0, ay = 0;
            if (h > c->hl_hsing[i]) {
                ax = sqrt(qx * qx + qy * qy) * c->hl_c1[i] / h;
                ay = sqrt(qx * qx + qy * qy) * c->hl_c2[i] / h;
            }
            qx = qx * exp(-ax * dt);
            qy = qy * exp(-ay * dt);
        } else {
            qx = 0.;
            qy = 0.;
        }
        s->qx[i] = qx * ca - qy * sa;
        s->qy[i] = qx * sa + qy * ca;
    }
}

/* swParent::interpolateTimeSeries -- swParent.cpp:331-387: clamped at both
 * ends, else a sweep from the first record that stops at the first record
 * k >= 1 with at <= t[k] (it visits record 1 twice, without effect), then
 * linear interpolation between the records around it.  A one-record table
 * returns v[0] in the sliver (t[0] - small_real, t[0]) where the reference
 * reads past the table. */
static double interp_series(const double *t, const double *v, int64_t n, double at) {
    if (at <= t[0] - SMALL_REAL) return v[0];
    if (at >= t[n - 1]) return v[n - 1];
    if (n < 2) return v[0];
    int64_t k = 1;
    while (at > t[k] && k < n - 1) k++;
    return v[k - 1] + (v[k] - v[k - 1]) * (at - t[k - 1]) / (t[k] - t[k - 1]);
}

double orc_interp_series(const double *t, const double *v, int64_t n, double at) {
    return interp_series(t, v, n, at);
}

/* swSp::parabComputeScalarExplicit -- swSpParab.cpp:48-96: substeps of
 * dt / n on one discharge component, fluxes from the values at the start of
 * the substep, scattered L then R in edge order */
static void parab_scalar(orc_solver *s, const double *D, double *v, double dt) {
    const orc_case *c = &s->c;
    double dtp = c->t_max - c->t0;
    for (int64_t e = 0; e < c->n_int; e++) {
        const int64_t L = c->int_L[e], R = c->int_R[e];
        dtp = smin(dtp, 0.5 * c->cell_area[L] * c->phi[L] / D[e]);
        dtp = smin(dtp, 0.5 * c->cell_area[R] * c->phi[R] / D[e]);
    }
    dtp = smin(dtp, dt);
    const int64_t n = (int64_t)(int)(dt / dtp);
    dtp = dt / (double)n;
    for (int64_t k = 0; k < n; k++) {
        for (int64_t e = 0; e < c->n_int; e++) {
            const int64_t L = c->int_L[e], R = c->int_R[e];
            s->pflux[e] = D[e] * s->pfac[e] * (v[L] / s->h[L] - v[R] / s->h[R]);
        }
        for (int64_t e = 0; e < c->n_int; e++) {
            const int64_t L = c->int_L[e], R = c->int_R[e];
            v[L] -= dtp * s->pflux[e] / c->cell_area[L] / smax(SMALL_REAL, c->phi[L]);
            v[R] += dtp * s->pflux[e] / c->cell_area[R] / smax(SMALL_REAL, c->phi[R]);
        }
    }
}

/* swSp::parabComputeExplicit -- swSpParab.cpp:10-46, coefficient 1e-6 (swSp.h:244) */
static void source_parab(orc_solver *s, double dt) {
    const orc_case *c = &s->c;
    for (int64_t e = 0; e < c->n_int; e++) {
        const int64_t L = c->int_L[e], R = c->int_R[e];
        const double dL = MOM_DIFF * s->h[L] * c->phi[L];
        const double dR = MOM_DIFF * s->h[R] * c->phi[R];
        s->pcoef[e] = (smax(dL, dR) > SMALL_REAL) ? 2.0 * dL * dR / (dL + dR) : 0.0;
    }
    parab_scalar(s, s->pcoef, s->qx, dt);
    parab_scalar(s, s->pcoef, s->qy, dt);
}

/* swSp::sourceRunoffEverywhere -- swSpSourceRunoffInfiltration.cpp:43-162,
 * swSp::sourceRunoffWateronly -- :164-263 */
static void source_runoff(orc_solver *s, double dt) {
    const orc_case *c = &s->c;
    const int64_t nt = c->n_rain_times;
    const double at = s->t + 0.5 * dt;
    for (int32_t k = 0; k < c->n_rain_station; k++)
        s->rain_now[k] = interp_series(c->rain_t, c->rain_v + k * nt, nt, at);
    for (int64_t i = 0; i < c->n_cell; i++) {
        const double rate = s->rain_now[c->rain_code[i] - 1];
        const double h0 = s->h[i];
        double dp;
        if (c->rain_water_only) {
            dp = dt * rate;
        } else {
            dp = 0.0;
            if (c->phi[i] > SMALL_REAL) {
                dp = dt * rate;
                /* the runoff-coefficient model divides rain only, not evaporation */
                if (c->runoff_model != 3 || dp >= 0.0) dp /= c->phi[i];
            }
        }
        if (dp > 0.) {
            if (c->runoff_model == 1) s->h[i] += dp;
            else if (c->runoff_model == 2) s->h[i] += (1.0 - c->runoff_coeff[i]) * dp;
            else s->h[i] += c->runoff_coeff[i] * dp;
        } else {
            s->h[i] = smax(0.0, h0 + dp);
            const double ratio = s->h[i] / smax(h0, SMALL_REAL);
            s->qx[i] *= ratio;
            s->qy[i] *= ratio;
        }
    }
}

/* swSp::sourceInfiltration, Horton -- swSpSourceRunoffInfiltration.cpp:12-40 */
static void source_infiltration(orc_solver *s, double dt) {
    const orc_case *c = &s->c;
    for (int64_t i = 0; i < c->n_cell; i++) {
        const double h0 = s->h[i];
        const double dh = -c->infil_rate[i] * dt;
        s->h[i] = smax(0.0, h0 + dh);
        const double ratio = s->h[i] / smax(h0, SMALL_REAL);
        s->qx[i] *= ratio;
        s->qy[i] *= ratio;
    }
}

/* swSp::sourceWind + sourceWindSmithBanke -- swSpSourceWind.cpp:11-85 */
static void source_wind(orc_solver *s, double dt) {
    const orc_case *c = &s->c;
    const int64_t nt = c->n_wind_times;
    const double at = s->t + 0.5 * dt;
    for (int32_t k = 0; k < c->n_wind_station; k++) {
        s->wvx[k] = interp_series(c->wind_t, c->wind_x + k * nt, nt, at);
        s->wvy[k] = interp_series(c->wind_t, c->wind_y + k * nt, nt, at);
    }
    for (int64_t i = 0; i < c->n_cell; i++) {
        const int32_t st = c->wind_code[i] - 1;
        const double h = s->h[i];
        if (!(h > SMALL_REAL)) continue;
        const double u = s->qx[i] / h, v = s->qy[i] / h;
        const double rx = s->wvx[st] - u, ry = s->wvy[st] - v;
        const double water_speed = sqrt(u * u + v * v);
        const double rel = sqrt(rx * rx + ry * ry);
        const double drag = 1e-3 * (0.63 + 0.066 * water_speed);
        const double m = smin(h, c->h_min) / smax(c->h_min, SMALL_REAL);
        const double tx = RHO_AIR * drag * rel * rx, ty = RHO_AIR * drag * rel * ry;
        s->qx[i] = s->qx[i] + tx / RHO_WATER * dt * m;
        s->qy[i] = s->qy[i] + ty / RHO_WATER * dt * m;
    }
}

/* swSp::correctVelocity -- swSp.cpp:447-479 */
static void correct_velocity(orc_solver *s) {
    const orc_case *c = &s->c;
    ORC_PAR
    for (int64_t i = 0; i < c->n_cell; i++) {
        double h = s->h[i];
        if (h > c->h_min) {
            double q = s->qx[i], r = s->qy[i];
            double qn = sqrt(q * q + r * r);
            double cc = sqrt(GRAVITY * h);
            double Fr = qn / cc / h;
            double fact = smax(qn / (h * c->v_max), Fr / c->fr_max);
            if (fact > 1.0) {
                s->qx[i] = q / fact;
                s->qy[i] = r / fact;
            }
        }
    }
}

/* one iteration of swSp::execute's time loop -- swSp.cpp:522-593, with
 * swSp::hypCompute (swSpHyp.cpp:137-211) inlined */
int orc_step(orc_solver *s) {
    const orc_case *c = &s->c;
    ORC_AUDIT_STEP(s->n_audit++);
    if (c->scheme == ORC_MUSCL_EVR)
        lr_muscl(s);
    else
        lr_godunov(s);
    double dt = time_step(s);
    s->dt = dt;
    int st = fluxes(s);
    if (st) return st;
    balance(s, dt);
    int neg = 0;
#ifdef ORC_OMP
#pragma omp parallel for schedule(static) reduction(| : neg)
#endif
    for (int64_t n = 0; n < c->n_cell; ++n) {
        if (s->h[n] < 0.) {
            if (c->neg_depth_detection) {
                neg = 1;
                continue;
            }
            s->h[n] = 0.;
            s->qx[n] = 0.;
            s->qy[n] = 0.;
        }
    }
    if (neg) return ORC_ERR_NEG_DEPTH;
    if (c->momentum_diffusion) source_parab(s, dt);
    if (c->friction != ORC_FRIC_NONE) source_friction(s, dt);
    if (c->head_loss) source_head_loss(s, dt);
    if (c->runoff_model) source_runoff(s, dt);
    if (c->infil_model) source_infiltration(s, dt);
    if (c->n_wind_station > 0) source_wind(s, dt);
    correct_velocity(s);
    s->t += dt;
    return ORC_OK;
}

int orc_run(orc_solver *s, double t_end, int64_t max_steps, int64_t *n_done) {
    int64_t n = 0;
    int st = ORC_OK;
    while (s->t < s->c.t_max && s->t < t_end && n < max_steps) {
        st = orc_step(s);
        if (st) break;
        n++;
    }
    if (n_done) *n_done = n;
    return st;
}

static void *xcopy(const void *p, size_t bytes) {
    if (!p) return NULL;
    void *q = malloc(bytes ? bytes : 1);
    memcpy(q, p, bytes);
    return q;
}

/* ---- math audit (see orc_math.h) ---- */
#ifdef ORC_SW_MATH
int64_t orc_audit_step = -1;
static int64_t g_audit_count = 0, g_audit_step = -1;
static int g_audit_fn = -1;
static double g_audit_x = 0.0, g_audit_g = 0.0, g_audit_s = 0.0;
void orc_audit_note(int fn, double x, double glibc, double sw) {
    if (g_audit_count++ == 0) {
        g_audit_step = orc_audit_step;
        g_audit_fn = fn;
        g_audit_x = x;
        g_audit_g = glibc;
        g_audit_s = sw;
    }
}
#endif

void orc_math_audit_reset(void) {
#ifdef ORC_SW_MATH
    g_audit_count = 0;
    g_audit_step = -1;
    g_audit_fn = -1;
#endif
}

/* 1 in the ORC_SW_MATH build: the first call (step, function 0 exp / 1 log
 * / 2 cbrt, argument, glibc's and the kernels' result) whose results differ
 * since the last reset, and how many did; 0 (nothing set) otherwise. */
int orc_math_audit(int64_t *step, int32_t *fn, double *x, double *glibc, double *sw, int64_t *count) {
#ifdef ORC_SW_MATH
    if (step) *step = g_audit_step;
    if (fn) *fn = g_audit_fn;
    if (x) *x = g_audit_x;
    if (glibc) *glibc = g_audit_g;
    if (sw) *sw = g_audit_s;
    if (count) *count = g_audit_count;
    return 1;
#else
    (void)step, (void)fn, (void)x, (void)glibc, (void)sw, (void)count;
    return 0;
#endif
}

/* The oracle's exp (0), log (1) or cbrt (2) of n values */
void orc_math_eval(int32_t fn, int64_t n, const double *x, double *y) {
    for (int64_t i = 0; i < n; ++i) y[i] = fn == 0 ? exp(x[i]) : fn == 1 ? log(x[i]) : cbrt(x[i]);
}

orc_solver *orc_create(const orc_case *c, const double *h0, const double *qx0, const double *qy0) {
    if (!c || c->n_cell <= 0) return NULL;
    orc_solver *s = (orc_solver *)calloc(1, sizeof(orc_solver));
    s->c = *c;
    int64_t nc = c->n_cell, nb = c->n_bnd, ni = c->n_int;
    size_t dc = sizeof(double) * (size_t)nc, db = sizeof(double) * (size_t)(nb ? nb : 1),
           di = sizeof(double) * (size_t)(ni ? ni : 1);
    /* own every input array so callers may free theirs */
    s->c.cell_area = xcopy(c->cell_area, dc);
    s->c.cell_x = xcopy(c->cell_x, dc);
    s->c.cell_y = xcopy(c->cell_y, dc);
    s->c.cell_zb = xcopy(c->cell_zb, dc);
    s->c.phi = xcopy(c->phi, dc);
    s->c.beta = xcopy(c->beta, dc);
    s->c.fric_c1 = xcopy(c->fric_c1, dc);
    s->c.fric_c2 = xcopy(c->fric_c2, dc);
    s->c.fric_alpha_deg = xcopy(c->fric_alpha_deg, dc);
    s->c.hl_hsing = xcopy(c->hl_hsing, dc);
    s->c.hl_c1 = xcopy(c->hl_c1, dc);
    s->c.hl_c2 = xcopy(c->hl_c2, dc);
    s->c.hl_alpha_deg = xcopy(c->hl_alpha_deg, dc);
    s->c.bnd_cell = xcopy(c->bnd_cell, sizeof(int64_t) * (size_t)nb);
    s->c.bnd_len = xcopy(c->bnd_len, db);
    s->c.bnd_nx = xcopy(c->bnd_nx, db);
    s->c.bnd_ny = xcopy(c->bnd_ny, db);
    s->c.bnd_series = xcopy(c->bnd_series, sizeof(int32_t) * (size_t)nb);
    s->c.int_L = xcopy(c->int_L, sizeof(int64_t) * (size_t)ni);
    s->c.int_R = xcopy(c->int_R, sizeof(int64_t) * (size_t)ni);
    s->c.int_len = xcopy(c->int_len, di);
    s->c.int_nx = xcopy(c->int_nx, di);
    s->c.int_ny = xcopy(c->int_ny, di);
    s->c.int_xm = xcopy(c->int_xm, di);
    s->c.int_ym = xcopy(c->int_ym, di);
    s->c.series_type = xcopy(c->series_type, sizeof(int32_t) * (size_t)c->n_series);
    s->c.series_t = xcopy(c->series_t, sizeof(double) * (size_t)c->n_times);
    s->c.series_v = xcopy(c->series_v, sizeof(double) * (size_t)c->n_times * (size_t)c->n_series);
    {
        const size_t nw = (size_t)(c->n_wind_station > 0 ? c->n_wind_station : 0);
        const size_t wt = (size_t)(c->n_wind_station > 0 ? c->n_wind_times : 0);
        s->c.wind_code = nw ? xcopy(c->wind_code, sizeof(int32_t) * (size_t)nc) : NULL;
        s->c.wind_t = nw ? xcopy(c->wind_t, sizeof(double) * wt) : NULL;
        s->c.wind_x = nw ? xcopy(c->wind_x, sizeof(double) * nw * wt) : NULL;
        s->c.wind_y = nw ? xcopy(c->wind_y, sizeof(double) * nw * wt) : NULL;
        s->wvx = (double *)calloc(nw ? nw : 1, sizeof(double));
        s->wvy = (double *)calloc(nw ? nw : 1, sizeof(double));
        const size_t nr = (size_t)(c->runoff_model ? c->n_rain_station : 0);
        const size_t rt = (size_t)(c->runoff_model ? c->n_rain_times : 0);
        s->c.rain_code = nr ? xcopy(c->rain_code, sizeof(int32_t) * (size_t)nc) : NULL;
        s->c.rain_t = nr ? xcopy(c->rain_t, sizeof(double) * rt) : NULL;
        s->c.rain_v = nr ? xcopy(c->rain_v, sizeof(double) * nr * rt) : NULL;
        s->c.runoff_coeff = (nr && c->runoff_model != 1) ? xcopy(c->runoff_coeff, dc) : NULL;
        s->c.infil_rate = c->infil_model ? xcopy(c->infil_rate, dc) : NULL;
        s->rain_now = (double *)calloc(nr ? nr : 1, sizeof(double));
        s->pfac = (double *)calloc((size_t)(ni ? ni : 1), sizeof(double));
        s->pcoef = (double *)calloc((size_t)(ni ? ni : 1), sizeof(double));
        s->pflux = (double *)calloc((size_t)(ni ? ni : 1), sizeof(double));
        /* swSp::setParabParameters -- swSp.cpp:1236-1246 */
        for (int64_t e = 0; e < ni; e++) {
            const int64_t L = c->int_L[e], R = c->int_R[e];
            const double dx = c->cell_x[R] - c->cell_x[L];
            const double dy = c->cell_y[R] - c->cell_y[L];
            const double d = fabs(dx * c->int_nx[e] + dy * c->int_ny[e]);
            s->pfac[e] = c->int_len[e] / smax(d, SMALL_REAL);
        }
    }

    s->h = xcopy(h0, dc);
    s->qx = xcopy(qx0, dc);
    s->qy = xcopy(qy0, dc);
#define ALLOC(p, bytes) (p) = (double *)calloc(1, (bytes))
    ALLOC(s->bf1, db); ALLOC(s->bf2F2, db); ALLOC(s->bf2F3, db); ALLOC(s->bf2S2, db);
    ALLOC(s->bf3F2, db); ALLOC(s->bf3F3, db); ALLOC(s->bf3S2, db);
    ALLOC(s->fL1, di); ALLOC(s->fR1, di); ALLOC(s->f2F2L, di); ALLOC(s->f2F2R, di);
    ALLOC(s->f2F3L, di); ALLOC(s->f2F3R, di); ALLOC(s->f2S2L, di); ALLOC(s->f2S2R, di);
    ALLOC(s->f3F2L, di); ALLOC(s->f3F2R, di); ALLOC(s->f3F3L, di); ALLOC(s->f3F3R, di);
    ALLOC(s->f3S2L, di); ALLOC(s->f3S2R, di);
    ALLOC(s->zrb, db); ALLOC(s->qrb, db); ALLOC(s->rrb, db); ALLOC(s->lpb, db); ALLOC(s->lmb, db);
    ALLOC(s->lmaxb, db);
    ALLOC(s->zri, 2 * di); ALLOC(s->qri, 2 * di); ALLOC(s->rri, 2 * di); ALLOC(s->lmi, 2 * di);
    ALLOC(s->lpi, 2 * di); ALLOC(s->lmaxi, di);
    ALLOC(s->F1b, db); ALLOC(s->F2b, db); ALLOC(s->F3b, db); ALLOC(s->S2Rb, db);
    ALLOC(s->F1i, di); ALLOC(s->F2i, di); ALLOC(s->F3i, di); ALLOC(s->S2Li, di); ALLOC(s->S2Ri, di);
    ALLOC(s->zc, dc); ALLOC(s->dzdx, dc); ALLOC(s->dzdy, dc); ALLOC(s->swept, dc);
    ALLOC(s->fcos, dc); ALLOC(s->fsin, dc); ALLOC(s->hl_alpha, dc);
    size_t ds = sizeof(double) * (size_t)(c->n_series ? c->n_series : 1);
    ALLOC(s->bc_vinf, ds); ALLOC(s->bc_vsup, ds); ALLOC(s->bc_tinf, ds); ALLOC(s->bc_tsup, ds);
    ALLOC(s->bc_val, ds);
#undef ALLOC
    s->bc_rank = (int64_t *)calloc((size_t)(c->n_series ? c->n_series : 1), sizeof(int64_t));
    s->bc_atend = (int *)calloc((size_t)(c->n_series ? c->n_series : 1), sizeof(int));

    /* neighbour lists in ascending internal-edge order -- swParent.cpp:476-486 */
    s->nb_ptr = (int64_t *)calloc((size_t)nc + 1, sizeof(int64_t));
    for (int64_t e = 0; e < ni; e++) {
        s->nb_ptr[c->int_L[e] + 1]++;
        s->nb_ptr[c->int_R[e] + 1]++;
    }
    for (int64_t i = 0; i < nc; i++) s->nb_ptr[i + 1] += s->nb_ptr[i];
    s->nb_idx = (int64_t *)calloc((size_t)(2 * ni + 1), sizeof(int64_t));
    int64_t *fill = (int64_t *)calloc((size_t)nc, sizeof(int64_t));
    for (int64_t e = 0; e < ni; e++) {
        int64_t L = c->int_L[e], R = c->int_R[e];
        s->nb_idx[s->nb_ptr[L] + fill[L]++] = R;
        s->nb_idx[s->nb_ptr[R] + fill[R]++] = L;
    }
    free(fill);
#ifdef ORC_OMP
    orc_inc_build(&s->inc, nc, ni, c->n_bnd, c->int_L, c->int_R, c->bnd_cell);
#endif

    /* friction angle: degrees -> radians, then cos/sin -- swSp.cpp:1062-1070 */
    if (c->friction != ORC_FRIC_NONE && c->fric_alpha_deg) {
        for (int64_t i = 0; i < nc; i++) {
            double a = c->fric_alpha_deg[i] * atan(1.0) / 45.0;
            s->fcos[i] = cos(a);
            s->fsin[i] = sin(a);
        }
    }
    /* head-loss angle -- swSp.cpp:1360-1366 */
    if (c->head_loss && c->hl_alpha_deg) {
        for (int64_t i = 0; i < nc; i++) s->hl_alpha[i] = c->hl_alpha_deg[i] * atan(1.0) / 45.0;
    }
    edge_factors(s);
    bc_init(s);
    s->t = c->t0;
    s->dt = 0.;
    return s;
}

void orc_destroy(orc_solver *s) {
    if (!s) return;
    void *owned[] = {
        (void *)s->c.cell_area, (void *)s->c.cell_x, (void *)s->c.cell_y, (void *)s->c.cell_zb,
        (void *)s->c.phi, (void *)s->c.beta, (void *)s->c.fric_c1, (void *)s->c.fric_c2,
        (void *)s->c.fric_alpha_deg, (void *)s->c.hl_hsing, (void *)s->c.hl_c1, (void *)s->c.hl_c2,
        (void *)s->c.hl_alpha_deg, (void *)s->c.bnd_cell, (void *)s->c.bnd_len, (void *)s->c.bnd_nx,
        (void *)s->c.bnd_ny, (void *)s->c.bnd_series, (void *)s->c.int_L, (void *)s->c.int_R,
        (void *)s->c.int_len, (void *)s->c.int_nx, (void *)s->c.int_ny, (void *)s->c.int_xm,
        (void *)s->c.int_ym, (void *)s->c.series_type, (void *)s->c.series_t, (void *)s->c.series_v,
        s->h, s->qx, s->qy, s->bf1, s->bf2F2, s->bf2F3, s->bf2S2, s->bf3F2, s->bf3F3, s->bf3S2,
        s->fL1, s->fR1, s->f2F2L, s->f2F2R, s->f2F3L, s->f2F3R, s->f2S2L, s->f2S2R, s->f3F2L,
        s->f3F2R, s->f3F3L, s->f3F3R, s->f3S2L, s->f3S2R, s->zrb, s->qrb, s->rrb, s->lpb, s->lmb,
        s->lmaxb, s->zri, s->qri, s->rri, s->lmi, s->lpi, s->lmaxi, s->F1b, s->F2b, s->F3b, s->S2Rb,
        s->F1i, s->F2i, s->F3i, s->S2Li, s->S2Ri, s->zc, s->dzdx, s->dzdy, s->nb_ptr, s->nb_idx,
        s->swept, s->fcos, s->fsin, s->hl_alpha, s->bc_vinf, s->bc_vsup, s->bc_tinf, s->bc_tsup,
        s->bc_val, s->bc_rank, s->bc_atend, (void *)s->c.wind_code, (void *)s->c.wind_t,
        (void *)s->c.wind_x, (void *)s->c.wind_y, s->wvx, s->wvy, (void *)s->c.rain_code,
        (void *)s->c.rain_t, (void *)s->c.rain_v, (void *)s->c.runoff_coeff,
        (void *)s->c.infil_rate, s->rain_now, s->pfac, s->pcoef, s->pflux};
    for (size_t i = 0; i < sizeof(owned) / sizeof(owned[0]); i++) free(owned[i]);
#ifdef ORC_OMP
    orc_inc_free(&s->inc);
#endif
    free(s);
}

/* threads of the OpenMP build (1 in the serial one) */
int orc_threads(void) {
#ifdef ORC_OMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_get_state(const orc_solver *s, double *h, double *qx, double *qy) {
    size_t b = sizeof(double) * (size_t)s->c.n_cell;
    if (h) memcpy(h, s->h, b);
    if (qx) memcpy(qx, s->qx, b);
    if (qy) memcpy(qy, s->qy, b);
}

void orc_set_state(orc_solver *s, const double *h, const double *qx, const double *qy) {
    size_t b = sizeof(double) * (size_t)s->c.n_cell;
    if (h) memcpy(s->h, h, b);
    if (qx) memcpy(s->qx, qx, b);
    if (qy) memcpy(s->qy, qy, b);
}

double orc_time(const orc_solver *s) { return s->t; }
double orc_last_dt(const orc_solver *s) { return s->dt; }

void orc_get_bnd_f1(const orc_solver *s, double *f1) {
    memcpy(f1, s->F1b, sizeof(double) * (size_t)s->c.n_bnd);
}

void orc_hllc(int blending, double h_min, double phi_min, double bL, double bR, double phiL,
              double phiR, double zbL, double zbR, double zL, double qL, double rL, double zR,
              double qR, double rR, double lmL, double lmR, double lpL, double lpR, double out[5],
              int *status) {
    int st = hllc(blending, h_min, phi_min, bL, bR, phiL, phiR, zbL, zbR, zL, qL, rL, zR, qR, rR,
                  lmL, lmR, lpL, lpR, out);
    if (status) *status = st;
}

void orc_bnd_flux(double h_min, double bR, double phiR, double zbR, double zR, double qR, double rR,
                  double lpR, int bc_type, double bc_val, double out[4], int *status) {
    int st = bnd_flux(h_min, bR, phiR, zbR, zR, qR, rR, lpR, bc_type, bc_val, out);
    if (status) *status = st;
}

double orc_friction_factor(int model, double coeff, double h, double u_norm, double dt) {
    return friction_factor(model, coeff, h, u_norm, dt);
}
