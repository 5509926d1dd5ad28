// sp_physics.h -- scalar physics of the SW2D single-porosity model, shared by
// the HIP kernels (device) and the host tooling.  FP64 throughout; every
// expression keeps the reference's left-to-right evaluation order and the
// library is compiled with -ffp-contract=off, so results are bit-identical
// to the reference CPU solver wherever only +,-,*,/ and sqrt are involved
// (sqrt and division are correctly rounded in IEEE FP64 on both sides); exp,
// log and cbrt are the project's own (sw_math.h), equal to glibc's for exp and
// cbrt and correctly rounded for log.
#pragma once

#ifdef __HIPCC__
#define SPD __host__ __device__ __forceinline__
#else
#define SPD inline
#endif

#include <math.h>

#include "sw_math.h"

namespace sp {

constexpr double kSmall = 1e-16;  // src/common/param.h:2 (small_real)
constexpr double kG = 9.81;       // src/common/param.h:3 (gravity)
constexpr double kRhoAir = 1.2;     // src/common/param.h:4 (rho_air)
constexpr double kRhoWater = 1.0e3; // src/common/param.h:5 (rho_water)
constexpr double kMomDiff = 1e-6;   // swSp.h:244 (momentum_diffusion_coefficient, never read from input)

enum BcType { BC_FR = 0, BC_H = 1, BC_Z = 2, BC_Q = 3 };
enum Friction { FRIC_NONE = 0, FRIC_MANNING = 1, FRIC_STRICKLER = 2, FRIC_CHEZY = 3 };
enum Status { ST_OK = 0, ST_NAN = 2, ST_NEG_DEPTH = 3, ST_FLOW_CONFIG = 4, ST_BC = 5,
              ST_PARAB_SUBSTEPS = 6 };

// std::max / std::min semantics (first argument wins unless strictly beaten)
SPD double smax(double a, double b) { return (a < b) ? b : a; }
SPD double smin(double a, double b) { return (b < a) ? b : a; }

// lambda-/+ : src/sw2dSolver/models/sp/swSp.cpp:972-995
SPD double lambda_minus(double b, double c, double u) {
    return b * u - sqrt((b - 1.0) * b * u * u + c * c);
}
SPD double lambda_plus(double b, double c, double u) {
    return b * u + sqrt((b - 1.0) * b * u * u + c * c);
}
// phi * (beta h u^2 + g h^2 / 2) : swSp.cpp:1388-1399
SPD double specific_force(double b, double phi, double h, double u) {
    return phi * (b * h * u * u + 0.5 * kG * h * h);
}

// Boundary flux, domain on the right of the edge: swSpHyp.cpp:13-135, with
// hypHCritical at swSpHyp.cpp:651-661.  f = {F1, F2, F3}; s2R is always 0.
SPD int bnd_flux(double h_min, double bR, double phiR, double zbR, double zR, double qR, double rR,
                 double lambda_plus_R, int bc_type, double bc_val, double *f) {
    double hR = zR - zbR;
    double q1 = qR, r1 = rR;
    if (hR < h_min) {
        q1 = 0.;
        r1 = 0.;
        hR = 0.;
    }
    double cR = sqrt(kG * smax(hR, kSmall));
    double uR = q1 / smax(hR, kSmall);
    double L1 = smax(0.0, lambda_plus_R);
    double flux2R = specific_force(bR, phiR, hR, uR);
    double flux1 = 0.0, flux2 = 0.0, flux3;
    if (bc_type == BC_FR) {
        if ((bc_val < 0.) | (bc_val > 1.)) return ST_BC;
        double c = (uR - 2. * cR) / (-bc_val - 2.);
        c = smax(c, 0.);
        double u = -bc_val * c;
        double h = c * c / kG;
        flux1 = phiR * u * h;
        flux2 = flux2R + (flux1 - q1 * phiR) * L1;
    } else if (bc_type == BC_H || bc_type == BC_Z) {
        double h_bc = (bc_type == BC_H) ? smax(bc_val, 0.) : smax(bc_val - zbR, 0.);
        if (hR > h_min) {
            flux1 = phiR * (q1 + L1 * (h_bc - hR));
            flux2 = flux2R + (flux1 - q1 * phiR) * L1;
        } else {
            double c = sqrt(kG * h_bc);
            flux1 = phiR * h_bc * c;
            flux2 = specific_force(bR, phiR, h_bc, c);
        }
    } else if (bc_type == BC_Q) {
        flux1 = bc_val;
        double hc = sw_exp((1. / 3.) * sw_log((bc_val / smax(phiR, kSmall)) / kG));
        if (hc >= hR) flux1 = phiR * 1.5 * sqrt(kG * hc) * hc;
        flux2 = flux2R + (flux1 - q1 * phiR) * L1;
    } else {
        return ST_BC;
    }
    if (flux1 > 0) {
        flux3 = 0;
    } else {
        double vR = r1 / smax(hR, kSmall);
        flux3 = flux1 * vR;
    }
    f[0] = flux1;
    f[1] = flux2;
    f[2] = flux3;
    return ST_OK;
}

// HLLC Riemann solver with the 4 flow regimes: discontinuous topography
// (swSpHyp.cpp:664-821) or blending (swSpHyp.cpp:824-998).
// f = {F1, F2, F3, s2L, s2R}.
SPD int hllc(bool blending, double h_min, double phi_min, double bL, double bR, double phiL,
             double phiR, double zbL, double zbR, double zL, double qL, double rL, double zR,
             double qR, double rR, double lmL, double lmR, double lpL, double lpR, double *f) {
    double hL, hR;
    if (blending) {
        hL = zL - zbL;
        hR = zR - zbR;
    } else {
        hL = smax(0.0, zL - zbL);
        hR = smax(0.0, zR - zbR);
    }
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
    if ((int)both + (int)lhs + (int)rhs + (int)nowhere != 1) return ST_FLOW_CONFIG;
    double f1 = 0., f2 = 0., f3 = 0., sl = 0., sr = 0.;
    if (both) {
        double uL, uR;
        if (blending) {
            uL = qL / smax(hL, kSmall);
            uR = qR / smax(hR, kSmall);
        } else {
            uL = qL / smax(kSmall, hL);
            uR = qR / smax(kSmall, hR);
        }
        double L1 = smin(lmL, smin(lmR, 0.));
        double L2 = smax(lpL, smax(lpR, 0.));
        double den = smax(L2 - L1, kSmall);
        double s2;
        double phiLR = smin(phiL, phiR);
        if (!blending) {
            f1 = phiLR * (L2 * qL - L1 * qR + L1 * L2 * (zR - zL)) / den;
            double fact = 0.5 * smin(phiL, phiR) * (hL + hR);
            s2 = (phiR * hR * hR - phiL * hL * hL) * 0.5 + fact * (zL - zR);
        } else {
            double alpha_h = 0.0;
            if (fabs(zbL - zbR) > kSmall) alpha_h = smax(0.0, smin(1.0, (zL - zR) / (zbL - zbR)));
            double F1h = phiLR * (L2 * qL - L1 * qR + L1 * L2 * (hR - hL)) / den;
            double F1z = phiLR * (L2 * qL - L1 * qR + L1 * L2 * (zR - zL)) / den;
            f1 = alpha_h * F1h + (1.0 - alpha_h) * F1z;
            double fact = 0.5 * smin(phiL, phiR) * (hL + hR);
            double s2z = (phiR * hR * hR - phiL * hL * hL) * 0.5 + fact * (zL - zR);
            double s2h = (zbL > zbR) ? phiR * hR * (zbL - zbR) : phiL * hL * (zbL - zbR);
            s2 = alpha_h * s2h + (1.0 - alpha_h) * s2z;
        }
        s2 = s2 * kG / den;
        sl = -L1 * s2;
        sr = L2 * s2;
        double FL = specific_force(bL, phiL, hL, uL);
        double FR = specific_force(bR, phiR, hR, uR);
        f2 = (L2 * (FL + sl) - L1 * (FR - sr) + L1 * L2 * (phiR * qR - phiL * qL)) / den;
        double vL = rL / smax(hL, kSmall);
        double vR = rR / smax(hR, kSmall);
        f3 = (f1 >= 0) ? bL * f1 * vL : bR * f1 * vR;
    } else if (rhs) {
        sr = phiR * 0.5 * kG * hR * hR;
    } else if (lhs) {
        sl = -phiL * 0.5 * kG * hL * hL;
    }
    f[0] = f1;
    f[1] = f2;
    f[2] = f3;
    f[3] = sl;
    f[4] = sr;
    return ST_OK;
}

// h^(-4/3) as the reference's friction laws form it, exp(-4/3 log h)
SPD double friction_hpow(double h) { return sw_exp(-4.0 / 3.0 * sw_log(h)); }

// velocity damping factor over one step: swSourceFriction.cpp:7-79.  hp:
// friction_hpow(h) for Manning and Strickler (formed once per cell by the
// caller: both principal directions use the same value)
SPD double friction_factor(int model, double coeff, double h, double u_norm, double dt, double hp) {
    double factor = 1.0;
    if (model == FRIC_MANNING) {
        if (h > kSmall) {
            double arg = -kG * u_norm * coeff * coeff * hp;
            factor = sw_exp(arg * dt);
        } else {
            factor = 0.0;
        }
    } else if (model == FRIC_STRICKLER) {
        if ((h > kSmall) && (coeff > kSmall)) {
            double arg = -kG * u_norm / coeff / coeff * hp;
            factor = sw_exp(arg * dt);
        } else {
            factor = 0.0;
        }
    } else if (model == FRIC_CHEZY) {
        if ((h > kSmall) && (coeff > kSmall)) {
            double arg = -kG * u_norm / coeff / coeff / h;
            factor = sw_exp(arg * dt);
        } else {
            factor = 0.0;
        }
    }
    return factor;
}

}  // namespace sp
