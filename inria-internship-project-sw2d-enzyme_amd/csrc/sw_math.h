/*
 * sw_math.h -- the FP64 exp, log and cbrt of the solver's kernels, written
 * so that the device and a host C build compute the same bits (explicit fma,
 * no contraction, only correctly rounded + - * / and bit operations).
 *
 * The reference runs on glibc (2.35 here, FMA variants selected on x86-64),
 * and its wet/dry fronts amplify any 1-ulp difference of these functions, so
 * they follow glibc where the method is published:
 *   sw_exp   the table-driven exp of P. T. P. Tang (ACM TOMS 15, 1989) in the
 *            form of ARM's optimized-routines that glibc ships (N = 128 table
 *            2^(k/N) = H (1 + T), degree-5 polynomial, the contraction into
 *            fma that glibc's FMA build performs); bitwise equal to glibc's
 *            exp on 8.5e7 random and rounding-boundary arguments;
 *   sw_cbrt  glibc's own cube root (frexp reduction, degree-6 polynomial, one
 *            rational Halley step, exponent by thirds); bitwise equal to
 *            glibc's cbrt on 2e7 arguments;
 *   sw_log   a double-double log (512-point reduction table, atanh series,
 *            error below 2^-74 before the final rounding): correctly rounded
 *            except in rare cases.  glibc's log uses a tuned table that cannot
 *            be regenerated from its published description, so the two differ
 *            where glibc itself is not correctly rounded (about 1e-4 of
 *            arguments).
 * Tables: sw_math_tables.h (gen_sw_math_tables.py).
 */
#pragma once
#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define SWM_FN __host__ __device__ static inline
#define SWM_TAB static constexpr
#else
#define SWM_FN static inline
#define SWM_TAB static const
#endif

#include "sw_math_tables.h"

SWM_FN uint64_t swm_bits(double x) {
    uint64_t u;
    __builtin_memcpy(&u, &x, 8);
    return u;
}

SWM_FN double swm_dbl(uint64_t u) {
    double x;
    __builtin_memcpy(&x, &u, 8);
    return x;
}

/* exp(x); glibc's special cases (tiny, huge, inf, nan) included */
SWM_FN double sw_exp(double x) {
    const double InvLn2N = 0x1.71547652b82fep0 * 128.0, NegLn2hiN = -0x1.62e42fefa0000p-8,
                 NegLn2loN = -0x1.cf79abc9e3b3ap-47, Shift = 0x1.8p52;
    const double C2 = 0x1.ffffffffffdbdp-2, C3 = 0x1.555555555543cp-3, C4 = 0x1.55555cf172b91p-5,
                 C5 = 0x1.1111167a4d017p-7;
    uint32_t abstop = (uint32_t)(swm_bits(x) >> 52) & 0x7ff;
    const uint32_t t_tiny = 0x3c9, t_big = 0x408, t_huge = 0x409, t_inf = 0x7ff;  /* top12 of 2^-54, 512, 1024, inf */
    if (abstop - t_tiny >= t_big - t_tiny) {
        if ((uint32_t)(abstop - t_tiny) >= 0x80000000u) return 1.0 + x;  /* |x| < 2^-54 */
        if (abstop >= t_huge) {
            if (swm_bits(x) == 0xfff0000000000000ULL) return 0.0;
            if (abstop >= t_inf) return 1.0 + x;
            return (swm_bits(x) >> 63) ? 0.0 : swm_dbl(0x7ff0000000000000ULL);
        }
        abstop = 0;  /* |x| in [512, 1024): scaled below */
    }
    double kd = fma(InvLn2N, x, Shift);
    const uint64_t ki = swm_bits(kd);
    kd -= Shift;
    const double r = fma(kd, NegLn2loN, fma(kd, NegLn2hiN, x));
    const uint64_t idx = 2 * (ki % 128), top = ki << 45;
    const double tail = swm_dbl(SWM_EXP_TAB[idx]);
    uint64_t sbits = SWM_EXP_TAB[idx + 1] + top;
    const double r2 = r * r;
    const double tmp = fma(r2 * r2, fma(r, C5, C4), fma(r2, fma(r, C3, C2), tail + r));
    if (abstop == 0) {
        if ((ki & 0x80000000ULL) == 0) {  /* k > 0: the exponent may overflow by <= 460 */
            sbits -= 1009ULL << 52;
            const double scale = swm_dbl(sbits);
            return 0x1p1009 * fma(scale, tmp, scale);
        }
        sbits += 1022ULL << 52;  /* k < 0: round once before scaling into the subnormals */
        const double scale = swm_dbl(sbits);
        const double st = scale * tmp;
        double y = scale + st;
        if (y < 1.0) {
            double lo = scale - y + st;
            const double hi = 1.0 + y;
            lo = 1.0 - hi + y + lo;
            y = (hi + lo) - 1.0;
            if (y == 0.0) y = 0.0;
        }
        return 0x1p-1022 * y;
    }
    const double scale = swm_dbl(sbits);
    return fma(scale, tmp, scale);
}

/* cbrt(x) */
SWM_FN double sw_cbrt(double x) {
    const double f[5] = {1.0 / 1.5874010519681994748, 1.0 / 1.2599210498948731648, 1.0,
                         1.2599210498948731648, 1.5874010519681994748};
    if (x == 0.0 || !(x == x) || fabs(x) == swm_dbl(0x7ff0000000000000ULL)) return x + x;
    int xe;
    const double xm = frexp(fabs(x), &xe);
    const double u =
        (0.354895765043919860 +
         ((1.50819193781584896 +
           ((-2.11499494167371287 +
             ((2.44693122563534430 +
               ((-1.83469277483613086 + (0.784932344976639262 - 0.145263899385486377 * xm) * xm) * xm)) *
              xm)) *
            xm)) *
          xm));
    const double t2 = u * u * u;
    const double ym = u * (t2 + 2.0 * xm) / (2.0 * t2 + xm) * f[2 + xe % 3];
    return ldexp(x > 0.0 ? ym : -ym, xe / 3);
}

/* log(x): double-double evaluation, then one rounding */
SWM_FN double sw_log(double x) {
    if (!(x > 0.0)) return x == 0.0 ? -swm_dbl(0x7ff0000000000000ULL) : (x - x) / (x - x);
    uint64_t ix = swm_bits(x);
    if (ix >= 0x7ff0000000000000ULL) return x;  /* +inf */
    int k = 0;
    if (ix < 0x0010000000000000ULL) {  /* subnormal */
        x *= 0x1p54;
        ix = swm_bits(x);
        k = -54;
    }
    k += (int)(ix >> 52) - 1023;
    double f = swm_dbl((ix & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);  /* [1, 2) */
    if (f > 0x1.6a09e667f3bcdp0) {  /* f in [sqrt(1/2), sqrt(2)) */
        f *= 0.5;
        k += 1;
    }
    const int j = (int)floor((f - 1.0) * 512.0 + 0.5);  /* exact: f - 1 and the scaling are */
    const double c = 1.0 + (double)j * 0x1p-9;
    const double num = f - c;  /* exact (Sterbenz) */
    /* s = (f - c) / (f + c) as a double-double */
    const double dh = f + c, bb = dh - f, dl = (f - (dh - bb)) + (c - bb);
    const double sh = num / dh;
    const double rr = fma(-sh, dh, num) - sh * dl;
    const double sl = rr / dh;
    /* log(f) = log(c) + 2 atanh(s), the series tail from the head of s */
    const double z = sh * sh;
    const double T =
        2.0 * sh * z *
        (0x1.5555555555555p-2 +
         z * (0x1.999999999999ap-3 + z * (0x1.2492492492492p-3 + z * 0x1.c71c71c71c71cp-4)));
    const double a = (double)k * SWM_LN2_HI;  /* exact */
    const double Lh = SWM_LOG_HI[j + 150], Ll = SWM_LOG_LO[j + 150];
    double h1 = a + Lh, b1 = h1 - a, e1 = (a - (h1 - b1)) + (Lh - b1);
    const double s2 = 2.0 * sh;
    double h2 = h1 + s2, b2 = h2 - h1, e2 = (h1 - (h2 - b2)) + (s2 - b2);
    const double lo = e1 + e2 + ((double)k * SWM_LN2_LO + Ll + 2.0 * sl + T);
    return h2 + lo;
}
