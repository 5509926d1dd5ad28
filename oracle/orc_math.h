/*
 * orc_math.h -- the transcendental functions of the oracle (test
 * infrastructure).
 *
 * Default build (liboracle.so): glibc's exp / log / cbrt, the functions the
 * reference runs on, so the oracle reproduces the reference bit for bit.
 *
 * ORC_SW_MATH build (liboracle_dm.so): the kernels' own functions
 * (inria-internship-project-sw2d-enzyme_amd/csrc/sw_math.h), so the oracle
 * reproduces the GPU bit for bit; every call is also evaluated with glibc and
 * the first call whose two results differ is recorded with the step it
 * happened in (orc_math_audit).  The two builds run the same restatement of
 * the reference's loops, so where they part, a transcendental result is the
 * only difference.
 */
#pragma once
#include <math.h>
#include <stdint.h>

#ifdef ORC_SW_MATH
#include "../inria-internship-project-sw2d-enzyme_amd/csrc/sw_math.h"

extern int64_t orc_audit_step;   /* step being computed (set by the step loops) */
void orc_audit_note(int fn, double x, double glibc, double sw);

static inline double orc_exp(double x) {
    const double a = sw_exp(x), g = exp(x);
    if (swm_bits(a) != swm_bits(g) && !(a != a && g != g)) orc_audit_note(0, x, g, a);
    return a;
}
static inline double orc_log(double x) {
    const double a = sw_log(x), g = log(x);
    if (swm_bits(a) != swm_bits(g) && !(a != a && g != g)) orc_audit_note(1, x, g, a);
    return a;
}
static inline double orc_cbrt(double x) {
    const double a = sw_cbrt(x), g = cbrt(x);
    if (swm_bits(a) != swm_bits(g) && !(a != a && g != g)) orc_audit_note(2, x, g, a);
    return a;
}
#define exp(x) orc_exp(x)
#define log(x) orc_log(x)
#define cbrt(x) orc_cbrt(x)
#define ORC_AUDIT_STEP(n) (orc_audit_step = (int64_t)(n))
#else
#define ORC_AUDIT_STEP(n) ((void)0)
#endif
