// sp_adjoint_host.cpp -- host build of the local _E functions and their
// hand-written reverse-mode derivatives (sp_adjoint.h), exported for the
// per-function derivative tests.  Not part of the public C-ABI.
#include "sp_adjoint.h"

extern "C" {

// p = {h, qx, qy, zb, b, nx, ny}; out = {z, q, r, lm, lp}
void sw2dt_lr_fwd(const double *p, double h_min, double *out) {
    const spe::Side s = spe::lr_fwd(p[0], p[1], p[2], p[3], p[4], p[5], p[6], h_min);
    out[0] = s.z;
    out[1] = s.q;
    out[2] = s.r;
    out[3] = s.lm;
    out[4] = s.lp;
}
// dout[5] -> d[3] = adjoints of h, qx, qy
void sw2dt_lr_vjp(const double *p, double h_min, const double *dout, double *d) {
    d[0] = d[1] = d[2] = 0.0;
    spe::lr_vjp(p[0], p[1], p[2], p[4], p[5], p[6], h_min, dout[0], dout[1], dout[2], dout[3],
                dout[4], d[0], d[1], d[2]);
}

// p = {bL, bR, phiL, phiR, zbL, zbR, zL, qL, rL, zR, qR, rR, lmL, lmR, lpL, lpR}
void sw2dt_hllc_fwd(const double *p, double h_min, double phi_min, double *f) {
    spe::hllc_fwd(h_min, phi_min, p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], p[9], p[10],
                  p[11], p[12], p[13], p[14], p[15], f);
}
void sw2dt_hllc_vjp(const double *p, double h_min, double phi_min, const double *df, double *d) {
    for (int k = 0; k < 12; ++k) d[k] = 0.0;
    spe::hllc_vjp(h_min, phi_min, p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], p[9], p[10],
                  p[11], p[12], p[13], p[14], p[15], df, d);
}

// p = {bR, phiR, zbR, zR, qR, rR, lpR, bc}
void sw2dt_bnd_fwd(const double *p, int type, double h_min, double phi_min, double *f) {
    spe::bnd_fwd(h_min, phi_min, p[0], p[1], p[2], p[3], p[4], p[5], p[6], type, p[7], f);
}
// d[6] = adjoints of zR qR rR lpR phiR bc
void sw2dt_bnd_vjp(const double *p, int type, double h_min, double phi_min, const double *df,
                   double *d) {
    for (int k = 0; k < 6; ++k) d[k] = 0.0;
    spe::bnd_vjp(h_min, phi_min, p[0], p[1], p[2], p[3], p[4], p[5], p[6], type, p[7], df, d);
}

// p = {h, qx, qy, ca, sa, K1, K2, dt}; out = {qx, qy}
void sw2dt_fric_fwd(const double *p, double h_min, double *out) {
    double qx = p[1], qy = p[2];
    spe::fric_fwd(h_min, p[0], qx, qy, p[3], p[4], p[5], p[6], p[7]);
    out[0] = qx;
    out[1] = qy;
}
// dout[2] -> d[5] = adjoints of h, qx, qy, K1, K2
void sw2dt_fric_vjp(const double *p, double h_min, const double *dout, double *d) {
    double dh = 0.0, dqx = dout[0], dqy = dout[1], dk1 = 0.0, dk2 = 0.0;
    spe::fric_vjp(h_min, p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], dh, dqx, dqy, dk1, dk2);
    d[0] = dh;
    d[1] = dqx;
    d[2] = dqy;
    d[3] = dk1;
    d[4] = dk2;
}

// p = {h, qx, qy, ca, sa, hsing, k1, k2, dt}; out = {qx, qy}
void sw2dt_hl_fwd(const double *p, double h_min, double *out) {
    double qx = p[1], qy = p[2];
    spe::hl_fwd(h_min, p[0], qx, qy, p[3], p[4], p[5], p[6], p[7], p[8]);
    out[0] = qx;
    out[1] = qy;
}
void sw2dt_hl_vjp(const double *p, double h_min, const double *dout, double *d) {
    double dh = 0.0, dqx = dout[0], dqy = dout[1];
    spe::hl_vjp(h_min, p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], dh, dqx, dqy);
    d[0] = dh;
    d[1] = dqx;
    d[2] = dqy;
}

// p = {h, qx, qy, zb, b, sx, sy, nx, ny, dxm, dym, dt}; out = {z, q, r, lm, lp}
void sw2dt_recon_fwd(const double *p, int side, double *out) {
    const spe::Side s = spe::recon_muscl(side, p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8],
                                         p[9], p[10], p[11]);
    out[0] = s.z;
    out[1] = s.q;
    out[2] = s.r;
    out[3] = s.lm;
    out[4] = s.lp;
}
// dout[5] -> d[5] = adjoints of h, qx, qy, sx, sy
void sw2dt_recon_vjp(const double *p, int side, const double *dout, double *d) {
    for (int k = 0; k < 5; ++k) d[k] = 0.0;
    spe::recon_muscl_vjp(side, p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], p[9], p[10],
                         p[11], dout[0], dout[1], dout[2], dout[3], dout[4], d[0], d[1], d[2], d[3],
                         d[4]);
}

// p = {h, qx, qy, zb, b, nx, ny}; out = {z, q, r, lp}
void sw2dt_bnd_muscl_fwd(const double *p, double *out) {
    const spe::Side s = spe::lr_bnd_muscl(p[0], p[1], p[2], p[3], p[4], p[5], p[6]);
    out[0] = s.z;
    out[1] = s.q;
    out[2] = s.r;
    out[3] = s.lp;
}
// dout[4] -> d[3] = adjoints of h, qx, qy
void sw2dt_bnd_muscl_vjp(const double *p, const double *dout, double *d) {
    d[0] = d[1] = d[2] = 0.0;
    spe::lr_bnd_muscl_vjp(p[0], p[1], p[2], p[4], p[5], p[6], dout[0], dout[1], dout[2], dout[3],
                          d[0], d[1], d[2]);
}

// one cell's slope over deg incidences: nb = deg x {zj, xj, yj, xm, ym, side}
struct HostNb {
    const double *v;
    double zj(int k) const { return v[6 * k]; }
    double xj(int k) const { return v[6 * k + 1]; }
    double yj(int k) const { return v[6 * k + 2]; }
    double xm(int k) const { return v[6 * k + 3]; }
    double ym(int k) const { return v[6 * k + 4]; }
    int side(int k) const { return (int)v[6 * k + 5]; }
};
struct HostOut {
    double *d;
    void set(int k, double x) const { d[k] = x; }
    void add(int k, double x) const { d[k] += x; }
    void put(int k, double x) const { d[k] = x; }
};
// c = {zi, xi, yi}; out = {sx, sy}
void sw2dt_slope_fwd(const double *c, const double *nb, int deg, double *out) {
    spe::slope_fwd(HostNb{nb}, deg, c[0], c[1], c[2], out[0], out[1]);
}
// g = adjoint of (sx, sy) -> d[0] = adjoint of zi, d[1 + k] = adjoint of zj(k)
void sw2dt_slope_vjp(const double *c, const double *nb, int deg, const double *g, double *d) {
    double dzi = 0.0;
    HostOut o{d + 1};
    spe::slope_vjp(HostNb{nb}, deg, c[0], c[1], c[2], g[0], g[1], dzi, o);
    d[0] = dzi;
}
// the same through the register variant (cells of at most 6 incidences);
// returns 0 when deg is beyond that
int sw2dt_slope_vjp_reg(const double *c, const double *nb, int deg, const double *g, double *d) {
    double dzi = 0.0;
    const HostOut o{d + 1};
    const bool ok = spe::slope_vjp_reg<6>(HostNb{nb}, deg, c[0], c[1], c[2], g[0], g[1], dzi, o);
    d[0] = dzi;
    return ok ? 1 : 0;
}

}  // extern "C"
